/*
 * FP64 CPU oracle for the NUFFT pair (test infrastructure only: imported by
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py; never
 * by the product path).
 *
 * Conventions follow the reference's CUDANufft/Nufft wrappers of (cu)FINUFFT:
 *   - coords is [ndim][M], row d pairs with image dim d (lib/fft/nufft.cpp:194-197,
 *     579-583 pass x=row ndim-1, y=row ndim-2, z=row 0 with nmodes flipped, 47-49);
 *   - mode index n of an image dim of size N is frequency k = n - floor(N/2)
 *     (FINUFFT default modeord=0, opts untouched at nufft.cpp:169, 552);
 *   - type-2: c_j = sum_k f_k exp(isign i k.x_j); type-1: f_k = sum_j c_j exp(isign i k.x_j)
 *     (isign from NufftOptions::get_positive, lib/fft/nufft.ixx:40-64).
 * The 1/sqrt(prod nmodes) scaling of nufft.cpp:487-506 is applied by the Python
 * wrapper, not here.
 *
 * ora_nudft_*: direct (brute-force) sums, the mathematical definition.
 * ora_spread / ora_interp: exponential-of-semicircle (ES) kernel spreading and
 *   interpolation on a periodic fine grid, the published FINUFFT method
 *   (Barnett, Magland, af Klinteberg, SIAM J. Sci. Comput. 41(5) 2019), used as a
 *   fast FP64 oracle at sizes where the direct sum is too slow; it is itself checked
 *   against ora_nudft_* in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void mode_offsets(int ndim, const int64_t* nmodes, int64_t* n, int64_t* off)
{
    for (int d = 0; d < 3; ++d) {
        n[d] = d < ndim ? nmodes[d] : 1;
        off[d] = d < ndim ? nmodes[d] / 2 : 0;
    }
}

/* type-2: out[j] = sum_k in[k] exp(isign i k.x_j). in: prod(nmodes) complex, out: M complex. */
void ora_nudft_type2(int ndim, const int64_t* nmodes, int64_t M, const double* coords,
                     const double* in, double* out, int isign, int nthreads)
{
    int64_t n[3], off[3];
    mode_offsets(ndim, nmodes, n, off);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double* e = (double*)malloc(sizeof(double) * 2 * (n[0] + n[1] + n[2]));
#pragma omp for schedule(static)
        for (int64_t j = 0; j < M; ++j) {
            double* ed[3] = {e, e + 2 * n[0], e + 2 * (n[0] + n[1])};
            for (int d = 0; d < 3; ++d) {
                double x = d < ndim ? coords[d * M + j] : 0.0;
                for (int64_t q = 0; q < n[d]; ++q) {
                    double ph = isign * (double)(q - off[d]) * x;
                    ed[d][2 * q] = cos(ph);
                    ed[d][2 * q + 1] = sin(ph);
                }
            }
            double sr = 0.0, si = 0.0;
            for (int64_t a = 0; a < n[0]; ++a) {
                double ar = ed[0][2 * a], ai = ed[0][2 * a + 1];
                for (int64_t b = 0; b < n[1]; ++b) {
                    double br = ed[1][2 * b], bi = ed[1][2 * b + 1];
                    double abr = ar * br - ai * bi, abi = ar * bi + ai * br;
                    const double* f = in + 2 * ((a * n[1] + b) * n[2]);
                    double tr = 0.0, ti = 0.0;
                    for (int64_t c = 0; c < n[2]; ++c) {
                        double cr = ed[2][2 * c], ci = ed[2][2 * c + 1];
                        double fr = f[2 * c], fi = f[2 * c + 1];
                        tr += fr * cr - fi * ci;
                        ti += fr * ci + fi * cr;
                    }
                    sr += tr * abr - ti * abi;
                    si += tr * abi + ti * abr;
                }
            }
            out[2 * j] = sr;
            out[2 * j + 1] = si;
        }
        free(e);
    }
}

/* type-1: out[k] = sum_j in[j] exp(isign i k.x_j). Parallel over the slowest mode index. */
void ora_nudft_type1(int ndim, const int64_t* nmodes, int64_t M, const double* coords,
                     const double* in, double* out, int isign, int nthreads)
{
    int64_t n[3], off[3];
    mode_offsets(ndim, nmodes, n, off);
    memset(out, 0, sizeof(double) * 2 * n[0] * n[1] * n[2]);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double* e = (double*)malloc(sizeof(double) * 2 * (n[1] + n[2]));
#pragma omp for schedule(dynamic, 1)
        for (int64_t a = 0; a < n[0]; ++a) {
            double* plane = out + 2 * a * n[1] * n[2];
            for (int64_t j = 0; j < M; ++j) {
                double x0 = ndim > 0 ? coords[0 * M + j] : 0.0;
                double ph0 = isign * (double)(a - off[0]) * x0;
                double ar = cos(ph0), ai = sin(ph0);
                double cr0 = in[2 * j], ci0 = in[2 * j + 1];
                double wr = cr0 * ar - ci0 * ai, wi = cr0 * ai + ci0 * ar;
                for (int d = 1; d < 3; ++d) {
                    double x = d < ndim ? coords[d * M + j] : 0.0;
                    double* ed = d == 1 ? e : e + 2 * n[1];
                    for (int64_t q = 0; q < n[d]; ++q) {
                        double ph = isign * (double)(q - off[d]) * x;
                        ed[2 * q] = cos(ph);
                        ed[2 * q + 1] = sin(ph);
                    }
                }
                for (int64_t b = 0; b < n[1]; ++b) {
                    double br = e[2 * b], bi = e[2 * b + 1];
                    double vr = wr * br - wi * bi, vi = wr * bi + wi * br;
                    double* f = plane + 2 * b * n[2];
                    const double* ec = e + 2 * n[1];
                    for (int64_t c = 0; c < n[2]; ++c) {
                        double cr = ec[2 * c], ci = ec[2 * c + 1];
                        f[2 * c] += vr * cr - vi * ci;
                        f[2 * c + 1] += vr * ci + vi * cr;
                    }
                }
            }
        }
        free(e);
    }
}

/* ES kernel phi(z) = exp(beta (sqrt(1 - (2z/ns)^2) - 1)) for |z| < ns/2, z in fine-grid units. */
static inline double es_kernel(double z, int ns, double beta)
{
    double u = 2.0 * z / ns;
    double t = 1.0 - u * u;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

/* Fine-grid position of x (any real) on a periodic grid of nf cells covering [0, 2pi). */
static inline double fine_pos(double x, int64_t nf)
{
    const double twopi = 6.283185307179586476925286766559;
    double u = x / twopi;
    u -= floor(u);
    double t = u * (double)nf;
    if (t >= (double)nf) t -= (double)nf;
    return t;
}

/* Weights and first cell index along one dim for a point at fine position t. */
static inline int64_t point_weights(double t, int ns, double beta, double* w)
{
    int64_t l0 = (int64_t)ceil(t - 0.5 * ns);
    for (int q = 0; q < ns; ++q) w[q] = es_kernel((double)(l0 + q) - t, ns, beta);
    return l0;
}

static inline int64_t wrap(int64_t i, int64_t n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

/*
 * Spread M complex strengths onto a zeroed periodic fine grid nf[0] x nf[1] x nf[2]
 * (nf[d] = 1 for unused dims). Parallel over slabs of the slowest fine-grid dim:
 * each thread owns a range of planes and adds only there, so the sum order per cell
 * is the point order (deterministic).
 */
void ora_spread(int ndim, const int64_t* nf_in, int64_t M, const double* coords, const double* c,
                double* grid, int ns, double beta, int nthreads)
{
    int64_t nf[3];
    for (int d = 0; d < 3; ++d) nf[d] = d < ndim ? nf_in[d] : 1;
    int nsd[3];
    for (int d = 0; d < 3; ++d) nsd[d] = d < ndim ? ns : 1;
    memset(grid, 0, sizeof(double) * 2 * nf[0] * nf[1] * nf[2]);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int64_t p0 = nf[0] * tid / nth, p1 = nf[0] * (tid + 1) / nth;
        double w[3][64];
        for (int64_t j = 0; j < M; ++j) {
            int64_t l0[3];
            for (int d = 0; d < 3; ++d) {
                if (d < ndim) {
                    l0[d] = point_weights(fine_pos(coords[d * M + j], nf[d]), ns, beta, w[d]);
                } else {
                    l0[d] = 0;
                    w[d][0] = 1.0;
                }
            }
            double cr = c[2 * j], ci = c[2 * j + 1];
            for (int a = 0; a < nsd[0]; ++a) {
                int64_t ia = wrap(l0[0] + a, nf[0]);
                if (ia < p0 || ia >= p1) continue;
                for (int b = 0; b < nsd[1]; ++b) {
                    int64_t ib = wrap(l0[1] + b, nf[1]);
                    double wab = w[0][a] * w[1][b];
                    double* row = grid + 2 * ((ia * nf[1] + ib) * nf[2]);
                    for (int q = 0; q < nsd[2]; ++q) {
                        int64_t ic = wrap(l0[2] + q, nf[2]);
                        double ww = wab * w[2][q];
                        row[2 * ic] += ww * cr;
                        row[2 * ic + 1] += ww * ci;
                    }
                }
            }
        }
    }
}

/* Interpolate the periodic fine grid at M points with the ES kernel. */
void ora_interp(int ndim, const int64_t* nf_in, int64_t M, const double* coords, const double* grid,
                double* c, int ns, double beta, int nthreads)
{
    int64_t nf[3];
    for (int d = 0; d < 3; ++d) nf[d] = d < ndim ? nf_in[d] : 1;
    int nsd[3];
    for (int d = 0; d < 3; ++d) nsd[d] = d < ndim ? ns : 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        double w[3][64];
        int64_t l0[3];
        for (int d = 0; d < 3; ++d) {
            if (d < ndim) {
                l0[d] = point_weights(fine_pos(coords[d * M + j], nf[d]), ns, beta, w[d]);
            } else {
                l0[d] = 0;
                w[d][0] = 1.0;
            }
        }
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < nsd[0]; ++a) {
            int64_t ia = wrap(l0[0] + a, nf[0]);
            for (int b = 0; b < nsd[1]; ++b) {
                int64_t ib = wrap(l0[1] + b, nf[1]);
                double wab = w[0][a] * w[1][b];
                const double* row = grid + 2 * ((ia * nf[1] + ib) * nf[2]);
                for (int q = 0; q < nsd[2]; ++q) {
                    int64_t ic = wrap(l0[2] + q, nf[2]);
                    double ww = wab * w[2][q];
                    sr += ww * row[2 * ic];
                    si += ww * row[2 * ic + 1];
                }
            }
        }
        c[2 * j] = sr;
        c[2 * j + 1] = si;
    }
}
