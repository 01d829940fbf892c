/* The oracle's C restatement (oracle/nufft_oracle.c, oracle/llr_shifts.cpp) under
 * AddressSanitizer + UndefinedBehaviorSanitizer, with 4 OpenMP threads, in 1, 2 and 3
 * dimensions and at odd and even sizes. Each run also checks an identity that an
 * indexing mistake would break:
 *   - type-2 and type-1 direct sums are adjoint with opposite signs:
 *     <c, T2(f, +1)> = <T1(c, -1), f>;
 *   - ES spreading and interpolation are adjoint: <c, interp(g)> = <spread(c), g>;
 *   - the LLR shift sequence starts at zero and stays within [0, block_shape[d]].
 * Exits 0 when every check holds. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void ora_nudft_type2(int ndim, const int64_t* nmodes, int64_t M, const double* coords, const double* in,
                     double* out, int isign, int nthreads);
void ora_nudft_type1(int ndim, const int64_t* nmodes, int64_t M, const double* coords, const double* in,
                     double* out, int isign, int nthreads);
void ora_spread(int ndim, const int64_t* nf, int64_t M, const double* coords, const double* c, double* grid,
                int ns, double beta, int nthreads);
void ora_interp(int ndim, const int64_t* nf, int64_t M, const double* coords, const double* grid, double* c,
                int ns, double beta, int nthreads);
void ora_llr_shifts(int64_t block_iter, const int64_t* block_shape, int32_t* shifts);

static const double kPi = 3.14159265358979323846;
static int failures = 0;

static double uniform(uint64_t* s)
{
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) / 9007199254740992.0;
}

static void fill(double* a, int64_t n, double lo, double hi, uint64_t* s)
{
    for (int64_t i = 0; i < n; ++i) a[i] = lo + (hi - lo) * uniform(s);
}

/* <a, b> = sum conj(a) b over n complex values: re and im parts */
static void inner(const double* a, const double* b, int64_t n, double* re, double* im)
{
    double r = 0.0, i = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        r += a[2 * k] * b[2 * k] + a[2 * k + 1] * b[2 * k + 1];
        i += a[2 * k] * b[2 * k + 1] - a[2 * k + 1] * b[2 * k];
    }
    *re = r;
    *im = i;
}

static void expect_equal(const char* what, int ndim, double ar, double ai, double br, double bi)
{
    const double scale = fabs(ar) + fabs(ai) + fabs(br) + fabs(bi) + 1e-300;
    const double err = (fabs(ar - br) + fabs(ai - bi)) / scale;
    if (!(err < 1e-12)) {
        ++failures;
        fprintf(stderr, "%s, ndim %d: relative mismatch %.3e\n", what, ndim, err);
    }
}

static void run_nudft(int ndim, const int64_t* nmodes, int64_t M, uint64_t* s)
{
    int64_t N = 1;
    for (int d = 0; d < ndim; ++d) N *= nmodes[d];
    double* coords = malloc(sizeof(double) * ndim * M);
    double* f = malloc(sizeof(double) * 2 * N);
    double* c = malloc(sizeof(double) * 2 * M);
    double* t2 = malloc(sizeof(double) * 2 * M);
    double* t1 = malloc(sizeof(double) * 2 * N);
    fill(coords, ndim * M, -kPi, kPi, s);
    fill(f, 2 * N, -1.0, 1.0, s);
    fill(c, 2 * M, -1.0, 1.0, s);
    ora_nudft_type2(ndim, nmodes, M, coords, f, t2, +1, 4);
    ora_nudft_type1(ndim, nmodes, M, coords, c, t1, -1, 4);
    double ar, ai, br, bi;
    inner(c, t2, M, &ar, &ai);
    inner(t1, f, N, &br, &bi);
    expect_equal("nudft type-2 / type-1 adjoint", ndim, ar, ai, br, bi);
    free(coords);
    free(f);
    free(c);
    free(t2);
    free(t1);
}

static void run_spread(int ndim, const int64_t* nf, int64_t M, int ns, uint64_t* s)
{
    int64_t G = 1;
    for (int d = 0; d < ndim; ++d) G *= nf[d];
    const double beta = 2.30 * ns;
    double* coords = malloc(sizeof(double) * ndim * M);
    double* c = malloc(sizeof(double) * 2 * M);
    double* g = malloc(sizeof(double) * 2 * G);
    double* sp = malloc(sizeof(double) * 2 * G);
    double* ip = malloc(sizeof(double) * 2 * M);
    fill(coords, ndim * M, -kPi, kPi, s);
    /* points on the periodic boundary exercise the wrap-around */
    coords[0] = -kPi;
    coords[1] = kPi * (1.0 - 1e-15);
    fill(c, 2 * M, -1.0, 1.0, s);
    fill(g, 2 * G, -1.0, 1.0, s);
    ora_spread(ndim, nf, M, coords, c, sp, ns, beta, 4);
    ora_interp(ndim, nf, M, coords, g, ip, ns, beta, 4);
    double ar, ai, br, bi;
    inner(c, ip, M, &ar, &ai);
    inner(sp, g, G, &br, &bi);
    expect_equal("ES spread / interp adjoint", ndim, ar, ai, br, bi);
    free(coords);
    free(c);
    free(g);
    free(sp);
    free(ip);
}

int main(void)
{
    uint64_t seed = 12345;
    const int64_t modes[3][3] = {{17, 1, 1}, {8, 11, 1}, {6, 7, 8}};
    const int64_t fine[3][3] = {{34, 1, 1}, {16, 22, 1}, {12, 14, 16}};
    for (int ndim = 1; ndim <= 3; ++ndim) {
        run_nudft(ndim, modes[ndim - 1], 97, &seed);
        run_spread(ndim, fine[ndim - 1], 211, 4, &seed);
        run_spread(ndim, fine[ndim - 1], 64, 7, &seed);
    }
    const int64_t block[3] = {16, 8, 5};
    int32_t shifts[3 * 9];
    ora_llr_shifts(9, block, shifts);
    for (int d = 0; d < 3; ++d) {
        if (shifts[d * 9] != 0) ++failures;
        for (int b = 0; b < 9; ++b)
            if (shifts[d * 9 + b] < 0 || shifts[d * 9 + b] > block[d]) ++failures;
    }
    printf("oracle_stress: %d failures\n", failures);
    return failures ? 1 : 0;
}
