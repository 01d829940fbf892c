// ES-kernel NUFFT type-1 / type-2 plans on gfx950 (the MI355X replacement for the
// cufinufft plan/setpts/execute/destroy calls of lib/fft/nufft.cpp:537-768).
#pragma once

#include "common.hpp"

namespace hasty_hip {

struct NufftPlan {
    int type = 1;       // 1: points -> modes, 2: modes -> points
    int dim = 3;
    int isign = 1;
    int ntransf = 1;
    int fp64 = 0;
    double tol = 1e-6;
    double scale = 1.0;  // applied to every output (1/sqrt(prod modes) in the reference)
    int64_t n[3] = {1, 1, 1};   // modes, slowest dim first (torch order)
    int64_t nf[3] = {1, 1, 1};  // fine grid
    int ns = 2;
    double beta = 0;
    int bs[3] = {1, 1, 1};
    int64_t nbins[3] = {1, 1, 1};
    // device buffers owned by the plan
    void* fine = nullptr;
    void* tmp = nullptr;
    void* phiinv[3] = {nullptr, nullptr, nullptr};
    // points (borrowed: must outlive the plan, as with cufinufft)
    int64_t M = 0;
    const void* pts[3] = {nullptr, nullptr, nullptr};
    int* counts = nullptr;    // nbins
    int* binstart = nullptr;  // nbins + 1
    int* rank = nullptr;      // M
    int* sortidx = nullptr;   // M
    int* subbin = nullptr;    // nsub
    int* suboff = nullptr;    // nsub
    int* nsubbin = nullptr;   // nbins + 1 (scan)
    int64_t cap_M = 0;
    int64_t cap_sub = 0;
    int64_t nsub = 0;
    int device = 0;
    int spread_variant = -1;  // 3-D spreader of this plan (-1: the library option)
    int locality = 3;         // bit 0 super-tile bin order, bit 1 XCD-aware blocks (fixed at makeplan)
    int chunk = 256;          // points per spreader subproblem at most (fixed at makeplan)
    // output-driven 3-D spreader (variant 6): points binned by 16^3-cell cores,
    // built on the first execute after setpts
    int gather_built = 0;
    int* g_counts = nullptr;    // points per core bin
    int* g_binstart = nullptr;  // core bins + 1 (scan)
    int* g_l0 = nullptr;        // [3][M] first window cell per dim, in core-bin order
    void* g_frac = nullptr;     // [3][M] fine-grid position minus that cell (plan precision)
    int* g_idx = nullptr;       // [M] point index, in core-bin order
    int* g_rank = nullptr;      // [M] rank within the core bin
    int* g_pcount = nullptr;    // (point, core) pairs per (core, neighbour bin): 27 per core
    int* g_pstart = nullptr;    // 27 cores + 1 (scan)
    int* g_pairs = nullptr;     // sorted point of each pair, grouped by core
    int* g_prel = nullptr;      // window corner relative to the core
    int* g_list = nullptr;      // (core, part) entries, then the crowded cores
    int64_t g_cap_M = 0;
    int64_t g_cap_pairs = 0;
    int64_t g_cap_list = 0;
    int64_t g_first = 0;        // (core, part) entries of the spreading launch
    int64_t g_rest = 0;         // of them the crowded cores' parts, first, adding with atomics
    int64_t g_nzero = 0;        // crowded cores (listed after the entries), zeroed beforehand
};

// nmodes in FINUFFT order (fastest-varying first), as the reference passes
// _nmodes_flipped (lib/fft/nufft.cpp:430-432).
NufftPlan* nufft_makeplan(int type, int dim, const int64_t* nmodes, int iflag, int ntransf, double tol,
                          int fp64);
// x pairs with nmodes[0] (fastest), y with nmodes[1], z with nmodes[2].
void nufft_setpts(NufftPlan* p, int64_t M, const void* x, const void* y, const void* z, hipStream_t s);
// type 1: c [ntransf][M] -> f [ntransf][modes]; type 2: f -> c.
void nufft_execute(NufftPlan* p, void* c, void* f, hipStream_t s);
void nufft_destroy(NufftPlan* p);
// 3-D spreader: 0 one lane per point with LDS atomics; window rows with
// 1 one wave per tile, 2 four waves and LDS atomics, 3 four waves clearing and
// flushing the tile while the first wave accumulates without atomics; 4 and 5 four and
// eight waves each owning the tile planes congruent to its index; 6 (default) output-driven:
// one workgroup per 16^3-cell core spreads the points whose windows overlap it and writes
// the core once (plans whose fine-grid dims are multiples of 16 and at least 48; other
// plans take 3)
void nufft_set_spread_variant(int v);
// whether the output-driven spreader (variant 6) serves the plan
bool nufft_gather_fits(const NufftPlan* p);
int nufft_get_spread_variant();
// Point ordering of new plans: bit 0 groups bins into 4^3 super-tiles, bit 1 maps the
// spreading and interpolation workgroups to XCDs in contiguous runs; 4 (default): 3 for
// type-1 plans, 0 for type-2 plans (measured best for spreading and for interpolation).
void nufft_set_locality(int v);
int nufft_get_locality();
// 3-D tile edge including the halo (0: 20 cells) and points per spreader subproblem
// (0: 128 for complex64, 1024 for complex128) of new plans
void nufft_set_tile(int v);
int nufft_get_tile();
void nufft_set_chunk(int v);
int nufft_get_chunk();
// Re-bin a 3-D plan's points (as last given to setpts) with another tile edge (halo
// included) and subproblem size; synchronises the stream.
void nufft_retile(NufftPlan* p, int tile, int chunk, hipStream_t s);
int nufft_tile_edge(const NufftPlan* p);

// Time the 3-D type-1 spreaders on the plan's points (c -> f, f overwritten) and keep
// the fastest for this plan; returns it and its execute time in *ms. Other plans: no-op
// returning the current spreader.
int nufft_autotune(NufftPlan* p, void* c, void* f, hipStream_t s, double* ms);

// ES kernel parameters for upsampling factor 2.
void es_params(double tol, int* ns, double* beta);

// Fill a complex buffer with zeros (a kernel, so it is safe under graph capture).
template <typename R> void zero_fill(cplx<R>* p, int64_t n, hipStream_t s);

}  // namespace hasty_hip
