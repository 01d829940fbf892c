/*
 * libhasty_hip: C ABI of the MI355X (gfx950) implementation of HastyCompute's batched
 * SENSE hot path. Plain pointers and sizes only; every pointer named "device" or any
 * data array is HIP device memory; `stream` is a hipStream_t (NULL = default stream).
 *
 * Every function returns 0 on success and a nonzero code on error (1 invalid
 * argument, 2 HIP runtime error, 3 matrix not positive definite, 4 unsupported,
 * 5 out of memory); hasty_last_error() returns the message of the calling thread's
 * last failure. No exception crosses this boundary (the reference's C++ layer throws
 * std::runtime_error and its FFI converts through future_catcher,
 * lib/torch_util.cpp:160-200; the C++ layer over this ABI rethrows the message).
 *
 * Complex data is interleaved (re, im) float32 ("complex64") unless a function says
 * otherwise. Multi-dimensional arrays are row-major with the slowest dimension first,
 * as torch tensors are.
 */
#ifndef HASTY_HIP_H
#define HASTY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" if none). */
const char* hasty_last_error(void);
/* ABI version (major * 100 + minor). */
int hasty_version(void);

/* ---------------------------------------------------------------------------------
 * NUFFT plans. Same call shape as the cufinufft API the reference binds:
 * cufinufft(f)_makeplan / _setpts / _execute / _destroy at lib/fft/nufft.cpp:537-637
 * (makeplan + setpts), 639-768 (execute), 441-485 (destroy); plus an explicit stream.
 * nmodes is in FINUFFT order (fastest-varying first), as the reference passes
 * _nmodes_flipped (nufft.cpp:430-432); x pairs with nmodes[0] (the reference passes
 * coords row ndim-1 as x, nufft.cpp:579-583). Frequencies use modeord 0 (mode index n
 * is k = n - floor(N/2)); type 1 (points -> modes) and type 2 (modes -> points) sums
 * use sign iflag (>= 0: +1). fp64 selects complex128/float64 data and coordinates.
 * Coordinates are borrowed: they must stay valid until setpts is called again or
 * the plan is destroyed (as with cufinufft).
 * ------------------------------------------------------------------------------- */
typedef struct hasty_nufft_plan_s* hasty_nufft_plan;

int hasty_nufft_makeplan(int type, int dim, const int64_t* nmodes, int iflag, int ntransf, double tol, int fp64,
                         hasty_nufft_plan* plan);
int hasty_nufft_setpts(hasty_nufft_plan plan, int64_t M, const void* x, const void* y, const void* z, void* stream);
/* type 1: c [ntransf][M] -> f [ntransf][modes]; type 2: f -> c. Unscaled, like cufinufft. */
int hasty_nufft_execute(hasty_nufft_plan plan, void* c, void* f, void* stream);
/* Multiply every output by `scale` inside the pipeline (the reference divides by
 * sqrt(prod nmodes) after each transform, nufft.cpp:487-506). Default 1. */
int hasty_nufft_set_scale(hasty_nufft_plan plan, double scale);
int hasty_nufft_destroy(hasty_nufft_plan plan);
/* Fine grid (FINUFFT order), kernel width and ES shape of a plan. */
int hasty_nufft_info(hasty_nufft_plan plan, int64_t* nf, int* ns, double* beta);

/* ---------------------------------------------------------------------------------
 * Toeplitz-embedded normal operator.
 * ------------------------------------------------------------------------------- */
/* Replaces CUDANormalNufftToeplitz::build_diagonal (lib/fft/nufft.cpp:932-1053).
 * coords [ndim][M] (float32, or float64 if coord_fp64), nmodes: image dims (slowest
 * first), weights [M] real float32 or NULL (= 1). Built with an FP64 type-1 NUFFT at
 * `tol` on the GPU. Outputs (either may be NULL):
 *   diag_ref: complex64 [2N...], the reference's diagonal fftn(ifftshift(psf)) with the
 *             doubled-grid scaling (2N)^-d;
 *   kernel:   float32 [2N...], the real kernel of hasty_sense_normal_toeplitz. */
int hasty_toeplitz_build(const void* coords, int coord_fp64, int64_t M, int ndim, const int64_t* nmodes,
                         const float* weights, void* diag_ref, float* kernel, double tol, void* stream);

/* Replaces CUDANormalNufftToeplitz::apply / apply_add / apply_addcmul / apply_inplace
 * (nufft.cpp:1163-1197): res = ifftn(fftn(in, s=2N) * diag)[:N];
 * mode 0: out = res; 1: out += res; 2: out += res * mul; 3: out = res (out may be in).
 * in/out/mul [ntransf][N...], diag [ntransf][2N...], storage1/2 each
 * >= ntransf * N0 * (2 N1) * (2 N2) complex64. */
int hasty_toeplitz_apply(const void* in, void* out, const void* mul, const void* diag, int ntransf, int ndim,
                         const int64_t* nmodes, int mode, void* storage1, void* storage2, void* stream);

/* ---------------------------------------------------------------------------------
 * Toeplitz SENSE normal operator (3-D), batched over problems.
 * Replaces the coil loop of sense_normal_apply (lib/mri/sense.ixx:500-578) as used by
 * CUDASenseNormal::apply and BatchedSenseNormal::apply_outer_batch
 * (lib/mri/batch_sense.cpp:383-450): for each problem p
 *   out[p] = sum_{c in coils} conj(S_c) N1(w N2(S_c x[p])) - sub[p]
 * with N1 N2 through the problem's Toeplitz kernel (built with weights w).
 * x, out, sub: [P][N0][N1][N2] complex64 (sub may be NULL); smaps [C][N0][N1][N2];
 * coils: DEVICE int32 array of ncoil indices into smaps; kernels: float32
 * [nk][2N0][2N1][2N2] as written by hasty_sense_normal_prepare with `layout`;
 * kernel_index: HOST int32 [P], kernel of each problem. coil_batch coils share one
 * oversampled-grid workspace; problem_batch problems go through each launch.
 * Layouts 1 and 2 (N1 == N2 in {32, 64, 96, 128, 160}, 2 N0 2,3,5-smooth with its
 * z-convolution in LDS: any 2 N0 = 64 k up to 320, others up to 315) run the fused
 * three-kernel path (2: the register plane kernels at N1 = N2 = 128, 1: the LDS or
 * coset plane kernels; at 160 always the coset kernels); layout 0 the generic
 * five-pass path. The apply follows the kernels' layout, whatever
 * the sense_xy_variant option is at that time.
 * ------------------------------------------------------------------------------- */
size_t hasty_sense_normal_workspace_size(const int64_t* nmodes3, int coil_batch, int problem_batch);
/* Convert nk kernels from hasty_toeplitz_build's natural layout into the operator's
 * layout (*layout receives 0, 1 or 2, chosen from the sense_fused and sense_xy_variant
 * options at this call; prepared may equal natural only for layout 0). */
int hasty_sense_normal_prepare(const float* natural, float* prepared, int64_t nk, const int64_t* nmodes3,
                               int* layout, void* stream);
int hasty_sense_normal_toeplitz(const void* x, void* out, const void* smaps, const int32_t* coils, int ncoil,
                                const float* kernels, int nkernels, int layout, const int32_t* kernel_index,
                                const void* sub, int nproblems, const int64_t* nmodes3, void* workspace,
                                size_t workspace_bytes, int coil_batch, void* stream);
/* The same for problems that do not share coil maps or coil counts (the batched
 * reconstruction server): problem p uses smaps rows coils[i] + p * coil_problem_stride
 * (a pool of nsmaps rows; nsmaps < 0 skips the range check) for i < problem_ncoil[p]
 * (1 <= problem_ncoil[p] <= ncoil; given as a HOST array, plus the same values in a
 * DEVICE array for the fused layout; both NULL: every problem uses all ncoil).
 * hasty_sense_normal_toeplitz is this call with stride 0 and no counts. The reference
 * runs one operator per problem with its own maps (batch_sense.cpp:383-450); this lets
 * problems with different maps and coil counts share launches. */
int hasty_sense_normal_toeplitz_ex(const void* x, void* out, const void* smaps, int64_t nsmaps, const int32_t* coils,
                                   int ncoil, int coil_problem_stride, const int32_t* problem_ncoil_host,
                                   const int32_t* problem_ncoil_dev, const float* kernels, int nkernels, int layout,
                                   const int32_t* kernel_index, const void* sub, int nproblems,
                                   const int64_t* nmodes3, void* workspace, size_t workspace_bytes, int coil_batch,
                                   void* stream);

/* Library options (no counterpart in the reference; tuning switches):
 *   "sense_fused" (default 1): hasty_sense_normal_prepare may choose the fused layout;
 *   "sense_xy_variant" 0..3 (default 2): plane-kernel family of the fused normal operator
 *     (0 LDS Stockham, 1 coset lane-DFT, 2 register radix-16 at 128 x 128 planes and
 *     coset elsewhere, 3 coset with an LDS-DMA inverse), set before preparing kernels;
 *   "sense_zconv_batch" 1..16 (default 16): problems per z-convolution launch of the fused
 *     normal operator at N0 = 128.
 *   "nufft_locality" 0..4 (default 4): point ordering of plans made after the call: bit 0
 *     numbers the bins by 4^3 super-tiles, bit 1 hands the spreading and interpolation
 *     workgroups to the XCDs in contiguous runs; 4 picks 3 for type-1 and 0 for type-2
 *     plans.
 *   "sense_zconv_variant" 0..3 (default 2): z-convolution kernel at N0 = 128 (0: 16
 *     columns per workgroup, one LDS transpose buffer; 1: 8 columns; 2: 16 columns, two
 *     buffers; 3: 8 columns, two buffers) and at N0 = 160 (0: 20 x 16 register FFT, one
 *     buffer; 2: the same with two buffers; 1, 3: LDS Stockham passes).
 *   "llr_max_sweeps" 1..15 (default 15): cap on the Jacobi sweeps of the LLR eigensolver.
 *   "llr_eig_threads" 256, 512 or 768 (default 512): the LLR eigensolver's workgroup.
 *   "nufft_spread_tile" 0 or 12..28 (default 0 = 20): 3-D spreading tile edge in fine-grid
 *     cells, halo included, of new plans (the tile must fit in LDS: 20 for complex128).
 *   "nufft_spread_chunk" 0 or 32..8192 (default 0 = 128 for complex64, 1024 for
 *     complex128): points per spreading workgroup of new plans.
 *   "nufft_spread_variant" 0..6 (default 6): 3-D type-1 spreader; 0 one lane per
 *     point with LDS atomics, 1 window rows with one wave per tile, 2 window rows with four
 *     waves and LDS atomics, 3 four waves clearing and flushing the tile while one wave
 *     accumulates without atomics, 4 and 5 four and eight waves each owning the tile
 *     planes congruent to its index, 6 output-driven: one workgroup per 16^3-cell core
 *     of the fine grid spreads the points whose windows overlap it and stores the core
 *     once (no zero pass, no global atomics outside crowded cores; plans whose fine-grid
 *     dims are multiples of 16 and at least 48, binned at setpts; other plans run 3).
 *   "fft_r16" 0 or 1 (default 1): complex64 axis FFT passes of length 128, 192, 256 or
 *     320 (the NUFFT at 64^3 .. 160^3 modes) without coil multipliers, accumulation or
 *     convolution run as a 16 x R four-step FFT in registers; 0 keeps them on the LDS
 *     Stockham passes.
 *   "fft_strided_lines" 0..64 (default 0 = auto): lines per workgroup of the axis FFT
 *     passes whose lines are strided (the NUFFT's outer axes, the Toeplitz kernel build,
 *     the generic normal operator); auto gives every pass 2048 / L (complex64) or
 *     1024 / L (complex128) lines, 1..64, and strided complex128 passes at least 4.
 * Returns 0, or 1 for an unknown name or a value out of range. */
int hasty_set_option(const char* name, int64_t value);
/* Current value of an option of hasty_set_option (0, or 1 for an unknown name). */
int hasty_get_option(const char* name, int64_t* value);

/* Autotuner (no counterpart in the reference, whose launch shapes are fixed inside
 * cuFINUFFT, cuFFT and cuSOLVER; it replaces the hand-set options above for a given
 * shape). Candidates are timed with HIP events on `stream` in a fixed coordinate-wise
 * order starting from the defaults; a candidate must be 2 % faster to replace the
 * incumbent. Results are cached per shape for the process. None of these may run
 * while `stream` is being captured.
 * hasty_autotune_sense_normal: the batched Toeplitz SENSE normal operator at image size
 *   nmodes3 with ncoil coils and nproblems problems (at most 8 timed), on synthetic
 *   data: plane kernels, problems per launch, problems per z-convolution launch and
 *   coils per workspace, and at N0 = 128 the z-convolution kernel. Sets
 *   sense_xy_variant, sense_zconv_variant and sense_zconv_batch to the winners (prepare
 *   kernels after this call) and writes choice[5] = {sense_xy_variant,
 *   sense_zconv_batch, problem_batch, coil_batch, sense_zconv_variant} (size the
 *   workspace with problem_batch and coil_batch) and the time per problem-apply.
 * hasty_nufft_autotune: the 3-D type-1 spreaders (nufft_spread_variant 0..3) of `plan`
 *   on its own points, executing c -> f (f is overwritten), then (complex64 plans) the
 *   spreading tile edge and points per workgroup (the points are re-binned for each;
 *   the stream is synchronised); the plan keeps the fastest. Other plans keep theirs.
 *   *spread_variant receives the spreader, *ms the execute time;
 *   hasty_nufft_plan_shape reports a plan's tile edge, chunk and spreader.
 * hasty_autotune_llr: the Gram row split of hasty_llr_normal_blocks for this image and
 *   block grid (threshold 0 on synthetic data); later calls with the same number of
 *   blocks per pass, rows and frames use it (size workspaces after this call); then the
 *   eigensolver's workgroup size, which it sets as the llr_eig_threads option. */
int hasty_autotune_sense_normal(const int64_t* nmodes3, int ncoil, int nproblems, int64_t* choice,
                                double* ms_per_problem, void* stream);
int hasty_nufft_autotune(hasty_nufft_plan plan, void* c, void* f, int64_t* spread_variant, double* ms, void* stream);
int hasty_nufft_plan_shape(hasty_nufft_plan plan, int64_t* tile, int64_t* chunk, int64_t* spread_variant);
int hasty_autotune_llr(const int64_t* dims5, const int64_t* block_strides, const int64_t* block_shape,
                       int64_t* nsplit, double* ms, void* stream);

/* Kernel timing (no counterpart in the reference; a measurement hook for benchmarks).
 * While enabled, each instrumented launch is bracketed by two HIP events on its own
 * stream, and the durations are summed per kernel name. hasty_timing_enable clears the
 * statistics; hasty_timing_entries waits for the recorded events and returns the number
 * of kernel names (or -1 on error); hasty_timing_read copies entry `index` (names in
 * sorted order, NUL-terminated, truncated to name_capacity). */
int hasty_timing_enable(int on);
int hasty_timing_entries(void);
int hasty_timing_read(int index, char* name, int name_capacity, double* total_ms, int64_t* count);

/* Streaming-bandwidth probe (no counterpart in the reference; a measurement hook for
 * the roofline's ceiling): reads nread device arrays src[0..nread) and writes nwrite
 * arrays dst[0..nwrite) of `bytes` each (a multiple of 16), 16 bytes per lane, with
 * non-temporal loads and stores if `nontemporal`. Mixes (nread:nwrite): 1:0, 2:0, 0:1,
 * 0:2, 1:1, 2:1, 1:2, 1:4, 4:1. A read-only mix needs dst[0], a 16-byte sink that is
 * not written for finite data. Timed as "stream_probe" by hasty_timing_*. */
int hasty_stream_probe(int nread, int nwrite, int nontemporal, const void* const* src, void* const* dst,
                       int64_t bytes, void* stream);

/* ---------------------------------------------------------------------------------
 * HIP graphs of library launch sequences. The reference drives each frame's operator
 * from a host worker thread per device context (ContextThreadPool,
 * lib/threading/thread_pool.ixx:110-259, fanned out by batch_sense.cpp:349-450 and
 * BatchedConjugateGradient, lib/op/opalgs.ixx:84-136); here a whole batched iteration
 * (normal operator over every problem + the device CG step) is recorded once and
 * replayed as one graph launch.
 * hasty_graph_begin starts capturing `stream` (a created stream, not NULL; thread-local
 * capture mode); every library call made on that stream by this thread until
 * hasty_graph_end is recorded instead of run. The recorded calls must not allocate or
 * synchronise: call each entry point once on the same shapes before capturing (that
 * creates its cached tables), and keep every buffer alive and unmoved while the graph
 * exists; host arrays (kernel_index, coil counts) are read at capture time.
 * Recorded launches are not timed (hasty_timing_*). hasty_graph_launch replays the
 * sequence on `stream`; hasty_graph_nodes returns its number of nodes (-1 for NULL).
 * ------------------------------------------------------------------------------- */
typedef struct hasty_graph_s* hasty_graph;

int hasty_graph_begin(void* stream);
int hasty_graph_end(void* stream, hasty_graph* graph);
int hasty_graph_launch(hasty_graph graph, void* stream);
int64_t hasty_graph_nodes(hasty_graph graph);
int hasty_graph_destroy(hasty_graph graph);

/* ---------------------------------------------------------------------------------
 * Device-side batched conjugate gradient, the loop body of ConjugateGradient::run
 * (lib/op/opalgs.cpp:50-96) for P problems of n complex64 unknowns with FP64 scalars
 * kept on the device (no host synchronisation per iteration).
 * state: device double [P][8]: rzold, alpha, beta, done, not_positive_definite, pAp,
 * iterations. The caller computes Ax / Ap with the normal operator between calls:
 *   hasty_cg_init(b, A x0) -> r, p; loop { Ap = A p; hasty_cg_iterate(...) }.
 * A problem stops (alpha = 0) once sqrt(rzold) < tol or when pAp <= 0, which also
 * sets its not_positive_definite flag (the reference throws "A was not positive
 * definite", opalgs.cpp:71-72; the C++ layer throws the same after reading state).
 * ------------------------------------------------------------------------------- */
size_t hasty_cg_workspace_size(int64_t nproblems, int64_t n);
int hasty_cg_init(int64_t nproblems, int64_t n, const void* b, const void* Ax, void* r, void* p, double* state,
                  void* workspace, void* stream);
int hasty_cg_iterate(int64_t nproblems, int64_t n, void* x, void* r, void* p, const void* Ap, double* state,
                     void* workspace, double tol, void* stream);
/* The iteration of ConjugateGradient::run on l2_reg_linop(A, lambda) = A + lambda I
 * (python/python/hastypy/base/opalg.py:626-627, the reference's Tikhonov-regularised
 * normal operator): the caller passes Ap = A p and lambda p is added inside the CG
 * kernels, so the regularisation costs no pass over the data. Initialise with
 * hasty_cg_init(b, A x0 + lambda x0) (for x0 = 0, A x0 alone). lambda >= 0. */
int hasty_cg_iterate_l2(int64_t nproblems, int64_t n, void* x, void* r, void* p, const void* Ap, double* state,
                        void* workspace, double tol, double lambda, void* stream);
/* The preconditioned loop of ConjugateGradient::run (the _P branch, opalgs.cpp:56-90),
 * z = P r computed by the caller between the calls:
 *   hasty_cg_init(b, A x0) -> r; z = P r; hasty_cg_init_precond(r, z) -> p = z, rzold = <r, z>;
 *   loop { Ap = A p; hasty_cg_iterate_xr(...) -> x, r; z = P r; hasty_cg_iterate_pz(r, z) -> p }.
 * Stopping and the positive-definiteness flag as in hasty_cg_iterate (resid = sqrt(<r, z>)). */
int hasty_cg_init_precond(int64_t nproblems, int64_t n, const void* r, const void* z, void* p, double* state,
                          void* workspace, void* stream);
int hasty_cg_iterate_xr(int64_t nproblems, int64_t n, void* x, void* r, const void* p, const void* Ap, double* state,
                        void* workspace, double tol, void* stream);
int hasty_cg_iterate_pz(int64_t nproblems, int64_t n, const void* r, const void* z, void* p, double* state,
                        void* workspace, void* stream);

/* ---------------------------------------------------------------------------------
 * DCT and the L1-DCT prox.
 * ------------------------------------------------------------------------------- */
/* In place, on batch volumes [batch][D0][D1][D2] complex64, along the axes whose bit is
 * set in axis_mask (bit d: axis Dd, each <= 512): the unnormalised DCT-II
 * y_k = 2 sum_n x_n cos(pi k (2n+1) / 2N) or its inverse (scipy.fft.dctn / idctn with
 * norm=None, which ProximalL1DCT calls through cupyx, python/python/hastypy/base/
 * proximal.py:148-152). */
int hasty_dct3(void* data, int64_t batch, const int64_t* dims3, int axis_mask, int inverse, void* stream);
/* Replaces ProximalL1DCT.tensor_apply (proximal.py:139-163): dct, then
 * c <- exp(i angle c) max(|c| - thresh, 0), then the inverse dct, in place. */
int hasty_l1dct_prox(void* data, int64_t batch, const int64_t* dims3, int axis_mask, double thresh, void* stream);

/* ---------------------------------------------------------------------------------
 * Circulant preconditioner.
 * ------------------------------------------------------------------------------- */
/* Replaces CirculantPreconditionerOp::build_diagonal (lib/mri/precond.cpp:65-117, after
 * sigpy.mri.circulant_precond): diag = 1 / p, p the eigenvalues of the circulant matrix
 * closest in Frobenius norm to sum_c S_c^H N^H diag(sqrt(w)) N S_c, plus lambda (p == 0
 * -> 1). smaps [ncoil][N...] complex64, coords [ndim][M] (float32, or float64 if
 * coord_fp64), weights [M] or NULL (= 1); N = nmodes (slowest first, 2N 2,3,5-smooth).
 * FP64 on the GPU (type-1 NUFFT of the PSF at `tol`); diag: complex64 [N...], in the
 * reference's centred order (frequency n - N/2 at index n) if centered, else FFT order. */
int hasty_circulant_precond_build(const void* smaps, int64_t ncoil, const void* coords, int coord_fp64, int64_t M,
                                  int ndim, const int64_t* nmodes, const float* weights, double lambda, int centered,
                                  void* diag, double tol, void* stream);
/* Replaces CirculantPreconditionerOp::apply (precond.cpp:27-50) for a batch:
 * out[p] = ifftn(d * fftn(in[p])), or with adjoint fftn(conj(d) * ifftn(in[p])), with
 * d = diag[diag_index[p]] (HOST int32 [nproblems], or NULL: diag p). The operator is a
 * circular convolution, so `centered` (fftc / ifftc in the reference) only says in
 * which order diag is stored. in / out [nproblems][N...] (out may equal in), diag
 * [ndiag][N...]; workspace of hasty_circulant_precond_workspace_size bytes. */
size_t hasty_circulant_precond_workspace_size(int64_t ndiag, int ndim, const int64_t* nmodes, int centered, int adjoint);
int hasty_circulant_precond_apply(const void* in, void* out, const void* diag, int64_t ndiag, int64_t nproblems,
                                  const int32_t* diag_index, int ndim, const int64_t* nmodes, int centered,
                                  int adjoint, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------
 * Locally-low-rank SVT.
 * ------------------------------------------------------------------------------- */
/* The block shifts of OldNormal3DBlocksSVT::apply (lib/op/svt.cpp:378-391): a default-
 * constructed std::default_random_engine, uniform_int_distribution(0, block_shape[d]).
 * shifts_out: int32 [3][block_iter]. */
int hasty_llr_shifts(int64_t block_iter, const int64_t* block_shape, int32_t* shifts_out);

size_t hasty_llr_workspace_size(const int64_t* dims5, const int64_t* block_strides, const int64_t* block_shape);

/* Replaces OldNormal3DBlocksSVT::apply (svt.cpp:354-449, the TorchScript target of
 * HastySVT.Normal3DBlocksSVT, python/cpp/py_svt.cpp:47-60) on device data
 * [F][E][X][Y][Z] complex64 (dims5), in place: block_iter passes over the strided block
 * grid (shifts from hasty_llr_shifts if shifts == NULL), skipping blocks whose end
 * reaches the image edge (svt.cpp:404); each block's Casorati matrix (E*bx*by*bz x F,
 * svt.cpp:37-60) is replaced by svt_soft(., thresh) or svt_hard(., int(thresh + 0.5)).
 * Blocks of one pass run as one batched launch when they are disjoint. */
int hasty_llr_normal_blocks(void* data, const int64_t* dims5, const int64_t* block_strides,
                            const int64_t* block_shape, int64_t block_iter, const int32_t* shifts, double thresh,
                            int soft, void* workspace, size_t workspace_bytes, void* stream);

/* Threshold an explicit list of disjoint blocks (corners: HOST int32 [nblk][3]); the
 * batched core of Random3DBlocksSVT::apply (svt.cpp:269-343) and of the above. */
int hasty_llr_blocks(void* data, const int64_t* dims5, const int64_t* block_shape, const int32_t* corners,
                     int64_t nblk, double thresh, int soft, void* workspace, size_t workspace_bytes, void* stream);
size_t hasty_llr_blocks_workspace_size(const int64_t* dims5, const int64_t* block_shape, int64_t nblk);

/* The blocks one pass of hasty_llr_normal_blocks visits for one shift (shift3: int32
 * [3]): *count blocks, their corners written to corners_out (HOST int32 [capacity][3])
 * when it is not NULL; no GPU work. With hasty_llr_blocks_dev it splits the prox into a
 * host step done once (the grid of each shift of a reconstruction) and launches that a
 * HIP graph can capture. */
int hasty_llr_grid_corners(const int64_t* dims5, const int64_t* block_strides, const int64_t* block_shape,
                           const int32_t* shift3, int32_t* corners_out, int64_t capacity, int64_t* count);
/* hasty_llr_blocks with the corners already on the device (DEVICE int32 [nblk][3],
 * disjoint blocks): no host copy and no synchronisation, so it can run under graph
 * capture. A block that does not lie inside the image is skipped (left unchanged). */
int hasty_llr_blocks_dev(void* data, const int64_t* dims5, const int64_t* block_shape, const int32_t* corners_dev,
                         int64_t nblk, double thresh, int soft, void* workspace, size_t workspace_bytes, void* stream);

/* svt_soft / svt_hard (svt.cpp:62-169) on one row-major [R][F] complex64 matrix, in
 * place. workspace >= hasty_svt_workspace_size(R, F). */
size_t hasty_svt_workspace_size(int64_t R, int64_t F);
int hasty_svt_matrix(void* A, int64_t R, int64_t F, double thresh, int soft, void* workspace, size_t workspace_bytes,
                     void* stream);

/* hasty_svt_matrix without a workspace argument and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * result as hasty_svt_matrix (same kernels). Its scratch (Gram partial sums, W,
 * recorded rotations) is a static 4 MiB device array per GPU, enough for any R and
 * F <= 128 (3.13 MB at F = 128); larger needs are rejected with an error. Nothing in
 * it needs to start at zero (every part is written before it is read), and no host
 * copy or synchronisation is made. All calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads) race and must be serialised by the
 * caller. */
int hasty_svt_matrix_noalloc(void* A, int64_t R, int64_t F, double thresh, int soft, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HASTY_HIP_H */
