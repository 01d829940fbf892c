// C++ operator layer with the reference's names and semantics, running on
// libhasty_hip.so through its C ABI (include/hasty_hip.h). Only torch tensors cross
// this layer's interface; every GPU call goes through the ABI on the current stream.
//
// Mirrors: lib/fft/nufft.ixx (NufftOptions, CUDANufft, CUDANufftNormal,
// CUDANormalNufftToeplitz), lib/mri/sense.ixx (CoilManipulator, CUDASense,
// CUDASenseAdjoint, CUDASenseNormal), lib/mri/batch_sense.ixx (BatchedSense,
// BatchedSenseAdjoint, BatchedSenseNormal), lib/op/svt.ixx (svt_soft/hard,
// extract/insert_block, Random3DBlocksSVT, OldNormal3DBlocksSVT), lib/op/opalgs.ixx
// (ConjugateGradient, BatchedConjugateGradient).
#pragma once

#include <torch/torch.h>

#include <array>
#include <functional>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "hasty_hip.h"

namespace hasty {

namespace detail {
void check(int rc);
void* stream();
}  // namespace detail

namespace fft {

enum NufftType { eType1 = 1, eType2 = 2, eType3 = 3 };

// lib/fft/nufft.ixx:26-66
class NufftOptions {
public:
    NufftOptions() = default;
    NufftOptions(NufftType intype, const at::optional<bool>& inpositive, const at::optional<double>& intol) : type(intype)
    {
        if (inpositive.has_value()) positive = *inpositive;
        if (intol.has_value()) tol = *intol;
    }
    static NufftOptions type1() { return {eType1, true, 1e-5}; }
    static NufftOptions type2() { return {eType2, false, 1e-5}; }
    static NufftOptions type1(double tol) { return {eType1, true, tol}; }
    static NufftOptions type2(double tol) { return {eType2, false, tol}; }

    NufftType type = eType1;
    bool positive = true;
    double tol = 1e-5;
    const NufftType& get_type() const { return type; }
    int get_positive() const { return positive ? 1 : -1; }
    double get_tol() const { return tol; }
};

at::Tensor allocate_out(const at::Tensor& coords, int ntransf);
at::Tensor allocate_adjoint_out(const at::Tensor& coords, const std::vector<int64_t>& nmodes);

// lib/fft/nufft.ixx:130-175, nufft.cpp:410-768: type 1 or 2 on the GPU, scaled by
// 1/sqrt(prod nmodes[1:]) (nufft.cpp:487-506). Coordinates are borrowed.
class CUDANufft {
public:
    CUDANufft(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const NufftOptions& opts = NufftOptions{});
    ~CUDANufft();
    CUDANufft(const CUDANufft&) = delete;
    CUDANufft& operator=(const CUDANufft&) = delete;
    void close();
    void apply(const at::Tensor& in, at::Tensor& out) const;
    at::ScalarType coord_type() const { return _type; }
    at::ScalarType complex_type() const;
    int32_t nfreq() const { return (int32_t)_nfreq; }
    int32_t ndim() const { return _ndim; }
    const std::vector<int64_t>& nmodes() const { return _nmodes; }

protected:
    void check_type1(const at::Tensor& in, const at::Tensor& out) const;
    void check_type2(const at::Tensor& in, const at::Tensor& out) const;

    at::ScalarType _type;
    int32_t _ndim;
    int32_t _ntransf;
    int64_t _nfreq;
    const at::Tensor _coords;
    std::vector<int64_t> _nmodes;
    NufftOptions _opts;
    bool _closed = false;
    hasty_nufft_plan _plan = nullptr;
    int64_t _nvoxels;
};

at::Tensor allocate_normal_out(const at::Tensor& coords, const std::vector<int64_t>& nmodes);
at::Tensor allocate_normal_storage(const at::Tensor& coords, int ntransf);

// nufft.cpp:846-920: type 2, optional func_between on the k-space storage, type 1.
class CUDANufftNormal {
public:
    CUDANufftNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const NufftOptions& forward_ops,
                    const NufftOptions& backward_ops);
    void apply(const at::Tensor& in, at::Tensor& out, at::Tensor& storage,
               at::optional<std::function<void(at::Tensor&)>> func_between) const;
    void apply_inplace(at::Tensor& in, at::Tensor& storage,
                       at::optional<std::function<void(at::Tensor&)>> func_between) const;
    void apply_forward(const at::Tensor& in, at::Tensor& out);
    void apply_backward(const at::Tensor& in, at::Tensor& out);
    at::ScalarType coord_type() const { return _forward.coord_type(); }
    at::ScalarType complex_type() const { return _forward.complex_type(); }
    int32_t nfreq() const { return _forward.nfreq(); }
    int32_t ndim() const { return _forward.ndim(); }
    const CUDANufft& get_forward() const { return _forward; }
    const CUDANufft& get_backward() const { return _backward; }

protected:
    CUDANufft _forward;
    CUDANufft _backward;
};

// nufft.cpp:932-1197. The diagonal is built in FP64 on the GPU (the tol argument is not
// used for it, as in the reference, nufft.cpp:937). apply = ifftn(fftn(x, s=2N) diag)[:N].
class CUDANormalNufftToeplitz {
public:
    static void build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double tol, at::Tensor& diagonal);
    static void build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double tol, at::Tensor& diagonal,
                               at::Tensor& storage, bool storage_is_frequency);
    static void build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double tol, at::Tensor& diagonal,
                               at::Tensor& frequency_storage, at::Tensor& input_storage);

    CUDANormalNufftToeplitz(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                            const at::optional<double>& tol,
                            const at::optional<std::reference_wrapper<at::Tensor>>& diagonal,
                            const at::optional<std::reference_wrapper<at::Tensor>>& frequency_storage,
                            const at::optional<std::reference_wrapper<at::Tensor>>& input_storage);
    CUDANormalNufftToeplitz(at::Tensor&& diagonal, const std::vector<int64_t>& nmodes);

    at::Tensor apply(const at::Tensor& in, at::Tensor& storage1, at::Tensor& storage2) const;
    void apply_add(const at::Tensor& in, at::Tensor& add_to, at::Tensor& storage1, at::Tensor& storage2) const;
    void apply_addcmul(const at::Tensor& in, at::Tensor& add_to, const at::Tensor& mul, at::Tensor& storage1,
                       at::Tensor& storage2) const;
    void apply_inplace(at::Tensor& in, at::Tensor& storage1, at::Tensor& storage2) const;
    at::Tensor get_diagonal() { return _diagonal; }

private:
    void run(const at::Tensor& in, at::Tensor& out, const at::Tensor* mul, int mode, at::Tensor& st1,
             at::Tensor& st2) const;
    std::vector<int64_t> _nmodes;
    at::Tensor _diagonal;
};

}  // namespace fft

namespace mri {

using CoilApplier = std::function<void(at::Tensor&, int32_t)>;

// lib/mri/sense.ixx:21-48
struct CoilManipulator {
    CoilManipulator& setPreApply(const CoilApplier& apply);
    CoilManipulator& setMidApply(const CoilApplier& apply);
    CoilManipulator& setPostApply(const CoilApplier& apply);
    at::optional<CoilApplier> preapplier;
    at::optional<CoilApplier> midapplier;
    at::optional<CoilApplier> postapplier;
};

// sense.ixx:87-112, 348-423
class CUDASense {
public:
    CUDASense(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
              const at::optional<fft::NufftOptions>& opts = at::nullopt);
    void apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const at::optional<at::Tensor>& imspace_storage, const at::optional<at::Tensor>& kspace_storage,
               const at::optional<CoilApplier>& premanip, const at::optional<CoilApplier>& postmanip);

private:
    fft::CUDANufft _nufft;
    std::vector<int64_t> _nmodes;
};

// sense.ixx:150-177, 426-497
class CUDASenseAdjoint {
public:
    CUDASenseAdjoint(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                     const at::optional<fft::NufftOptions>& opts = at::nullopt);
    void apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const at::optional<at::Tensor>& imspace_storage, const at::optional<at::Tensor>& kspace_storage,
               const at::optional<CoilApplier>& premanip, const at::optional<CoilApplier>& postmanip);

private:
    fft::CUDANufft _nufft;
    std::vector<int64_t> _nmodes;
};

// sense.ixx:262-327, 500-693: explicit type-2 / type-1 normal operator per coil.
class CUDASenseNormal {
public:
    CUDASenseNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                    const at::optional<fft::NufftOptions>& forward_opts = at::nullopt,
                    const at::optional<fft::NufftOptions>& backward_opts = at::nullopt);
    void apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const at::optional<at::Tensor>& imspace_storage = at::nullopt,
               const at::optional<at::Tensor>& kspace_storage = at::nullopt,
               const at::optional<CoilApplier>& premanip = at::nullopt,
               const at::optional<CoilApplier>& midmanip = at::nullopt,
               const at::optional<CoilApplier>& postmanip = at::nullopt);
    void apply_forward(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
                       const at::optional<at::Tensor>& imspace_storage = at::nullopt,
                       const at::optional<at::Tensor>& kspace_storage = at::nullopt);
    void apply_backward(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                        const std::vector<int64_t>& coils, const at::optional<at::Tensor>& kspace_storage = at::nullopt,
                        const at::optional<at::Tensor>& imspace_storage = at::nullopt);

private:
    fft::CUDANufftNormal _normal_nufft;
    std::vector<int64_t> _nmodes;
};

// The Toeplitz-embedded SENSE normal operator (3-D): equal to CUDASenseNormal::apply
// (with weights w in the middle when given) through the fused kernels. Batched over
// problems that share coil maps.
class CUDASenseNormalToeplitz {
public:
    CUDASenseNormalToeplitz(const std::vector<at::Tensor>& coords, const std::vector<int64_t>& nmodes,
                            const at::optional<std::vector<at::Tensor>>& weights = at::nullopt);
    // in/out [P, N0, N1, N2]; kernel_index[p] selects the trajectory of problem p.
    void apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const std::vector<int32_t>& kernel_index, const at::optional<at::Tensor>& sub = at::nullopt,
               int coil_batch = 16, int problem_batch = 4);
    int64_t nkernels() const { return _kernels.size(0); }

private:
    std::vector<int64_t> _nmodes;  // image dims
    at::Tensor _kernels;           // prepared layout
    int _layout = 0;
    at::Tensor _workspace;
    int _ws_cb = 0, _ws_pb = 0;
};

// lib/mri/batch_sense.ixx:161-279. Contexts/streams are accepted for interface parity;
// one GPU runs everything on the current stream. Inputs may live on CPU (as in the
// reference) or on the GPU.
class BatchedSenseBase {
public:
    BatchedSenseBase(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                     const at::optional<std::vector<at::Tensor>>& kdata,
                     const at::optional<std::vector<at::Tensor>>& weights);
    int nouter_batches() const { return (int)_coords.size(); }
    int ncoils() const { return (int)_smaps.size(0); }
    bool has_kdata() const { return !_kdata.empty(); }
    bool has_weights() const { return !_weights.empty(); }

protected:
    at::Device _device;
    std::vector<at::Tensor> _coords;   // on the GPU
    at::Tensor _smaps;                 // on the GPU
    std::vector<at::Tensor> _kdata;    // on the GPU
    std::vector<at::Tensor> _weights;  // on the GPU
    std::vector<int64_t> _nmodes;      // [1, N...]
};

class BatchedSense : public BatchedSenseBase {
public:
    using BatchedSenseBase::BatchedSenseBase;
    // in [F, E, N...] -> out list of [E, C, M_f]; manipulators as batch_sense.cpp:153-198
    void apply(const at::Tensor& in, std::vector<at::Tensor>& out, const std::vector<std::vector<int64_t>>& coils);
};

class BatchedSenseAdjoint : public BatchedSenseBase {
public:
    using BatchedSenseBase::BatchedSenseBase;
    // in list of [E, C, M_f] -> out [F, E, N...]; manipulators as batch_sense.cpp:300-336
    void apply(const std::vector<at::Tensor>& in, at::Tensor& out, const std::vector<std::vector<int64_t>>& coils);
};

class BatchedSenseNormal : public BatchedSenseBase {
public:
    BatchedSenseNormal(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                       const at::optional<std::vector<at::Tensor>>& kdata,
                       const at::optional<std::vector<at::Tensor>>& weights);
    // in/out [F, E, N...]: out = A^H W (A x - y) per (frame, encoding), batch_sense.cpp:349-486,
    // evaluated as T_w x - A^H W y with one Toeplitz kernel per frame.
    void apply(const at::Tensor& in, at::Tensor& out, const std::vector<std::vector<int64_t>>& coils);

private:
    void ensure_rhs(const std::vector<int64_t>& coils, int E);
    std::unique_ptr<CUDASenseNormalToeplitz> _toep;
    at::Tensor _rhs;  // [F, E, N...] = A^H W y (when kdata is given)
    std::vector<int64_t> _rhs_coils;
};

// lib/mri/precond.ixx:13-38, precond.cpp:11-117: P = ifft(diag * fft(x)) (fftc / ifftc when
// centered), the adjointified form fft(conj(diag) * ifft(x)). apply transforms the last
// diag.dim() axes; leading axes are a batch. The norm does not change the product of the
// forward and inverse transform and is kept only for the interface.
class CirculantPreconditionerOp {
public:
    static std::unique_ptr<CirculantPreconditionerOp> Create(at::Tensor diag, bool centered,
                                                             const std::optional<std::string>& norm, bool adjointify);
    // smaps [C, N...] complex64, coord [ndim, M] float32/float64 CUDA, weights [M] float32;
    // returns the complex64 diagonal [N...] in centred order (hasty_circulant_precond_build).
    static at::Tensor build_diagonal(at::Tensor smaps, at::Tensor coord,
                                     const at::optional<at::Tensor>& weights = at::nullopt,
                                     const at::optional<double>& lambda = 0.0);
    at::Tensor apply(const at::Tensor& in) const;
    std::shared_ptr<CirculantPreconditionerOp> to_device(at::Stream stream) const;
    std::shared_ptr<CirculantPreconditionerOp> adjoint() const;
    const at::Tensor& diag() const { return _diag; }

protected:
    CirculantPreconditionerOp(at::Tensor diag, bool centered, const std::optional<std::string>& norm,
                              bool adjointify);

private:
    at::Tensor _diag;
    bool _centered;
    std::optional<std::string> _norm;
    bool _adjointify;
};

}  // namespace mri

namespace svt {

template <size_t N> struct Block {
    std::array<int64_t, N> first_corner;
    std::array<int64_t, N> second_corner;
};

// svt.cpp:37-60, 208-243
at::Tensor extract_block(const at::Tensor& in, const Block<3>& block, const std::optional<std::vector<int64_t>>& perms,
                         int flatten_split, bool transpose);
void insert_block(at::Tensor& in, const Block<3>& block, at::Tensor& block_tensor,
                  const std::optional<std::vector<int64_t>>& perms, bool transpose);
// svt.cpp:62-169 on the GPU (Gram/eigen form)
at::Tensor svt_hard(const at::Tensor& in, int rank);
void svt_hard_inplace(at::Tensor& in, int rank);
at::Tensor svt_soft(const at::Tensor& in, float lambda);
void svt_soft_inplace(at::Tensor& in, float lambda);

class BlocksSVTBase {
public:
    explicit BlocksSVTBase(int ncontexts = 1) : _nctxt(ncontexts) {}

protected:
    int _nctxt;
    at::Tensor _workspace;
    at::Tensor& workspace(size_t bytes, const at::Device& dev);
};

// svt.cpp:264-343: nblocks random blocks (each corner uniform in [0, size - shape)),
// thresholded sequentially in draw order.
class Random3DBlocksSVT : public BlocksSVTBase {
public:
    using BlocksSVTBase::BlocksSVTBase;
    void apply(at::Tensor in, int64_t nblocks, std::array<int64_t, 3> block_shape, double thresh, bool soft);
    // corners [nblocks][3] (x, y, z) of the last apply's blocks, in the order they were
    // thresholded (no counterpart in the reference: the draws are not reproducible, so
    // this is how a caller can replay them)
    const std::vector<int32_t>& last_corners() const { return _last; }

private:
    std::vector<int32_t> _last;
};

// svt.cpp:354-449 (the TorchScript target of HastySVT.Normal3DBlocksSVT).
class OldNormal3DBlocksSVT : public BlocksSVTBase {
public:
    using BlocksSVTBase::BlocksSVTBase;
    void apply(at::Tensor in, const std::array<int64_t, 3>& block_strides, const std::array<int64_t, 3>& block_shape,
               int64_t block_iter, double thresh, bool soft);
};
using Normal3DBlocksSVT = OldNormal3DBlocksSVT;

}  // namespace svt

namespace op {

// lib/op/mriop.cpp:156-266: the SENSE normal operator as an operator object that holds
// its trajectory, coil maps and coil list. apply = A^H A x (same shape as x),
// apply_forward = A x (k-space [coils.size(), M]), apply_backward = A^H y (image
// [1, N...], summed over coils); the list forms map over the children, as the reference's
// nested Vector does. The reference's holder (mriop.cpp:165-188) never stores the
// coordinates or mode counts and forces both NUFFTs to type 2; here they are kept and the
// backward NUFFT is type 1 (the options give the tolerances; types and signs are fixed,
// forward type 2 with e^{-i}, backward type 1 with e^{+i}). The reference sizes the forward output by nmodes[0] and the backward output
// by nmodes, which is the same when one transform slot is used; here the forward output
// has one row per listed coil and the backward output is the coil-combined image.
class SenseNOp {
public:
    SenseNOp(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const at::Tensor& smaps,
             const std::vector<int64_t>& coils, const at::optional<fft::NufftOptions>& forward_opts = at::nullopt,
             const at::optional<fft::NufftOptions>& backward_opts = at::nullopt);
    at::Tensor apply(const at::Tensor& in) const;
    at::Tensor apply_forward(const at::Tensor& in) const;
    at::Tensor apply_backward(const at::Tensor& in) const;
    std::vector<at::Tensor> apply(const std::vector<at::Tensor>& in) const;
    std::vector<at::Tensor> apply_forward(const std::vector<at::Tensor>& in) const;
    std::vector<at::Tensor> apply_backward(const std::vector<at::Tensor>& in) const;
    // the normal operator is self-adjoint: the adjoint shares this operator's state
    std::shared_ptr<SenseNOp> adjoint() const;
    const std::vector<int64_t>& nmodes() const { return _h->nmodes; }

private:
    struct Holder {
        at::Tensor coords;
        std::vector<int64_t> nmodes;
        at::Tensor smaps;
        std::vector<int64_t> coils;
        std::unique_ptr<mri::CUDASenseNormal> sense;
    };
    explicit SenseNOp(std::shared_ptr<Holder> h) : _h(std::move(h)) {}
    std::shared_ptr<Holder> _h;
};

// lib/op/opalgs.cpp:29-44 on tensors: per iteration v = A v, lambda = |v|, v /= lambda;
// returns lambda (|v| if iters < 1). v is updated in place.
class PowerIteration {
public:
    using Op = std::function<at::Tensor(const at::Tensor&)>;
    PowerIteration() = default;
    at::Tensor run(const Op& A, at::Tensor& v, int iters = 30);
};

// lib/op/opalgs.cpp:50-96 on tensors: A is a callable, P an optional preconditioner.
class ConjugateGradient {
public:
    using Op = std::function<at::Tensor(const at::Tensor&)>;
    ConjugateGradient(Op A, at::Tensor b, at::optional<Op> P = at::nullopt);
    void run(at::Tensor& x, int iter, double tol = 0.0);

private:
    Op _A;
    at::Tensor _b;
    at::optional<Op> _P;
};

// Independent CG problems in one batch ([P, ...] tensors) with the device CG: FP64
// scalars on the GPU, no host synchronisation per iteration. A(v, out) writes A v.
class BatchedConjugateGradient {
public:
    using BatchOp = std::function<void(const at::Tensor&, at::Tensor&)>;
    BatchedConjugateGradient(BatchOp A, at::Tensor b);
    // returns per-problem state [P, 8] (rzold, alpha, beta, done, not_pd, pAp, iters)
    at::Tensor run(at::Tensor& x, int iter, double tol = 0.0);

private:
    BatchOp _A;
    at::Tensor _b;
};

}  // namespace op

}  // namespace hasty
