// TorchScript classes with the reference's module / class / method names:
//   torch.classes.HastyNufft.{NufftOptions, Nufft, NufftNormal}        python/cpp/py_nufft.cpp:83-106
//   torch.classes.HastySense.{Sense, SenseAdjoint, SenseNormal}        python/cpp/py_sense.cpp:83-96
//   torch.classes.HastyBatchedSense.{BatchedSense, BatchedSenseAdjoint, BatchedSenseNormal}
//                                                                      python/cpp/py_batched_sense.cpp:171-190
//   torch.classes.HastySVT.{Random3DBlocksSVT, Normal3DBlocksSVT}      python/cpp/py_svt.cpp:64-73
//   torch.classes.HastyInterface.FunctionLambda                        python/cpp/py_interface.cpp:37-42
//   torch.classes.HastyOp.SenseNOp                                     lib/op/mriop.cpp:156-266 (not bound there)
// Errors surface as RuntimeError carrying the operator's message (torch_util.cpp:181-200).
#include <torch/custom_class.h>
#include <torch/jit.h>
#include <torch/script.h>

#include <numeric>

#include "hasty_ops.hpp"

namespace hasty {
namespace ffi {

namespace {
template <typename F> void catcher(F&& f)
{
    try {
        f();
    } catch (const c10::Error& e) {
        throw std::runtime_error(e.what());
    }
}

std::vector<std::vector<int64_t>> all_coils(int nouter, int ncoil)
{
    std::vector<int64_t> spec(ncoil);
    std::iota(spec.begin(), spec.end(), 0);
    return std::vector<std::vector<int64_t>>(nouter, spec);
}

using Streams = c10::optional<std::vector<c10::Stream>>;

void check_streams(const Streams& streams)
{
    // one GPU per process: streams on other devices cannot be honoured
    if (!streams.has_value()) return;
    for (const auto& s : *streams)
        if (s.device_type() != c10::DeviceType::CUDA)
            throw std::runtime_error("streams must be GPU streams");
}
}  // namespace

class FunctionLambda : public torch::CustomClassHolder {
public:
    FunctionLambda(const std::string& script, const std::string& entry, std::vector<at::Tensor> captures)
        : _cunit(torch::jit::compile(script)), _entry(entry), _captures(std::move(captures))
    {
    }
    void apply(std::vector<at::Tensor> inout) const { _cunit->run_method(_entry, inout, _captures); }

private:
    std::shared_ptr<torch::jit::CompilationUnit> _cunit;
    std::string _entry;
    std::vector<at::Tensor> _captures;
};

class NufftOptions : public torch::CustomClassHolder {
public:
    NufftOptions(int64_t type, const c10::optional<bool>& positive, const c10::optional<double>& tol)
    {
        if (type > 3 || type < 1) throw std::runtime_error("Not available type");
        _opts = fft::NufftOptions((fft::NufftType)type, positive, tol);
    }
    const fft::NufftOptions& get() const { return _opts; }

private:
    fft::NufftOptions _opts;
};

namespace {
void require_cuda_coords(const at::Tensor& coords)
{
    // the reference's CPU (finufft) operators are not part of this library
    if (!coords.is_cuda()) throw std::runtime_error("Coordinates did not reside in cuda device");
}
}  // namespace

class Nufft : public torch::CustomClassHolder {
public:
    Nufft(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const NufftOptions& opts)
    {
        require_cuda_coords(coords);
        _op = std::make_unique<fft::CUDANufft>(coords, nmodes, opts.get());
    }
    void apply(const at::Tensor& in, at::Tensor out) const
    {
        catcher([&] { _op->apply(in, out); });
    }

private:
    std::unique_ptr<fft::CUDANufft> _op;
};

class NufftNormal : public torch::CustomClassHolder {
public:
    NufftNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const NufftOptions& fwd,
                const NufftOptions& bwd)
    {
        require_cuda_coords(coords);
        _op = std::make_unique<fft::CUDANufftNormal>(coords, nmodes, fwd.get(), bwd.get());
    }
    void apply(const at::Tensor& in, at::Tensor out, at::Tensor storage,
               const c10::optional<c10::intrusive_ptr<FunctionLambda>>& func_between) const
    {
        at::optional<std::function<void(at::Tensor&)>> lambda;
        if (func_between.has_value()) {
            auto fl = *func_between;
            lambda = std::function<void(at::Tensor&)>([fl](at::Tensor& t) { fl->apply({t}); });
        }
        catcher([&] { _op->apply(in, out, storage, lambda); });
    }

private:
    std::unique_ptr<fft::CUDANufftNormal> _op;
};

class Sense : public torch::CustomClassHolder {
public:
    Sense(const at::Tensor& coords, const std::vector<int64_t>& nmodes)
    {
        require_cuda_coords(coords);
        _op = std::make_unique<mri::CUDASense>(coords, nmodes);
    }
    void apply(const at::Tensor& in, at::Tensor out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const c10::optional<at::Tensor>& imspace_storage, const c10::optional<at::Tensor>& kspace_storage)
    {
        catcher([&] { _op->apply(in, out, smaps, coils, imspace_storage, kspace_storage, at::nullopt, at::nullopt); });
    }

private:
    std::unique_ptr<mri::CUDASense> _op;
};

class SenseAdjoint : public torch::CustomClassHolder {
public:
    SenseAdjoint(const at::Tensor& coords, const std::vector<int64_t>& nmodes)
    {
        require_cuda_coords(coords);
        _op = std::make_unique<mri::CUDASenseAdjoint>(coords, nmodes);
    }
    void apply(const at::Tensor& in, at::Tensor out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const c10::optional<at::Tensor>& imspace_storage, const c10::optional<at::Tensor>& kspace_storage)
    {
        catcher([&] { _op->apply(in, out, smaps, coils, imspace_storage, kspace_storage, at::nullopt, at::nullopt); });
    }

private:
    std::unique_ptr<mri::CUDASenseAdjoint> _op;
};

class SenseNormal : public torch::CustomClassHolder {
public:
    SenseNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes)
    {
        require_cuda_coords(coords);
        _op = std::make_unique<mri::CUDASenseNormal>(coords, nmodes);
    }
    void apply(const at::Tensor& in, at::Tensor out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
               const c10::optional<at::Tensor>& imspace_storage, const c10::optional<at::Tensor>& kspace_storage)
    {
        catcher([&] { _op->apply(in, out, smaps, coils, imspace_storage, kspace_storage); });
    }

private:
    std::unique_ptr<mri::CUDASenseNormal> _op;
};

namespace {
c10::optional<std::vector<at::Tensor>> opt_list(const c10::optional<std::vector<at::Tensor>>& l) { return l; }
}  // namespace

class BatchedSense : public torch::CustomClassHolder {
public:
    BatchedSense(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                 const c10::optional<std::vector<at::Tensor>>& kdata,
                 const c10::optional<std::vector<at::Tensor>>& weights, const Streams& streams)
    {
        check_streams(streams);
        _bs = std::make_unique<mri::BatchedSense>(coords, smaps, opt_list(kdata), opt_list(weights));
    }
    void apply(const at::Tensor& in, std::vector<at::Tensor> out,
               const c10::optional<std::vector<std::vector<int64_t>>>& coils)
    {
        catcher([&] { _bs->apply(in, out, coils.has_value() ? *coils : all_coils(_bs->nouter_batches(), _bs->ncoils())); });
    }

private:
    std::unique_ptr<mri::BatchedSense> _bs;
};

class BatchedSenseAdjoint : public torch::CustomClassHolder {
public:
    BatchedSenseAdjoint(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                        const c10::optional<std::vector<at::Tensor>>& kdata,
                        const c10::optional<std::vector<at::Tensor>>& weights, const Streams& streams)
    {
        check_streams(streams);
        _bs = std::make_unique<mri::BatchedSenseAdjoint>(coords, smaps, opt_list(kdata), opt_list(weights));
    }
    void apply(std::vector<at::Tensor> in, at::Tensor out,
               const c10::optional<std::vector<std::vector<int64_t>>>& coils)
    {
        catcher([&] { _bs->apply(in, out, coils.has_value() ? *coils : all_coils(_bs->nouter_batches(), _bs->ncoils())); });
    }

private:
    std::unique_ptr<mri::BatchedSenseAdjoint> _bs;
};

class BatchedSenseNormal : public torch::CustomClassHolder {
public:
    BatchedSenseNormal(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                       const c10::optional<std::vector<at::Tensor>>& kdata,
                       const c10::optional<std::vector<at::Tensor>>& weights, const Streams& streams)
    {
        check_streams(streams);
        _bs = std::make_unique<mri::BatchedSenseNormal>(coords, smaps, opt_list(kdata), opt_list(weights));
    }
    void apply(const at::Tensor& in, at::Tensor out, const c10::optional<std::vector<std::vector<int64_t>>>& coils)
    {
        catcher([&] { _bs->apply(in, out, coils.has_value() ? *coils : all_coils(_bs->nouter_batches(), _bs->ncoils())); });
    }
    mri::BatchedSenseNormal& impl() { return *_bs; }

private:
    std::unique_ptr<mri::BatchedSenseNormal> _bs;
};

class Random3DBlocksSVT : public torch::CustomClassHolder {
public:
    explicit Random3DBlocksSVT(const Streams& streams)
    {
        check_streams(streams);
        _op = std::make_unique<svt::Random3DBlocksSVT>();
    }
    void apply(at::Tensor in, int64_t nblocks, const std::array<int64_t, 3>& block_shape, double thresh, bool soft)
    {
        catcher([&] { _op->apply(in, nblocks, block_shape, thresh, soft); });
    }
    at::Tensor last_corners() const
    {
        const auto& c = _op->last_corners();
        std::vector<int64_t> v(c.begin(), c.end());
        return torch::tensor(v, torch::kInt64).reshape({(int64_t)c.size() / 3, 3});
    }

private:
    std::unique_ptr<svt::Random3DBlocksSVT> _op;
};

class Normal3DBlocksSVT : public torch::CustomClassHolder {
public:
    explicit Normal3DBlocksSVT(const Streams& streams)
    {
        check_streams(streams);
        _op = std::make_unique<svt::OldNormal3DBlocksSVT>();
    }
    void apply(at::Tensor in, const std::array<int64_t, 3>& block_strides, const std::array<int64_t, 3>& block_shape,
               int64_t block_iter, double thresh, bool soft)
    {
        catcher([&] { _op->apply(in, block_strides, block_shape, block_iter, thresh, soft); });
    }

private:
    std::unique_ptr<svt::OldNormal3DBlocksSVT> _op;
};

// op::SenseNOp (lib/op/mriop.cpp:156-266); the reference has no Python binding for it.
class SenseNOp : public torch::CustomClassHolder {
public:
    SenseNOp(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const at::Tensor& smaps,
             const std::vector<int64_t>& coils, const c10::optional<c10::intrusive_ptr<NufftOptions>>& forward_opts,
             const c10::optional<c10::intrusive_ptr<NufftOptions>>& backward_opts)
    {
        require_cuda_coords(coords);
        at::optional<fft::NufftOptions> fwd, bwd;
        if (forward_opts.has_value()) fwd = (*forward_opts)->get();
        if (backward_opts.has_value()) bwd = (*backward_opts)->get();
        catcher([&] { _op = std::make_shared<op::SenseNOp>(coords, nmodes, smaps, coils, fwd, bwd); });
    }
    explicit SenseNOp(std::shared_ptr<op::SenseNOp> o) : _op(std::move(o)) {}
    at::Tensor apply(const at::Tensor& in) const
    {
        at::Tensor r;
        catcher([&] { r = _op->apply(in); });
        return r;
    }
    at::Tensor apply_forward(const at::Tensor& in) const
    {
        at::Tensor r;
        catcher([&] { r = _op->apply_forward(in); });
        return r;
    }
    at::Tensor apply_backward(const at::Tensor& in) const
    {
        at::Tensor r;
        catcher([&] { r = _op->apply_backward(in); });
        return r;
    }
    std::vector<at::Tensor> apply_list(const std::vector<at::Tensor>& in) const
    {
        std::vector<at::Tensor> r;
        catcher([&] { r = _op->apply(in); });
        return r;
    }
    c10::intrusive_ptr<SenseNOp> adjoint() const { return c10::make_intrusive<SenseNOp>(_op->adjoint()); }

private:
    std::shared_ptr<op::SenseNOp> _op;
};

}  // namespace ffi
}  // namespace hasty

using namespace hasty::ffi;

TORCH_LIBRARY(HastyInterface, m)
{
    m.class_<FunctionLambda>("FunctionLambda")
        .def(torch::init<const std::string&, const std::string&, std::vector<at::Tensor>>())
        .def("apply", &FunctionLambda::apply);
    m.def("doc", []() -> std::string {
        return "FunctionLambda(script, entry, captures).apply(inout): runs TorchScript function `entry` of `script` "
               "as entry(inout, captures).";
    });
}

TORCH_LIBRARY(HastyNufft, m)
{
    m.class_<NufftOptions>("NufftOptions")
        .def(torch::init<int64_t, const c10::optional<bool>&, const c10::optional<double>&>());
    m.class_<Nufft>("Nufft")
        .def(torch::init([](const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                            const c10::intrusive_ptr<NufftOptions>& opts) {
            return c10::make_intrusive<Nufft>(coords, nmodes, *opts);
        }))
        .def("apply", &Nufft::apply);
    m.class_<NufftNormal>("NufftNormal")
        .def(torch::init([](const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                            const c10::intrusive_ptr<NufftOptions>& fwd, const c10::intrusive_ptr<NufftOptions>& bwd) {
            return c10::make_intrusive<NufftNormal>(coords, nmodes, *fwd, *bwd);
        }))
        .def("apply", &NufftNormal::apply);
    m.def("doc", []() -> std::string {
        return "Nufft(coords [ndim, M], nmodes [ntransf, N...], NufftOptions(type, positive, tol)).apply(in, out); "
               "NufftNormal(coords, nmodes, fwd, bwd).apply(in, out, storage, func_between). Both scale by "
               "1/sqrt(prod N).";
    });
}

TORCH_LIBRARY(HastySense, m)
{
    m.class_<Sense>("Sense")
        .def(torch::init<const at::Tensor&, const std::vector<int64_t>&>())
        .def("apply", &Sense::apply);
    m.class_<SenseAdjoint>("SenseAdjoint")
        .def(torch::init<const at::Tensor&, const std::vector<int64_t>&>())
        .def("apply", &SenseAdjoint::apply);
    m.class_<SenseNormal>("SenseNormal")
        .def(torch::init<const at::Tensor&, const std::vector<int64_t>&>())
        .def("apply", &SenseNormal::apply);
    m.def("doc", []() -> std::string {
        return "Sense / SenseAdjoint / SenseNormal(coords, nmodes).apply(in, out, smaps, coils, imspace_storage, "
               "kspace_storage), all tensors on the GPU.";
    });
}

TORCH_LIBRARY(HastyBatchedSense, m)
{
    m.class_<BatchedSense>("BatchedSense")
        .def(torch::init<const std::vector<at::Tensor>&, const at::Tensor&, const c10::optional<std::vector<at::Tensor>>&,
                         const c10::optional<std::vector<at::Tensor>>&, const Streams&>())
        .def("apply", &BatchedSense::apply);
    m.class_<BatchedSenseAdjoint>("BatchedSenseAdjoint")
        .def(torch::init<const std::vector<at::Tensor>&, const at::Tensor&, const c10::optional<std::vector<at::Tensor>>&,
                         const c10::optional<std::vector<at::Tensor>>&, const Streams&>())
        .def("apply", &BatchedSenseAdjoint::apply);
    m.class_<BatchedSenseNormal>("BatchedSenseNormal")
        .def(torch::init<const std::vector<at::Tensor>&, const at::Tensor&, const c10::optional<std::vector<at::Tensor>>&,
                         const c10::optional<std::vector<at::Tensor>>&, const Streams&>())
        .def("apply", &BatchedSenseNormal::apply);
    m.def("doc", []() -> std::string {
        return "BatchedSense{,Adjoint,Normal}(coords list, smaps, kdata list?, weights list?, streams?).apply(in, out, "
               "coils?). Tensors may live on the CPU or the GPU; BatchedSenseNormal runs the Toeplitz-embedded "
               "normal operator, one kernel per outer batch.";
    });
}

TORCH_LIBRARY(HastySVT, m)
{
    m.class_<Random3DBlocksSVT>("Random3DBlocksSVT")
        .def(torch::init<const Streams&>())
        .def("apply", &Random3DBlocksSVT::apply)
        .def("last_corners", &Random3DBlocksSVT::last_corners);
    m.class_<Normal3DBlocksSVT>("Normal3DBlocksSVT").def(torch::init<const Streams&>()).def("apply", &Normal3DBlocksSVT::apply);
    m.def("doc", []() -> std::string {
        return "Random3DBlocksSVT(streams?).apply(in [F,E,X,Y,Z], nblocks, block_shape, thresh, soft); "
               "Normal3DBlocksSVT(streams?).apply(in, block_strides, block_shape, block_iter, thresh, soft).";
    });
}

// The C++ iterative algorithms over the operators above (no TorchScript binding in the
// reference; they let Python drive op::PowerIteration / ConjugateGradient /
// BatchedConjugateGradient, lib/op/opalgs.cpp:29-96, opalgs.ixx:84-136).
class PowerIteration : public torch::CustomClassHolder {
public:
    at::Tensor run(const c10::intrusive_ptr<SenseNOp>& A, at::Tensor v, int64_t iters)
    {
        at::Tensor lambda;
        catcher([&] {
            at::Tensor w = v;
            lambda = hasty::op::PowerIteration().run([&](const at::Tensor& t) { return A->apply(t); }, w, (int)iters);
            v.copy_(w);
        });
        return lambda;
    }
};

class ConjugateGradient : public torch::CustomClassHolder {
public:
    ConjugateGradient(const c10::intrusive_ptr<SenseNOp>& A, const at::Tensor& b, double l2) : _A(A), _b(b), _l2(l2) {}
    void run(at::Tensor x, int64_t iter, double tol)
    {
        catcher([&] {
            auto A = _A;
            const double l2 = _l2;
            hasty::op::ConjugateGradient cg([A, l2](const at::Tensor& t) { return l2 != 0.0 ? A->apply(t).add_(t, l2) : A->apply(t); },
                                     _b);
            cg.run(x, (int)iter, tol);
        });
    }

private:
    c10::intrusive_ptr<SenseNOp> _A;
    at::Tensor _b;
    double _l2;
};

class BatchedConjugateGradient : public torch::CustomClassHolder {
public:
    // A: a BatchedSenseNormal without k-space data (its apply is then the normal
    // operator); b, x: [F, E, N...] complex64 on the GPU, solved as F * E problems.
    BatchedConjugateGradient(const c10::intrusive_ptr<BatchedSenseNormal>& A, const at::Tensor& b) : _A(A), _b(b) {}
    at::Tensor run(at::Tensor x, int64_t iter, double tol)
    {
        at::Tensor state;
        catcher([&] {
            if (!x.is_cuda() || !_b.is_cuda() || x.dim() < 3 || x.sizes() != _b.sizes())
                throw std::runtime_error("BatchedConjugateGradient takes x and b of one [F, E, N...] shape on the GPU");
            const auto shape = x.sizes().vec();
            std::vector<int64_t> flat = {shape[0] * shape[1]};
            flat.insert(flat.end(), shape.begin() + 2, shape.end());
            auto A = _A;
            auto coils = all_coils((int)shape[0], A->impl().ncoils());
            hasty::op::BatchedConjugateGradient cg(
                [A, shape, coils](const at::Tensor& v, at::Tensor& out) {
                    at::Tensor o = out.view(shape);
                    A->impl().apply(v.view(shape), o, coils);
                },
                _b.reshape(flat));
            at::Tensor xf = x.view(flat);
            state = cg.run(xf, (int)iter, tol);
        });
        return state;
    }

private:
    c10::intrusive_ptr<BatchedSenseNormal> _A;
    at::Tensor _b;
};

TORCH_LIBRARY(HastyOp, m)
{
    m.class_<SenseNOp>("SenseNOp")
        .def(torch::init<const at::Tensor&, const std::vector<int64_t>&, const at::Tensor&, const std::vector<int64_t>&,
                         const c10::optional<c10::intrusive_ptr<NufftOptions>>&,
                         const c10::optional<c10::intrusive_ptr<NufftOptions>>&>())
        .def("apply", &SenseNOp::apply)
        .def("apply_forward", &SenseNOp::apply_forward)
        .def("apply_backward", &SenseNOp::apply_backward)
        .def("apply_list", &SenseNOp::apply_list)
        .def("adjoint", &SenseNOp::adjoint);
    // registered after SenseNOp, whose type their signatures use
    m.class_<PowerIteration>("PowerIteration").def(torch::init<>()).def("run", &PowerIteration::run);
    m.class_<ConjugateGradient>("ConjugateGradient")
        .def(torch::init<const c10::intrusive_ptr<SenseNOp>&, const at::Tensor&, double>())
        .def("run", &ConjugateGradient::run);
    m.class_<BatchedConjugateGradient>("BatchedConjugateGradient")
        .def(torch::init<const c10::intrusive_ptr<BatchedSenseNormal>&, const at::Tensor&>())
        .def("run", &BatchedConjugateGradient::run);
    m.def("doc", []() -> std::string {
        return "SenseNOp(coords, nmodes, smaps, coils, forward_opts?, backward_opts?): apply(x) = A^H A x, "
               "apply_forward(x) = A x [coils, M], apply_backward(y) = A^H y [1, N...], apply_list maps apply over a "
               "list, adjoint() is the operator itself.";
    });
}
