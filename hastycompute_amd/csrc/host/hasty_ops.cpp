#include "hasty_ops.hpp"

#include <c10/hip/HIPStream.h>

#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>

namespace hasty {

namespace detail {

void check(int rc)
{
    if (rc != 0) throw std::runtime_error(hasty_last_error());
}

void* stream() { return (void*)c10::hip::getCurrentHIPStream().stream(); }

at::ScalarType complex_of(at::ScalarType t)
{
    switch (t) {
    case at::ScalarType::Float: return at::ScalarType::ComplexFloat;
    case at::ScalarType::Double: return at::ScalarType::ComplexDouble;
    default: throw std::runtime_error("Coordinates must be real!");
    }
}

}  // namespace detail

using detail::check;

namespace fft {

at::Tensor allocate_out(const at::Tensor& coords, int ntransf)
{
    return at::empty({ntransf, coords.size(1)}, coords.options().dtype(detail::complex_of(coords.scalar_type())));
}

at::Tensor allocate_adjoint_out(const at::Tensor& coords, const std::vector<int64_t>& nmodes)
{
    return at::empty(nmodes, coords.options().dtype(detail::complex_of(coords.scalar_type())));
}

at::Tensor allocate_normal_out(const at::Tensor& coords, const std::vector<int64_t>& nmodes)
{
    return allocate_adjoint_out(coords, nmodes);
}

at::Tensor allocate_normal_storage(const at::Tensor& coords, int ntransf) { return allocate_out(coords, ntransf); }

CUDANufft::CUDANufft(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const NufftOptions& opts)
    : _coords(coords), _nmodes(nmodes), _opts(opts)
{
    _type = _coords.scalar_type();
    if (_type != at::ScalarType::Float && _type != at::ScalarType::Double)
        throw std::runtime_error("Coordinates must be real!");
    if (_opts.get_type() == eType3) throw std::runtime_error("Type 3 Nufft is not yet supported");
    _ntransf = (int32_t)_nmodes.at(0);
    _ndim = (int32_t)_coords.size(0);
    _nfreq = _coords.size(1);
    if (_ndim + 1 != (int32_t)_nmodes.size()) throw std::runtime_error("_ndim + 1 must match number of nmodes given");
    if (_ndim < 1 || _ndim > 3) throw std::runtime_error("Dimension must be 1,2 or 3");
    if (!_coords.is_contiguous()) throw std::runtime_error("coords must be contiguous");
    if (!_coords.is_cuda()) throw std::runtime_error("Coordinates did not reside in cuda device");
    _nvoxels = 1;
    for (int i = 0; i < _ndim; ++i) _nvoxels *= _nmodes[i + 1];
    std::vector<int64_t> flipped(_ndim);
    for (int i = 0; i < _ndim; ++i) flipped[i] = _nmodes[_ndim - i];
    const int fp64 = _type == at::ScalarType::Double ? 1 : 0;
    check(hasty_nufft_makeplan((int)_opts.get_type(), _ndim, flipped.data(), _opts.get_positive(), _ntransf,
                               _opts.get_tol(), fp64, &_plan));
    check(hasty_nufft_set_scale(_plan, 1.0 / std::sqrt((double)_nvoxels)));
    const char* base = static_cast<const char*>(_coords.data_ptr());
    const int64_t es = _coords.element_size();
    const void* rows[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < _ndim; ++i) rows[i] = base + (int64_t)(_ndim - 1 - i) * _nfreq * es;  // x = last row
    check(hasty_nufft_setpts(_plan, _nfreq, rows[0], rows[1], rows[2], detail::stream()));
}

void CUDANufft::close()
{
    if (!_closed) {
        _closed = true;
        if (_plan == nullptr) throw std::runtime_error("Tried to close on Nufft with _plan == nullptr");
        check(hasty_nufft_destroy(_plan));
        _plan = nullptr;
    }
}

CUDANufft::~CUDANufft()
{
    try {
        close();
    } catch (...) {
    }
}

at::ScalarType CUDANufft::complex_type() const { return detail::complex_of(_type); }

namespace {
void common_checks(const at::Tensor& coords, const at::Tensor& in, const at::Tensor& out, at::ScalarType type)
{
    if (coords.get_device() != in.get_device() || coords.get_device() != out.get_device())
        throw std::runtime_error("All tensors must reside on same device");
    if (!out.is_contiguous()) throw std::runtime_error("out must be contiguous");
    if (!in.is_contiguous()) throw std::runtime_error("in must be contiguous");
    const bool f = type == at::ScalarType::Float;
    for (const at::Tensor* t : {&in, &out}) {
        if (t->scalar_type() == at::ScalarType::ComplexFloat && !f)
            throw std::runtime_error("if type(in) is complex float type(coords) must be float");
        if (t->scalar_type() == at::ScalarType::ComplexDouble && f)
            throw std::runtime_error("if type(in) is complex double type(coords) must be double");
        if (t->scalar_type() != at::ScalarType::ComplexFloat && t->scalar_type() != at::ScalarType::ComplexDouble)
            throw std::runtime_error("data must be complex");
    }
}
}  // namespace

void CUDANufft::check_type1(const at::Tensor& in, const at::Tensor& out) const
{
    common_checks(_coords, in, out, _type);
    if (out.sizes() != at::IntArrayRef(_nmodes)) throw std::runtime_error("out.sizes() must equal nmodes given at construct");
    if (in.sizes() != at::IntArrayRef({_ntransf, _nfreq}))
        throw std::runtime_error("in tensor must match ntransf in first dim and nfreq in second dim");
}

void CUDANufft::check_type2(const at::Tensor& in, const at::Tensor& out) const
{
    common_checks(_coords, in, out, _type);
    if (in.sizes() != at::IntArrayRef(_nmodes)) throw std::runtime_error("in.sizes() must equal nmodes given at construct");
    if (out.sizes() != at::IntArrayRef({_ntransf, _nfreq}))
        throw std::runtime_error("out tensor must match ntransf in first dim and nfreq in second dim");
}

void CUDANufft::apply(const at::Tensor& in, at::Tensor& out) const
{
    if (_closed) throw std::runtime_error("Nufft plan is closed");
    switch (_opts.get_type()) {
    case eType1:
        check_type1(in, out);
        check(hasty_nufft_execute(_plan, in.data_ptr(), out.data_ptr(), detail::stream()));
        break;
    case eType2:
        check_type2(in, out);
        check(hasty_nufft_execute(_plan, out.data_ptr(), in.data_ptr(), detail::stream()));
        break;
    default: throw std::runtime_error("Only Nufft Type 1 and Type 2 is supported");
    }
}

CUDANufftNormal::CUDANufftNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                                 const NufftOptions& forward_ops, const NufftOptions& backward_ops)
    : _forward(coords, nmodes, forward_ops), _backward(coords, nmodes, backward_ops)
{
}

void CUDANufftNormal::apply(const at::Tensor& in, at::Tensor& out, at::Tensor& storage,
                            at::optional<std::function<void(at::Tensor&)>> func_between) const
{
    _forward.apply(in, storage);
    if (func_between.has_value()) (*func_between)(storage);
    _backward.apply(storage, out);
}

void CUDANufftNormal::apply_inplace(at::Tensor& in, at::Tensor& storage,
                                    at::optional<std::function<void(at::Tensor&)>> func_between) const
{
    _forward.apply(in, storage);
    if (func_between.has_value()) (*func_between)(storage);
    _backward.apply(storage, in);
}

void CUDANufftNormal::apply_forward(const at::Tensor& in, at::Tensor& out) { _forward.apply(in, out); }
void CUDANufftNormal::apply_backward(const at::Tensor& in, at::Tensor& out) { _backward.apply(in, out); }

namespace {
std::vector<int64_t> doubled(const std::vector<int64_t>& nmodes)
{
    std::vector<int64_t> ns(nmodes.size());
    for (size_t i = 0; i < nmodes.size(); ++i) ns[i] = i == 0 ? nmodes[i] : 2 * nmodes[i];
    return ns;
}
}  // namespace

void CUDANormalNufftToeplitz::build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double,
                                             at::Tensor& diagonal)
{
    if (!coords.is_contiguous()) throw std::runtime_error("coords must be contiguous");
    if (coords.size(0) + 1 != (int64_t)nmodes.size())
        throw std::runtime_error("coords.size(0) must match number of nmodes given");
    const auto ns = doubled(nmodes);
    if (!diagonal.defined() || diagonal.sizes() != at::IntArrayRef(ns))
        throw std::runtime_error("diagonal must have the doubled nmodes shape");
    if (diagonal.scalar_type() != at::ScalarType::ComplexFloat)
        throw std::runtime_error("the Toeplitz diagonal is kept in complex64");
    std::vector<int64_t> img(nmodes.begin() + 1, nmodes.end());
    int64_t per = 1;
    for (size_t i = 1; i < ns.size(); ++i) per *= ns[i];
    at::Tensor d0 = diagonal.select(0, 0);
    check(hasty_toeplitz_build(coords.data_ptr(), coords.scalar_type() == at::ScalarType::Double ? 1 : 0,
                               coords.size(1), (int)img.size(), img.data(), nullptr, d0.data_ptr(), nullptr, 1e-10,
                               detail::stream()));
    for (int64_t t = 1; t < ns[0]; ++t) diagonal.select(0, t).copy_(d0);
    (void)per;
}

void CUDANormalNufftToeplitz::build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double tol,
                                             at::Tensor& diagonal, at::Tensor&, bool)
{
    build_diagonal(coords, nmodes, tol, diagonal);
}

void CUDANormalNufftToeplitz::build_diagonal(const at::Tensor& coords, std::vector<int64_t> nmodes, double tol,
                                             at::Tensor& diagonal, at::Tensor&, at::Tensor&)
{
    build_diagonal(coords, nmodes, tol, diagonal);
}

CUDANormalNufftToeplitz::CUDANormalNufftToeplitz(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                                                 const at::optional<double>& tol,
                                                 const at::optional<std::reference_wrapper<at::Tensor>>& diagonal,
                                                 const at::optional<std::reference_wrapper<at::Tensor>>&,
                                                 const at::optional<std::reference_wrapper<at::Tensor>>&)
    : _nmodes(nmodes)
{
    if (coords.size(0) + 1 != (int64_t)nmodes.size())
        throw std::runtime_error("coords.size(0) must match number of nmodes given");
    if (!coords.is_contiguous()) throw std::runtime_error("coords must be contiguous");
    _diagonal = diagonal.has_value() ? diagonal->get()
                                     : at::empty(doubled(nmodes), coords.options().dtype(at::kComplexFloat));
    build_diagonal(coords, nmodes, tol.value_or(1e-6), _diagonal);
}

CUDANormalNufftToeplitz::CUDANormalNufftToeplitz(at::Tensor&& diagonal, const std::vector<int64_t>& nmodes)
    : _nmodes(nmodes), _diagonal(std::move(diagonal))
{
}

void CUDANormalNufftToeplitz::run(const at::Tensor& in, at::Tensor& out, const at::Tensor* mul, int mode, at::Tensor& st1,
                                  at::Tensor& st2) const
{
    if (in.scalar_type() != at::kComplexFloat) throw std::runtime_error("Toeplitz apply takes complex64 data");
    if (!in.is_contiguous() || !out.is_contiguous()) throw std::runtime_error("in/out must be contiguous");
    if (in.sizes() != at::IntArrayRef(_nmodes)) throw std::runtime_error("in.sizes() must equal nmodes");
    const auto ns = doubled(_nmodes);
    int64_t need = 1;
    for (auto v : ns) need *= v;
    if (st1.numel() < need / 2 || st2.numel() < need / 2 || !st1.is_contiguous() || !st2.is_contiguous())
        throw std::runtime_error("storage tensors too small");
    std::vector<int64_t> img(_nmodes.begin() + 1, _nmodes.end());
    check(hasty_toeplitz_apply(in.data_ptr(), out.data_ptr(), mul ? mul->data_ptr() : nullptr, _diagonal.data_ptr(),
                               (int)_nmodes[0], (int)img.size(), img.data(), mode, st1.data_ptr(), st2.data_ptr(),
                               detail::stream()));
}

at::Tensor CUDANormalNufftToeplitz::apply(const at::Tensor& in, at::Tensor& storage1, at::Tensor& storage2) const
{
    at::Tensor out = at::empty_like(in);
    run(in, out, nullptr, 0, storage1, storage2);
    return out;
}

void CUDANormalNufftToeplitz::apply_add(const at::Tensor& in, at::Tensor& add_to, at::Tensor& storage1,
                                        at::Tensor& storage2) const
{
    run(in, add_to, nullptr, 1, storage1, storage2);
}

void CUDANormalNufftToeplitz::apply_addcmul(const at::Tensor& in, at::Tensor& add_to, const at::Tensor& mul,
                                            at::Tensor& storage1, at::Tensor& storage2) const
{
    at::Tensor m = mul.contiguous();
    run(in, add_to, &m, 2, storage1, storage2);
}

void CUDANormalNufftToeplitz::apply_inplace(at::Tensor& in, at::Tensor& storage1, at::Tensor& storage2) const
{
    run(in, in, nullptr, 3, storage1, storage2);
}

}  // namespace fft

namespace mri {

CoilManipulator& CoilManipulator::setPreApply(const CoilApplier& apply)
{
    if (preapplier.has_value()) throw std::runtime_error("Tried to set non nullopt applier");
    preapplier = apply;
    return *this;
}
CoilManipulator& CoilManipulator::setMidApply(const CoilApplier& apply)
{
    if (midapplier.has_value()) throw std::runtime_error("Tried to set non nullopt applier");
    midapplier = apply;
    return *this;
}
CoilManipulator& CoilManipulator::setPostApply(const CoilApplier& apply)
{
    if (postapplier.has_value()) throw std::runtime_error("Tried to set non nullopt applier");
    postapplier = apply;
    return *this;
}

namespace {
void check_coils(const std::vector<int64_t>& coils, int64_t ntransf, const at::Tensor& smaps)
{
    if (coils.size() % ntransf != 0)
        throw std::runtime_error(
            "Number of coils used in Sense::apply must be divisible by number of ntransf used for nufft");
    for (auto c : coils)
        if (c < 0 || c >= smaps.size(0)) throw std::runtime_error("coil index out of range");
}
}  // namespace

CUDASense::CUDASense(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                     const at::optional<fft::NufftOptions>& opts)
    : _nufft(coords, nmodes, opts.has_value() ? *opts : fft::NufftOptions::type2()), _nmodes(nmodes)
{
}

// sense.ixx:348-423. As there, the coil map of transform idx is smaps[idx] (the position in
// the coil list, not coils[idx]); the two agree for the default coil list 0..C-1.
void CUDASense::apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps, const std::vector<int64_t>& coils,
                      const at::optional<at::Tensor>& imspace_storage, const at::optional<at::Tensor>& kspace_storage,
                      const at::optional<CoilApplier>& premanip, const at::optional<CoilApplier>& postmanip)
{
    c10::InferenceMode guard;
    check_coils(coils, _nmodes[0], smaps);
    if ((int64_t)coils.size() > smaps.size(0)) throw std::runtime_error("more coils listed than coil maps given");
    const int ntransf = (int)_nmodes[0];
    const int batch_runs = (int)coils.size() / ntransf;
    out.zero_();
    at::Tensor imstore = imspace_storage.has_value() ? *imspace_storage : at::empty(_nmodes, in.options());
    at::Tensor kstore = kspace_storage.has_value() ? *kspace_storage : at::empty({ntransf, _nufft.nfreq()}, in.options());
    const bool accumulate = out.size(0) == 1;
    for (int brun = 0; brun < batch_runs; ++brun) {
        if (premanip.has_value()) {
            for (int j = 0; j < ntransf; ++j) imstore.select(0, j).copy_(in.select(0, brun * ntransf + j));
            (*premanip)(imstore, brun);
            for (int j = 0; j < ntransf; ++j) imstore.select(0, j).mul_(smaps.select(0, brun * ntransf + j));
        } else {
            for (int j = 0; j < ntransf; ++j) {
                at::Tensor dst = imstore.select(0, j);
                at::mul_out(dst, in.select(0, 0), smaps.select(0, brun * ntransf + j));
            }
        }
        _nufft.apply(imstore, kstore);
        if (postmanip.has_value()) (*postmanip)(kstore, brun);
        for (int j = 0; j < ntransf; ++j) {
            if (accumulate)
                out.add_(kstore.select(0, j));
            else
                out.select(0, brun * ntransf + j).unsqueeze(0).add_(kstore.select(0, j));
        }
    }
}

CUDASenseAdjoint::CUDASenseAdjoint(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                                   const at::optional<fft::NufftOptions>& opts)
    : _nufft(coords, nmodes, opts.has_value() ? *opts : fft::NufftOptions::type1()), _nmodes(nmodes)
{
}

// sense.ixx:426-497
void CUDASenseAdjoint::apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                             const std::vector<int64_t>& coils, const at::optional<at::Tensor>& imspace_storage,
                             const at::optional<at::Tensor>& kspace_storage, const at::optional<CoilApplier>& premanip,
                             const at::optional<CoilApplier>& postmanip)
{
    c10::InferenceMode guard;
    check_coils(coils, _nmodes[0], smaps);
    const int ntransf = (int)_nmodes[0];
    const int batch_runs = (int)coils.size() / ntransf;
    out.zero_();
    at::Tensor imstore = imspace_storage.has_value() ? *imspace_storage : at::empty(_nmodes, out.options());
    at::Tensor kstore = kspace_storage.has_value() ? *kspace_storage : at::empty({ntransf, _nufft.nfreq()}, in.options());
    const bool accumulate = out.size(0) == 1;
    for (int brun = 0; brun < batch_runs; ++brun) {
        for (int j = 0; j < ntransf; ++j) kstore.select(0, j).copy_(in.select(0, brun * ntransf + j));
        if (premanip.has_value()) (*premanip)(kstore, brun);
        _nufft.apply(kstore, imstore);
        for (int j = 0; j < ntransf; ++j)
            imstore.select(0, j).mul_(smaps.select(0, coils[brun * ntransf + j]).conj());
        if (postmanip.has_value()) (*postmanip)(imstore, brun);
        for (int j = 0; j < ntransf; ++j) {
            if (accumulate)
                out.add_(imstore.select(0, j));
            else
                out.select(0, brun * ntransf + j).add_(imstore.select(0, j));
        }
    }
}

CUDASenseNormal::CUDASenseNormal(const at::Tensor& coords, const std::vector<int64_t>& nmodes,
                                 const at::optional<fft::NufftOptions>& forward_opts,
                                 const at::optional<fft::NufftOptions>& backward_opts)
    : _normal_nufft(coords, nmodes, forward_opts.has_value() ? *forward_opts : fft::NufftOptions::type2(),
                    backward_opts.has_value() ? *backward_opts : fft::NufftOptions::type1()),
      _nmodes(nmodes)
{
}

// sense.ixx:500-578 (the conj multiply is applied per transform slot)
void CUDASenseNormal::apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                            const std::vector<int64_t>& coils, const at::optional<at::Tensor>& imspace_storage,
                            const at::optional<at::Tensor>& kspace_storage, const at::optional<CoilApplier>& premanip,
                            const at::optional<CoilApplier>& midmanip, const at::optional<CoilApplier>& postmanip)
{
    c10::InferenceMode guard;
    check_coils(coils, _nmodes[0], smaps);
    const int ntransf = (int)_nmodes[0];
    const int batch_runs = (int)coils.size() / ntransf;
    out.zero_();
    at::Tensor imstore = imspace_storage.has_value() ? *imspace_storage : at::empty(_nmodes, in.options());
    at::Tensor kstore =
        kspace_storage.has_value() ? *kspace_storage : at::empty({ntransf, _normal_nufft.nfreq()}, in.options());
    for (int brun = 0; brun < batch_runs; ++brun) {
        if (premanip.has_value()) {
            for (int j = 0; j < ntransf; ++j) imstore.select(0, j).copy_(in.select(0, 0));
            (*premanip)(imstore, brun);
            for (int j = 0; j < ntransf; ++j) imstore.select(0, j).mul_(smaps.select(0, coils[brun * ntransf + j]));
        } else {
            for (int j = 0; j < ntransf; ++j) {
                at::Tensor dst = imstore.select(0, j);
                at::mul_out(dst, in.select(0, 0), smaps.select(0, coils[brun * ntransf + j]));
            }
        }
        if (midmanip.has_value()) {
            const CoilApplier& mm = *midmanip;
            _normal_nufft.apply_inplace(imstore, kstore, std::function<void(at::Tensor&)>(
                                                             [&mm, brun](at::Tensor& t) { mm(t, brun); }));
        } else {
            _normal_nufft.apply_inplace(imstore, kstore, at::nullopt);
        }
        for (int j = 0; j < ntransf; ++j) imstore.select(0, j).mul_(smaps.select(0, coils[brun * ntransf + j]).conj());
        if (postmanip.has_value()) (*postmanip)(imstore, brun);
        for (int j = 0; j < ntransf; ++j) out.add_(imstore.select(0, j));
    }
}

void CUDASenseNormal::apply_forward(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                                    const std::vector<int64_t>& coils, const at::optional<at::Tensor>& imspace_storage,
                                    const at::optional<at::Tensor>& kspace_storage)
{
    c10::InferenceMode guard;
    check_coils(coils, _nmodes[0], smaps);
    const int ntransf = (int)_nmodes[0];
    const int batch_runs = (int)coils.size() / ntransf;
    out.zero_();
    at::Tensor imstore = imspace_storage.has_value() ? *imspace_storage : at::empty(_nmodes, in.options());
    at::Tensor kstore =
        kspace_storage.has_value() ? *kspace_storage : at::empty({ntransf, _normal_nufft.nfreq()}, in.options());
    for (int brun = 0; brun < batch_runs; ++brun) {
        for (int j = 0; j < ntransf; ++j) {
            at::Tensor dst = imstore.select(0, j);
            at::mul_out(dst, in.select(0, 0), smaps.select(0, coils[brun * ntransf + j]));
        }
        _normal_nufft.apply_forward(imstore, kstore);
        for (int j = 0; j < ntransf; ++j) out.select(0, brun * ntransf + j).unsqueeze(0).add_(kstore.select(0, j));
    }
}

void CUDASenseNormal::apply_backward(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                                     const std::vector<int64_t>& coils, const at::optional<at::Tensor>& kspace_storage,
                                     const at::optional<at::Tensor>& imspace_storage)
{
    c10::InferenceMode guard;
    check_coils(coils, _nmodes[0], smaps);
    const int ntransf = (int)_nmodes[0];
    const int batch_runs = (int)coils.size() / ntransf;
    out.zero_();
    at::Tensor imstore = imspace_storage.has_value() ? *imspace_storage : at::empty(_nmodes, out.options());
    at::Tensor kstore =
        kspace_storage.has_value() ? *kspace_storage : at::empty({ntransf, _normal_nufft.nfreq()}, in.options());
    const bool accumulate = out.size(0) == 1;
    for (int brun = 0; brun < batch_runs; ++brun) {
        for (int j = 0; j < ntransf; ++j) kstore.select(0, j).copy_(in.select(0, brun * ntransf + j));
        _normal_nufft.apply_backward(kstore, imstore);
        for (int j = 0; j < ntransf; ++j)
            imstore.select(0, j).mul_(smaps.select(0, coils[brun * ntransf + j]).conj());
        for (int j = 0; j < ntransf; ++j) {
            if (accumulate)
                out.add_(imstore.select(0, j));
            else
                out.select(0, brun * ntransf + j).add_(imstore.select(0, j));
        }
    }
}

CUDASenseNormalToeplitz::CUDASenseNormalToeplitz(const std::vector<at::Tensor>& coords,
                                                 const std::vector<int64_t>& nmodes,
                                                 const at::optional<std::vector<at::Tensor>>& weights)
    : _nmodes(nmodes)
{
    if (nmodes.size() != 3) throw std::runtime_error("the Toeplitz SENSE normal operator is 3-D");
    if (coords.empty()) throw std::runtime_error("need at least one trajectory");
    const auto dev = coords[0].device();
    const int64_t kvox = 8 * nmodes[0] * nmodes[1] * nmodes[2];
    at::Tensor natural = at::empty({(int64_t)coords.size(), kvox}, at::TensorOptions(dev).dtype(at::kFloat));
    for (size_t i = 0; i < coords.size(); ++i) {
        at::Tensor co = coords[i].contiguous();
        if (co.size(0) != 3) throw std::runtime_error("coords must be [3, M]");
        at::Tensor w;
        if (weights.has_value()) w = (*weights)[i].to(dev, at::kFloat).contiguous();
        check(hasty_toeplitz_build(co.data_ptr(), co.scalar_type() == at::kDouble ? 1 : 0, co.size(1), 3,
                                   nmodes.data(), w.defined() ? w.data_ptr<float>() : nullptr, nullptr,
                                   natural[i].data_ptr<float>(), 1e-8, detail::stream()));
    }
    _kernels = at::empty_like(natural);
    check(hasty_sense_normal_prepare(natural.data_ptr<float>(), _kernels.data_ptr<float>(), natural.size(0),
                                     nmodes.data(), &_layout, detail::stream()));
}

void CUDASenseNormalToeplitz::apply(const at::Tensor& in, at::Tensor& out, const at::Tensor& smaps,
                                    const std::vector<int64_t>& coils, const std::vector<int32_t>& kernel_index,
                                    const at::optional<at::Tensor>& sub, int coil_batch, int problem_batch)
{
    c10::InferenceMode guard;
    const int64_t nvox = _nmodes[0] * _nmodes[1] * _nmodes[2];
    if (in.scalar_type() != at::kComplexFloat || !in.is_contiguous() || !out.is_contiguous() ||
        in.numel() % nvox != 0 || out.sizes() != in.sizes())
        throw std::runtime_error("in/out must be contiguous complex64 [P, N0, N1, N2]");
    const int P = (int)(in.numel() / nvox);
    if ((int)kernel_index.size() != P) throw std::runtime_error("kernel_index needs one entry per problem");
    at::Tensor S = smaps.contiguous();
    if (S.scalar_type() != at::kComplexFloat || S.numel() != S.size(0) * nvox)
        throw std::runtime_error("smaps must be complex64 [C, N0, N1, N2]");
    for (auto c : coils)
        if (c < 0 || c >= S.size(0)) throw std::runtime_error("coil index out of range");
    std::vector<int32_t> c32(coils.begin(), coils.end());
    at::Tensor coils_dev = at::tensor(c32, at::TensorOptions().dtype(at::kInt)).to(in.device());
    const int cb = std::max(1, std::min<int>(coil_batch, (int)coils.size()));
    const int pb = std::max(1, problem_batch);
    if (!_workspace.defined() || _ws_cb != cb || _ws_pb != pb) {
        const size_t ws = hasty_sense_normal_workspace_size(_nmodes.data(), cb, pb);
        _workspace = at::empty({(int64_t)ws}, in.options().dtype(at::kByte));
        _ws_cb = cb;
        _ws_pb = pb;
    }
    at::Tensor sb;
    if (sub.has_value()) {
        sb = sub->contiguous();
        if (sb.sizes() != in.sizes()) throw std::runtime_error("sub must match in");
    }
    check(hasty_sense_normal_toeplitz(in.data_ptr(), out.data_ptr(), S.data_ptr(), coils_dev.data_ptr<int32_t>(),
                                      (int)coils.size(), _kernels.data_ptr<float>(), (int)_kernels.size(0), _layout,
                                      kernel_index.data(), sb.defined() ? sb.data_ptr() : nullptr, P, _nmodes.data(),
                                      _workspace.data_ptr(), (size_t)_workspace.numel(), cb, detail::stream()));
}

BatchedSenseBase::BatchedSenseBase(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                                   const at::optional<std::vector<at::Tensor>>& kdata,
                                   const at::optional<std::vector<at::Tensor>>& weights)
    : _device(smaps.is_cuda() ? smaps.device() : at::Device(at::kCUDA, c10::hip::current_device()))
{
    _smaps = smaps.to(_device).contiguous();
    for (const auto& c : coords) _coords.push_back(c.to(_device).contiguous());
    if (kdata.has_value())
        for (const auto& k : *kdata) _kdata.push_back(k.to(_device).contiguous());
    if (weights.has_value())
        for (const auto& w : *weights) _weights.push_back(w.to(_device).contiguous());
    _nmodes.resize(_smaps.dim());
    _nmodes[0] = 1;
    for (int i = 1; i < _smaps.dim(); ++i) _nmodes[i] = _smaps.size(i);
}

// batch_sense.cpp:47-151 with the kdata / weights post-manipulators (153-198)
void BatchedSense::apply(const at::Tensor& in, std::vector<at::Tensor>& out,
                         const std::vector<std::vector<int64_t>>& coils)
{
    c10::InferenceMode guard;
    if (in.dim() != (int64_t)_nmodes.size() + 1) throw std::runtime_error("For a ND-image input to apply should be (N+2)D tensor");
    for (int f = 0; f < in.size(0); ++f) {
        CUDASense sense(_coords[f], _nmodes);
        for (int e = 0; e < in.size(1); ++e) {
            at::Tensor x = in[f][e].unsqueeze(0).to(_device).contiguous();
            at::Tensor o = at::empty({(int64_t)coils[f].size(), _coords[f].size(1)}, x.options());
            at::Tensor kd = has_kdata() ? _kdata[f][e] : at::Tensor();
            at::Tensor w = has_weights() ? _weights[f] : at::Tensor();
            at::optional<CoilApplier> post;
            if (kd.defined() || w.defined())
                post = CoilApplier([&](at::Tensor& t, int32_t coil) {
                    if (kd.defined()) t.sub_(kd.select(0, coil).unsqueeze(0));
                    if (w.defined()) t.mul_(w);
                });
            sense.apply(x, o, _smaps, coils[f], at::nullopt, at::nullopt, at::nullopt, post);
            out[f][e].copy_(o);
        }
    }
}

// batch_sense.cpp:200-336 with the kdata / weights pre-manipulators
void BatchedSenseAdjoint::apply(const std::vector<at::Tensor>& in, at::Tensor& out,
                                const std::vector<std::vector<int64_t>>& coils)
{
    c10::InferenceMode guard;
    for (int f = 0; f < (int)in.size(); ++f) {
        CUDASenseAdjoint sense(_coords[f], _nmodes);
        for (int e = 0; e < in[f].size(0); ++e) {
            at::Tensor y = in[f][e].to(_device).contiguous();
            at::Tensor o = at::empty(_nmodes, y.options());
            at::Tensor kd = has_kdata() ? _kdata[f][e] : at::Tensor();
            at::Tensor w = has_weights() ? _weights[f] : at::Tensor();
            at::optional<CoilApplier> pre;
            if (kd.defined() || w.defined())
                pre = CoilApplier([&](at::Tensor& t, int32_t coil) {
                    if (kd.defined()) t.sub_(kd.select(0, coil).unsqueeze(0));
                    if (w.defined()) t.mul_(w);
                });
            sense.apply(y, o, _smaps, coils[f], at::nullopt, at::nullopt, pre, at::nullopt);
            out[f][e].copy_(o[0]);
        }
    }
}

BatchedSenseNormal::BatchedSenseNormal(const std::vector<at::Tensor>& coords, const at::Tensor& smaps,
                                       const at::optional<std::vector<at::Tensor>>& kdata,
                                       const at::optional<std::vector<at::Tensor>>& weights)
    : BatchedSenseBase(coords, smaps, kdata, weights)
{
    std::vector<int64_t> img(_nmodes.begin() + 1, _nmodes.end());
    _toep = std::make_unique<CUDASenseNormalToeplitz>(_coords, img,
                                                      has_weights() ? at::optional<std::vector<at::Tensor>>(_weights)
                                                                    : at::nullopt);
}

void BatchedSenseNormal::ensure_rhs(const std::vector<int64_t>& coils, int E)
{
    if (!has_kdata() || (_rhs.defined() && _rhs_coils == coils && _rhs.size(1) == E)) return;
    const int F = nouter_batches();
    std::vector<int64_t> shp = {F, E};
    for (size_t i = 1; i < _nmodes.size(); ++i) shp.push_back(_nmodes[i]);
    _rhs = at::empty(shp, _smaps.options());
    for (int f = 0; f < F; ++f) {
        CUDASenseAdjoint adj(_coords[f], _nmodes);
        for (int e = 0; e < E; ++e) {
            at::Tensor y = _kdata[f][e];
            at::Tensor o = at::empty(_nmodes, _smaps.options());
            at::Tensor w = has_weights() ? _weights[f] : at::Tensor();
            at::optional<CoilApplier> pre;
            if (w.defined()) pre = CoilApplier([&](at::Tensor& t, int32_t) { t.mul_(w); });
            // kdata row i pairs with the i-th listed coil, as in the reference's
            // mid-manipulators (batch_sense.cpp:452-486)
            adj.apply(y, o, _smaps, coils, at::nullopt, at::nullopt, pre, at::nullopt);
            _rhs[f][e].copy_(o[0]);
        }
    }
    _rhs_coils = coils;
}

void BatchedSenseNormal::apply(const at::Tensor& in, at::Tensor& out, const std::vector<std::vector<int64_t>>& coils)
{
    c10::InferenceMode guard;
    if (out.dim() != (int64_t)_nmodes.size() + 1) throw std::runtime_error("For a ND-image output to apply should be (N+2)D tensor");
    const int F = (int)in.size(0), E = (int)in.size(1);
    if (F != nouter_batches()) throw std::runtime_error("one outer batch per trajectory");
    bool same = true;
    for (int f = 1; f < F; ++f) same = same && coils[f] == coils[0];
    at::Tensor x = in.to(_device, at::kComplexFloat).contiguous();
    at::Tensor o = out.is_cuda() && out.is_contiguous() ? out : at::empty_like(x);
    const int64_t nvox = x.numel() / (F * (int64_t)E);
    auto run = [&](int f0, int f1, const std::vector<int64_t>& cl) {
        // the reference's kdata manipulator subtracts kdata row `coil` for the i-th
        // listed coil (i = brun); the constant term uses the same pairing
        ensure_rhs(cl, E);
        std::vector<int32_t> kidx;
        for (int f = f0; f < f1; ++f)
            for (int e = 0; e < E; ++e) kidx.push_back(f);
        at::Tensor xs = x.narrow(0, f0, f1 - f0).reshape({-1, _nmodes[1], _nmodes[2], _nmodes[3]});
        at::Tensor os = o.narrow(0, f0, f1 - f0).view({-1, _nmodes[1], _nmodes[2], _nmodes[3]});
        at::optional<at::Tensor> sub;
        if (has_kdata()) sub = _rhs.narrow(0, f0, f1 - f0).reshape({-1, _nmodes[1], _nmodes[2], _nmodes[3]});
        std::vector<int32_t> all(kidx.begin(), kidx.end());
        _toep->apply(xs, os, _smaps, cl, all, sub);
        (void)nvox;
    };
    if (same) {
        run(0, F, coils[0]);
    } else {
        for (int f = 0; f < F; ++f) run(f, f + 1, coils[f]);
    }
    if (!o.is_same(out)) out.copy_(o);
}

// ---------------------------------------------------------------------------------
// CirculantPreconditionerOp (lib/mri/precond.cpp:11-117)
// ---------------------------------------------------------------------------------

std::unique_ptr<CirculantPreconditionerOp> CirculantPreconditionerOp::Create(at::Tensor diag, bool centered,
                                                                             const std::optional<std::string>& norm,
                                                                             bool adjointify)
{
    return std::unique_ptr<CirculantPreconditionerOp>(
        new CirculantPreconditionerOp(std::move(diag), centered, norm, adjointify));
}

CirculantPreconditionerOp::CirculantPreconditionerOp(at::Tensor diag, bool centered,
                                                     const std::optional<std::string>& norm, bool adjointify)
    : _diag(std::move(diag)), _centered(centered), _norm(norm), _adjointify(adjointify)
{
    if (_norm.has_value() && *_norm != "ortho" && *_norm != "forward" && *_norm != "backward")
        throw std::runtime_error("Invalid normalization mode: " + *_norm);
}

at::Tensor CirculantPreconditionerOp::build_diagonal(at::Tensor smaps, at::Tensor coord,
                                                     const at::optional<at::Tensor>& weights,
                                                     const at::optional<double>& lambda)
{
    if (!smaps.is_cuda() || !coord.is_cuda()) throw std::runtime_error("Coordinates did not reside in cuda device");
    if (smaps.dim() < 2 || smaps.dim() > 4) throw std::runtime_error("smaps must be [C, N...] with 1 to 3 image dims");
    const int ndim = (int)smaps.dim() - 1;
    if (coord.dim() != 2 || coord.size(0) != ndim) throw std::runtime_error("coord must be [ndim, M]");
    at::Tensor S = smaps.to(at::kComplexFloat).contiguous();
    at::Tensor co = coord.scalar_type() == at::kDouble ? coord.contiguous() : coord.to(at::kFloat).contiguous();
    at::Tensor w;
    if (weights.has_value()) w = weights->to(at::kFloat).contiguous();
    std::vector<int64_t> nm(S.sizes().begin() + 1, S.sizes().end());
    at::Tensor d = at::empty(nm, S.options());
    check(hasty_circulant_precond_build(S.data_ptr(), S.size(0), co.data_ptr(), co.scalar_type() == at::kDouble ? 1 : 0,
                                        co.size(1), ndim, nm.data(), w.defined() ? w.data_ptr<float>() : nullptr,
                                        lambda.has_value() ? *lambda : 0.0, 1, d.data_ptr(), 1e-10, detail::stream()));
    return d;
}

at::Tensor CirculantPreconditionerOp::apply(const at::Tensor& in) const
{
    const int ndim = (int)_diag.dim();
    if (ndim < 1 || ndim > 3 || in.dim() < ndim) throw std::runtime_error("input does not match the diagonal");
    for (int d = 0; d < ndim; ++d)
        if (in.size(in.dim() - ndim + d) != _diag.size(d)) throw std::runtime_error("input does not match the diagonal");
    at::Tensor x = in.to(at::kComplexFloat).contiguous();
    at::Tensor dg = _diag.to(at::kComplexFloat).contiguous();
    std::vector<int64_t> nm(dg.sizes().begin(), dg.sizes().end());
    const int64_t nb = x.numel() / dg.numel();
    std::vector<int32_t> idx((size_t)nb, 0);
    at::Tensor out = at::empty_like(x);
    const size_t wsb = hasty_circulant_precond_workspace_size(1, ndim, nm.data(), _centered ? 1 : 0, _adjointify ? 1 : 0);
    at::Tensor ws = at::empty({(int64_t)std::max<size_t>(wsb, 1)}, x.options().dtype(at::kByte));
    check(hasty_circulant_precond_apply(x.data_ptr(), out.data_ptr(), dg.data_ptr(), 1, nb, idx.data(), ndim,
                                        nm.data(), _centered ? 1 : 0, _adjointify ? 1 : 0, ws.data_ptr(), wsb,
                                        detail::stream()));
    return out.to(in.scalar_type());
}

std::shared_ptr<CirculantPreconditionerOp> CirculantPreconditionerOp::to_device(at::Stream stream) const
{
    return std::shared_ptr<CirculantPreconditionerOp>(
        Create(_diag.to(stream.device()), _centered, _norm, _adjointify).release());
}

std::shared_ptr<CirculantPreconditionerOp> CirculantPreconditionerOp::adjoint() const
{
    return std::shared_ptr<CirculantPreconditionerOp>(Create(_diag, _centered, _norm, !_adjointify).release());
}

}  // namespace mri

namespace svt {

at::Tensor extract_block(const at::Tensor& in, const Block<3>& block, const std::optional<std::vector<int64_t>>& perms,
                         int flatten_split, bool transpose)
{
    using namespace at::indexing;
    at::Tensor b = in.index({"...", Slice(block.first_corner[0], block.second_corner[0]),
                             Slice(block.first_corner[1], block.second_corner[1]),
                             Slice(block.first_corner[2], block.second_corner[2])});
    if (perms.has_value()) b = b.permute(*perms).contiguous();
    b = b.flatten(0, flatten_split).flatten(1).contiguous();
    if (transpose) b = b.transpose(0, 1);
    return b;
}

void insert_block(at::Tensor& in, const Block<3>& block, at::Tensor& block_tensor,
                  const std::optional<std::vector<int64_t>>& perms, bool transpose)
{
    using namespace at::indexing;
    std::vector<int64_t> view_lens = in.sizes().vec();
    for (int i = 0; i < 3; ++i)
        view_lens[view_lens.size() - 3 + i] = block.second_corner[i] - block.first_corner[i];
    at::Tensor bt = block_tensor.contiguous();
    if (transpose) bt = bt.transpose(0, 1).contiguous();
    if (perms.has_value()) throw std::runtime_error("permuted insert_block is not supported");
    in.index_put_({"...", Slice(block.first_corner[0], block.second_corner[0]),
                   Slice(block.first_corner[1], block.second_corner[1]),
                   Slice(block.first_corner[2], block.second_corner[2])},
                  bt.view(view_lens));
}

namespace {
at::Tensor svt_run(const at::Tensor& in, double thresh, bool soft)
{
    if (in.dim() != 2) throw std::runtime_error("svt takes a 2-D (Casorati) matrix");
    const auto dev = in.is_cuda() ? in.device() : at::Device(at::kCUDA, c10::hip::current_device());
    at::Tensor A = in.to(dev, at::kComplexFloat).contiguous().clone();
    const int64_t R = A.size(0), F = A.size(1);
    const size_t ws = hasty_svt_workspace_size(R, F);
    at::Tensor w = at::empty({(int64_t)ws}, A.options().dtype(at::kByte));
    check(hasty_svt_matrix(A.data_ptr(), R, F, thresh, soft ? 1 : 0, w.data_ptr(), ws, detail::stream()));
    return A.to(in.device(), in.scalar_type());
}
}  // namespace

at::Tensor svt_hard(const at::Tensor& in, int rank) { return svt_run(in, (double)rank, false); }
void svt_hard_inplace(at::Tensor& in, int rank) { in.copy_(svt_run(in, (double)rank, false)); }
at::Tensor svt_soft(const at::Tensor& in, float lambda) { return svt_run(in, lambda, true); }
void svt_soft_inplace(at::Tensor& in, float lambda) { in.copy_(svt_run(in, lambda, true)); }

at::Tensor& BlocksSVTBase::workspace(size_t bytes, const at::Device& dev)
{
    if (!_workspace.defined() || (size_t)_workspace.numel() < bytes || _workspace.device() != dev)
        _workspace = at::empty({(int64_t)bytes}, at::TensorOptions(dev).dtype(at::kByte));
    return _workspace;
}

namespace {
// Run on the GPU; CPU inputs (the reference's layout) are copied there and back.
template <typename Fn> void on_gpu(at::Tensor& in, Fn&& fn)
{
    if (in.dim() != 5) throw std::runtime_error("input must be [F, E, X, Y, Z]");
    if (in.is_cuda() && in.is_contiguous() && in.scalar_type() == at::kComplexFloat) {
        fn(in);
        return;
    }
    at::Tensor g = in.to(at::Device(at::kCUDA, c10::hip::current_device()), at::kComplexFloat).contiguous();
    fn(g);
    in.copy_(g);
}
}  // namespace

void Random3DBlocksSVT::apply(at::Tensor in, int64_t nblocks, std::array<int64_t, 3> block_shape, double thresh,
                              bool soft)
{
    c10::InferenceMode guard;
    // Block<3>::randomize (block.ixx:20-32): corner d uniform in [0, size_d - shape_d - 1],
    // from a random_device-seeded mt19937 (so, as in the reference, not reproducible).
    static std::random_device dev;
    static std::mt19937 rng(dev());
    std::vector<int32_t> corners;
    for (int64_t i = 0; i < nblocks; ++i) {
        for (int d = 0; d < 3; ++d) {
            const int64_t bound = in.size(2 + d) - block_shape[d];
            if (bound < 1) throw std::runtime_error("block shape must be smaller than the image");
            std::uniform_int_distribution<std::mt19937::result_type> dist(0, bound - 1);
            corners.push_back((int32_t)dist(rng));
        }
    }
    _last = corners;
    on_gpu(in, [&](at::Tensor& g) {
        std::vector<int64_t> dims = g.sizes().vec();
        std::vector<int64_t> shp(block_shape.begin(), block_shape.end());
        const size_t ws = hasty_llr_blocks_workspace_size(dims.data(), shp.data(), 1);
        at::Tensor& w = workspace(ws, g.device());
        // blocks may overlap: threshold them one at a time in draw order
        for (int64_t b = 0; b < nblocks; ++b)
            check(hasty_llr_blocks(g.data_ptr(), dims.data(), shp.data(), corners.data() + 3 * b, 1, thresh,
                                   soft ? 1 : 0, w.data_ptr(), (size_t)w.numel(), detail::stream()));
    });
}

void OldNormal3DBlocksSVT::apply(at::Tensor in, const std::array<int64_t, 3>& block_strides,
                                 const std::array<int64_t, 3>& block_shape, int64_t block_iter, double thresh,
                                 bool soft)
{
    c10::InferenceMode guard;
    if (block_iter < 1) throw std::runtime_error("Can't perform Normal3DBlocksSVT with block_iter < 1");
    on_gpu(in, [&](at::Tensor& g) {
        std::vector<int64_t> dims = g.sizes().vec();
        std::vector<int64_t> st(block_strides.begin(), block_strides.end());
        std::vector<int64_t> shp(block_shape.begin(), block_shape.end());
        const size_t ws = hasty_llr_workspace_size(dims.data(), st.data(), shp.data());
        at::Tensor& w = workspace(ws, g.device());
        check(hasty_llr_normal_blocks(g.data_ptr(), dims.data(), st.data(), shp.data(), block_iter, nullptr, thresh,
                                      soft ? 1 : 0, w.data_ptr(), (size_t)w.numel(), detail::stream()));
    });
}

}  // namespace svt

namespace op {

SenseNOp::SenseNOp(const at::Tensor& coords, const std::vector<int64_t>& nmodes, const at::Tensor& smaps,
                   const std::vector<int64_t>& coils, const at::optional<fft::NufftOptions>& forward_opts,
                   const at::optional<fft::NufftOptions>& backward_opts)
    : _h(std::make_shared<Holder>())
{
    TORCH_CHECK(coords.is_cuda(), "Coordinates did not reside in cuda device");
    TORCH_CHECK(!nmodes.empty() && nmodes[0] >= 1, "nmodes must start with the transform count");
    // tolerances from the options; types and signs fixed so that apply_backward is the
    // adjoint of apply_forward
    const fft::NufftOptions fwd = fft::NufftOptions::type2(forward_opts.has_value() ? forward_opts->tol : 1e-5);
    const fft::NufftOptions bwd = fft::NufftOptions::type1(backward_opts.has_value() ? backward_opts->tol : 1e-5);
    _h->coords = coords;
    _h->nmodes = nmodes;
    _h->smaps = smaps;
    _h->coils = coils;
    _h->sense = std::make_unique<mri::CUDASenseNormal>(coords, nmodes, fwd, bwd);
}

at::Tensor SenseNOp::apply(const at::Tensor& in) const
{
    at::Tensor out = at::empty_like(in);
    _h->sense->apply(in, out, _h->smaps, _h->coils);
    return out;
}

at::Tensor SenseNOp::apply_forward(const at::Tensor& in) const
{
    at::Tensor out = fft::allocate_out(_h->coords, (int)_h->coils.size());
    _h->sense->apply_forward(in, out, _h->smaps, _h->coils);
    return out;
}

at::Tensor SenseNOp::apply_backward(const at::Tensor& in) const
{
    std::vector<int64_t> img = _h->nmodes;
    img[0] = 1;
    at::Tensor out = fft::allocate_adjoint_out(_h->coords, img);
    _h->sense->apply_backward(in, out, _h->smaps, _h->coils);
    return out;
}

std::vector<at::Tensor> SenseNOp::apply(const std::vector<at::Tensor>& in) const
{
    std::vector<at::Tensor> out;
    out.reserve(in.size());
    for (const auto& t : in) out.push_back(apply(t));
    return out;
}

std::vector<at::Tensor> SenseNOp::apply_forward(const std::vector<at::Tensor>& in) const
{
    std::vector<at::Tensor> out;
    out.reserve(in.size());
    for (const auto& t : in) out.push_back(apply_forward(t));
    return out;
}

std::vector<at::Tensor> SenseNOp::apply_backward(const std::vector<at::Tensor>& in) const
{
    std::vector<at::Tensor> out;
    out.reserve(in.size());
    for (const auto& t : in) out.push_back(apply_backward(t));
    return out;
}

std::shared_ptr<SenseNOp> SenseNOp::adjoint() const { return std::shared_ptr<SenseNOp>(new SenseNOp(_h)); }

// opalgs.cpp:29-44
at::Tensor PowerIteration::run(const Op& A, at::Tensor& v, int iters)
{
    at::Tensor lambda = at::linalg_vector_norm(v);
    for (int i = 0; i < iters; ++i) {
        v = A(v);
        lambda = at::linalg_vector_norm(v);
        v.div_(lambda);
    }
    return lambda;
}

ConjugateGradient::ConjugateGradient(Op A, at::Tensor b, at::optional<Op> P)
    : _A(std::move(A)), _b(std::move(b)), _P(std::move(P))
{
}

// opalgs.cpp:50-96
void ConjugateGradient::run(at::Tensor& x, int iter, double tol)
{
    if (iter < 1) return;
    at::Tensor r = _b - _A(x);
    at::Tensor z = _P.has_value() ? (*_P)(r) : r;
    at::Tensor p = z.clone();
    at::Tensor rzold = at::real(at::vdot(r.flatten(), z.flatten()));
    double resid = std::sqrt(rzold.item<double>());
    for (int i = 0; i < iter; ++i) {
        if (resid < tol) return;
        at::Tensor Ap = _A(p);
        at::Tensor pAp = at::real(at::vdot(p.flatten(), Ap.flatten()));
        if (pAp.item<double>() <= 0.0) throw std::runtime_error("A was not positive definite");
        at::Tensor alpha = rzold / pAp;
        x.add_(p * alpha);
        r.sub_(Ap * alpha);
        z = _P.has_value() ? (*_P)(r) : r;
        at::Tensor rznew = at::real(at::vdot(r.flatten(), z.flatten()));
        p.mul_(rznew / rzold);
        p.add_(z);
        rzold = rznew;
        resid = std::sqrt(rzold.item<double>());
    }
}

BatchedConjugateGradient::BatchedConjugateGradient(BatchOp A, at::Tensor b) : _A(std::move(A)), _b(std::move(b)) {}

at::Tensor BatchedConjugateGradient::run(at::Tensor& x, int iter, double tol)
{
    const int64_t P = _b.size(0);
    const int64_t n = _b.numel() / P;
    if (_b.scalar_type() != at::kComplexFloat || !_b.is_contiguous() || !x.is_contiguous() || x.sizes() != _b.sizes())
        throw std::runtime_error("batched CG takes contiguous complex64 [P, ...] tensors");
    at::Tensor state = at::zeros({P, 8}, _b.options().dtype(at::kDouble));
    if (iter < 1) return state;
    at::Tensor r = at::empty_like(_b), p = at::empty_like(_b), Ap = at::empty_like(_b);
    const size_t ws = hasty_cg_workspace_size(P, n);
    at::Tensor w = at::empty({(int64_t)ws}, _b.options().dtype(at::kByte));
    // A 0 is exactly 0: a zero start gives r = b to the bit without an apply
    if (at::count_nonzero(at::view_as_real(x)).item<int64_t>() == 0)
        Ap.zero_();
    else
        _A(x, Ap);
    check(hasty_cg_init(P, n, _b.data_ptr(), Ap.data_ptr(), r.data_ptr(), p.data_ptr(), state.data_ptr<double>(),
                        w.data_ptr(), detail::stream()));
    for (int i = 0; i < iter; ++i) {
        _A(p, Ap);
        check(hasty_cg_iterate(P, n, x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(), state.data_ptr<double>(),
                               w.data_ptr(), tol, detail::stream()));
    }
    return state;
}

}  // namespace op

}  // namespace hasty
