"""hasty_svt_matrix_noalloc (the entry point that allocates nothing, with its scratch in a
static device array) against hasty_svt_matrix (the same SVT with a caller workspace,
svt_soft / svt_hard of lib/op/svt.cpp:62-169), through the C ABI: two calls in a row,
the second on a smaller matrix, soft and hard thresholds; the first also against the
FP64 oracle (LAPACK SVD) at the LLR tolerance 1e-4."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import _capi

    return _capi


def _svt_ws(C, A, thresh, soft):
    R, F = A.shape
    ws = torch.empty(C.lib().hasty_svt_workspace_size(R, F), dtype=torch.uint8, device=A.device)
    C.call("hasty_svt_matrix", A.data_ptr(), R, F, float(thresh), int(soft), ws.data_ptr(), ws.numel(),
           C.stream_ptr())
    return A


def _svt_noalloc(C, A, thresh, soft):
    R, F = A.shape
    C.call("hasty_svt_matrix_noalloc", A.data_ptr(), R, F, float(thresh), int(soft), C.stream_ptr())
    return A


def _lowrank(rng, R, F, rank):
    a = rng.standard_normal((R, rank)) + 1j * rng.standard_normal((R, rank))
    b = rng.standard_normal((rank, F)) + 1j * rng.standard_normal((rank, F))
    return a @ b + 0.05 * (rng.standard_normal((R, F)) + 1j * rng.standard_normal((R, F)))


@pytest.mark.parametrize("soft", [True, False])
def test_noalloc_svt_matches_workspace_svt(lib, soft):
    rng = np.random.default_rng(11 + soft)
    for call, (R, F, rank) in enumerate([(4096, 100, 5), (1000, 37, 3)]):
        A = _lowrank(rng, R, F, rank)
        s = np.linalg.svd(A, compute_uv=False)
        thresh = 0.01 * s[0] if soft else float(rank)
        Ag = torch.as_tensor(A).to(torch.complex64).cuda()
        a = _svt_noalloc(lib, Ag.clone(), thresh, soft)
        b = _svt_ws(lib, Ag.clone(), thresh, soft)
        torch.cuda.synchronize()
        e = O.rel_l2(a.cpu().numpy(), b.cpu().numpy())
        print(f"call {call + 1} ({R} x {F}, {'soft' if soft else 'hard'}): no-alloc vs workspace rel L2 {e:.2e}, "
              f"bit-equal {torch.equal(a, b)}", flush=True)
        assert e < 1e-6
        if call == 0:
            A32 = Ag.cpu().numpy().astype(np.complex128)
            ref = O.svt_soft(A32, thresh) if soft else O.svt_hard(A32, rank)
            eo = O.rel_l2(a.cpu().numpy(), ref)
            print(f"   no-alloc vs FP64 oracle {eo:.2e}", flush=True)
            assert eo < 1e-4


def test_noalloc_svt_rejects_too_many_columns(lib):
    A = torch.zeros((64, 129), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="1..128 columns"):
        _svt_noalloc(lib, A, 0.1, True)
