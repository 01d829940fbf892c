"""HIP-graph replay of one batched CG-SENSE iteration (the Toeplitz SENSE normal operator
over every problem, then the device CG step) against the same iterations run eagerly.
The reference runs these iterations from host worker threads (ConjugateGradient::run,
lib/op/opalgs.cpp:50-96, per frame through batch_sense.cpp:349-450); the graph replays
the whole batched iteration as one launch. Every replay is checked bit for bit against
the eager iterate, with eager library work and synchronisation between replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def N():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hastycompute_amd import native

    return native


def _problem(N, shape, C, P, seed):
    rng = np.random.default_rng(seed)
    kern = []
    for _ in range(2):
        co = torch.as_tensor(rng.uniform(-np.pi, np.pi, (3, 4000)).astype(np.float32)).cuda()
        _, k = N.toeplitz_build(co, list(shape), diag=False)
        kern.append(k)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    S = torch.randn((C,) + shape, dtype=torch.complex64, device="cuda", generator=g)
    S /= torch.sqrt((S.abs() ** 2).sum(0, keepdim=True))
    b = torch.randn((P,) + shape, dtype=torch.complex64, device="cuda", generator=g)
    return torch.stack(kern), S, b


@pytest.mark.parametrize("shape,C,P,pbatch,cbatch", [((32, 64, 64), 3, 4, 2, 2), ((128, 32, 32), 4, 5, 4, 4)])
def test_graph_replays_match_eager_iterations(N, shape, C, P, pbatch, cbatch):
    K, S, b = _problem(N, shape, C, P, seed=sum(shape) + C)
    op = N.SenseNormalToeplitz(S, K, coil_batch=cbatch, problem_batch=pbatch)
    kidx = [p % 2 for p in range(P)]
    n = b[0].numel()
    iters = 6

    def run(graph):
        x = torch.zeros_like(b)
        r, p, Ap = torch.empty_like(b), torch.empty_like(b), torch.empty_like(b)
        cg = N.DeviceCG(P, n, b.device)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        iterates = []
        other = torch.randn_like(b)
        with torch.cuda.stream(side):
            op.apply(x, out=Ap, kernel_index=kidx)
            cg.init(b, Ap, r, p)

            def step():
                op.apply(p, out=Ap, kernel_index=kidx)
                cg.iterate(x, r, p, Ap, 0.0)

            step()
            side.synchronize()
            iterates.append(x.clone())
            G = N.Graph(step, side) if graph else None
            for _ in range(iters - 1):
                G.replay() if graph else step()
                side.synchronize()
                iterates.append(x.clone())
                # eager work of the program between replays
                op.apply(other, out=other.clone(), kernel_index=kidx)
                side.synchronize()
            if G is not None:
                print(f"graph of one iteration: {G.nodes} nodes", flush=True)
                G.close()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        return iterates, cg.state.clone()

    eager, st_e = run(False)
    graph, st_g = run(True)
    for k, (a, g) in enumerate(zip(eager, graph)):
        assert torch.equal(a, g), f"iteration {k + 1} differs"
    assert torch.equal(st_e, st_g)
    assert bool((st_g[:, 6] == iters).all())


def test_cg_solve_graph_equals_eager(N):
    K, S, b = _problem(N, (32, 64, 64), 4, 3, seed=5)
    op = N.SenseNormalToeplitz(S, K, coil_batch=4, problem_batch=3)
    kidx = [1, 0, 1]
    A = lambda v, o: op.apply(v, out=o, kernel_index=kidx)
    xe, se = N.cg_solve(A, b, torch.zeros_like(b), 8)
    xg, sg = N.cg_solve(A, b, torch.zeros_like(b), 8, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(xe, xg) and torch.equal(se, sg)


def test_cg_solve_zero_start_skips_the_first_apply(N):
    # A 0 is exactly 0 through the fused operator, so a zero start needs no apply for
    # r = b: the solve makes `iters` applies instead of iters + 1 and gives the same bits
    # as a solve that applies A to the zero start
    K, S, b = _problem(N, (32, 64, 64), 4, 3, seed=6)
    op = N.SenseNormalToeplitz(S, K, coil_batch=4, problem_batch=3)
    kidx = [1, 0, 1]
    calls = [0]

    def A(v, o):
        calls[0] += 1
        op.apply(v, out=o, kernel_index=kidx)

    zero = torch.zeros_like(b)
    assert torch.count_nonzero(torch.view_as_real(op.apply(zero, kernel_index=kidx))) == 0
    for l2 in (0.0, 0.5):
        calls[0] = 0
        x_skip, st_skip = N.cg_solve(A, b, zero, 6, l2=l2)
        assert calls[0] == 6
        # the same solve with the first apply made: a start that is zero but is not
        # recognised as one (the skip tests the values, so this goes through a wrapper
        # that always applies)
        Ap = torch.empty_like(b)
        r = torch.empty_like(b)
        p = torch.empty_like(b)
        x = zero.clone()
        cg = N.DeviceCG(b.shape[0], b[0].numel(), b.device)
        A(x, Ap)
        if l2:
            Ap.add_(x, alpha=l2)
        cg.init(b, Ap, r, p)
        for _ in range(6):
            A(p, Ap)
            cg.iterate(x, r, p, Ap, 0.0, l2)
        torch.cuda.synchronize()
        assert torch.equal(x, x_skip) and torch.equal(cg.state, st_skip)


def test_llr_prox_graph_replays_equal_eager(N):
    # a prox-gradient step recorded once as a HIP graph: the normal operator (gradient
    # T x - b through the sub path), the step and the LLR prox on device-resident block
    # corners (hasty_llr_blocks_dev) of two shifts; every replay equals the same step run
    # eagerly through hasty_llr_normal_blocks; a corner outside the image is skipped
    F, n = 6, 32
    shape = (n, n, n)
    K, S, b = _problem(N, shape, 3, F, seed=12)
    op = N.SenseNormalToeplitz(S, K[:1], coil_batch=3, problem_batch=F)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(13)
    x0 = torch.randn((F,) + shape, dtype=torch.complex64, device="cuda", generator=gen)
    blk = [8, 8, 8]
    shifts = [[0, 3], [0, 5], [0, 2]]
    corner_sets = [N.llr_grid_corners([F, 1, n, n, n], blk, blk, [shifts[d][it] for d in range(3)]) for it in range(2)]
    cdev = [torch.tensor(c, dtype=torch.int32, device="cuda") for c in corner_sets]
    ws = N.llr_blocks_workspace(x0.view(F, 1, n, n, n), blk, max(len(c) for c in corner_sets))
    thresh, step = 0.5, 1e-3
    kidx = [0] * F

    def eager(x):
        grad = op.apply(x, kernel_index=kidx, sub=b)
        y = x - step * grad
        for it in range(2):
            N.llr_normal_blocks(y.view(F, 1, n, n, n), blk, blk, 1, thresh, True,
                                shifts=[[shifts[d][it]] for d in range(3)])
        return y

    x = x0.clone()
    grad = torch.empty_like(x)

    def record():
        op.apply(x, out=grad, kernel_index=kidx, sub=b)
        x.sub_(grad, alpha=step)
        for it in range(2):
            N.llr_blocks_dev(x.view(F, 1, n, n, n), blk, cdev[it], thresh, True, ws)

    record()  # eager run on the same buffers first
    ref = eager(x0.clone())
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    gr = N.Graph(record)
    for _ in range(3):
        ref = eager(ref)
        gr.replay()
        gr.stream.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(x, ref)
    gr.close()
    # a block outside the image is left unchanged, the one inside is thresholded to zero
    y = x0.clone().view(F, 1, n, n, n)
    before = y.clone()
    bad = torch.tensor([[0, 0, 0], [n - 4, 0, 0]], dtype=torch.int32, device="cuda")
    N.llr_blocks_dev(y, blk, bad, 1e6, True, ws)
    torch.cuda.synchronize()
    assert torch.count_nonzero(y[:, :, :8, :8, :8]) == 0
    assert torch.equal(y[:, :, 8:], before[:, :, 8:])


def test_nufft_pair_graph_replays_equal_eager(N):
    # a type-2 and a type-1 NUFFT (64^3 modes: register FFT passes, the output-driven
    # spreader with atomic parts for the crowded k-space centre) recorded as one HIP
    # graph after setpts; before each replay the outputs are overwritten, so a replay must
    # write every cell (no zero pass); type 2 equals the eager result bit for bit, type 1
    # (atomic adds of the crowded cores' further parts, in no fixed order) to rounding
    from oracle import oracle as O
    import numpy as np

    n = 64
    co = torch.as_tensor(O.kooshball(300, 2 * n).astype(np.float32)).cuda()
    M = co.shape[1]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(21)
    f = torch.randn((1, n, n, n), dtype=torch.complex64, device="cuda", generator=gen)
    c = torch.randn((1, M), dtype=torch.complex64, device="cuda", generator=gen)
    p2 = N.NufftPlan(co, [n, n, n], 2, -1, 1, 1e-6)
    p1 = N.NufftPlan(co, [n, n, n], 1, +1, 1, 1e-6)
    kout = torch.empty((1, M), dtype=torch.complex64, device="cuda")
    img = torch.empty((1, n, n, n), dtype=torch.complex64, device="cuda")

    def record():
        p2.execute(kout, f)
        p1.execute(c, img)

    record()
    torch.cuda.synchronize()
    ref2, ref1 = kout.clone(), img.clone()
    gr = N.Graph(record)
    for _ in range(3):
        kout.fill_(3.0)
        img.fill_(7.0)
        p1.execute(c, torch.empty_like(img))  # eager work between the replays
        torch.cuda.synchronize()
        gr.replay()
        gr.stream.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(kout, ref2)
        e = (torch.linalg.vector_norm(img - ref1) / torch.linalg.vector_norm(ref1)).item()
        assert e < 1e-6, e
    gr.close()
    p1.close()
    p2.close()
