"""Batched persistent IC(0)-PCG and the differentiable solves, through the C
ABI, against the oracle.  Bars: FP64 per-system iterations within +-2 of the
oracle's FlexibleCG with the same IC(0) factor and x within 1e-6; mixed
precision true relative residual < 1e-8 and x within 1e-6 of the FP64 oracle;
gradients within 1e-5 (relative) of central finite differences (h = 1e-6, FP64)
and within 1e-6 of the FP64 CPU adjoint (oracle CG on A^T)."""
import numpy as np
import pytest

from oracle import ppo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def pp(torch):
    import prispack_amd
    return prispack_amd


def make_systems(nsys, lo, hi, seed):
    """Uneven 2D variable-coefficient diffusion systems: side ~ U[lo, hi],
    kappa = exp(z), z ~ N(0, 1) per node."""
    sides = [lo + int((hi - lo + 1) * u) for u in (ppo.uniform(nsys, seed) + 1) / 2]
    systems = []
    for s, side in enumerate(sides):
        kappa = np.exp(ppo.normal(side * side, 1000 + s))
        systems.append(ppo.vardiff2d(side, kappa))
    return sides, systems


def test_batched_fp64_and_mixed_against_oracle(pp, torch):
    sides, systems = make_systems(12, 6, 40, 3)
    print("[batch] sides", sides, flush=True)
    B = pp.BatchedSystems(systems)
    bs = [ppo.spmv(A, ppo.uniform(len(A[0]) - 1, 50 + s)) for s, A in enumerate(systems)]
    b = torch.tensor(np.concatenate(bs), device="cuda")
    x64 = torch.zeros_like(b)
    its = B.solve(b, x64, precision="fp64", tol=1e-8)
    assert B.report.converged and B.report.true_relres < 1e-8
    xm = torch.zeros_like(b)
    B.solve(b, xm, precision="mixed", tol=1e-8)
    assert B.report.converged and B.report.true_relres < 1e-8
    off = B.row_off
    for s, A in enumerate(systems):
        o = ppo.cg(A, bs[s], ppo.Prec(A, "ic0"), tol=1e-8, maxit=5000)
        xs = x64[off[s]:off[s + 1]].cpu().numpy()
        assert abs(int(its[s]) - o["iters"]) <= 2, (s, its[s], o["iters"])
        assert np.linalg.norm(xs - o["x"]) / np.linalg.norm(o["x"]) <= 1e-6
        xms = xm[off[s]:off[s + 1]].cpu().numpy()
        assert np.linalg.norm(xms - o["x"]) / np.linalg.norm(o["x"]) <= 1e-6


def test_batched_mixed_reduced_and_flexible_inner_agree(pp, torch):
    # the same colour-ordered batch: reduced inner solves (every system two
    # colours) against FlexibleCG inner solves (one 9-point system among
    # them has three colours, so the batch keeps the sweep-based inner solve)
    from mc_order import greedy_order
    _, systems = make_systems(16, 8, 72, 6)
    rp, ci, v = ppo.laplacian(9, 9, 1)
    rows = [dict(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist())) for i in range(81)]
    for i in range(81):
        x, y = i % 9, i // 9
        for dx, dy in ((1, 1), (-1, -1), (1, -1), (-1, 1)):
            if 0 <= x + dx < 9 and 0 <= y + dy < 9:
                rows[i][(y + dy) * 9 + x + dx] = -0.25
        rows[i][i] += 1.0
    nine = (np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint32),
            np.array([c for r in rows for c in sorted(r)], np.uint32),
            np.array([r[c] for r in rows for c in sorted(r)]))
    assert greedy_order(nine)[1] > 2
    bs = [ppo.uniform(len(A[0]) - 1, 90 + s) for s, A in enumerate(systems)]
    xs = {}
    for name, batch in (("reduced", systems), ("flexible", systems + [nine])):
        B = pp.BatchedSystems(batch, ordering=1)
        bb = bs + ([ppo.uniform(81, 99)] if name == "flexible" else [])
        b = torch.tensor(np.concatenate(bb), device="cuda")
        x = torch.zeros_like(b)
        B.solve(b, x, precision="mixed", tol=1e-8)
        print(f"[batch] {name}: variant {B.report.cg_variant}, inner {B.report.inner_iters}", flush=True)
        assert B.report.converged and B.report.true_relres < 1e-8
        assert B.report.cg_variant == (3 if name == "reduced" else 1)
        xs[name] = x[:B.row_off[len(systems)]].cpu().numpy()
    assert np.linalg.norm(xs["reduced"] - xs["flexible"]) / np.linalg.norm(xs["flexible"]) <= 1e-6


def test_batched_autotune_and_graph_replays(pp, torch):
    _, systems = make_systems(40, 8, 64, 4)
    B = pp.BatchedSystems(systems)
    bsz = B.tune()
    assert bsz in (256, 512, 1024)
    b = torch.tensor(np.concatenate([ppo.uniform(len(A[0]) - 1, 7) for A in systems]), device="cuda")
    x1 = torch.zeros_like(b)
    B.solve(b, x1, precision="mixed")
    for _ in range(3):  # replays give the same answer
        x2 = torch.zeros_like(b)
        B.solve(b, x2, precision="mixed")
        assert torch.equal(x1, x2)
    assert B.report.true_relres < 1e-8


def test_batched_gradients(pp, torch):
    from prispack_amd.autograd import batched_solve
    sides, systems = make_systems(3, 5, 8, 9)
    B = pp.BatchedSystems(systems)
    n = B.n
    b0 = ppo.uniform(n, 11)
    xt = ppo.uniform(n, 12)
    vals0 = B.v.copy()

    def loss_np(vals, b, tol=1e-13):
        B.set_values(vals)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        B.solve(torch.tensor(b, device="cuda"), x, precision="fp64", tol=tol)
        x = x.cpu().numpy()
        return float(np.sum((x - xt) ** 2)), x

    values = torch.tensor(vals0, device="cuda", requires_grad=True)
    bt = torch.tensor(b0, device="cuda", requires_grad=True)
    x = batched_solve(values, bt, B, precision="fp64", tol=1e-13)
    L = torch.sum((x - torch.tensor(xt, device="cuda")) ** 2)
    L.backward()
    gv, gb = values.grad.cpu().numpy(), bt.grad.cpu().numpy()

    # FP64 CPU adjoint with the oracle: A^T lambda = 2 (x - x_t), per system
    _, xs = loss_np(vals0, b0)
    off, rp, ci = B.row_off, B.rp, B.ci
    lam = np.zeros(n)
    for s, A in enumerate(systems):
        r0, r1 = off[s], off[s + 1]
        g = 2 * (xs[r0:r1] - xt[r0:r1])
        lam[r0:r1] = ppo.cg(ppo.transpose(A), g, ppo.Prec(A, "ic0"), tol=1e-13, maxit=5000)["x"]
    rows = np.repeat(np.arange(n), np.diff(rp))
    gv_cpu = -lam[rows] * xs[ci]
    np.testing.assert_allclose(gb, lam, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gv, gv_cpu, rtol=1e-6, atol=1e-9)

    # central finite differences, h = 1e-6: b entries, diagonal entries and
    # symmetric off-diagonal pairs (the perturbed matrix stays SPD)
    h = 1e-6
    for i in (0, n // 2, n - 1):
        e = np.zeros(n)
        e[i] = h
        fd = (loss_np(vals0, b0 + e)[0] - loss_np(vals0, b0 - e)[0]) / (2 * h)
        assert abs(fd - gb[i]) <= 1e-5 * max(1.0, abs(gb[i])), (i, fd, gb[i])
    diag = [k for k in range(len(ci)) if ci[k] == rows[k]]
    off_pairs = []
    for k in range(len(ci)):
        if ci[k] > rows[k]:
            kt = [q for q in range(rp[ci[k]], rp[ci[k] + 1]) if ci[q] == rows[k]][0]
            off_pairs.append((k, kt))
        if len(off_pairs) >= 3:
            break
    for ks in [[diag[0]], [diag[len(diag) // 2]], *[list(p) for p in off_pairs]]:
        vp, vm = vals0.copy(), vals0.copy()
        vp[ks] += h
        vm[ks] -= h
        fd = (loss_np(vp, b0)[0] - loss_np(vm, b0)[0]) / (2 * h)
        an = sum(gv[k] for k in ks)
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (ks, fd, an)


def test_sparse_solve_gradient(pp, torch):
    from prispack_amd.autograd import sparse_solve
    A = ppo.upwind(6, 5, 4, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    b0, xt = ppo.uniform(n, 1), ppo.uniform(n, 2)
    values = torch.tensor(A[2], device="cuda", requires_grad=True)
    bt = torch.tensor(b0, device="cuda", requires_grad=True)
    x = sparse_solve(values, bt, M, lambda m: pp.ILU0(m), gmres=True, tol=1e-12)
    L = torch.sum((x - torch.tensor(xt, device="cuda")) ** 2)
    L.backward()
    # dense FP64 check of the adjoint formulas
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(A[0]))
    D[rows, A[1]] = A[2]
    xs = np.linalg.solve(D, b0)
    lam = np.linalg.solve(D.T, 2 * (xs - xt))
    np.testing.assert_allclose(bt.grad.cpu().numpy(), lam, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(values.grad.cpu().numpy(), -lam[rows] * xs[A[1]], rtol=1e-6, atol=1e-9)


def test_batched_multicolor_against_oracle(pp, torch):
    # every system in colour order: the oracle's FlexibleCG + IC(0) on P A P^T
    from mc_order import greedy_order, permute
    sides, systems = make_systems(10, 6, 48, 4)
    B = pp.BatchedSystems(systems, ordering=1)
    bs = [ppo.spmv(A, ppo.uniform(len(A[0]) - 1, 70 + s)) for s, A in enumerate(systems)]
    b = torch.tensor(np.concatenate(bs), device="cuda")
    x64 = torch.zeros_like(b)
    its = B.solve(b, x64, precision="fp64", tol=1e-8)
    assert B.report.converged and B.report.true_relres < 1e-8
    assert B.report.cg_variant == 1  # FP64: the reference's FlexibleCG per system
    xm = torch.zeros_like(b)
    B.solve(b, xm, precision="mixed", tol=1e-8)
    assert B.report.converged and B.report.true_relres < 1e-8
    # two-colour systems: the FP32 inner solves run the reduced CG
    assert B.report.cg_variant == 3
    off = B.row_off
    for s, A in enumerate(systems):
        order, _ = greedy_order(A)
        Ap = permute(A, order)
        o = ppo.cg(Ap, bs[s][order], ppo.Prec(Ap, "ic0"), tol=1e-8, maxit=5000)
        xo = np.empty_like(o["x"])
        xo[order] = o["x"]
        assert abs(int(its[s]) - o["iters"]) <= 2, (s, its[s], o["iters"])
        for xx in (x64, xm):
            xs = xx[off[s]:off[s + 1]].cpu().numpy()
            assert np.linalg.norm(xs - xo) / np.linalg.norm(xo) <= 1e-6
    # new values on the same pattern keep the ordering
    B.set_values(B.v * 1.5)
    x2 = torch.zeros_like(b)
    B.solve(b, x2, precision="fp64", tol=1e-8)
    np.testing.assert_allclose(x2.cpu().numpy() * 1.5, x64.cpu().numpy(), rtol=1e-6, atol=1e-9)


def test_gradients_of_two_forwards_before_backward(pp, torch):
    # two forwards on one matrix with different values, then one backward:
    # each gradient is the one of its own forward's values (dense FP64 check)
    from prispack_amd.autograd import sparse_solve
    A = ppo.upwind(6, 5, 4, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    rows = np.repeat(np.arange(n), np.diff(A[0]))
    b0, xt = ppo.uniform(n, 1), ppo.uniform(n, 2)
    v1 = np.asarray(A[2], dtype=np.float64)
    v2 = v1 * 1.3
    v2[np.asarray(A[1]) == rows] += 0.5
    t1 = torch.tensor(v1, device="cuda", requires_grad=True)
    t2 = torch.tensor(v2, device="cuda", requires_grad=True)
    bt = torch.tensor(b0, device="cuda")
    x1 = sparse_solve(t1, bt, M, lambda m: pp.ILU0(m), gmres=True, tol=1e-12)
    x2 = sparse_solve(t2, bt, M, lambda m: pp.ILU0(m), gmres=True, tol=1e-12)
    xtt = torch.tensor(xt, device="cuda")
    (torch.sum((x1 - xtt) ** 2) + torch.sum((x2 - xtt) ** 2)).backward()
    for vals, t in ((v1, t1), (v2, t2)):
        D = np.zeros((n, n))
        D[rows, A[1]] = vals
        xs = np.linalg.solve(D, b0)
        lam = np.linalg.solve(D.T, 2 * (xs - xt))
        np.testing.assert_allclose(t.grad.cpu().numpy(), -lam[rows] * xs[A[1]], rtol=1e-6, atol=1e-9)

    # the batched form: two forwards with different coefficient values
    from prispack_amd.autograd import batched_solve
    sides, systems = make_systems(3, 5, 8, 21)
    B = pp.BatchedSystems(systems)
    nb = B.n
    w1 = B.v.copy()
    w2 = w1 * 2.0
    bb = torch.tensor(ppo.uniform(nb, 5), device="cuda")
    xtb = torch.tensor(ppo.uniform(nb, 6), device="cuda")
    s1 = torch.tensor(w1, device="cuda", requires_grad=True)
    s2 = torch.tensor(w2, device="cuda", requires_grad=True)
    y1 = batched_solve(s1, bb, B, precision="fp64", tol=1e-13)
    y2 = batched_solve(s2, bb, B, precision="fp64", tol=1e-13)
    (torch.sum((y1 - xtb) ** 2) + torch.sum((y2 - xtb) ** 2)).backward()
    # A scaled by 2: x2 = x1 / 2, lambda2 = A^-T 2 (x1/2 - xt) / 2
    rb = np.repeat(np.arange(nb), np.diff(B.rp))
    for vals, t in ((w1, s1), (w2, s2)):
        D = np.zeros((nb, nb))
        D[rb, B.ci] = vals
        xs = np.linalg.solve(D, bb.cpu().numpy())
        lam = np.linalg.solve(D.T, 2 * (xs - xtb.cpu().numpy()))
        np.testing.assert_allclose(t.grad.cpu().numpy(), -lam[rb] * xs[B.ci], rtol=1e-6, atol=1e-9)
