"""BiCGSTAB driver (the Krylov use of the HMVM): eager and hipGraph-replayed
solves give the same bits, follow a CPU BiCGSTAB built on the oracle, and
actually solve the system."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cpu_bicgstab(O, h, b, u, iters, history=None):
    """The iteration of m_HACApK_solve.f90:306-380 in numpy, with the oracle HMVM;
    ``history`` collects ||r||/||b|| after every iteration."""
    def A(v):
        return O.seq_2(h, v)
    r = b - A(u)
    sh = r.copy()
    p = np.zeros_like(b)
    ap = np.zeros_like(b)
    beta = zeta = 0.0
    for _ in range(iters):
        p = r + beta * (p - zeta * ap)
        ap = A(p)
        znom = sh @ r
        alpha = znom / (sh @ ap)
        t = r - alpha * ap
        at = A(t)
        zeta = (at @ t) / (at @ at)
        u = u + alpha * p + zeta * t
        r = t - zeta * at
        beta = alpha / zeta * (sh @ r) / znom
        if history is not None:
            history.append(np.linalg.norm(r) / np.linalg.norm(b))
    return u, np.linalg.norm(r) / np.linalg.norm(b)


def shifted(H, n, geometry, sigma):
    """The generated H-matrix plus sigma on the diagonal.  The point kernel
    matrix itself is indefinite and ill-conditioned (BiCGSTAB stagnates on it,
    on the CPU as well); the shift makes the solve a meaningful check."""
    h = H.generate(n, geometry=geometry, eps=1e-6)
    a = h.arrays()
    for i in range(h.mat2.nlf):
        if a["ltmtx"][i] == 2 and a["nstrtl"][i] == a["nstrtt"][i]:
            ndl, ndt = a["ndl"][i], a["ndt"][i]
            for il in range(ndl):
                a["rowmat"][a["a1"][i] + il * ndt + il] += sigma
                a["rowmat_t"][a["a1"][i] + il * ndt + il] += sigma
    return h


@pytest.mark.parametrize("mode", [0, 1])
def test_bicgstab_graph_equals_eager_and_cpu(gpu, H, O, mode):
    h = shifted(H, 3000, "sphere", 8.0)
    b = np.random.default_rng(21).uniform(-1, 1, h.nd)
    p = H.Plan(h, mode=mode)
    assert p.info()["mode"] == mode  # mode 1: the persistent launch runs inside the solve
    u_e, r_e = p.bicgstab(b, max_iter=8, tol=0.0, use_graph=False)
    u_g, r_g = p.bicgstab(b, max_iter=8, tol=0.0, use_graph=True, check_every=2)
    assert r_e["iters"] == r_g["iters"] == 8 and r_g["graph_used"] == 1 and r_g["hmvms"] == 17
    assert np.array_equal(u_e, u_g)
    u_c, rel_c = cpu_bicgstab(O, h, b, np.zeros(h.nd), 8)
    # eight BiCGSTAB steps amplify the ~1e-16 HMVM summation-order difference
    # (observed 2.5e-9 on MI355X); a wrong step of the recurrence gives O(1)
    assert np.linalg.norm(u_e - u_c) <= 1e-7 * np.linalg.norm(u_c)
    assert abs(r_e["relres"] - rel_c) <= 1e-4 * max(rel_c, 1e-30) + 1e-14  # observed 1.3e-5 relative
    p.close()


def test_bicgstab_converges(gpu, H, O):
    h = shifted(H, 2500, "cube", 8.0)
    b = np.random.default_rng(22).uniform(-1, 1, h.nd)
    p = H.Plan(h)
    u, res = p.bicgstab(b, max_iter=200, tol=1e-10, use_graph=True, check_every=5)
    assert res["relres"] < 1e-10 and res["iters"] < 200
    true_res = np.linalg.norm(b - O.seq_2(h, u)) / np.linalg.norm(b)
    assert true_res < 1e-8
    p.close()


def test_bicgstab_stops_at_first_iteration_below_tol(gpu, H, O):
    """The reference leaves the loop at the first iteration whose residual is
    below eps (m_HACApK_solve.f90:342).  The device stop word gives that
    iteration and iterate whatever the host's read-back interval: checked
    every iteration, every 7, only at the end, eagerly and as graph replays,
    the solves agree bit for bit and stop where the CPU recurrence first
    crosses tol."""
    h = shifted(H, 2500, "sphere", 8.0)
    b = np.random.default_rng(23).uniform(-1, 1, h.nd)
    hist = []
    cpu_bicgstab(O, h, b, np.zeros(h.nd), 12, hist)
    # the first iteration k >= 5 whose residual is clearly below every earlier
    # one; tol halfway (geometrically) between that residual and the smallest
    # earlier one, so that k is the first iteration below tol with a margin
    k = next(j for j in range(5, 13) if hist[j - 1] < 0.7 * min(hist[:j - 1]))
    tol = float(np.sqrt(min(hist[:k - 1]) * hist[k - 1]))
    p = H.Plan(h)
    runs = [p.bicgstab(b, max_iter=40, tol=tol, use_graph=g, check_every=c)
            for g, c in ((False, 1), (False, 7), (True, 7), (True, 0))]
    for u, res in runs:
        assert res["iters"] == k and res["hmvms"] == 2 * k + 1
        assert res["relres"] < tol
        assert np.array_equal(u, runs[0][0])
    assert abs(runs[0][1]["relres"] - hist[k - 1]) <= 1e-6 * hist[k - 1]
    p.close()


def test_bicgstab_breakdown_is_reported(gpu, H):
    """b = 0 with tol = 0: r = 0, so alpha = 0/0 in the first iteration.  The
    solve stops there and reports a breakdown instead of returning NaNs as a
    result (the host forces the condition; every kernel exits normally)."""
    h = shifted(H, 1500, "cube", 8.0)
    p = H.Plan(h)
    for g in (False, True):
        with pytest.raises(H.KrylovBreakdown) as ei:
            p.bicgstab(np.zeros(h.nd), max_iter=10, tol=0.0, use_graph=g, check_every=3)
        assert ei.value.result["iters"] == 0
        assert np.all(ei.value.u == 0.0)
    # with tol > 0 the zero residual is already converged: no iteration at all
    u, res = p.bicgstab(np.zeros(h.nd), max_iter=10, tol=1e-12)
    assert res["iters"] == 0 and np.all(u == 0.0)
    p.close()
