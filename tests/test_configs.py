"""Parity at the benchmark workloads themselves, not only at small sizes.

Each test builds the exact input of one benchmark configuration and checks the
HIP path against the oracle (the restatement of the reference's hmvm_seq_2 /
hmvm_omp_2, src_c/hmvm_seq.c:113-166, src_c/hmvm_omp.c:147-217):

  C2  N=101,250 cube surface, leaf 32, eta 2, ACA eps 1e-4, one right-hand
      side: the drop-in entry points 1/1t/2/2t after a hmatrix_1.bin round
      trip, and the FP64/FP32/BF16 plans
  C3  the same matrix, 32 right-hand sides on the matrix cores
  C4  64 independent matrices (N log-uniform in [2,000, 50,000], sphere/cube,
      eps in {1e-4, 1e-6}) in one batched FP32 plan, replayed as a hipGraph
  C5  N=1,048,576 sphere, eps 1e-6, leaf 64: more than 2^31 stored elements
      (int64 offsets), FP64/FP32/BF16 plans and a captured BiCGSTAB loop

Bars (relative l2 against the FP64 oracle): FP64 1e-13; FP32 1e-6 for one
right-hand side (the FP32-storage target) and 2e-6 on the matrix cores; BF16
is reported against the ACA eps the matrix was built with and bounded at
1e-2 (BF16 keeps 8 significant bits, so its error exceeds eps = 1e-4 by an
order of magnitude; that is stated, not hidden).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(*a):
    print(*a, flush=True)


# ---- C2 ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(gpu, H, tmp_path_factory):
    g = H.generate(101250, geometry="cube", eps=1e-4, leaf=32)
    fname = str(tmp_path_factory.mktemp("c2") / "hmatrix_1.bin")
    H.writeHmatrix1(fname, g)
    g.free()
    h = H.loadHmatrix(fname)  # matrix + matrix2, as the reference drivers load it
    os.remove(fname)
    say("C2: N=%d nlf=%d len=%d ktmax=%d" % (h.nd, h.mat2.nlf, h.mat2.len, h.mat2.ktmax))
    yield h
    h.free()


@pytest.fixture(scope="module")
def c2_x_y(c2, O):
    x = np.random.default_rng(20261202).uniform(-1.0, 1.0, c2.nd)
    return x, O.seq_2(c2, x)


@pytest.mark.timeout(300)
def test_c2_dropin_variants(H, O, c2, c2_x_y):
    """hmvm_hip_1/1t/2/2t at N=101,250 equal hmvm_seq_* to 1e-13 (and the
    OpenMP baseline agrees with the sequential oracle)."""
    x, y_or = c2_x_y
    assert rel(O.omp_2(c2, x), y_or) <= 1e-13
    for name, fn, seq in (("1", H.hmvm_hip_1, O.seq_1), ("1t", H.hmvm_hip_1t, O.seq_1t),
                          ("2", H.hmvm_hip_2, O.seq_2), ("2t", H.hmvm_hip_2t, O.seq_2t)):
        say("C2 drop-in hmvm_hip_%s" % name)
        v = np.zeros(c2.nd)
        fn(v, c2, x)
        ref = y_or if name == "2" else seq(c2, x)
        err = rel(v, ref)
        say("C2 hmvm_hip_%s rel err %.3e" % (name, err))
        assert err <= 1e-13
    H.release(c2)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 1e-6), ("bf16", 1e-2)])
def test_c2_plans(H, c2, c2_x_y, storage, tol):
    x, y_or = c2_x_y
    say("C2 plan %s" % storage)
    p = H.Plan(c2, storage=storage)
    y = p.apply_host(x)
    err = rel(y, y_or)
    say("C2 %s rel err %.3e = %.3g x ACA eps 1e-4" % (storage, err, err / 1e-4))
    assert err <= tol
    # accumulate form, and back-to-back applies give the same bits
    y2 = y.copy()
    p.apply_host(x, y2, accumulate=True)
    assert rel(y2, 2 * y) <= 1e-15
    assert np.array_equal(p.apply_host(x), y)
    p.close()


# ---- C3 ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3_X_Y(c2, O):
    X = np.random.default_rng(20261204).uniform(-1.0, 1.0, (c2.nd, 32))
    Y = np.stack([O.omp_2(c2, np.ascontiguousarray(X[:, c])) for c in range(32)], axis=1)
    return X, Y


@pytest.mark.timeout(300)
@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 2e-6), ("bf16", 1e-2)])
def test_c3_multi_rhs(H, c2, c3_X_Y, storage, tol):
    X, Y = c3_X_Y
    say("C3 32 RHS %s" % storage)
    p = H.Plan(c2, storage=storage)
    got = p.apply_multi_host(X)
    errs = [rel(got[:, c], Y[:, c]) for c in range(32)]
    say("C3 %s max column rel err %.3e (%.3g x eps)" % (storage, max(errs), max(errs) / 1e-4))
    assert max(errs) <= tol
    p.close()


# ---- C4 ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4(gpu, H):
    rng = np.random.default_rng(20261205)
    ns = np.exp(rng.uniform(np.log(2000), np.log(50000), 64)).astype(int)
    eps = rng.choice([1e-4, 1e-6], 64)
    mats = [H.generate(int(n), geometry=("sphere", "cube")[i % 2], eps=float(e), seed=3000 + i)
            for i, (n, e) in enumerate(zip(ns, eps))]
    say("C4: 64 matrices, N %d..%d, sum len %d" % (ns.min(), ns.max(), sum(m.mat2.len for m in mats)))
    yield mats, eps
    for m in mats:
        m.free()


@pytest.mark.timeout(600)
def test_c4_batch_graph(H, O, c4):
    """One FP32 plan over all 64 matrices: each matrix's rows of y equal its
    own oracle product to the FP32 bar, and hipGraph replays (with eager
    work and synchronisation between them) reproduce the eager apply bit for
    bit."""
    mats, eps = c4
    say("C4 batch plan f32")
    p = H.Plan(mats, storage="f32")
    offs = p.row_offsets
    x = np.random.default_rng(20261206).uniform(-1.0, 1.0, p.nd)
    y = p.apply_host(x)
    worst = 0.0
    for q, m in enumerate(mats):
        xs = np.ascontiguousarray(x[offs[q]:offs[q + 1]])
        err = rel(y[offs[q]:offs[q + 1]], O.omp_2(m, xs))
        worst = max(worst, err)
        assert err <= 1e-6, "matrix %d (N=%d, eps %g): %.3e" % (q, m.nd, eps[q], err)
    say("C4 worst matrix rel err %.3e" % worst)
    dx, dy = H.DeviceBuffer(8 * p.nd), H.DeviceBuffer(8 * p.nd)
    dx.upload(x)
    g = p.graph(dx.value, dy.value)
    for rep in range(3):
        g.launch()
        H.lib.hmvm_device_synchronize()
        got = np.empty(p.nd)
        dy.download(got)
        assert np.array_equal(got, y), "replay %d" % rep
        assert np.array_equal(p.apply_host(x), y)  # eager work between replays
    g.close()
    dx.close()
    dy.close()
    p.close()


# ---- C5 ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5(gpu, H, O):
    h = H.generate64(1048576, geometry="sphere", eps=1e-6, leaf=64)
    say("C5: N=%d nlf=%d len=%d (%.1f GB FP64)" % (h.nd, h.mat2.nlf, h.mat2.len, 8e-9 * h.mat2.len))
    x = np.random.default_rng(20261207).uniform(-1.0, 1.0, h.nd)
    y = O.omp_2_64(h, x)
    yield h, x, y
    h.free()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 1e-6), ("bf16", 1e-2)])
def test_c5_int64_plans(H, c5, storage, tol):
    h, x, y_or = c5
    assert h.mat2.len > 2 ** 31  # the int64 form is really exercised
    say("C5 plan %s" % storage)
    p = H.Plan(h, storage=storage)
    y = p.apply_host(x)
    err = rel(y, y_or)
    say("C5 %s rel err %.3e = %.3g x ACA eps 1e-6" % (storage, err, err / 1e-6))
    assert err <= tol
    p.close()


@pytest.mark.timeout(600)
def test_c5_bicgstab_graph_loop(H, O, c5):
    """The C5 usage pattern: BiCGSTAB iterations captured as one hipGraph.
    The replayed and the eager solves agree bit for bit, and two iterations
    match the CPU recurrence built on the oracle."""
    h, _, _ = c5
    b = np.random.default_rng(20261208).uniform(-1.0, 1.0, h.nd)
    say("C5 BiCGSTAB f64")
    p = H.Plan(h)
    u_g, r_g = p.bicgstab(b, max_iter=2, tol=0.0, use_graph=True, check_every=0)
    u_e, r_e = p.bicgstab(b, max_iter=2, tol=0.0, use_graph=False)
    p.close()
    assert r_g["graph_used"] == 1 and r_g["iters"] == r_e["iters"] == 2
    assert np.array_equal(u_g, u_e)

    def A(v):
        return O.omp_2_64(h, v)
    r = b - A(np.zeros(h.nd))
    sh, pp, ap, u = r.copy(), np.zeros(h.nd), np.zeros(h.nd), np.zeros(h.nd)
    beta = zeta = 0.0
    for _ in range(2):  # m_HACApK_solve.f90:341-372
        pp = r + beta * (pp - zeta * ap)
        ap = A(pp)
        znom = sh @ r
        alpha = znom / (sh @ ap)
        t = r - alpha * ap
        at = A(t)
        zeta = (at @ t) / (at @ at)
        u = u + alpha * pp + zeta * t
        r = t - zeta * at
        beta = alpha / zeta * (sh @ r) / znom
    err = rel(u_g, u)
    say("C5 BiCGSTAB u rel err vs CPU recurrence %.3e" % err)
    assert err <= 1e-8  # summation-order differences (~1e-16) amplified by two steps; a wrong step is O(1)
