"""hmvm_bicgstab_static, the allocation-free BiCGSTAB entry point, called
through the C ABI on device vectors: the same iterates, iteration counts and
residuals, bit for bit, as the allocating hmvm_bicgstab (eager), on two calls
in a row, the second with a smaller system; a plan beyond its fixed work
space is refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def shifted(H, n, geometry, sigma):
    """Generated H-matrix plus sigma on the diagonal (a well-posed solve)."""
    h = H.generate(n, geometry=geometry, eps=1e-6)
    a = h.arrays()
    for i in range(h.mat2.nlf):
        if a["ltmtx"][i] == 2 and a["nstrtl"][i] == a["nstrtt"][i]:
            for il in range(a["ndl"][i]):
                a["rowmat"][a["a1"][i] + il * a["ndt"][i] + il] += sigma
                a["rowmat_t"][a["a1"][i] + il * a["ndt"][i] + il] += sigma
    return h


def static_solve(H, p, b, u0, max_iter, tol, check_every):
    n = p.nd
    db, du = H.DeviceBuffer(8 * n), H.DeviceBuffer(8 * n)
    db.upload(b)
    du.upload(u0)
    res = H.hmvm_krylov_result()
    rc = H.lib.hmvm_bicgstab_static(p._p, db.value, du.value, max_iter, tol, check_every, None, ctypes.byref(res))
    u = np.empty(n)
    du.download(u)
    db.close()
    du.close()
    return rc, u, res.as_dict()


def test_bicgstab_static_equals_allocating(gpu, H):
    for n, geo, seed in ((3000, "sphere", 51), (1800, "cube", 52)):  # the second call is smaller
        h = shifted(H, n, geo, 8.0)
        b = np.random.default_rng(seed).uniform(-1, 1, h.nd)
        u0 = np.random.default_rng(seed + 1).uniform(-0.1, 0.1, h.nd)
        p = H.Plan(h)
        for max_iter, tol, every in ((12, 0.0, 5), (60, 1e-9, 3)):
            u_ref, r_ref = p.bicgstab(b, u0=u0, max_iter=max_iter, tol=tol, use_graph=False, check_every=every)
            rc, u, r = static_solve(H, p, b, u0, max_iter, tol, every)
            assert rc == 0, H.lib.hmvm_hip_last_error_message().decode()
            assert r["iters"] == r_ref["iters"] and r["hmvms"] == r_ref["hmvms"]
            assert r["relres"] == r_ref["relres"]
            assert np.array_equal(u, u_ref)
            if tol > 0:
                assert r["relres"] < tol and r["iters"] < max_iter
        p.close()
        h.free()


def test_bicgstab_static_rejects_too_many_unknowns(gpu, H):
    """A block-diagonal batch of 16,385 dense 64 x 64 blocks has nd =
    1,048,640 > HMVM_BICGSTAB_STATIC_MAX_N (2^20); the call returns HMVM_ERR_ARG
    before touching the device."""
    d = H.dummymatrix(32, 32, 2)  # nd = 64: four dense 32 x 32 leaves
    p = H.Plan([d] * 16385, storage="bf16")
    assert p.nd == 1048640
    rc, _, _ = static_solve(H, p, np.ones(p.nd), np.zeros(p.nd), 2, 0.0, 1)
    assert rc == H.HMVM_ERR_ARG
    assert "1048576" in H.lib.hmvm_hip_last_error_message().decode()
    p.close()
