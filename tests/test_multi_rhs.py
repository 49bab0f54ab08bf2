"""Several right-hand sides on the matrix cores (MFMA): each column of A X
must equal the oracle's A x (FP64 storage to 1e-13; FP32 storage to the FP32
bar; BF16 factors and BF16-rounded vectors to a few 1e-3)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def hm(H):
    return H.generate(3000, geometry="cube", eps=1e-4)


@pytest.fixture(scope="module")
def XY(O, hm):
    X = np.random.default_rng(31).uniform(-1, 1, (hm.nd, 32))
    Y = np.stack([O.seq_2(hm, np.ascontiguousarray(X[:, c])) for c in range(32)], axis=1)
    return X, Y


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 2e-6), ("bf16", 2e-2)])
def test_multi_rhs_matches_oracle(gpu, H, hm, XY, storage, tol):
    X, Y = XY
    p = H.Plan(hm, storage=storage)
    got = p.apply_multi_host(X)
    errs = [rel(got[:, c], Y[:, c]) for c in range(32)]
    print(storage, "max column rel err", max(errs))
    assert max(errs) <= tol
    # 16 and 64 columns, and accumulation
    got16 = p.apply_multi_host(np.ascontiguousarray(X[:, :16]))
    assert max(rel(got16[:, c], Y[:, c]) for c in range(16)) <= tol
    X64 = np.concatenate([X, X], axis=1)
    got64 = p.apply_multi_host(X64)
    assert np.array_equal(got64[:, :32], got64[:, 32:])
    acc = got.copy()
    p.apply_multi_host(X, acc, accumulate=True)
    assert rel(acc, 2 * got) <= 1e-14
    p.close()


def test_multi_rhs_equals_single_rhs_f64(gpu, H, hm, XY):
    X, _ = XY
    p = H.Plan(hm)
    got = p.apply_multi_host(X)
    for c in (0, 7, 31):
        y1 = p.apply_host(np.ascontiguousarray(X[:, c]))
        assert rel(got[:, c], y1) <= 1e-14
    t = p.time_multi(32, 1, 3)
    assert t["ms_stream"] > 0
    p.close()


def test_multi_rhs_rank_rows_straddle(gpu, H, O, tmp_path):
    """The non-fused path (materialised t) with several right-hand sides."""
    from test_schedule_replay import big_rank_matrix
    h = big_rank_matrix(H, str(tmp_path / "big.bin"))
    X = np.random.default_rng(32).uniform(-1, 1, (h.nd, 16))
    p = H.Plan(h)
    got = p.apply_multi_host(X)
    for c in (0, 5, 15):
        assert rel(got[:, c], O.seq_2(h, np.ascontiguousarray(X[:, c]))) <= 1e-13
    p.close()


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 2e-6), ("bf16", 2e-2)])
@pytest.mark.parametrize("t_mode", [0, 2, 3])
def test_multi_rhs_partial_slot_panels(gpu, H, hm, XY, storage, tol, t_mode):
    """Stage-2 panels that sum stage-1 partial slots (t_mode 0 and 2) take the
    kernel instantiation with the summing loop."""
    X, Y = XY
    p = H.Plan(hm, storage=storage, t_mode=t_mode)
    got = p.apply_multi_host(X)
    assert max(rel(got[:, c], Y[:, c]) for c in range(32)) <= tol
    got16 = p.apply_multi_host(np.ascontiguousarray(X[:, :16]))
    assert max(rel(got16[:, c], Y[:, c]) for c in range(16)) <= tol
    p.close()


def test_multi_rhs_device_pointers_any_alignment(gpu, H, hm, XY):
    """hmvm_plan_apply_multi on caller device buffers: a Y that is only
    8-byte aligned takes the one-column y sum, an X that is only 8-byte
    aligned is copied before the 16-byte operand loads, and every case equals
    the host entry point bit for bit (same sums in the same order).  Device
    buffers come from the HIP runtime the library itself links."""
    import ctypes
    with open("/proc/self/maps") as f:  # the HIP runtime the library loaded (not torch's copy)
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln and "torch" not in ln}
    assert paths, "the HIP runtime is not loaded"
    hip = ctypes.CDLL(sorted(paths)[0])
    X, _ = XY
    n = hm.nd * 32
    dxbuf, dbuf = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dxbuf), ctypes.c_size_t(8 * (n + 1))) == 0
    assert hip.hipMalloc(ctypes.byref(dbuf), ctypes.c_size_t(8 * (n + 1))) == 0
    try:
        Xc = np.ascontiguousarray(X)
        for storage in ("f64", "f32"):
            p = H.Plan(hm, storage=storage)
            want = p.apply_multi_host(X)
            assert H.lib.hmvm_plan_reserve_multi(p._p, 32) == 0
            for xoff in (0, 1):
                dX = ctypes.c_void_p(dxbuf.value + 8 * xoff)
                assert hip.hipMemcpy(dX, Xc.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * n), 1) == 0  # H2D
                for off in (0, 1):
                    dY = ctypes.c_void_p(dbuf.value + 8 * off)
                    assert (dY.value % 16 == 0) == (off == 0) and (dX.value % 16 == 0) == (xoff == 0)
                    assert H.lib.hmvm_plan_apply_multi(p._p, dX, dY, 32, 0, None) == 0
                    assert hip.hipDeviceSynchronize() == 0
                    got = np.empty((hm.nd, 32))
                    assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), dY, ctypes.c_size_t(8 * n), 2) == 0
                    assert np.array_equal(got, want), (storage, xoff, off)
            p.close()
    finally:
        hip.hipFree(dxbuf)
        hip.hipFree(dbuf)
