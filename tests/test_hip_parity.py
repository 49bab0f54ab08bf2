"""GPU parity of the HIP HMVM against the CPU oracle (hmvm_seq restated).

Bar: FP64 y within 1e-13 relative l2 of the sequential
FP64 oracle; the FP32-stored form within 1e-6 of it.  Every test calls the
library through its C ABI (ctypes)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def sphere4096(H):
    h = H.generate(4096, geometry="sphere", eps=1e-4)
    h.build_aos()
    return h


@pytest.fixture(scope="module")
def cube3000(H):
    h = H.generate(3000, geometry="cube", eps=1e-6)
    h.build_aos()
    return h


@pytest.mark.parametrize("variant", ["1", "1t", "2", "2t"])
@pytest.mark.parametrize("mname", ["sphere4096", "cube3000"])
def test_variants_match_oracle(gpu, H, O, request, variant, mname):
    h = request.getfixturevalue(mname)
    rng = np.random.default_rng(11)
    b = rng.uniform(-1, 1, h.nd)
    ref = O.seq_2(h, b)
    v = np.zeros(h.nd)
    getattr(H, "hmvm_hip_" + variant)(v, h, b)
    assert rel(v, ref) <= 1e-13


def test_reference_input_vector(gpu, H, O, sphere4096):
    h = sphere4096
    b = np.sin(np.arange(1, h.nd + 1, dtype=np.float64))  # hmvm1.c:32-35
    ref = O.seq_2(h, b)
    v = np.zeros(h.nd)
    H.hmvm_hip_2(v, h, b)
    assert rel(v, ref) <= 1e-13


def test_accumulates(gpu, H, O, sphere4096):
    h = sphere4096
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, h.nd)
    y0 = rng.uniform(-1, 1, h.nd)
    v = y0.copy()
    H.hmvm_hip_2(v, h, b)
    ref = O.seq_2(h, b, y0.copy())
    assert rel(v, ref) <= 1e-13


def test_deterministic(gpu, H, sphere4096):
    h = sphere4096
    b = np.random.default_rng(5).uniform(-1, 1, h.nd)
    runs = []
    for _ in range(3):
        v = np.zeros(h.nd)
        H.hmvm_hip_2(v, h, b)
        runs.append(v)
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])


def test_linearity(gpu, H, cube3000):
    h = cube3000
    rng = np.random.default_rng(9)
    b1, b2 = rng.uniform(-1, 1, h.nd), rng.uniform(-1, 1, h.nd)
    y = [np.zeros(h.nd) for _ in range(3)]
    H.hmvm_hip_2(y[0], h, b1)
    H.hmvm_hip_2(y[1], h, b2)
    H.hmvm_hip_2(y[2], h, 2.0 * b1 - 3.0 * b2)
    assert rel(y[2], 2.0 * y[0] - 3.0 * y[1]) <= 1e-13


def test_all_dense_closed_form(gpu, H):
    """dummymatrix (loadmatrix.c:240): an all-dense grid equals a plain GEMV."""
    h = H.dummymatrix(24, 24, 9)
    b = np.random.default_rng(1).uniform(-1, 1, h.nd)
    A = np.zeros((h.nd, h.nd))
    m = h.mat
    for i in range(m.nlf):
        s = m.submat[i]
        blk = np.ctypeslib.as_array(s.a1, (s.ndl * s.ndt,)).reshape(s.ndl, s.ndt)
        A[s.nstrtl - 1:s.nstrtl - 1 + s.ndl, s.nstrtt - 1:s.nstrtt - 1 + s.ndt] = blk
    for fn in (H.hmvm_hip_1, H.hmvm_hip_2):
        v = np.zeros(h.nd)
        fn(v, h, b)
        assert rel(v, A @ b) <= 1e-13


@pytest.mark.parametrize("storage,tol", [("f32", 1e-6), ("bf16", 2e-2)])
def test_reduced_storage(gpu, H, O, sphere4096, storage, tol):
    h = sphere4096
    b = np.random.default_rng(2).uniform(-1, 1, h.nd)
    ref = O.seq_2(h, b)
    p = H.Plan(h, storage=storage)
    y = p.apply_host(b)
    err = rel(y, ref)
    print("storage", storage, "rel err vs FP64 oracle", err)
    assert err <= tol
    p.close()


@pytest.mark.parametrize("w", [64, 1000, 100000])
def test_tile_sizes(gpu, H, O, cube3000, w):
    h = cube3000
    b = np.random.default_rng(4).uniform(-1, 1, h.nd)
    p = H.Plan(h, w1=w, w2=w)
    assert rel(p.apply_host(b), O.seq_2(h, b)) <= 1e-13
    p.close()


def test_stage2_width_apart_from_dense(gpu, H, O, cube3000):
    """w3 sets the stage-2 tile width apart from w2 (dense tiles): launch A
    keeps its tiles, launch B gets others, and the product stays exact."""
    h = cube3000
    b = np.random.default_rng(5).uniform(-1, 1, h.nd)
    ref = O.seq_2(h, b)
    base = H.Plan(h, w1=4096, w2=4096)
    n0 = (base.info()["ntiles_a"], base.info()["ntiles_b"])
    base.close()
    for w3 in (256, 1 << 20):
        p = H.Plan(h, w1=4096, w2=4096, w3=w3)
        na, nb = p.info()["ntiles_a"], p.info()["ntiles_b"]
        assert na == n0[0] and (nb > n0[1] if w3 == 256 else nb < n0[1])
        assert rel(p.apply_host(b), ref) <= 1e-13
        p.close()


def test_plan_sources_agree(gpu, H, sphere4096):
    h = sphere4096
    b = np.random.default_rng(6).uniform(-1, 1, h.nd)
    ys = [H.Plan(h, source=s).apply_host(b) for s in ("rowmat", "rowmat_t", "a2", "a2t")]
    for y in ys[1:]:
        assert np.array_equal(y, ys[0])  # same device layout -> same bits


def test_golden_fixtures(gpu, H):
    """Committed matrices + y produced by the compiled reference hmvm_seq_2."""
    for name in sorted(os.listdir(GOLDEN)):
        if not name.endswith("_1.bin"):
            continue
        stem = name[:-6]
        h = H.loadHmatrix(os.path.join(GOLDEN, name))
        h.build_aos()
        g = np.load(os.path.join(GOLDEN, stem + "_y.npz"))
        for key in ("sin", "rand"):
            b, y = g["x_" + key], g["y_" + key]
            v = np.zeros(h.nd)
            H.hmvm_hip_2(v, h, b)
            assert rel(v, y) <= 1e-13, (stem, key)


def test_release_and_reupload(gpu, H, O):
    h = H.generate(1500, geometry="cube", eps=1e-4)
    b = np.random.default_rng(8).uniform(-1, 1, h.nd)
    v = np.zeros(h.nd)
    H.hmvm_hip_2(v, h, b)
    a = h.arrays()
    a["rowmat"][:] *= 2.0  # low-rank blocks scale by 4 (both factors), dense by 2
    a["rowmat_t"][:] *= 2.0
    w = np.zeros(h.nd)
    H.hmvm_hip_2(w, h, b)  # same struct, same pointers: the cached copy is used
    assert np.array_equal(w, v)
    H.release(h)  # contents changed: drop the cached device copy
    w = np.zeros(h.nd)
    H.hmvm_hip_2(w, h, b)
    assert rel(w, O.seq_2(h, b)) <= 1e-13
    assert rel(w, 2.0 * v) > 1e-3


def test_dump_and_bench(gpu, H, O, tmp_path, capfd, sphere4096):
    h = sphere4096
    b = np.sin(np.arange(1, h.nd + 1, dtype=np.float64))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        H.hmvm_hip(h, b, 1)
        H.hmvm_hip_bench(h, b)
    finally:
        os.chdir(cwd)
    ref = O.seq_2(h, b)
    for var in ("1", "1t", "2", "2t"):
        lines = open(tmp_path / ("result_hip_%s_d.txt" % var)).read().split()
        assert len(lines) == h.nd
        assert lines[0] == "%.3E" % ref[0]
    out = capfd.readouterr().out
    assert out.count("TIME 10 hmvm_hip_") == 4


def test_plan_time(gpu, H, sphere4096):
    p = H.Plan(sphere4096)
    t = p.time(warmup=2, iters=5)
    assert t["ms_avg"] > 0 and t["ms_a"] > 0 and t["ms_b"] > 0
    info = p.info()
    assert info["len"] == sphere4096.mat2.len
    p.close()


def test_bad_arguments_report_errors(gpu, H):
    h = H.generate(600, geometry="sphere")
    with pytest.raises(H.HMVMError):
        H.hmvm_hip_2(np.zeros(10), h, np.zeros(h.nd))
    with pytest.raises(H.HMVMError):
        H.Plan(h, storage=7)


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
def test_kernel_variants_bit_identical(gpu, H, cube3000, storage):
    """Every tile-kernel variant keeps the per-lane summation order (the
    staged vector entries of the grouped path included), so all of them (and
    every workgroup size and both tile orders) give the same bits."""
    h = cube3000
    b = np.random.default_rng(12).uniform(-1, 1, h.nd)
    p = H.Plan(h, storage=storage)
    y0 = p.apply_host(b)
    for k in range(8):
        for bt in (64, 128, 256):
            p.set_launch(k, bt)
            assert np.array_equal(p.apply_host(b), y0), (k, bt)
    p.close()
    q = H.Plan(h, storage=storage, order=1)
    assert np.array_equal(q.apply_host(b), y0)
    q.close()


def test_rank_rows_straddle_row_groups(gpu, H, O, tmp_path):
    """Leaves of rank 300 exceed one tile's 256 rank rows: stage 2 reads the
    materialised t (extra segment-sum launch)."""
    from test_schedule_replay import big_rank_matrix
    h = big_rank_matrix(H, str(tmp_path / "big.bin"))
    h.build_aos()
    b = np.random.default_rng(13).uniform(-1, 1, h.nd)
    ref = O.seq_2(h, b)
    for fn in (H.hmvm_hip_1, H.hmvm_hip_1t, H.hmvm_hip_2, H.hmvm_hip_2t):
        v = np.zeros(h.nd)
        fn(v, h, b)
        assert rel(v, ref) <= 1e-13
    for storage, tol in (("f32", 1e-5), ("bf16", 3e-2)):
        p = H.Plan(h, storage=storage)
        assert rel(p.apply_host(b), ref) <= tol
        p.close()


@pytest.mark.parametrize("t_mode", [0, 1, 3])
@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
def test_persistent_mode_matches_launches(gpu, H, sphere4096, storage, t_mode):
    """The single persistent launch (tiles dealt to the waves, grid barriers
    between the stages) computes the same bits as the launches, repeatedly
    (its counters reset themselves), with stage 2 summing the partial slots
    (t_mode 0), reading the t its own t-sum stage wrote (t_mode 1), or the t
    that the stage-1 tiles of the same launch wrote (t_mode 3)."""
    h = sphere4096
    b = np.random.default_rng(14).uniform(-1, 1, h.nd)
    p = H.Plan(h, storage=storage, t_mode=t_mode)
    assert (t_mode == 0) == bool(p.info()["t_fused"]) and (t_mode == 3) == bool(p.info()["t_direct"])
    y0 = p.apply_host(b)
    p.set_mode(1)
    assert p.info()["mode"] == 1  # the persistent launch is what runs
    for k in (0, 6):
        p.set_launch(k, 256)
        for _ in range(3):
            assert np.array_equal(p.apply_host(b), y0)
    y1 = y0.copy()
    p.apply_host(b, y1, accumulate=True)
    assert np.allclose(y1, 2 * y0, rtol=1e-14, atol=0)
    assert p.fault() == 0
    t = p.time(2, 5)
    assert t["ms_stream"] > 0 and p.fault() == 0
    p.close()


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 1e-6), ("bf16", 2e-2)])
def test_per_tile_paths(gpu, H, O, sphere4096, storage, tol):
    """Every tile of at most 32 lanes per column may run the lane-per-row path
    instead of the column-split one, chosen per tile: any mix gives the
    oracle's result to the storage's bar (the summation order inside a tile
    differs), the same bits on every repeat, and the flags and the path
    feature read back as set; tiles without the choice ignore the flag."""
    h = sphere4096
    b = np.random.default_rng(19).uniform(-1, 1, h.nd)
    ref = O.seq_2(h, b)
    p = H.Plan(h, storage=storage)
    y0 = p.apply_host(b)
    rng = np.random.default_rng(20)
    for which in (0, 1):
        f = p.tile_features(which)
        elig = f[:, 6] <= 5.0
        assert np.all(f[~elig, 17] == 1) and np.all(f[elig, 17] == 0)
    for mix in ("all", "random"):
        for which in (0, 1):
            n = p.info()["ntiles_a" if which == 0 else "ntiles_b"]
            fl = np.ones(n, np.int32) if mix == "all" else rng.integers(0, 2, n).astype(np.int32)
            p.set_tile_flags(which, fl)
            assert np.array_equal(p.tile_flags(which), fl)
            f = p.tile_features(which)
            elig = f[:, 6] <= 5.0
            assert np.array_equal(f[:, 17] == 1, ~elig | (fl == 1))
        y = p.apply_host(b)
        assert rel(y, ref) <= tol
        assert np.array_equal(p.apply_host(b), y)
    for which in (0, 1):
        p.set_tile_flags(which, np.zeros(p.info()["ntiles_a" if which == 0 else "ntiles_b"], np.int32))
    assert np.array_equal(p.apply_host(b), y0)
    with pytest.raises(H.HMVMError):
        p.set_tile_flags(0, np.full(p.info()["ntiles_a"], 2, np.int32))
    p.close()


def test_persistent_mode_default_plan(gpu, H, sphere4096):
    """Mode 1 at creation keeps the default t_mode (a materialised t, summed
    by a stage of the persistent launch) and gives the launches' bits; tile
    flags set after mode 1 is selected are dealt to the waves again."""
    h = sphere4096
    b = np.random.default_rng(17).uniform(-1, 1, h.nd)
    p = H.Plan(h)
    y0 = p.apply_host(b)
    p.close()
    p = H.Plan(h, mode=1)
    assert not p.info()["t_fused"] and not p.info()["t_direct"] and p.info()["mode"] == 1
    assert np.array_equal(p.apply_host(b), y0) and p.fault() == 0
    f = p.tile_features(0)
    p.set_tile_flags(0, (f[:, 6] <= 5.0).astype(np.int32))  # every eligible tile lane-per-row
    y1 = p.apply_host(b)
    assert rel(y1, y0) <= 1e-14 and p.fault() == 0
    p.set_mode(0)
    assert np.array_equal(p.apply_host(b), y1)  # the same tiles through the launches
    p.close()


def test_modes_regroup_launch_a(gpu, H, sphere4096):
    """The default launch A interleaves stage-1 and dense tiles (largest
    first); selecting mode 1 or 2 regroups it as stage-1 tiles, then dense,
    which those modes rely on, with the same bits."""
    h = sphere4096
    b = np.random.default_rng(16).uniform(-1, 1, h.nd)
    for mode, t_mode in ((2, None), (1, 0)):
        p = H.Plan(h, t_mode=t_mode)
        k = p.tile_features(0)[:, 0] > 0.5  # stage-1 tiles
        n1 = int(k.sum())
        assert 0 < n1 < len(k) and not np.all(k[:n1])
        y0 = p.apply_host(b)
        p.set_mode(mode)
        k = p.tile_features(0)[:, 0] > 0.5
        assert np.all(k[:n1]) and not np.any(k[n1:])
        assert np.array_equal(p.apply_host(b), y0)
        p.close()


@pytest.mark.parametrize("storage", ["f64", "bf16"])
def test_two_stream_overlap_matches_one_stream(gpu, H, sphere4096, storage):
    """Mode 2 runs the dense tiles on a second stream; mode 0 runs every launch
    on one stream: same tiles, same slots, same bits, also back to back and
    under graph replay."""
    h = sphere4096
    b = np.random.default_rng(15).uniform(-1, 1, h.nd)
    p = H.Plan(h, storage=storage)
    y0 = p.apply_host(b)
    p.set_mode(2)
    for _ in range(3):
        assert np.array_equal(p.apply_host(b), y0)
    t = p.time(2, 10)
    tg = p.time_graph(2, 2, 5)
    assert t["ms_stream"] > 0 and tg["ms_stream"] > 0
    assert np.array_equal(p.apply_host(b), y0)
    p.close()


_TALL = []


def tall_matrix(H):
    if not _TALL:
        _TALL.append(H.generate(12000, geometry="sphere", eps=1e-4))
    return _TALL[0]


@pytest.mark.parametrize("storage,max_rows,tol", [("f64", 383, 1e-13), ("f32", 765, 1e-5), ("bf16", 1529, 3e-2),
                                                  ("f32", 512, 1e-5), ("bf16", 1024, 3e-2)])
def test_tall_tiles(gpu, H, O, storage, max_rows, tol):
    """Tiles of up to 192 lanes x 16 bytes of rows (three row groups per lane)."""
    h = tall_matrix(H)  # row clusters of 753 rows: 2 or 3 row groups per lane
    x = np.random.default_rng(16).uniform(-1, 1, h.nd)
    p = H.Plan(h, storage=storage, max_rows=max_rows, w1=1 << 18, w2=1 << 18)
    y = p.apply_host(x)
    s = H.schedule(h, storage=storage, max_rows=max_rows, w1=1 << 18, w2=1 << 18)
    assert max(s["tiles_a"]["nrows"].max(), s["tiles_b"]["nrows"].max()) > 256
    assert rel(y, O.seq_2(h, x)) <= tol
    for k in (0, 6, 7):
        p.set_launch(k, 256)
        assert np.array_equal(p.apply_host(x), y)
    p.close()


def test_time_flushed(gpu, H, cube3000):
    """Per-apply timing with the caches flushed first: positive times, ordered
    statistics, and the plan's results untouched."""
    b = np.random.default_rng(14).uniform(-1, 1, cube3000.nd)
    p = H.Plan(cube3000)
    y0 = p.apply_host(b)
    t = p.time_flushed(iters=5, flush_mb=64)
    assert t["iters"] == 5 and 0 < t["ms_min"] <= t["ms_median"] and t["ms_avg"] > 0
    assert np.array_equal(p.apply_host(b), y0)
    p.close()


@pytest.mark.parametrize("storage,max_rows,max_rows_b,tol", [("f64", 256, 64, 1e-13), ("f64", 128, 383, 1e-13),
                                                             ("f32", 0, 96, 1e-5), ("bf16", 512, 200, 3e-2)])
def test_stage2_tile_height(gpu, H, O, storage, max_rows, max_rows_b, tol):
    """Stage-2 tiles cut to their own height (max_rows_b), independent of the
    launch-A tiles."""
    h = tall_matrix(H)
    x = np.random.default_rng(17).uniform(-1, 1, h.nd)
    p = H.Plan(h, storage=storage, max_rows=max_rows, max_rows_b=max_rows_b)
    s = H.schedule(h, storage=storage, max_rows=max_rows, max_rows_b=max_rows_b)
    assert s["tiles_b"]["nrows"].max() <= max_rows_b
    assert rel(p.apply_host(x), O.seq_2(h, x)) <= tol
    p.close()
