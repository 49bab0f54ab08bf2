"""Batched plans, hipGraph replays and schedule knobs, against the oracle.

A batch plan places several matrices on the diagonal of one block-diagonal
matrix; each segment of y must equal that matrix's own product (and the
oracle's, hmvm_seq_2 of src_c/hmvm_seq.c:113).  Graph replays, tile
reordering, per-launch kernel variants and the stage-2 vector modes must not
change a single bit of y, or stay within the FP64 tolerance of the oracle
where the summation order may differ.
"""
import numpy as np
import pytest


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def mats(H):
    return [H.generate(900, geometry="sphere", eps=1e-4),
            H.generate(1400, geometry="cube", eps=1e-6),
            H.generate(2100, geometry="sphere", eps=1e-6, leaf=24)]


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_batch_equals_individual(gpu, H, O, mats, storage):
    # Default tile widths follow the element count, which differs between the
    # batch and its members, so bit identity is checked at equal explicit
    # widths; the default-width batch is checked against the oracle.
    x = np.random.default_rng(11).uniform(-1, 1, 4400)
    pd = H.Plan(mats, storage=storage)
    yd = pd.apply_host(x)
    pd.close()
    pb = H.Plan(mats, storage=storage, w1=2048, w2=1024)
    offs = pb.row_offsets
    assert list(offs) == [0, 900, 2300, 4400]
    y = pb.apply_host(x)
    for q, m in enumerate(mats):
        xs = np.ascontiguousarray(x[offs[q]:offs[q + 1]])
        p = H.Plan(m, storage=storage, w1=2048, w2=1024)
        yq = p.apply_host(xs)
        p.close()
        assert np.array_equal(y[offs[q]:offs[q + 1]], yq)
        ref = O.seq_2(m, xs)
        tol = 1e-13 if storage == "f64" else 1e-5
        assert rel(yq, ref) <= tol
        assert rel(yd[offs[q]:offs[q + 1]], ref) <= tol
    pb.close()


@pytest.mark.gpu
def test_graph_replays_equal_eager(gpu, H, mats):
    m = mats[1]
    p = H.Plan(m)
    x = np.random.default_rng(12).uniform(-1, 1, m.nd)
    y_eager = p.apply_host(x)
    dx = H.DeviceBuffer(8 * m.nd)
    dy = H.DeviceBuffer(8 * m.nd)
    dx.upload(x)
    g = p.graph(dx.value, dy.value)
    for r in range(4):
        # eager work and a host synchronisation between replays
        assert np.array_equal(p.apply_host(x), y_eager)
        dy.upload(np.full(m.nd, np.nan))
        g.launch()
        H.lib.hmvm_device_synchronize()
        assert np.array_equal(dy.download(np.empty(m.nd)), y_eager), "replay %d" % r
    # accumulate form: 3 applies in one graph add 3 A x to y
    g3 = p.graph(dx.value, dy.value, accumulate=True, count=3)
    y0 = np.random.default_rng(13).uniform(-1, 1, m.nd)
    dy.upload(y0)
    g3.launch()
    H.lib.hmvm_device_synchronize()
    y3 = dy.download(np.empty(m.nd))
    assert rel(y3, y0 + 3 * y_eager) <= 1e-14
    g.close()
    g3.close()
    t = p.time_graph(count=2, warmup=1, iters=3)
    assert t["ms_avg"] > 0 and t["ms_stream"] > 0
    p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f64", "bf16"])
def test_batch_persistent_graph_equals_launches(gpu, H, mats, storage):
    """A batch plan in the persistent mode (one launch, grid barriers) gives
    the bits of the four launches, eagerly and as graph replays with eager
    work in between, and its fault word stays clear."""
    x = np.random.default_rng(18).uniform(-1, 1, 4400)
    p = H.Plan(mats, storage=storage)
    y0 = p.apply_host(x)
    p.set_mode(1)
    assert p.info()["mode"] == 1
    assert np.array_equal(p.apply_host(x), y0)
    dx, dy = H.DeviceBuffer(8 * 4400), H.DeviceBuffer(8 * 4400)
    dx.upload(x)
    g = p.graph(dx.value, dy.value)
    for r in range(3):
        dy.upload(np.full(4400, np.nan))
        g.launch()
        H.lib.hmvm_device_synchronize()
        assert np.array_equal(dy.download(np.empty(4400)), y0), "replay %d" % r
        assert np.array_equal(p.apply_host(x), y0)
    assert p.fault() == 0
    g.close()
    dx.close()
    dy.close()
    p.close()


@pytest.mark.gpu
def test_tile_order_and_phase_launch_keep_bits(gpu, H, mats):
    m = mats[2]
    p = H.Plan(m)
    x = np.random.default_rng(14).uniform(-1, 1, m.nd)
    y0 = p.apply_host(x)
    info = p.info()
    rng = np.random.default_rng(15)
    p.set_tile_order(0, rng.permutation(info["ntiles_a"]))
    p.set_tile_order(1, rng.permutation(info["ntiles_b"]))
    assert np.array_equal(p.apply_host(x), y0)
    for k in range(8):
        p.set_phase_launch(0, k, 64 * (1 + k % 4))
        p.set_phase_launch(1, 7 - k, 64 * (1 + (k + 1) % 4))
        assert np.array_equal(p.apply_host(x), y0), k
    with pytest.raises(H.HMVMError):
        p.set_tile_order(0, np.zeros(info["ntiles_a"], dtype=np.int32))
    p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("t_mode", [0, 1, 2, 3])
def test_t_modes(gpu, H, O, mats, t_mode):
    m = mats[2]
    x = np.random.default_rng(16).uniform(-1, 1, m.nd)
    p = H.Plan(m, t_mode=t_mode)
    y = p.apply_host(x)
    ref = O.seq_2(m, x)
    assert rel(y, ref) <= 1e-13
    if t_mode == 1:
        assert not p.info()["t_fused"]
    if t_mode == 3:  # t written by the stage-1 tiles: no partial slots
        assert p.info()["t_direct"] and p.info()["n_tslots"] == 0
    # the same t mode is bit-reproducible; modes differ in how panels merge
    assert np.array_equal(p.apply_host(x), y)
    p.close()


@pytest.mark.gpu
def test_profile_tiles(gpu, H, mats):
    m = mats[0]
    p = H.Plan(m)
    x = np.random.default_rng(17).uniform(-1, 1, m.nd)
    y0 = p.apply_host(x)
    for which in (0, 1):
        d, s0, lm = p.profile_tiles(which, iters=2)
        n = p.info()["ntiles_a" if which == 0 else "ntiles_b"]
        assert d.shape == (n,) and np.all(d > 0) and np.all(s0 >= 0) and lm > 0
        assert d.max() <= 1.5 * lm * 1e3 + 50.0  # no tile outlasts its launch (us)
        f = p.tile_features(which)
        assert f.shape == (n, H.HMVM_NFEAT) and np.all(np.isfinite(f))
        kind = f[:, 0:3].argmax(axis=1)
        assert np.all(kind == 2) if which == 1 else np.all(kind < 2)
    assert np.array_equal(p.apply_host(x), y0)
    p.close()
