"""Autotuner and learned cost model on generated H-matrices (GPU).

The network is trained on tile durations measured on two matrices and judged
on a third it has not seen: its duration predictions must beat the method
without it (a least-squares fit of the same features), and the schedule it
picks from predictions alone must run about as fast as the one the exhaustive
search finds by timing every candidate.
"""
import numpy as np
import pytest

from hacapk_hmvm_amd import autotune as AT


@pytest.fixture(scope="module")
def train_mats(H):
    return [H.generate(12000, geometry="sphere", eps=1e-4), H.generate(16000, geometry="cube", eps=1e-6)]


@pytest.fixture(scope="module")
def test_mat(H):
    return H.generate(14000, geometry="cube", eps=1e-4, seed=99)


@pytest.fixture(scope="module")
def big_mat(H):
    """Large enough (FP32 0.7 GB, more than twice the 256 MB Infinity Cache)
    that a schedule's time is stable to about a percent."""
    return H.generate(70000, geometry="cube", eps=1e-4, seed=98)


@pytest.mark.gpu
def test_schedule_features_match_plan(gpu, H, test_mat):
    for storage in ("f64", "bf16"):
        w1, w2 = AT.widths(storage, 1.0, 0.5)
        p = H.Plan(test_mat, storage=storage, w1=w1, w2=w2)
        p.set_phase_launch(0, 6, 128)
        p.set_phase_launch(1, 2, 64)
        sched = H.schedule(test_mat, storage=storage, w1=w1, w2=w2)
        for which, (k, bt) in ((0, (6, 128)), (1, (2, 64))):
            a = p.tile_features(which)
            b = AT.schedule_features(sched, which, storage, k, bt)
            assert a.shape == b.shape
            assert np.allclose(a, b, rtol=1e-6, atol=1e-6)
        p.close()


@pytest.mark.gpu
def test_learned_model_beats_baseline_and_tracks_exhaustive(gpu, H, train_mats, test_mat, big_mat):
    storage = "f32"
    Xs, ys, recs = [], [], []
    for i, m in enumerate(train_mats):
        X, y, r = AT.collect(m, storage, w_scales=((1.0, 1.0), (0.5, 2.0)), iters=2, max_tiles=8000, seed=i)
        Xs.append(X)
        ys.append(y)
        recs += r
    X, y = np.concatenate(Xs), np.concatenate(ys)
    model = AT.CostModel(seed=5)
    model.fit(X, y, recs, epochs=60)
    Xt, yt, _ = AT.collect(test_mat, storage, iters=2, max_tiles=8000, seed=7)
    err_mlp = np.mean(np.abs(model.predict_log2(Xt) - yt))
    err_lin = np.mean(np.abs(AT.linear_baseline(X, y)(Xt) - yt))
    print("held-out mean |log2 error|: mlp %.3f  linear %.3f" % (err_mlp, err_lin))
    assert err_mlp < err_lin
    # schedule quality on the large matrix: the learned schedule within 5% of
    # the exhaustive one, both timed afresh, interleaved, median of five
    best, _, _ = AT.exhaustive(big_mat, storage, warmup=2, iters=6)
    p_ex = H.Plan(big_mat, storage=storage, w1=best["w1"], w2=best["w2"])
    AT._apply_choice(p_ex, best)
    choice, p_l, _ = AT.learned(big_mat, storage, model)
    t_ex, t_l = [], []
    for _ in range(5):
        t_ex.append(p_ex.time(3, 20)["ms_stream"])
        t_l.append(p_l.time(3, 20)["ms_stream"])
    p_ex.close()
    p_l.close()
    ms_ex, ms_learned = float(np.median(t_ex)), float(np.median(t_l))
    print("exhaustive %.4f ms, learned %.4f ms (%s)" % (ms_ex, ms_learned, choice))
    assert ms_learned <= 1.05 * ms_ex
