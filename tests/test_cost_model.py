"""The learned cost model: HIP forward/backward/Adam kernels against the
float64 numpy oracle (oracle/cost_mlp.py), and the oracle against finite
differences of its own loss."""
import numpy as np
import pytest

from oracle import cost_mlp

NIN, NH = 15, 64


def rand_problem(n=700, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, NIN)).astype(np.float32) * rng.uniform(0.5, 3, NIN).astype(np.float32) + 2.0
    w = rng.normal(size=NIN)
    y = (np.tanh(X @ w / 6.0) + 0.1 * (X[:, 0] > 2)).astype(np.float32)
    return X, y


def rand_params(seed=1):
    rng = np.random.default_rng(seed)
    n = NH * NIN + NH + NH * NH + NH + NH + 1
    return (rng.normal(size=n) * 0.2).astype(np.float32)


# A 3 -> 2 -> 2 -> 1 network with hand-set weights and outputs worked out by
# hand: standardisation (x - mean)/stdev, ReLU after both hidden layers, W
# row-major as (out, in).  W2 is not symmetric and several units are cut by
# the ReLU, so a transposed weight or a missing/wrong activation changes the
# outputs (with W2 transposed the first output is 0.25, without ReLU 14.75).
HAND = {
    "nin": 3, "nh": 2,
    "params": [1, 0, 2, 0, -1, 1,   # W1 (2 x 3)
               0.5, -1,             # b1
               1, 2, -3, 1,         # W2 (2 x 2)
               0, 0.25,             # b2
               2, -1,               # w3
               0.5],                # b3
    "mean": [1, 0, -1], "stdev": [2, 1, 0.5],
    "X": [[3, 1, -0.5], [1, -2, -1], [-1, 3, -2], [1, 0, -1]],
    # x_s = (1,1,1): h1 = relu(3.5, -1) = (3.5, 0), h2 = relu(3.5, -10.25), out = 2*3.5 + 0.5
    # x_s = (0,-2,0): h1 = (0.5, 1), h2 = relu(2.5, -0.25), out = 5 + 0.5
    # x_s = (-1,3,-2): h1 = relu(-4.5, -6) = 0, h2 = (0, 0.25), out = -0.25 + 0.5
    # x_s = 0: h1 = (0.5, 0), h2 = relu(0.5, -1.25), out = 1 + 0.5
    "out": [7.5, 5.5, 0.25, 1.5],
}


def test_oracle_forward_closed_form():
    h = HAND
    out, _ = cost_mlp.forward(np.array(h["params"], float), h["mean"], h["stdev"], np.array(h["X"], float),
                              h["nin"], h["nh"])
    assert np.array_equal(out, np.array(h["out"]))


@pytest.mark.gpu
def test_hip_forward_closed_form(gpu, H):
    h = HAND
    m = H.CostMLP(h["nin"], h["nh"], seed=1)
    m.set(params=np.array(h["params"], np.float32), mean=np.array(h["mean"], np.float32),
          stdev=np.array(h["stdev"], np.float32))
    pred = m.predict(np.array(h["X"], np.float32))
    assert np.allclose(pred, h["out"], rtol=0, atol=1e-6)
    m.close()


def test_oracle_gradient_matches_finite_differences():
    X, y = rand_problem(60, seed=3)
    X = X.astype(np.float64)
    p = rand_params(4).astype(np.float64)
    mu, sd = X.mean(axis=0), X.std(axis=0)
    loss, g = cost_mlp.loss_grad(p, mu, sd, X, y.astype(np.float64), NIN, NH)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(NH * NIN, 20, replace=False),                 # W1
                          NH * NIN + rng.choice(NH, 5, replace=False),             # b1
                          NH * NIN + NH + rng.choice(NH * NH, 20, replace=False),  # W2
                          [p.size - 1 - NH - NH + 3, p.size - 1 - 5, p.size - 1]])  # b2, w3, b3
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = cost_mlp.loss_grad(p + e, mu, sd, X, y, NIN, NH)
        lm, _ = cost_mlp.loss_grad(p - e, mu, sd, X, y, NIN, NH)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


@pytest.mark.gpu
def test_hip_forward_and_gradient_match_oracle(gpu, H):
    X, y = rand_problem(700, seed=6)   # 700 samples: three partial chunks of 256
    m = H.CostMLP(NIN, NH, seed=9)
    p = rand_params(7)
    m.fit_scaler(X)
    _, mu, sd = m.get()
    assert np.allclose(mu, X.astype(np.float64).mean(axis=0), rtol=1e-5, atol=1e-5)
    m.set(params=p)
    pred = m.predict(X)
    ref, _ = cost_mlp.forward(p, mu, sd, X, NIN, NH)
    assert np.max(np.abs(pred - ref)) <= 1e-4 * max(1.0, np.max(np.abs(ref)))
    loss, g = m.loss_grad(X, y)
    loss_r, g_r = cost_mlp.loss_grad(p, mu, sd, X, y, NIN, NH)
    assert abs(loss - loss_r) <= 1e-5 * max(loss_r, 1e-6)
    assert np.linalg.norm(g - g_r) <= 1e-4 * np.linalg.norm(g_r)
    # deterministic: the same gradient bit for bit
    _, g2 = m.loss_grad(X, y)
    assert np.array_equal(g, g2)
    m.close()


@pytest.mark.gpu
def test_hip_training_fits(gpu, H):
    X, y = rand_problem(4000, seed=8)
    m = H.CostMLP(NIN, NH, seed=2)
    m.fit_scaler(X)
    l0, _ = m.loss_grad(X, y)
    lf = m.train(X, y, epochs=60, batch=256, lr=2e-3, seed=3)
    assert lf < 0.05 * l0
    p, mu, sd = m.get()
    ref, _ = cost_mlp.forward(p, mu, sd, X, NIN, NH)
    assert np.max(np.abs(m.predict(X) - ref)) <= 1e-4 * max(1.0, np.max(np.abs(ref)))
    m.close()


@pytest.mark.gpu
def test_hip_training_throughput(gpu, H):
    """A training pass costs two launches per minibatch with the work spread
    over one workgroup per 64 samples: 20 epochs over 200,000 samples
    (batch 4096) in well under the time of a benchmark run, and the loss
    drops."""
    import time
    X, y = rand_problem(200000, seed=10)
    m = H.CostMLP(NIN, NH, seed=4)
    m.fit_scaler(X)
    l0, _ = m.loss_grad(X, y)
    t0 = time.time()
    lf = m.train(X, y, epochs=20, batch=4096, lr=3e-3, seed=5)
    dt = time.time() - t0
    print("20 epochs x 200k samples: %.2f s, loss %.4f -> %.4f" % (dt, l0, lf))
    assert lf < 0.2 * l0
    assert dt < 10.0
    m.close()


def test_features_clamped_to_training_range(H):
    """Launch-level features beyond every training launch (a large batched
    plan) are clamped to the training range before prediction."""
    from hacapk_hmvm_amd import autotune as AT
    m = AT.CostModel.__new__(AT.CostModel)  # the clamp needs no network (nor a GPU)
    X = np.array([[0.0] * H.HMVM_NFEAT, [1.0] * H.HMVM_NFEAT], dtype=np.float32)
    X[1, 15] = 14.0
    m.lo, m.hi = X.min(axis=0), X.max(axis=0)
    Z = np.full((1, H.HMVM_NFEAT), 0.5, dtype=np.float32)
    Z[0, 15] = 18.0
    Z[0, 3] = -1.0
    c = m.clip(Z)
    assert c.dtype == np.float32
    assert c[0, 15] == 14.0 and c[0, 3] == 0.0 and c[0, 4] == 0.5
    lin = AT.linear_baseline(np.vstack([X, X + 0.25]).astype(np.float32), np.array([0.0, 1.0, 0.5, 1.5]))
    z2 = c.copy()
    z2[0, 15] = 14.25  # the linear fit's own training maximum of that feature
    assert np.allclose(lin(Z), lin(z2))
