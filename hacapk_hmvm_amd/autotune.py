"""Autotuner and learned cost model for the HIP HMVM schedule.

The reference picks its GPU kernel variant offline: a sweep over 96 variant
ids prints timings (src_cpp/hmvm1.cpp:82-96, hmvm_cuda1.cu:236-302) and a human
reads the CSV.  Here the schedule of a plan has these knobs:

  w1, w2          tile widths (target elements per stage-1 / dense+stage-2 tile)
  kernel_a/b      tile-kernel variant of launch A / launch B (0..7)
  bt_a/b          workgroup size of launch A / launch B (64, 128, 256)
  path, per tile  column-split (lane groups take different columns) or
                  lane-per-row (lanes own row groups, read whole columns),
                  for every tile of at most 32 lanes per column
  order           dispatch order of the tiles of each launch

Two ways to choose them:

exhaustive(...)  builds a plan for every (w1, w2), times every (kernel, bt)
                 and both uniform paths (every eligible tile column-split, or
                 every one lane-per-row) for each launch on the GPU and keeps
                 the fastest of each; this is the ground truth and the
                 fallback.  It cannot try per-tile mixes (2^tiles of them).
learned(...)     builds the candidate schedules on the host only, predicts
                 every tile's duration with the MLP (HIP inference kernel),
                 replays the hardware's greedy dispatch of the tiles (longest
                 predicted first) onto the kernel's resident wavefront slots
                 and picks the candidate with the smallest predicted launch
                 times; then it gives every eligible tile the path predicted
                 faster for it.  Only the chosen plans are uploaded.

The MLP is trained on tile durations measured in situ (a clock-stamping copy
of the tile kernel) on generated H-matrices, by HIP forward/backward/Adam
kernels.  Its target is log2 of the tile duration in microseconds.
"""
import heapq
import itertools
import time

import numpy as np

from . import CostMLP, HMVM_NFEAT, Plan, schedule, tile_occupancy

W1_CHOICES = (0.5, 1.0, 2.0)     # multiples of the base stage-1 tile width (32768 elements)
W2_CHOICES = (0.5, 1.0, 2.0)     # multiples of the base stage-2 tile width (16384 elements)
KERNELS = (2, 4, 6, 7)           # pipelined / non-temporal / batched-entry variants
BLOCK_THREADS = (64, 256)
EB = {"f64": 8, "f32": 4, "bf16": 2}


def widths(storage, a, b, length=None):
    """Tile widths (elements) for multiples a, b of the base widths.  Given the
    matrix's element count, the bases shrink as the library's default widths do
    for small matrices (to length / 4096, at least 1024), so that the candidates
    stay around one tile per resident wavefront slot."""
    b1, b2 = 32768, 16384
    if length is not None:
        fit = max(1024, int(length) // 4096)
        b1, b2 = min(b1, fit), min(b2, fit)
    return int(a * b1), int(b * b2)


def _length(h):
    m = getattr(h, "mat2", None)
    return getattr(m, "len", None)


# ---- features from an exported schedule (host only) ------------------------------
def schedule_features(sched, which, storage, kernel, block_threads, nrhs=1, row_per_lane=None):
    """Tile features of launch A (0) or B (1) of an exported schedule, as
    hmvm_plan_tile_features computes them for an uploaded plan.
    row_per_lane: per-tile path override (bool array; None: the schedule's
    tile flags)."""
    tiles = sched["tiles_a"] if which == 0 else sched["tiles_b"]
    pan = sched["panels"]
    V = 16 // EB[storage]
    n = len(tiles)
    f = np.zeros((n, HMVM_NFEAT), dtype=np.float32)
    if n == 0:
        return f
    p0 = tiles["panel0"].astype(np.int64)
    npan = tiles["npanel"].astype(np.int64)
    # per-tile sums over panels with a prefix sum over the panel array
    ncols = pan["ncols"].astype(np.float64)
    nch = np.where(pan["vec_sel"] == 2, (pan["vec_chunks"] >> 16).astype(np.float64) * ncols, 0.0)
    cc = np.concatenate([[0.0], np.cumsum(ncols)])
    cn = np.concatenate([[0.0], np.cumsum(nch)])
    cols = cc[p0 + npan] - cc[p0]
    chs = cn[p0 + npan] - cn[p0]
    shift = (pan["val_off"][p0] % V).astype(np.int64)
    nrows = tiles["nrows"].astype(np.int64)
    lpc = (nrows + shift + V - 1) // V
    kind = np.where(which == 1, 2, np.where(pan["factor"][p0] < 0, 0, 1))
    f[np.arange(n), kind] = 1.0
    f[:, 3] = np.log2(nrows)
    f[:, 4] = np.log2(1.0 + cols)
    f[:, 5] = np.log2(npan)
    f[:, 6] = np.log2(lpc)
    f[:, 7] = np.log2(1.0 + lpc * V * cols)
    f[:, 8] = np.log2(1.0 + np.where(cols > 0, chs / np.maximum(cols, 1), 0.0))
    f[:, 9] = np.log2(EB[storage])
    f[:, 10] = kernel & 1
    f[:, 11] = (kernel >> 1) & 1
    f[:, 12] = (kernel >> 2) & 1
    f[:, 13] = np.log2(block_threads // 64)
    f[:, 14] = np.log2(max(nrhs, 1))
    slots = max(occupancy(storage, kernel, block_threads), 1)
    f[:, 15] = np.log2(min(n, slots))  # saturation of the launch
    f[:, 16] = np.log2(slots)
    flag = (tiles["flags"] & 1) != 0 if row_per_lane is None else np.asarray(row_per_lane, dtype=bool)
    f[:, 17] = ((lpc > 32) | flag).astype(np.float32)
    return f


def eligible(X):
    """Tiles with the choice of path (at most 32 lanes per column), from features."""
    return X[:, 6] <= 5.0 + 1e-6


def path_flags(model, X):
    """Per-tile flags: HMVM_TILE_ROW_PER_LANE where the model predicts the
    lane-per-row path faster than the column-split one."""
    e = eligible(X)
    X0, X1 = X.copy(), X.copy()
    X0[:, 17] = np.where(e, 0.0, 1.0)
    X1[:, 17] = 1.0
    d0, d1 = model.tile_us(X0), model.tile_us(X1)
    return (e & (d1 < d0)).astype(np.int32)


# ---- training data -------------------------------------------------------------------
def collect(h, storage, variants=None, w_scales=((1.0, 1.0),), iters=2, max_tiles=20000, seed=0):
    """Measured tile durations of one matrix: (features, log2 us, per-launch
    records).  Each record holds the variant, the launch's features in
    dispatch order and its measured time, against which the launch-time model
    is calibrated."""
    rng = np.random.default_rng(seed)
    variants = variants or [(k, bt) for k in KERNELS for bt in BLOCK_THREADS]
    Xs, ys, recs = [], [], []
    for a, b in w_scales:
        # training tiles use the fixed base widths whatever the matrix size,
        # so that small matrices still contribute full-size tiles
        w1, w2 = widths(storage, a, b)
        p = Plan(h, storage=storage, w1=w1, w2=w2)
        for k, bt in variants:
            p.set_launch(k, bt)
            for which in (0, 1):
                # both paths: the default, then every eligible tile lane-per-row
                for rpl in (False, True):
                    f = p.tile_features(which)
                    if len(f) == 0 or (rpl and not eligible(f).any()):
                        continue
                    p.set_tile_flags(which, eligible(f).astype(np.int32) if rpl else np.zeros(len(f), np.int32))
                    d, _, lms = p.profile_tiles(which, iters=iters)
                    f = p.tile_features(which)
                    recs.append({"storage": storage, "which": which, "kernel": k, "bt": bt, "w1": w1, "w2": w2,
                                 "row_per_lane": rpl, "sum_us": float(d.sum()), "launch_ms": lms, "ntiles": len(d),
                                 "X": f})
                    sel = rng.choice(len(d), min(len(d), max_tiles), replace=False)
                    Xs.append(f[sel])
                    ys.append(np.log2(np.maximum(d[sel], 1e-2)).astype(np.float32))
                p.set_tile_flags(which, np.zeros(len(f), np.int32))
        p.close()
    return np.concatenate(Xs), np.concatenate(ys), recs


def makespan(d, slots):
    """Greedy list schedule of durations d (dispatch order) onto `slots`
    identical wavefront slots: each tile starts on the slot that frees first."""
    d = np.asarray(d, dtype=np.float64)
    if len(d) == 0:
        return 0.0
    if len(d) <= slots:
        return float(d.max())
    h = list(d[:slots])
    heapq.heapify(h)
    for x in d[slots:]:
        heapq.heapreplace(h, h[0] + x)
    return float(max(h))


_OCC = {}


def occupancy(storage, kernel, bt):
    key = (storage, kernel, bt)
    if key not in _OCC:
        _OCC[key] = tile_occupancy(storage, kernel, bt)
    return _OCC[key]


class CostModel:
    """MLP on tile features -> log2(tile duration us), and a launch-time model:
    the greedy dispatch of the predicted durations onto the kernel's resident
    wavefront slots, calibrated (ms = a * makespan + b per storage) against the
    measured launch times of the training runs."""

    def __init__(self, nh=64, seed=1):
        self.mlp = CostMLP(HMVM_NFEAT, nh, seed)
        self.calib = {}
        self.lo = self.hi = None

    def clip(self, X):
        """Features clamped to the range seen in training.  The launch-level
        features (tiles and elements per launch) of a large batched plan lie
        far beyond any training launch; past saturation of the chip they no
        longer change a tile's duration, and an MLP extrapolates them badly."""
        if self.lo is None:
            return X
        return np.clip(X, self.lo, self.hi).astype(np.float32)

    def fit(self, X, y, recs, epochs=80, batch=None, lr=2e-3, seed=3):
        """Trains the network on (features, log2 us) and calibrates the launch
        model.  batch=None: about 512 steps per epoch (512 to 8192 samples a
        step), so that millions of training tiles cost seconds, not minutes."""
        self.lo, self.hi = X.min(axis=0), X.max(axis=0)
        self.mlp.fit_scaler(X)
        if batch is None:
            batch = int(min(8192, max(512, len(y) // 512)))
        loss = self.mlp.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        by = {}
        for r in recs:
            sim = self.simulate_ms(r["X"], r["storage"], r["kernel"], r["bt"], reorder=False)
            by.setdefault(r["storage"], []).append((sim, r["launch_ms"]))
        for st, pts in by.items():
            a = np.array(pts)
            A = np.stack([a[:, 0], np.ones(len(a))], axis=1)
            coef, *_ = np.linalg.lstsq(A, a[:, 1], rcond=None)
            self.calib[st] = (float(coef[0]), float(coef[1]))
        return loss

    def predict_log2(self, X):
        return self.mlp.predict(self.clip(X))

    def tile_us(self, X):
        return np.exp2(self.predict_log2(X).astype(np.float64))

    def simulate_ms(self, X, storage, kernel, bt, reorder=True):
        d = self.tile_us(X)
        if reorder:
            d = np.sort(d)[::-1]
        return makespan(d, occupancy(storage, kernel, bt)) / 1e3

    def launch_ms(self, X, storage, which, kernel, bt):
        a, b = self.calib.get(storage, (1.0, 0.0))
        return a * self.simulate_ms(X, storage, kernel, bt) + b


def linear_baseline(X, y):
    """The method without the network: log-duration linear in the features
    (least squares), i.e. a roofline-style 'cost grows with elements' model.
    Inputs are clamped to the training range like the MLP's."""
    lo, hi = X.min(axis=0), X.max(axis=0)
    A = np.hstack([X.astype(np.float64), np.ones((X.shape[0], 1))])
    coef, *_ = np.linalg.lstsq(A, y.astype(np.float64), rcond=None)
    return lambda Z: np.hstack([np.clip(Z, lo, hi).astype(np.float64), np.ones((Z.shape[0], 1))]) @ coef


# ---- schedule search ---------------------------------------------------------------
def _apply_choice(p, choice):
    p.set_phase_launch(0, choice["kernel_a"], choice["bt_a"])
    p.set_phase_launch(1, choice["kernel_b"], choice["bt_b"])
    for which, key in ((0, "rpl_a"), (1, "rpl_b")):
        if key in choice:  # uniform path of the launch (exhaustive search)
            f = p.tile_features(which)
            p.set_tile_flags(which, (eligible(f) & bool(choice[key])).astype(np.int32))


def exhaustive(h, storage, warmup=2, iters=8, w_choices=None, kernels=KERNELS, bts=BLOCK_THREADS, paths=True):
    """Times every candidate on the GPU; returns (best choice, all results, seconds spent).
    paths: also time every launch with all eligible tiles lane-per-row."""
    t0 = time.time()
    w_choices = w_choices or list(itertools.product(W1_CHOICES, W2_CHOICES))
    results = []
    best = None
    for a, b in w_choices:
        w1, w2 = widths(storage, a, b, _length(h))
        p = Plan(h, storage=storage, w1=w1, w2=w2)
        per = {}
        elig = [eligible(p.tile_features(which)) for which in (0, 1)]
        for rpl in ((False, True) if paths else (False,)):
            for which in (0, 1):
                p.set_tile_flags(which, (elig[which] & rpl).astype(np.int32))
            for k in kernels:
                for bt in bts:
                    p.set_launch(k, bt)
                    t = p.time(warmup, iters)
                    per[(k, bt, rpl)] = t
                    results.append({"w1": w1, "w2": w2, "kernel": k, "bt": bt, "row_per_lane": rpl,
                                    "ms_a": t["ms_a"], "ms_b": t["ms_b"], "ms_other": t["ms_tsum"] + t["ms_ysum"],
                                    "ms_stream": t["ms_stream"]})
        ka = min(per, key=lambda v: per[v]["ms_a"])
        kb = min(per, key=lambda v: per[v]["ms_b"])
        choice = {"w1": w1, "w2": w2, "kernel_a": ka[0], "bt_a": ka[1], "kernel_b": kb[0], "bt_b": kb[1],
                  "rpl_a": ka[2], "rpl_b": kb[2]}
        _apply_choice(p, choice)
        choice["ms"] = p.time(warmup, iters)["ms_stream"]
        if best is None or choice["ms"] < best["ms"]:
            best = choice
        p.close()
    return best, results, time.time() - t0


def learned(h, storage, model, w_choices=None, kernels=KERNELS, bts=BLOCK_THREADS, reorder=True, verify=3):
    """Chooses the schedule from the cost model's predictions; returns (choice,
    plan, seconds spent).

    Every candidate (tile widths, then per launch the variant and workgroup
    size with the smallest predicted launch time) is predicted on the host.
    With verify > 0 the `verify` candidates predicted fastest and the library
    defaults are built and timed (a few applies each, with and without the
    predicted-duration tile order, and at each candidate's widths also with
    the library's launches) and the fastest is kept: the model prunes the
    search to a handful of runs instead of timing all of them.  verify=0
    takes the best prediction as it is."""
    t0 = time.time()
    w_choices = w_choices or list(itertools.product(W1_CHOICES, W2_CHOICES))
    cands = []
    for a, b in w_choices:
        w1, w2 = widths(storage, a, b, _length(h))
        sched = schedule(h, storage=storage, w1=w1, w2=w2)
        pred = {}
        for which in (0, 1):
            opts = []
            for k in kernels:
                for bt in bts:
                    X = schedule_features(sched, which, storage, k, bt)
                    opts.append((model.launch_ms(X, storage, which, k, bt), k, bt))
            pred[which] = min(opts)
        cands.append({"w1": w1, "w2": w2, "kernel_a": pred[0][1], "bt_a": pred[0][2], "kernel_b": pred[1][1],
                      "bt_b": pred[1][2], "pred_ms": pred[0][0] + pred[1][0]})
    cands.sort(key=lambda c: c["pred_ms"])

    def build(c, ordered, paths, default_launch=False):
        p = Plan(h, storage=storage) if c is None else Plan(h, storage=storage, w1=c["w1"], w2=c["w2"])
        if c is not None:
            if not default_launch:  # else the library's variants and workgroup sizes at these widths
                _apply_choice(p, c)
            for which in (0, 1):
                if paths:  # each eligible tile on the path predicted faster for it
                    X = p.tile_features(which)
                    if len(X):
                        p.set_tile_flags(which, path_flags(model, X))
                if ordered:
                    d = model.tile_us(p.tile_features(which))
                    if len(d):
                        p.set_tile_order(which, np.argsort(-d, kind="stable"))
        return p

    if verify <= 0:
        best = dict(cands[0], reordered=bool(reorder), tile_paths=True)
        return best, build(cands[0], reorder, True), time.time() - t0
    # per predicted width candidate: the predicted launches (in dispatch order,
    # in predicted-duration order, and with per-tile paths), and the
    # library's launches at those widths (the network ranks widths better
    # than variants and workgroup sizes, whose differences show mostly in
    # the memory system)
    trials = [(None, False, False, False)]
    for c in cands[:verify]:
        trials += [(c, False, False, False)]
        trials += [(c, True, False, False), (c, True, True, False)] if reorder else [(c, False, True, False)]
        trials += [(c, False, False, True)]
    best, best_ms, trial_ms = None, None, []
    for c, ordered, paths, dl in trials:
        p = build(c, ordered, paths, dl)
        # a fresh plan's first applies run slow: warm it up; median of three
        ms = float(np.median([p.time(3, 15)["ms_stream"] for _ in range(3)]))
        p.close()
        trial_ms.append(round(ms, 5))
        if best_ms is None or ms < best_ms:
            best, best_ms = (c, ordered, paths, dl), ms
    c, ordered, paths, dl = best
    choice = {"defaults": True} if c is None else dict(c)
    if dl:
        choice.update({"kernel_a": None, "bt_a": None, "kernel_b": None, "bt_b": None, "default_launch": True})
    choice.update({"reordered": ordered, "tile_paths": paths, "timed_trials": len(trials), "trial_ms": trial_ms,
                   "candidates": [(cd["w1"], cd["w2"], cd["kernel_a"], cd["bt_a"], cd["kernel_b"], cd["bt_b"],
                                   round(cd["pred_ms"], 5)) for cd in cands[:verify]]})
    p = build(c, ordered, paths, dl)
    if paths:
        choice["row_per_lane_tiles"] = [int(p.tile_flags(w).sum()) for w in (0, 1)]
    return choice, p, time.time() - t0


def default_plan(h, storage):
    """The schedule without any tuning (library defaults)."""
    return Plan(h, storage=storage)
