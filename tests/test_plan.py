"""Host logic of the HIP path (CPU only, no GPU calls): the C ABI exports, and
the schedule checker on generated and synthetic matrices."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    names = set()
    for h in ("hacapk.h", "hgen.h", "hmvm_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-zA-Z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return sorted(names)


def test_c_abi_exports_every_declared_symbol(H):
    names = declared_functions()
    assert "hmvm_hip_2" in names and "loadHmatrix" in names and "hgen_generate" in names
    lib = ctypes.CDLL(H.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("w", [64, 512, 16384, 1 << 20])
def test_schedule_check_generated(H, w):
    h = H.generate(3000, geometry="cube", eps=1e-4)
    info = H.plan_check(h, w1=w, w2=w)
    assert info["len"] == h.mat2.len
    assert info["elems_a"] + info["elems_b"] == h.mat2.len
    assert info["elems_stage1"] + info["elems_dense"] + info["elems_stage2"] == h.mat2.len
    assert info["device_elems"] >= h.mat2.len
    assert info["algorithmic_bytes"] == 8 * h.mat2.len + 16 * h.nd
    a = h.arrays()
    lr = a["ltmtx"] == 1
    assert info["nlowrank"] == lr.sum() and info["ndense"] == (~lr).sum()
    # t holds only the leaves whose stage 2 reads it; none when every leaf sums
    # its stage-1 partials on the fly, all with t_mode 1
    assert bool(info["t_fused"]) == (info["n_t"] == 0)
    info1 = H.plan_check(h, w1=w, w2=w, t_mode=1)
    assert info1["n_t"] == a["kt"][lr].sum() and not info1["t_fused"]
    assert info1["elems_a"] + info1["elems_b"] == h.mat2.len


def test_schedule_check_dense_grid(H):
    h = H.dummymatrix(40, 40, 5)  # rows of 40 -> one 40-row tile per row cluster
    info = H.plan_check(h)
    assert info["ntiles_b"] == 0 and info["elems_dense"] == h.mat2.len


def test_schedule_rejects_bad_descriptors(H):
    h = H.generate(500, geometry="sphere", leaf=16)
    a = h.arrays()
    saved = a["nstrtl"][3]
    a["nstrtl"][3] = h.nd  # leaf now runs past the last row
    with pytest.raises(H.HMVMError):
        H.plan_check(h)
    a["nstrtl"][3] = saved
    saved = a["a1"][5]
    a["a1"][5] = h.mat2.len  # factor outside the value buffer
    with pytest.raises(H.HMVMError):
        H.plan_check(h)
    a["a1"][5] = saved
    H.plan_check(h)


def test_stage2_width_option_validated(H):
    h = H.generate(400, geometry="sphere", leaf=16)
    for bad in (1, 63):
        with pytest.raises(H.HMVMError):
            H.plan_check(h, w3=bad)
    a = H.plan_check(h, w1=1024, w2=1024)
    b = H.plan_check(h, w1=1024, w2=1024, w3=64)
    assert b["ntiles_a"] == a["ntiles_a"] and b["ntiles_b"] > a["ntiles_b"]


def test_storage_option_validated(H):
    h = H.generate(400, geometry="sphere", leaf=16)
    assert H.plan_check(h, storage="bf16")["elem_bytes"] == 2
    assert H.plan_check(h, storage="f32")["algorithmic_bytes"] == 4 * h.mat2.len + 16 * h.nd


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("t_mode", [0, 1, 2])
def test_tile_panels_back_to_back(H, storage, t_mode):
    """The grouped kernel path walks a tile's panels as one column space: in
    every exported schedule the panels of a tile follow each other in memory
    with one column stride (the plan check enforces it as well)."""
    h = H.generate(6000, geometry="sphere", eps=1e-6, leaf=48)
    s = H.schedule(h, storage=storage, t_mode=t_mode)
    pan = s["panels"]
    multi = 0
    for key in ("tiles_a", "tiles_b"):
        for p0, n in zip(s[key]["panel0"], s[key]["npanel"]):
            multi += n > 1
            for q in range(p0 + 1, p0 + n):
                assert pan["ld"][q] == pan["ld"][q - 1]
                assert pan["val_off"][q] == pan["val_off"][q - 1] + pan["ncols"][q - 1] * pan["ld"][q - 1]
    assert multi > 0
    h.free()


def test_stage2_tile_height_option(H):
    """max_rows_b bounds the stage-2 tiles only; max_rows alone bounds both."""
    h = H.generate(6000, geometry="sphere", eps=1e-6, leaf=48)
    s = H.schedule(h, storage="f64", max_rows=256, max_rows_b=64)
    assert s["tiles_b"]["nrows"].max() <= 64 < s["tiles_a"]["nrows"].max()
    s = H.schedule(h, storage="f64", max_rows=96)
    assert max(s["tiles_a"]["nrows"].max(), s["tiles_b"]["nrows"].max()) <= 96
    with pytest.raises(H.HMVMError):
        H.schedule(h, storage="f64", max_rows_b=400)  # above 192*2-1 for FP64
    h.free()


def test_launch_a_order(H):
    """Largest first (the default): launch A's stage-1 and dense tiles in one
    order of non-increasing elements, interleaved; memory order: all stage-1
    tiles, then the dense ones.  Both orders hold the same tiles."""
    h = H.generate(4096, geometry="sphere", eps=1e-4)

    def kinds_and_work(s):
        t, pan = s["tiles_a"], s["panels"]
        cc = np.concatenate([[0], np.cumsum(pan["ncols"].astype(np.int64))])
        p0, n = t["panel0"].astype(np.int64), t["npanel"].astype(np.int64)
        return pan["factor"][p0] < 0, t["nrows"].astype(np.int64) * (cc[p0 + n] - cc[p0])

    s1 = H.schedule(h)
    k1, w1 = kinds_and_work(s1)
    assert np.all(np.diff(w1) <= 0)
    n1 = int(k1.sum())
    assert 0 < n1 < len(k1) and not np.all(k1[:n1])  # interleaved
    s0 = H.schedule(h, order=0)
    k0, _ = kinds_and_work(s0)
    assert np.all(k0[:n1]) and not np.any(k0[n1:])
    assert sorted(s0["tiles_a"]["panel0"].tolist()) == sorted(s1["tiles_a"]["panel0"].tolist())
    h.free()
