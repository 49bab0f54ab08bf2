"""The int64 extension (matrix2_64, format 2L) against the int32 reference structs.

CPU tests: the int64 loaders/writers/generator reproduce the matrix2 the
reference-format loaders build (src_c/loadmatrix.c:8-232) field by field, and
the 2L header is rejected by the int32 reader exactly as an unknown record is.
GPU tests: plans built from matrix2_64, from a streamed file (formats 1, 2 and
2L, with upload windows forced small) and the int64 drop-in entry points give
the same y as the int32 path, bit for bit.
"""
import os
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def small(H):
    return H.generate(1500, geometry="cube", eps=1e-4)


def same_descriptors(a, b):
    for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt", "a1", "a2"):
        assert np.array_equal(np.asarray(a[k], dtype=np.int64), np.asarray(b[k], dtype=np.int64)), k


def test_widen_is_exact(H, small):
    w = H.widen(small)
    a, b = small.arrays(), w.arrays()
    same_descriptors(a, b)
    assert w.mat2.len == small.mat2.len and w.mat2.nd == small.nd and w.mat2.ktmax == small.mat2.ktmax
    assert np.array_equal(a["rowmat"], b["rowmat"]) and np.array_equal(a["rowmat_t"], b["rowmat_t"])


@pytest.mark.parametrize("fmt", ["1", "2", "2L", "1_64"])
def test_load64_matches_int32_loader(H, small, tmp_path, fmt):
    f = str(tmp_path / ("m_%s.bin" % fmt))
    if fmt == "1":
        H.writeHmatrix1(f, small)
    elif fmt == "2":
        H.writeHmatrix2(f, small)
    elif fmt == "2L":
        H.writeHmatrix2_64(f, H.widen(small))
    else:
        H.writeHmatrix1_64(f, H.widen(small))
    w = H.loadHmatrix_64(f, with_rowmat_t=True)
    a, b = small.arrays(), w.arrays()
    same_descriptors(a, b)
    assert w.mat2.len == small.mat2.len
    assert np.array_equal(a["rowmat"], b["rowmat"])
    assert np.array_equal(a["rowmat_t"], b["rowmat_t"])
    w2 = H.loadHmatrix_64(f)
    assert w2.arrays()["rowmat_t"] is None and np.array_equal(w2.arrays()["rowmat"], a["rowmat"])


def test_format_2L_header(H, small, tmp_path):
    f = str(tmp_path / "m.bin")
    H.writeHmatrix2_64(f, H.widen(small))
    raw = open(f, "rb").read(28)
    m0, nd, nlf, ktmax, ln, m1 = struct.unpack("<iiiiqi", raw)
    assert m0 == m1 == 20 and nd == small.nd and nlf == small.mat2.nlf and ln == small.mat2.len
    # the int32 reader refuses the 20-byte header record as a marker mismatch
    with pytest.raises(H.HMVMError, match="marker"):
        H.loadHmatrix2(f)
    # and the 2L file equals the format-2 file after the header
    f2 = str(tmp_path / "m2.bin")
    H.writeHmatrix2(f2, small)
    assert open(f, "rb").read()[28:] == open(f2, "rb").read()[24:]


def test_scan_file_offsets(H, small, tmp_path):
    f = str(tmp_path / "m1.bin")
    H.writeHmatrix1(f, small)
    h, pos, fmt = H.scan_file(f)
    assert fmt == 1 and pos == -1
    same_descriptors(h.arrays(), small.arrays())
    assert not h.mat2.rowmat
    f2 = str(tmp_path / "m2.bin")
    H.writeHmatrix2(f2, small)
    h2, pos2, fmt2 = H.scan_file(f2)
    assert fmt2 == 2 and pos2 == 24 + 6 * (8 + 4 * small.mat2.nlf)


def test_load64_errors(H, small, tmp_path):
    f = str(tmp_path / "m.bin")
    H.writeHmatrix2_64(f, H.widen(small))
    data = open(f, "rb").read()
    t = str(tmp_path / "trunc.bin")
    open(t, "wb").write(data[:-100])
    with pytest.raises(H.HMVMError, match="truncated|marker"):
        H.loadHmatrix_64(t)
    bad = bytearray(data)
    struct.pack_into("<q", bad, 16, small.mat2.len + 1)  # stored len disagrees with the leaves
    b = str(tmp_path / "badlen.bin")
    open(b, "wb").write(bytes(bad))
    with pytest.raises(H.HMVMError, match="len"):
        H.loadHmatrix_64(b)
    with pytest.raises(H.HMVMError, match="open"):
        H.loadHmatrix_64(str(tmp_path / "missing.bin"))


def test_generate64_equals_generate(H, small):
    g = H.generate64(1500, geometry="cube", eps=1e-4, with_rowmat_t=True)
    a, b = small.arrays(), g.arrays()
    same_descriptors(a, b)
    assert np.array_equal(a["rowmat"], b["rowmat"]) and np.array_equal(a["rowmat_t"], b["rowmat_t"])
    a64, a32 = H.accuracy64(g, 3), H.accuracy(small, 3)  # OpenMP reduction: order varies
    assert abs(a64 - a32) <= 1e-10 * a32


# ---- GPU ------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
def test_plan64_equals_plan32(gpu, H, small, storage):
    x = np.random.default_rng(3).uniform(-1, 1, small.nd)
    p32 = H.Plan(small, storage=storage)
    p64 = H.Plan(H.widen(small), storage=storage)
    assert np.array_equal(p32.apply_host(x), p64.apply_host(x))
    p32.close()
    p64.close()


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["1", "2", "2L"])
@pytest.mark.parametrize("window", [None, "4096"])
def test_plan_from_file_streams(gpu, H, O, small, tmp_path, monkeypatch, fmt, window):
    f = str(tmp_path / ("m_%s.bin" % fmt))
    {"1": lambda: H.writeHmatrix1(f, small), "2": lambda: H.writeHmatrix2(f, small),
     "2L": lambda: H.writeHmatrix2_64(f, H.widen(small))}[fmt]()
    if window:
        monkeypatch.setenv("HMVM_UPLOAD_WINDOW", window)
    x = np.random.default_rng(4).uniform(-1, 1, small.nd)
    pf = H.Plan(f)
    pm = H.Plan(small)
    yf, ym = pf.apply_host(x), pm.apply_host(x)
    assert np.array_equal(yf, ym)
    ref = O.seq_2(small, x)
    assert np.linalg.norm(yf - ref) <= 1e-13 * np.linalg.norm(ref)
    pf.close()
    pm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f64", "bf16"])
def test_small_upload_windows_bit_identical(gpu, H, small, monkeypatch, storage):
    x = np.random.default_rng(5).uniform(-1, 1, small.nd)
    p0 = H.Plan(small, storage=storage)
    y0 = p0.apply_host(x)
    p0.close()
    monkeypatch.setenv("HMVM_UPLOAD_WINDOW", "2048")
    p1 = H.Plan(small, storage=storage)
    assert np.array_equal(p1.apply_host(x), y0)
    p1.close()


@pytest.mark.gpu
def test_drop_in_64(gpu, H, small):
    w = H.widen(small)
    b = np.sin(np.arange(1, small.nd + 1.0))
    v0 = np.linspace(-1, 1, small.nd)
    v32 = H.hmvm_hip_2(v0.copy(), small, b)
    v64 = H.hmvm_hip_2_64(v0.copy(), w, b)
    assert np.array_equal(v32, v64)
    v32t = H.hmvm_hip_2t(v0.copy(), small, b)
    v64t = H.hmvm_hip_2t_64(v0.copy(), w, b)
    assert np.array_equal(v32t, v64t)
    H.release(small)
    w.free()


def test_oracle_64_equals_oracle(H, O, small):
    w = H.widen(small)
    x = np.random.default_rng(6).uniform(-1, 1, small.nd)
    assert np.array_equal(O.seq_2_64(w, x), O.seq_2(small, x))
    ref = O.seq_2(small, x)
    assert np.linalg.norm(O.omp_2_64(w, x) - ref) <= 1e-14 * np.linalg.norm(ref)


def test_leaf_subset_view(H, O, small):
    a = small.arrays()
    rows = a["nstrtl"] <= small.nd // 2          # a row band, as one GPU's shard
    band = H.leaf_subset(small, np.nonzero(rows)[0])
    rest = H.leaf_subset(small, np.nonzero(~rows)[0])
    assert band.mat2.nlf + rest.mat2.nlf == small.mat2.nlf
    x = np.random.default_rng(7).uniform(-1, 1, small.nd)
    yb, yr = O.seq_2_64(band, x), O.seq_2_64(rest, x)
    ref = O.seq_2(small, x)
    assert np.linalg.norm(yb + yr - ref) <= 1e-14 * np.linalg.norm(ref)
    # the band only touches its rows
    top = a["nstrtl"][rows] + a["ndl"][rows] - 1
    assert np.all(yb[top.max():] == 0)


@pytest.mark.gpu
def test_plan_on_leaf_subset(gpu, H, O, small):
    a = small.arrays()
    rows = a["nstrtl"] <= small.nd // 2
    band = H.leaf_subset(small, np.nonzero(rows)[0])
    x = np.random.default_rng(8).uniform(-1, 1, small.nd)
    p = H.Plan(band)
    y = p.apply_host(x)
    ref = O.seq_2_64(band, x)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    p.close()
