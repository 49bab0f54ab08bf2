"""Loaders and writers of the dump formats (CPU).

Pins: round trips through both formats; field-by-field identity with the
reference loader compiled from src_c/loadmatrix.c (when the reference sources
are present, i.e. in the build container); the error codes the reference does
not have (truncated/corrupt files)."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, reference_path

GOLDEN = os.path.join(ROOT, "tests", "golden")


def arrays_equal(h1, h2):
    a, b = h1.arrays(), h2.arrays()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for f in ("nd", "nlf", "ktmax", "len"):
        assert getattr(h1.mat2, f) == getattr(h2.mat2, f), f


def test_struct_layout_matches_reference_header(tmp_path):
    """offsetof/sizeof of matrix, matrix2, submatrix: ours vs src_c/hacapk.h."""
    src = tmp_path / "lay.c"
    fields = {"submatrix": "ltmtx kt ndl ndt nstrtl nstrtt a1 a2 a2t",
              "matrix": "nd nlf ktmax nleavs submat b1 b2 b3 e1 e2 e3",
              "matrix2": "nd nlf ktmax len nleavs ltmtx kt ndl ndt nstrtl nstrtt a1 a2 rowmat rowmat_t"}
    body = "".join('printf("%s %%zu\\n", sizeof(%s));' % (s, s) +
                   "".join('printf("%s.%s %%zu\\n", offsetof(%s,%s));' % (s, f, s, f) for f in fl.split())
                   for s, fl in fields.items())
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "hacapk.h"\nint main(){%s return 0;}\n' % body)
    outs = []
    incs = [os.path.join(ROOT, "include")]
    ref = reference_path("src_hmvm", "src_c")
    if ref:
        incs.append(ref)
    for inc in incs:
        exe = tmp_path / ("lay%d" % len(outs))
        subprocess.run(["gcc", "-I" + inc, str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout)
    # ctypes mirror agrees with the C header too
    import hacapk_hmvm_amd as H
    lines = dict(l.split() for l in outs[0].split("\n") if l)
    assert int(lines["matrix2"]) == ctypes.sizeof(H.matrix2) == 4 * 4 + 11 * 8
    assert int(lines["matrix"]) == ctypes.sizeof(H.matrix)
    assert int(lines["submatrix"]) == ctypes.sizeof(H.submatrix)
    assert int(lines["matrix2.rowmat"]) == H.matrix2.rowmat.offset
    if len(outs) == 2:
        assert outs[0] == outs[1]


@pytest.mark.parametrize("geometry,eps", [("sphere", 1e-4), ("cube", 1e-6)])
def test_roundtrip_both_formats(H, tmp_path, geometry, eps):
    g = H.generate(700, geometry=geometry, eps=eps, leaf=16)
    H.writeHmatrix1(str(tmp_path / "h1.bin"), g)
    H.writeHmatrix2(str(tmp_path / "h2.bin"), g)
    h1 = H.loadHmatrix(str(tmp_path / "h1.bin"))
    h2 = H.loadHmatrix2(str(tmp_path / "h2.bin"))
    arrays_equal(g, h1)
    arrays_equal(g, h2)
    # matrix (AoS) form: a2t is the transpose of a2, a1 identical to rowmat
    a = h1.arrays()
    for i in range(h1.mat.nlf):
        s = h1.mat.submat[i]
        if s.ltmtx == 1:
            a2 = np.ctypeslib.as_array(s.a2, (s.kt * s.ndl,)).reshape(s.kt, s.ndl)
            a2t = np.ctypeslib.as_array(s.a2t, (s.kt * s.ndl,)).reshape(s.ndl, s.kt)
            assert np.array_equal(a2.T, a2t)
            off = a["a2"][i]
            assert np.array_equal(a["rowmat_t"][off:off + s.kt * s.ndl], a2t.ravel())


def _ref_load(R, H, fname, fmt):
    class M(ctypes.Structure):
        _fields_ = H.matrix._fields_

    class M2(ctypes.Structure):
        _fields_ = H.matrix2._fields_
    m, m2 = M(), M2()
    if fmt == 1:
        assert R.loadHmatrix(fname.encode(), ctypes.byref(m), ctypes.byref(m2)) == 0
    else:
        assert R.loadHmatrix2(fname.encode(), ctypes.byref(m2)) == 0
    return m, m2


def _cmp_with_ref(H, ours, m, m2, with_aos):
    assert (ours.mat2.nd, ours.mat2.nlf, ours.mat2.ktmax, ours.mat2.len) == (m2.nd, m2.nlf, m2.ktmax, m2.len)
    n, L = m2.nlf, m2.len
    a = ours.arrays()
    for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt", "a1", "a2"):
        assert np.array_equal(a[k], np.ctypeslib.as_array(getattr(m2, k), (n,))), k
    assert np.array_equal(a["rowmat"], np.ctypeslib.as_array(m2.rowmat, (L,)))
    assert np.array_equal(a["rowmat_t"], np.ctypeslib.as_array(m2.rowmat_t, (L,)))
    if with_aos:
        for i in range(n):
            s, r = ours.mat.submat[i], m.submat[i]
            assert (s.ltmtx, s.kt, s.ndl, s.ndt, s.nstrtl, s.nstrtt) == (r.ltmtx, r.kt, r.ndl, r.ndt, r.nstrtl, r.nstrtt)
            n1 = s.kt * s.ndt if s.ltmtx == 1 else s.ndl * s.ndt
            assert np.array_equal(np.ctypeslib.as_array(s.a1, (n1,)), np.ctypeslib.as_array(r.a1, (n1,)))
            if s.ltmtx == 1 and s.kt > 0:
                n2 = s.kt * s.ndl
                assert np.array_equal(np.ctypeslib.as_array(s.a2, (n2,)), np.ctypeslib.as_array(r.a2, (n2,)))
                assert np.array_equal(np.ctypeslib.as_array(s.a2t, (n2,)), np.ctypeslib.as_array(r.a2t, (n2,)))


def test_loader_matches_reference_loader(H, O, tmp_path):
    R = O.ref()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not present")
    files = [os.path.join(GOLDEN, f) for f in sorted(os.listdir(GOLDEN)) if f.endswith("_1.bin")]
    sample = reference_path("src_hdump", "hmatrix_sample.bin")
    if sample:
        files.append(sample)
    assert files
    for f in files:
        ours = H.loadHmatrix(f)
        m, m2 = _ref_load(R, H, f, 1)
        _cmp_with_ref(H, ours, m, m2, True)
        # format 2 written by us, read by the reference
        f2 = str(tmp_path / "two.bin")
        H.writeHmatrix2(f2, ours)
        _, r2 = _ref_load(R, H, f2, 2)
        ours2 = H.loadHmatrix2(f2)
        _cmp_with_ref(H, ours2, None, r2, False)


def test_sample_fixture_structure(H):
    """Structure of the reference's own fixture as recorded in its survey:
    nd=600, nlf=1426, ktmax=31, 210 low-rank + 1216 dense, len=248,662, and the
    leaves tile the 600 x 600 matrix."""
    sample = reference_path("src_hdump", "hmatrix_sample.bin")
    if not sample:
        pytest.skip("reference fixture not present")
    h = H.loadHmatrix(sample)
    a = h.arrays()
    assert (h.mat2.nd, h.mat2.nlf, h.mat2.ktmax, h.mat2.len) == (600, 1426, 31, 248662)
    assert int((a["ltmtx"] == 1).sum()) == 210 and int((a["ltmtx"] == 2).sum()) == 1216
    assert int((a["ndl"].astype(np.int64) * a["ndt"]).sum()) == 600 * 600


def test_dummymatrix_matches_reference(H, O):
    R = O.ref()
    if R is None:
        pytest.skip("reference build not present")
    ours = H.dummymatrix(7, 5, 4)

    class M(ctypes.Structure):
        _fields_ = H.matrix._fields_

    class M2(ctypes.Structure):
        _fields_ = H.matrix2._fields_
    m, m2 = M(), M2()
    assert R.dummymatrix(ctypes.byref(m), ctypes.byref(m2), 7, 5, 4) == 0
    _cmp_with_ref(H, ours, m, m2, True)


# ---- error paths -------------------------------------------------------------------
def _write_records(path, records):
    with open(path, "wb") as f:
        for payload in records:
            f.write(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))


def _leaf_file(path, nd=8, leaves=((2, 4, 4, 1, 1, 0),), ktmax=0, values=None):
    recs = [struct.pack("<3i", nd, len(leaves), ktmax)]
    for lt, ndl, ndt, nl, nt, kt in leaves:
        recs.append(struct.pack("<6i", lt, ndl, ndt, nl, nt, kt))
        if lt == 1:
            recs.append(np.ones(kt * ndt).tobytes())
            recs.append(np.ones(kt * ndl).tobytes())
        else:
            recs.append(np.ones(ndl * ndt).tobytes())
    _write_records(path, recs)


def test_error_codes(H, tmp_path):
    f = str(tmp_path / "m.bin")
    with pytest.raises(H.HMVMError, match="code -1"):
        H.loadHmatrix(str(tmp_path / "missing.bin"))
    _leaf_file(f)
    h = H.loadHmatrix(f)
    assert h.mat2.len == 16
    data = open(f, "rb").read()
    open(f, "wb").write(data[:-10])
    with pytest.raises(H.HMVMError, match="code -2"):
        H.loadHmatrix(f)
    bad = bytearray(data)
    bad[0] = 13  # header marker no longer matches 12 bytes
    open(f, "wb").write(bytes(bad))
    with pytest.raises(H.HMVMError, match="code -3"):
        H.loadHmatrix(f)
    _leaf_file(f, leaves=((2, 4, 4, 6, 1, 0),))  # rows 6..9 > nd=8
    with pytest.raises(H.HMVMError, match="code -5"):
        H.loadHmatrix(f)
    _leaf_file(f, leaves=((3, 4, 4, 1, 1, 0),))  # ltmtx 3
    with pytest.raises(H.HMVMError, match="code -5"):
        H.loadHmatrix(f)
    _leaf_file(f, leaves=((1, 4, 4, 1, 1, 5),), ktmax=3)  # kt > ktmax
    with pytest.raises(H.HMVMError, match="code -5"):
        H.loadHmatrix(f)
    _write_records(f, [struct.pack("<3i", -1, 1, 0)])
    with pytest.raises(H.HMVMError, match="code -4"):
        H.loadHmatrix(f)


def test_format2_len_mismatch(H, tmp_path):
    g = H.generate(300, geometry="sphere", leaf=16)
    f = str(tmp_path / "h2.bin")
    H.writeHmatrix2(f, g)
    data = bytearray(open(f, "rb").read())
    struct.pack_into("<i", data, 16, g.mat2.len + 1)  # header record: nd nlf ktmax len
    open(f, "wb").write(bytes(data))
    with pytest.raises(H.HMVMError, match="code -8"):
        H.loadHmatrix2(f)
