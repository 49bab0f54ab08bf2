"""Pins of the CPU oracle (the restated reference HMVM), CPU only.

- bit-exact against golden vectors the compiled reference produced
  (tests/golden, made by tests/golden/make_golden.py);
- bit-exact against the reference compiled from its own sources, on fresh
  generated matrices (build container only);
- closed form: an all-dense grid equals a plain dense GEMV;
- brute force: the assembled dense matrix times x, summed exactly (math.fsum);
- a planted mistake in the oracle (wrong a2 layout, dropped term) fails the
  golden check, so the pins have power."""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import ROOT, reference_path

GOLDEN = os.path.join(ROOT, "tests", "golden")
STEMS = sorted(f[:-6] for f in os.listdir(GOLDEN) if f.endswith("_1.bin"))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module", params=STEMS)
def golden(request, H):
    h = H.loadHmatrix(os.path.join(GOLDEN, request.param + "_1.bin"))
    return h, dict(np.load(os.path.join(GOLDEN, request.param + "_y.npz")))


@pytest.mark.parametrize("variant", ["1", "1t", "2", "2t"])
def test_oracle_bit_exact_vs_golden(O, golden, variant):
    h, g = golden
    fn = getattr(O, "seq_" + variant)
    for key in ("sin", "rand"):
        y = fn(h, g["x_" + key])
        assert np.array_equal(y, g["y_%s_%s" % (key, variant)]), key


def test_oracle_vs_reference_sample(H, O):
    """The reference's own fixture read by our loader, our oracle vs the y the
    reference computed from it (golden sample_y.npz)."""
    sample = reference_path("src_hdump", "hmatrix_sample.bin")
    if not sample:
        pytest.skip("reference fixture not present")
    h = H.loadHmatrix(sample)
    g = np.load(os.path.join(GOLDEN, "sample_y.npz"))
    for key in ("sin", "rand"):
        for var in ("1", "1t", "2", "2t"):
            assert np.array_equal(getattr(O, "seq_" + var)(h, g["x_" + key]), g["y_%s_%s" % (key, var)])


@pytest.mark.parametrize("geometry,n,eps", [("sphere", 1500, 1e-4), ("cube", 1200, 1e-6)])
def test_oracle_vs_compiled_reference(H, O, geometry, n, eps):
    R = O.ref()
    if R is None:
        pytest.skip("reference build not present")
    h = H.generate(n, geometry=geometry, eps=eps).build_aos()
    x = np.random.default_rng(1).uniform(-1, 1, n)
    dp = ctypes.POINTER(ctypes.c_double)
    for var, m in (("1", h.mat), ("1t", h.mat), ("2", h.mat2), ("2t", h.mat2)):
        v = np.zeros(n)
        getattr(R, "hmvm_seq_" + var)(v.ctypes.data_as(dp), ctypes.cast(ctypes.byref(m), ctypes.c_void_p),
                                       x.ctypes.data_as(dp))
        assert np.array_equal(getattr(O, "seq_" + var)(h, x), v), var
        w = np.zeros(n)
        getattr(R, "hmvm_omp_" + var)(w.ctypes.data_as(dp), ctypes.cast(ctypes.byref(m), ctypes.c_void_p),
                                       x.ctypes.data_as(dp))
        assert rel(getattr(O, "omp_" + var)(h, x), w) <= 1e-14


def test_fp32_oracle_vs_reference_template(H, O):
    """oracle_hmvm_seq_2_f32 against hmvm_seq_2<float,0> of src_cpp/hmvm_seq.cpp."""
    C = O.ref_cpp()
    if C is None:
        pytest.skip("reference C++ build not present")
    h = H.generate(900, geometry="sphere", leaf=16)
    a = h.arrays()
    r32 = a["rowmat"].astype(np.float32)
    x32 = np.sin(np.arange(1, h.nd + 1, dtype=np.float32))

    class M2F(ctypes.Structure):
        _fields_ = [f for f in H.matrix2._fields_ if f[0] not in ("rowmat", "rowmat_t")] + [
            ("rowmat", ctypes.POINTER(ctypes.c_float)), ("rowmat_t", ctypes.POINTER(ctypes.c_float)),
            ("napprox", ctypes.c_int), ("ndense", ctypes.c_int),
            ("approx", ctypes.POINTER(ctypes.c_int)), ("dense", ctypes.POINTER(ctypes.c_int))]
    m = M2F()
    for f, _ in H.matrix2._fields_:
        if f not in ("rowmat", "rowmat_t"):
            setattr(m, f, getattr(h.mat2, f))
    fp = ctypes.POINTER(ctypes.c_float)
    m.rowmat = r32.ctypes.data_as(fp)
    m.rowmat_t = r32.ctypes.data_as(fp)
    fn = getattr(C, "_Z10hmvm_seq_2IfLi0EEvPT_PK7matrix2IS0_EPKS0_")
    v = np.zeros(h.nd, dtype=np.float32)
    fn(v.ctypes.data_as(fp), ctypes.byref(m), x32.ctypes.data_as(fp))
    assert np.array_equal(O.seq_2_f32(h, x32, r32), v)


def test_closed_form_all_dense(H, O):
    h = H.dummymatrix(12, 12, 7)
    A = np.zeros((h.nd, h.nd))
    for i in range(h.mat.nlf):
        s = h.mat.submat[i]
        A[s.nstrtl - 1:s.nstrtl - 1 + s.ndl, s.nstrtt - 1:s.nstrtt - 1 + s.ndt] = \
            np.ctypeslib.as_array(s.a1, (s.ndl * s.ndt,)).reshape(s.ndl, s.ndt)
    x = np.random.default_rng(2).uniform(-1, 1, h.nd)
    for var in ("1", "1t", "2", "2t"):
        assert rel(getattr(O, "seq_" + var)(h, x), A @ x) <= 1e-14


def assemble(h):
    n = h.nd
    A = np.zeros((n, n))
    a = h.arrays()
    for i in range(h.mat2.nlf):
        ndl, ndt, kt = a["ndl"][i], a["ndt"][i], a["kt"][i]
        r0, c0 = a["nstrtl"][i] - 1, a["nstrtt"][i] - 1
        if a["ltmtx"][i] == 1:
            a1 = a["rowmat"][a["a1"][i]:a["a1"][i] + kt * ndt].reshape(kt, ndt)
            a2 = a["rowmat"][a["a2"][i]:a["a2"][i] + kt * ndl].reshape(kt, ndl)
            A[r0:r0 + ndl, c0:c0 + ndt] += a2.T @ a1
        else:
            A[r0:r0 + ndl, c0:c0 + ndt] += a["rowmat"][a["a1"][i]:a["a1"][i] + ndl * ndt].reshape(ndl, ndt)
    return A


def test_brute_force_assembled(H, O):
    h = H.generate(600, geometry="cube", eps=1e-4, leaf=16)
    A = assemble(h)
    x = np.random.default_rng(3).uniform(-1, 1, h.nd)
    exact = np.array([math.fsum(A[r] * x) for r in range(h.nd)])
    assert rel(O.seq_2(h, x), exact) <= 1e-14
    # and the assembled operator approximates the kernel matrix to about eps
    K = H.kernel_block(h.params, 0, h.nd, 0, h.nd)
    assert np.linalg.norm(A - K) / np.linalg.norm(K) <= 3e-4


def test_variants_agree_and_linear(H, O):
    h = H.generate(1000, geometry="sphere", eps=1e-6).build_aos()
    rng = np.random.default_rng(4)
    x1, x2 = rng.uniform(-1, 1, h.nd), rng.uniform(-1, 1, h.nd)
    y = O.seq_2(h, x1)
    for var in ("1", "1t", "2t"):
        assert rel(getattr(O, "seq_" + var)(h, x1), y) <= 1e-14
    assert rel(O.seq_2(h, 2 * x1 + x2), 2 * y + O.seq_2(h, x2)) <= 1e-14
    assert rel(O.omp_2(h, x1), y) <= 1e-14


def test_planted_mistakes_fail_golden(H, O):
    """Two plausible oracle mistakes planted on the CPU: the golden check must
    reject both."""
    h = H.loadHmatrix(os.path.join(GOLDEN, STEMS[0] + "_1.bin"))
    g = np.load(os.path.join(GOLDEN, STEMS[0] + "_y.npz"))
    x, y = g["x_rand"], g["y_rand"]
    a = h.arrays()
    # (1) a2 read with the transposed layout
    saved = a["rowmat"].copy()
    a["rowmat"][:] = a["rowmat_t"]
    assert rel(O.seq_2(h, x), y) > 1e-3
    a["rowmat"][:] = saved
    # (2) the last rank term of every low-rank leaf dropped
    for i in np.nonzero(a["ltmtx"] == 1)[0]:
        kt, ndt, ndl = a["kt"][i], a["ndt"][i], a["ndl"][i]
        a["rowmat"][a["a1"][i] + (kt - 1) * ndt:a["a1"][i] + kt * ndt] = 0.0
    assert not np.array_equal(O.seq_2(h, x), y)
    assert rel(O.seq_2(h, x), y) > 1e-8
    a["rowmat"][:] = saved
    assert np.array_equal(O.seq_2(h, x), y)


def test_convert_d2f_rounds_to_nearest_even(O):
    """oracle_convert_d2f against the IEEE 754 results fixed by the rules
    (ties to even, subnormals, overflow to infinity, signed zero), and against
    numpy's conversion on doubles spread over the whole range; NaN stays NaN.
    A truncating or ties-away conversion fails the tie cases."""
    from f32_rounding import case_bits, case_values, check, random_values
    got = O.convert_d2f(case_values())
    assert np.array_equal(got.view(np.uint32), case_bits())
    r = random_values(200000, 5)
    check(r, O.convert_d2f(r))
    n = O.convert_d2f(np.array([np.nan, -np.nan, 1.0]))
    assert np.isnan(n[0]) and np.isnan(n[1]) and n[2] == 1.0
