"""Writes the golden fixtures of tests/golden/ (run in the build container).

For each small generated H-matrix: the matrix itself in dump format 1
(<stem>_1.bin) and, in <stem>_y.npz, input vectors and the y the
REFERENCE's own sequential HMVM (src_hmvm/src_c/hmvm_seq.c, compiled by
oracle/Makefile into oracle/_ref/) produced for them:
    x_sin  = sin(i+1)  (the reference drivers' input, hmvm1.c:32-35)
    x_rand = U[-1,1] with numpy seed 20261202
    y_<x>_<variant> for the variants 1, 1t, 2, 2t (y_<x> is variant 2).
For the reference's own fixture src_hdump/hmatrix_sample.bin (not copied
here) only the output vectors are kept, in sample_y.npz.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import hacapk_hmvm_amd as H  # noqa: E402
import oracle as O  # noqa: E402

CASES = [
    ("sphere320", dict(n=320, geometry="sphere", leaf=16, eps=1e-4)),
    ("cube384", dict(n=384, geometry="cube", leaf=16, eps=1e-6)),
]
SAMPLE = "/root/reference/src_hdump/hmatrix_sample.bin"


def ref_y(R, h, x):
    out = {}
    dp = ctypes.POINTER(ctypes.c_double)
    for var, fn, m in (("1", "hmvm_seq_1", h.mat), ("1t", "hmvm_seq_1t", h.mat),
                       ("2", "hmvm_seq_2", h.mat2), ("2t", "hmvm_seq_2t", h.mat2)):
        v = np.zeros(h.nd)
        getattr(R, fn)(v.ctypes.data_as(dp), ctypes.cast(ctypes.byref(m), ctypes.c_void_p), x.ctypes.data_as(dp))
        out[var] = v
    return out


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("reference build missing: make -C oracle ref")
    for stem, kw in CASES:
        n = kw.pop("n")
        g = H.generate(n, **kw)
        H.writeHmatrix1(os.path.join(HERE, stem + "_1.bin"), g)
        h = H.loadHmatrix(os.path.join(HERE, stem + "_1.bin"))
        xs = {"sin": np.sin(np.arange(1, n + 1, dtype=np.float64)),
              "rand": np.random.default_rng(20261202).uniform(-1, 1, n)}
        data = {}
        for key, x in xs.items():
            ys = ref_y(R, h, x)
            data["x_" + key] = x
            data["y_" + key] = ys["2"]
            for var, y in ys.items():
                data["y_%s_%s" % (key, var)] = y
        np.savez(os.path.join(HERE, stem + "_y.npz"), **data)
        print(stem, "nd", h.nd, "nlf", h.mat2.nlf, "len", h.mat2.len)
    if os.path.exists(SAMPLE):
        # the reference's own loader on its own fixture, then its own HMVM
        class M(ctypes.Structure):
            _fields_ = H.matrix._fields_

        class M2(ctypes.Structure):
            _fields_ = H.matrix2._fields_
        m, m2 = M(), M2()
        assert R.loadHmatrix(SAMPLE.encode(), ctypes.byref(m), ctypes.byref(m2)) == 0
        n = m.nd
        xs = {"sin": np.sin(np.arange(1, n + 1, dtype=np.float64)),
              "rand": np.random.default_rng(20261202).uniform(-1, 1, n)}
        dp = ctypes.POINTER(ctypes.c_double)
        data = {"nd": n, "nlf": m.nlf, "ktmax": m.ktmax, "len": m2.len}
        for key, x in xs.items():
            data["x_" + key] = x
            for var, fn, mm in (("1", "hmvm_seq_1", m), ("1t", "hmvm_seq_1t", m),
                                ("2", "hmvm_seq_2", m2), ("2t", "hmvm_seq_2t", m2)):
                v = np.zeros(n)
                getattr(R, fn)(v.ctypes.data_as(dp), ctypes.cast(ctypes.byref(mm), ctypes.c_void_p),
                               x.ctypes.data_as(dp))
                data["y_%s_%s" % (key, var)] = v
        np.savez(os.path.join(HERE, "sample_y.npz"), **data)
        print("sample nd", n, "nlf", m.nlf, "len", m2.len)


if __name__ == "__main__":
    main()
