"""The drop-in from compiled callers: a C program with the reference's
hmvm1.c flow and a C++ program with the hmvm1.cpp flow, linked against the
library.  CPU: they compile and link.  GPU: their dumps match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hacapk_hmvm_amd")
DRV = os.path.join(ROOT, "tests", "drivers")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def build(tmp, name, compiler, std):
    exe = os.path.join(tmp, name)
    src = os.path.join(DRV, name + (".c" if compiler == "gcc" else ".cpp"))
    subprocess.run([compiler, std, "-O2", "-I" + INC, src, "-o", exe, "-L" + LIBDIR, "-lhmvm_hip",
                    "-Wl,-rpath," + LIBDIR, "-lm"], check=True)
    return exe


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("drv"))
    return {"c": build(tmp, "hmvm_c_driver", "gcc", "-std=c99"),
            "cpp": build(tmp, "hmvm_cpp_driver", "g++", "-std=c++14")}


def test_drivers_compile_and_link(drivers):
    for exe in drivers.values():
        assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_c_driver_dumps_match_oracle(gpu, H, O, drivers, tmp_path):
    f = os.path.join(GOLDEN, "sphere320_1.bin")
    r = subprocess.run([drivers["c"], f, "1"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("TIME 10 hmvm_hip_") == 4
    h = H.loadHmatrix(f)
    b = np.sin(np.arange(1, h.nd + 1, dtype=np.float64))
    ref = O.seq_2(h, b)
    for var in ("1", "1t", "2", "2t"):
        got = np.array([float(v) for v in open(tmp_path / ("result_hip_%s_d.txt" % var))])
        want = np.array([float("%.3E" % v) for v in ref])
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)) <= 1e-3
    twice = np.loadtxt(tmp_path / "result_hip_2x2_d.txt")
    assert np.linalg.norm(twice - 2 * ref) <= 1e-13 * np.linalg.norm(2 * ref)


@pytest.mark.gpu
def test_cpp_driver_dumps_match_oracle(gpu, H, O, drivers, tmp_path):
    f = os.path.join(GOLDEN, "cube384_1.bin")
    r = subprocess.run([drivers["cpp"], f, "1", "3"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("TIME 3 hmvm_hip_k") == 4  # k = -1, 2, 5 in FP64, -1 in FP32
    h = H.loadHmatrix(f)
    b = np.sin(np.arange(1, h.nd + 1, dtype=np.float64))
    ref = O.seq_2(h, b)
    for k in (-1, 2, 5):
        got = np.array([float(v) for v in open(tmp_path / ("result_hip_k%d_d.txt" % k))])
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-3  # %.3E dump
    gotf = np.array([float(v) for v in open(tmp_path / "result_hip_k-1_f.txt")])
    assert np.max(np.abs(gotf - ref) / np.abs(ref)) <= 2e-3


def read_convert_dump(path):
    recs = []
    with open(path, "rb") as f:
        count = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        for _ in range(count):
            n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            d = np.frombuffer(f.read(8 * n), dtype=np.float64)
            s = np.frombuffer(f.read(4 * n), dtype=np.float32)
            same = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
            recs.append((d, s, same))
    return recs


def test_convertD2F_rounds_every_element(H, tmp_path):
    """The product convertD2F (C++ API) on a matrix whose values are the
    IEEE tie, subnormal and overflow cases, and on the golden files (low-rank
    leaves: a1, a2 and a2t): every converted element of rowmat, rowmat_t and
    the per-leaf arrays is the double rounded to the nearest float32, and the
    descriptors and offsets are unchanged (src_cpp/loadmatrix.cpp:252-342)."""
    from f32_rounding import CASES, case_values, check, random_values
    exe = str(tmp_path / "convert_dump")
    subprocess.run(["g++", "-std=c++14", "-O2", "-I" + INC, os.path.join(DRV, "convert_dump.cpp"), "-o", exe,
                    "-L" + LIBDIR, "-lhmvm_hip", "-Wl,-rpath," + LIBDIR], check=True)
    d = H.dummymatrix(4, 4, 4)  # 16 dense 4 x 4 leaves, 256 values
    vals = np.concatenate([case_values(), random_values(256 - len(CASES), 9)])
    vals[-1] = np.nan
    a = d.arrays()
    a["rowmat"][:] = vals
    special = str(tmp_path / "special_1.bin")
    H.writeHmatrix1(special, d)
    files = [special, os.path.join(GOLDEN, "sphere320_1.bin"), os.path.join(GOLDEN, "cube384_1.bin")]
    for f in files:
        out = str(tmp_path / "conv.bin")
        r = subprocess.run([exe, f, out], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
        assert r.returncode == 0, r.stdout
        recs = read_convert_dump(out)
        assert len(recs) >= 2
        if f == special:  # the loader handed over the crafted values
            assert np.array_equal(recs[0][0].view(np.uint64)[:-1], vals.view(np.uint64)[:-1])
            assert np.isnan(recs[0][0][-1])
        for src, got, same in recs:
            assert same == 1
            check(src, got)
    recs = read_convert_dump(str(tmp_path / "conv.bin"))  # the golden file has low-rank leaves: 3 records each
    assert any(len(recs[i][0]) != len(recs[i + 1][0]) for i in range(2, len(recs) - 1))
