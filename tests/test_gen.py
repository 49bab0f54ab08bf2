"""Generator invariants (CPU): tiling, ordering, admissibility, ACA accuracy."""
import numpy as np
import pytest


@pytest.mark.parametrize("geometry,n,eps,leaf", [("sphere", 2000, 1e-4, 32), ("cube", 2400, 1e-6, 16),
                                                 ("cube", 1000, 1e-4, 32)])
def test_generator_invariants(H, geometry, n, eps, leaf):
    h = H.generate(n, geometry=geometry, eps=eps, leaf=leaf)
    a = h.arrays()
    nlf = h.mat2.nlf
    # the leaves tile the n x n matrix exactly once
    cover = np.zeros((n, n), dtype=np.int8)
    for i in range(nlf):
        r0, c0 = a["nstrtl"][i] - 1, a["nstrtt"][i] - 1
        cover[r0:r0 + a["ndl"][i], c0:c0 + a["ndt"][i]] += 1
    assert cover.min() == 1 and cover.max() == 1
    # sorted by (nstrtl, nstrtt)  (m_HACApK_base.f90:245-255)
    key = a["nstrtl"].astype(np.int64) * (n + 1) + a["nstrtt"]
    assert np.all(np.diff(key) > 0)
    # low-rank leaves: both sides larger than the leaf size; dense leaves have kt 0
    lr = a["ltmtx"] == 1
    assert np.all(a["ndl"][lr] > leaf) and np.all(a["ndt"][lr] > leaf)
    assert np.all(a["kt"][~lr] == 0) and np.all(a["kt"][lr] >= 1)
    assert h.mat2.ktmax == a["kt"].max()
    # offsets follow the reference packing (loadmatrix.c:98-128)
    off = 0
    for i in range(nlf):
        assert a["a1"][i] == off
        if lr[i]:
            off += a["kt"][i] * a["ndt"][i]
            assert a["a2"][i] == off
            off += a["kt"][i] * a["ndl"][i]
        else:
            off += a["ndl"][i] * a["ndt"][i]
    assert off == h.mat2.len
    # ACA accuracy (HACApK_accuracy_leafmtx): ||A - H||_F / ||A||_F near eps
    acc = H.accuracy(h)
    assert 0 < acc <= 3 * eps


def test_admissibility_rule(H):
    """Every low-rank leaf satisfies min(width_l, width_t) <= eta * dist of the
    bounding boxes of its point sets (m_HACApK_base.f90:870)."""
    n = 1500
    h = H.generate(n, geometry="sphere", eps=1e-4, leaf=16)
    pts = np.zeros(3 * n)
    diag = np.zeros(1)
    import ctypes
    H.lib.hgen_points(ctypes.byref(h.params), pts.ctypes.data_as(H.c_double_p), diag.ctypes.data_as(H.c_double_p))
    P = pts.reshape(n, 3)
    a = h.arrays()
    for i in np.nonzero(a["ltmtx"] == 1)[0]:
        L = P[a["nstrtl"][i] - 1:a["nstrtl"][i] - 1 + a["ndl"][i]]
        T = P[a["nstrtt"][i] - 1:a["nstrtt"][i] - 1 + a["ndt"][i]]
        wl = np.linalg.norm(L.max(0) - L.min(0))
        wt = np.linalg.norm(T.max(0) - T.min(0))
        gap = np.maximum(0, np.maximum(T.min(0) - L.max(0), L.min(0) - T.max(0)))
        assert min(wl, wt) <= 2.0 * np.linalg.norm(gap) + 1e-12
    assert abs(diag[0] - 1 / (4 * np.pi * np.sqrt(4 * np.pi / n))) < 1e-12


def test_generator_deterministic(H):
    ha, hb = H.generate(900, geometry="cube"), H.generate(900, geometry="cube")
    a, b = ha.arrays(), hb.arrays()  # views: keep ha/hb alive
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_cube_grid_shape(H):
    """N = 101,250 splits as 6 faces x 135 x 125 (distinct jittered points)."""
    import ctypes
    n = 6 * 135 * 125
    p = H.gen_params(n, geometry="cube")
    pts = np.zeros(3 * n)
    H.lib.hgen_points(ctypes.byref(p), pts.ctypes.data_as(H.c_double_p), None)
    P = pts.reshape(n, 3)
    on_face = np.isclose(P, 0.0) | np.isclose(P, 1.0)
    assert np.all(on_face.any(axis=1))
    assert len(np.unique(np.round(P, 12), axis=0)) == n
