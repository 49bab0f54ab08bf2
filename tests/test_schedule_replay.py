"""CPU replay of the exported tile schedule (host logic of the HIP path).

The schedule (repack records, tiles, panels, segment sums) is executed here in
numpy exactly as the kernels interpret it; the result must equal the oracle.
This checks the layout and the bookkeeping of the plan builder without a GPU."""
import struct

import numpy as np
import pytest


def replay(H, h, x, source="rowmat", multi=False, **kw):
    """multi: as the multi-RHS launches run it (stage-1 rows through
    map_off_multi; the t sum skips the entries those maps write)."""
    s = H.schedule(h, source=source, **kw)
    a = h.arrays()
    src = a["rowmat_t"] if source == "rowmat_t" else a["rowmat"]
    vals = np.full(s["nvals"], np.nan)  # padding must never be read
    for r in s["repack"]:
        rr, cc = np.meshgrid(np.arange(r["rows"]), np.arange(r["cols"]), indexing="ij")
        vals[r["dst_off"] + rr + cc * r["dst_cs"]] = src[r["src_off"] + rr * r["src_rs"] + cc * r["src_cs"]]
    slots = np.zeros(s["n_tslots"] + s["n_zslots"])
    t = np.zeros(max(s["n_t"], 1))
    P = s["panels"]

    def vector(p, ncols):
        j = np.arange(ncols)
        if p["vec_sel"] == 0:
            return x[p["vec_off"] + j]
        if p["vec_sel"] == 1:
            return t[p["vec_off"] + j]
        nch, stride = p["vec_chunks"] >> 16, p["vec_chunks"] & 0xFFFF
        return sum(slots[p["vec_off"] + c * stride + j] for c in range(nch))

    def run(tiles):
        for td in tiles:
            acc = np.zeros(td["nrows"])
            for p in P[td["panel0"]:td["panel0"] + td["npanel"]]:
                rr, jj = np.meshgrid(np.arange(td["nrows"]), np.arange(p["ncols"]), indexing="ij")
                blk = vals[p["val_off"] + rr + jj * p["ld"]]
                assert not np.isnan(blk).any()
                acc += blk @ vector(p, p["ncols"])
            m = td["map_off_multi"] if multi else td["map_off"]
            if m >= 0:  # stage-1 rows straight into t
                t[s["tmap"][m:m + td["nrows"]]] = acc
            else:
                slots[td["out_off"]:td["out_off"] + td["nrows"]] = acc

    def segsum(ss, out, skip=()):
        for i in range(ss["n"]):
            sg = ss["seg"][i]
            assert sg >= 0
            if i not in skip:
                out[i] = sum(slots[e + i] for e in ss["list"][ss["ptr"][sg]:ss["ptr"][sg + 1]])
        return out

    run(s["tiles_a"])
    if not s["t_fused"] and not s["t_direct"]:
        segsum(s["tsum"], t, set(s["tmap"].tolist()) if multi else ())
    run(s["tiles_b"])
    return segsum(s["ysum"], np.zeros(s["ysum"]["n"])), s


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kw", [{}, {"w1": 64, "w2": 64}, {"w1": 500, "w2": 3000, "order": 1},
                                {"w1": 1 << 20, "w2": 1 << 20}, {"w2": 3000, "w3": 64}])
@pytest.mark.parametrize("source", ["rowmat", "rowmat_t"])
@pytest.mark.parametrize("t_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("multi", [False, True])
def test_replay_matches_oracle(H, O, kw, source, t_mode, multi):
    h = H.generate(1500, geometry="cube", eps=1e-6)
    x = np.random.default_rng(3).uniform(-1, 1, h.nd)
    y, s = replay(H, h, x, source, multi=multi, t_mode=t_mode, **kw)
    if t_mode == 0:
        assert s["t_fused"]   # every stage-2 panel sums its partial slots
    if t_mode in (1, 3):
        assert not s["t_fused"] and s["n_t"] == int(h.arrays()["kt"][h.arrays()["ltmtx"] == 1].sum())
    if t_mode == 3:  # no partial slots, no t sum: every t entry written once by a stage-1 tile
        assert s["t_direct"] and s["n_tslots"] == 0 and s["tsum"]["n"] == 0
        assert np.array_equal(np.sort(s["tmap"]), np.arange(s["n_t"]))
    if t_mode in (1, 2) and not s["t_fused"]:
        # the one-vector launches write no t through a map; the multi-RHS row
        # maps name each t entry at most once
        assert all(td["map_off"] < 0 for td in s["tiles_a"])
        assert len(np.unique(s["tmap"])) == len(s["tmap"])
        if t_mode == 1 and not kw:  # default widths: the one-chunk row groups write t themselves
            assert 0 < len(s["tmap"]) < s["n_t"]
    assert rel(y, O.seq_2(h, x)) <= 1e-13


def test_replay_sphere_and_dense(H, O):
    for h in (H.generate(2500, geometry="sphere", eps=1e-4, leaf=16), H.dummymatrix(40, 40, 6)):
        x = np.random.default_rng(5).uniform(-1, 1, h.nd)
        y, _ = replay(H, h, x)
        assert rel(y, O.seq_2(h, x)) <= 1e-13


def write_format1(path, nd, leaves, rng):
    """leaves: (ltmtx, ndl, ndt, nstrtl, nstrtt, kt); random factor values."""
    ktmax = max([lf[5] for lf in leaves if lf[0] == 1] + [0])

    def rec(b):
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    out = [rec(struct.pack("<3i", nd, len(leaves), ktmax))]
    for lt, ndl, ndt, nl, nt, kt in leaves:
        out.append(rec(struct.pack("<6i", lt, ndl, ndt, nl, nt, kt)))
        if lt == 1:
            out.append(rec(rng.uniform(-1, 1, kt * ndt).tobytes()))
            out.append(rec(rng.uniform(-1, 1, kt * ndl).tobytes()))
        else:
            out.append(rec(rng.uniform(-1, 1, ndl * ndt).tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))


def big_rank_matrix(H, path):
    """Two off-diagonal low-rank leaves of rank 300 (> 256 rank rows of one
    tile) and two dense diagonal leaves: stage 2 must read a materialised t."""
    n = 700
    write_format1(path, n, [(2, 350, 350, 1, 1, 0), (1, 350, 350, 1, 351, 300),
                            (1, 350, 350, 351, 1, 300), (2, 350, 350, 351, 351, 0)], np.random.default_rng(0))
    return H.loadHmatrix(path)


def test_replay_rank_rows_straddle(H, O, tmp_path):
    h = big_rank_matrix(H, str(tmp_path / "big.bin"))
    x = np.random.default_rng(6).uniform(-1, 1, h.nd)
    y, s = replay(H, h, x)
    assert not s["t_fused"]
    assert rel(y, O.seq_2(h, x)) <= 1e-13
