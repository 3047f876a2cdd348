"""GPU hashing tables and jump maps: bit-exact against the oracle and the
golden fixtures, and through closed-form size identities and sampled brute
force at the full N=12/M=12 and N=16/M=12 shapes."""
import math
import os

import numpy as np
import pytest

import mcpi_amd
import mcpi_amd.reference as R
import oracle
from tests import packed
from tests.gpu_util import note

pytestmark = pytest.mark.gpu

SMALL = [(1, 1), (0, 3), (1, 4), (2, 3), (3, 3), (5, 5), (4, 6), (6, 4), (3, 7), (7, 3), (2, 9)]


@pytest.mark.parametrize("N,M", SMALL)
def test_tables_bit_exact(oracle_lib, N, M):
    note(f"basis tables N={N} M={M}")
    g = mcpi_amd.Basis(N, M, "full")
    o = oracle.Basis(N, M)
    assert g.nc == o.nc
    assert np.array_equal(g.ncmat(), o.NCmat)
    assert np.array_equal(g.IF(), o.IF)
    assert np.array_equal(g.Map(), o.Map)
    s, m = g.OT()
    assert np.array_equal(s, o.strideOT) and np.array_equal(m, o.MapOT)
    s, m = g.TT()
    assert np.array_equal(s, o.strideTT) and np.array_equal(m, o.MapTT)
    for maps in ("single", "none"):
        h = mcpi_amd.Basis(N, M, maps)
        assert np.array_equal(h.IF(), o.IF)
        if maps == "single":
            assert np.array_equal(h.Map(), o.Map)


@pytest.mark.parametrize("N,M", [(2, 3), (5, 5), (4, 6), (6, 4), (3, 7), (7, 3), (6, 9), (1, 4), (0, 3), (1, 1)])
def test_packed_double_jump_tables(oracle_lib, N, M):
    """Wave-packed OT/TT tables equal the oracle's reference-layout maps
    re-laid out (tests/packed.py), every entry and every offset."""
    note(f"packed tables N={N} M={M}")
    g = mcpi_amd.Basis(N, M, "full")
    o = oracle.Basis(N, M)
    nw = (g.nc + 63) // 64
    for tt, (Mp, st) in ((False, (o.MapOT, o.strideOT)), (True, (o.MapTT, o.strideTT))):
        pack, poff = packed.expected_packed(o.IF.reshape(-1, M), M, Mp, st, tt)
        name = "TT" if tt else "OT"
        got_off = g.copy_range("poff" + name, 0, nw + 1)
        assert np.array_equal(got_off, poff), name
        n = g.ptt_len if tt else g.pot_len
        assert n == pack.size
        assert np.array_equal(g.copy_range("pack" + name, 0, n), pack), name


def test_golden_tables(oracle_lib):
    gd = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1_perf.npz"))
    note("golden cfg1 tables")
    g = mcpi_amd.Basis(5, 5, "full")
    assert np.array_equal(g.IF(), gd["IF"]) and np.array_equal(g.Map(), gd["Map"])
    s, m = g.OT()
    assert np.array_equal(m, gd["MapOT"]) and np.array_equal(s, gd["strideOT"])
    s, m = g.TT()
    assert np.array_equal(m, gd["MapTT"]) and np.array_equal(s, gd["strideTT"])


def test_rank_unrank_api(oracle_lib):
    note("rank/unrank API N=8 M=9")
    N, M = 8, 9
    g = mcpi_amd.Basis(N, M, "none")
    rng = np.random.default_rng(0)
    idx = rng.integers(0, g.nc, 5000)
    v = g.unrank(idx)
    T = oracle.ncmat(N, M)
    for k, row in zip(idx[:300], v[:300]):
        assert np.array_equal(row, oracle.unrank(int(k), N, M))
    assert np.array_equal(g.rank(v), idx)


def test_reference_shaped_builders(oracle_lib):
    for (N, M) in [(5, 5), (4, 6), (1, 4), (6, 3)]:
        note(f"reference-shaped builders N={N} M={M}")
        r = R.RefBasis(N, M)
        o = oracle.Basis(N, M)
        for name in ("NCmat", "IF", "Map", "MapOT", "MapTT"):
            assert np.array_equal(getattr(r, name), getattr(o, name)), name
        assert np.array_equal(r.strideOT, o.strideOT) and np.array_equal(r.strideTT, o.strideTT)


def _sample_rows_check(g, N, M, rows, check_double=True):
    """Brute-force check of the map entries of a set of rows: every entry is
    the oracle rank of the moved occupation vector."""
    T = oracle.ncmat(N, M)
    nc = g.nc
    IF = np.concatenate([g.copy_range("IF", int(i) * M, M) for i in rows]).reshape(-1, M)
    for i, v in zip(rows, IF):
        assert np.array_equal(v, oracle.unrank(int(i), N, M))
        for k in range(M):
            col = [g.copy_range("Map", int(i) + (k + l * M) * nc, 1)[0] for l in range(M)]
            for l in range(M):
                if v[k] < 1:
                    assert col[l] == -1
                else:
                    w = v.copy(); w[k] -= 1; w[l] += 1
                    assert col[l] == oracle.rank(N, M, T, w)
        if not check_double:
            continue
        for name, chunk_of in (("OT", lambda v: [(k, k) for k in range(M) if v[k] >= 2]),
                               ("TT", lambda v: [(k, s) for k in range(M) for s in range(k + 1, M)
                                                 if v[k] >= 1 and v[s] >= 1])):
            st = g.copy_range("stride" + name, int(i), 2)
            chunks = chunk_of(v)
            assert st[1] - st[0] == len(chunks) * M * M
            data = g.copy_range("Map" + name, int(st[0]), int(st[1] - st[0]))
            for c, (k, s) in enumerate(chunks):
                for e in range(M * M):
                    l, q = e % M, e // M
                    w = v.copy(); w[k] -= 1; w[s] -= 1; w[l] += 1; w[q] += 1
                    assert data[c * M * M + e] == oracle.rank(N, M, T, w)


def test_full_maps_n12_m12(oracle_lib):
    """N=12, M=12 (nc 1,352,078): MapTT has 3.35e9 entries (beyond INT32)."""
    N, M = 12, 12
    note("full maps N=12 M=12 build")
    g = mcpi_amd.Basis(N, M, "full")
    nc2 = math.comb(N - 2 + M - 1, N - 2)
    assert g.nc == 1352078
    assert g.ot_len == M * nc2 * M * M == 609493248
    assert g.tt_len == math.comb(M, 2) * nc2 * M * M == 3352212864
    last = g.copy_range("strideTT", g.nc - 1, 2)
    assert last[0] == last[1] == g.tt_len
    rng = np.random.default_rng(12)
    rows = np.concatenate([[0, 1, g.nc - 2, g.nc - 1], rng.integers(0, g.nc, 40)])
    note("full maps N=12 M=12 sampled brute force")
    _sample_rows_check(g, N, M, rows)
    # wave-packed tables: sampled wavefronts re-laid out from the reference
    # layout maps (pinned above), and the totals
    note("packed tables N=12 M=12 sampled wavefronts")
    nw = (g.nc + 63) // 64
    for name, tt, n_dbl in (("OT", False, M * nc2), ("TT", True, math.comb(M, 2) * nc2)):
        poff = g.copy_range("poff" + name, 0, nw + 1)
        assert poff[0] == 0 and poff[-1] == (g.ptt_len if tt else g.pot_len)
        assert np.all(np.diff(poff) >= 0)
        assert poff[-1] == n_dbl * packed.entries(M, tt)   # one slot row per reference chunk
        for w in np.concatenate([[0, nw - 1], rng.integers(0, nw, 12)]):
            r0, r1 = 64 * int(w), min(64 * int(w) + 64, g.nc)
            rows_w = list(range(r0, r1))
            IFw = g.copy_range("IF", r0 * M, (r1 - r0) * M).reshape(-1, M)
            IFfull = np.zeros((r1, M), np.int32)
            IFfull[r0:r1] = IFw
            st = g.copy_range("stride" + name, r0, r1 - r0 + 1)
            Mp = g.copy_range("Map" + name, int(st[0]), int(st[-1] - st[0]))
            exp = packed.expected_wave(IFfull, M, rows_w, Mp, st - st[0], r0, tt)
            got = g.copy_range("pack" + name, int(poff[w]), int(poff[w + 1] - poff[w]))
            assert np.array_equal(got, exp), (name, w)


def test_single_map_n16_m12(oracle_lib):
    """N=16, M=12 (nc 13,037,895), Map only (7.5 GB)."""
    N, M = 16, 12
    note("single map N=16 M=12 build")
    g = mcpi_amd.Basis(N, M, "single")
    assert g.nc == 13037895 and g.map_len == g.nc * M * M
    rng = np.random.default_rng(16)
    rows = np.concatenate([[0, g.nc - 1], rng.integers(0, g.nc, 30)])
    _sample_rows_check(g, N, M, rows, check_double=False)
