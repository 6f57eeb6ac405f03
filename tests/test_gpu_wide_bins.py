"""Bin counts past the LDS kernels (PARENT.cpp:211-238 accepts any positive
count; the histogram is a long long histo[B_i][biggest], PARENT.cpp:790):
2D counts whose histogram does not fit LDS (from about 200 bins a side) and
counts above 255 (uint16 indices) run on the slab kernel.

  * the reference's own .par at 210 256 300 500 256 210 bins
    (tests/golden/chain12_wide.par, written by PARENT.x) through the engine;
  * bin indices (uint8 at 210, uint16 above 255) bit-exact vs the oracle;
  * counts of the slab kernel bit-exact vs the oracle's histograms, one slab
    and many slabs;
  * S2D / MI vs the oracle for the sweep and the batch paths at such bins.
Bars as in test_gpu_parity.py.
"""
import os

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

MI_REL, MI_ABS, S_ABS = 1e-6, 1e-11, 1e-10
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pa():
    import parent_amd
    return parent_amd


def _par_body(path, nD):
    D = 3 * nD + 3
    raw = open(path, "rb").read()
    n = D + D * (D - 1) // 2
    body = np.frombuffer(raw[len(raw) - 8 * n:], dtype=np.float64)
    return body[:D].copy(), body[D:].copy()


def _mi(nD, s1, s2):
    D = 3 * nD + 3
    gi, gj = np.triu_indices(D, 1)
    pos = np.array([O.lib().oracle_par_pos(nD, int(a), int(b)) for a, b in zip(gi, gj)])
    return s1[gi] + s1[gj] - s2[pos]


def _check(nD, s1, s2, s1o, s2o, label):
    d1, d2 = np.max(np.abs(s1 - s1o)), np.max(np.abs(s2 - s2o))
    mi, mio = _mi(nD, s1, s2), _mi(nD, s1o, s2o)
    err = np.abs(mi - mio)
    print(f"{label}: max|dS1D| {d1:.2e} max|dS2D| {d2:.2e} max|dMI| {err.max():.2e}", flush=True)
    assert d1 <= 1e-12 and d2 <= S_ABS
    assert np.all(err <= MI_REL * np.abs(mio) + MI_ABS)


def test_reference_par_at_wide_bins():
    """PARENT_hip's library path on the golden .bat at 210 256 300 500 256 210
    bins against the .par the reference PARENT.x wrote for the same command."""
    pa = _pa()
    bins = [210, 256, 300, 500, 256, 210]
    nD = 9
    s1g, s2g = _par_body(os.path.join(GOLD, "chain12_wide.par"), nD)
    with pa.Engine(0) as eng, pa.Bat(os.path.join(GOLD, "chain12_1000.bat")) as bat:
        eng.set_trajectory(bat.read_frames())
        eng.prepare(bins)
        eng.pairs(pa.METHOD_AUTO)
        assert eng.stats()["method"] == pa.METHOD_SLABS
        _check(nD, eng.s1d(), eng.s2d(), s1g, s2g, "wide golden")
        edges, mi, left, tot = eng.mist()
        eo, mo, lo, to = O.mist(nD, s1g, s2g)
        assert np.array_equal(edges, eo) and np.array_equal(left, lo)


@pytest.mark.parametrize("bins", [[210] * 6, [256] * 6, [7, 300, 13, 500, 29, 260], [1000, 60, 2000, 700, 40, 1200]])
def test_wide_indices_and_counts_bit_exact(bins):
    pa = _pa()
    atoms, frames, seed = 14, 3001, 12
    nD = atoms - 3
    D = 3 * nD + 3
    vals = pa.synth_values(atoms, frames, seed)
    s1o, s2o, mno, mxo, _, shifted = O.compute(nD, frames, vals.T.astype(np.float64), bins)
    rng = np.random.default_rng(seed)
    pairs = {(0, 1), (0, D - 1), (D - 2, D - 1), (1, nD + 2), (nD + 2, 2 * nD + 3), (2 * nD + 3, 2 * nD + 4)}
    while len(pairs) < 24:
        a, b = sorted(rng.choice(D, 2, replace=False))
        pairs.add((int(a), int(b)))
    pairs = sorted(pairs)
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.prepare(bins)
        idx = eng.bins16()
        for g in range(D):
            t = O.type_of(nD, g)
            ref, bad = O.bin_index(shifted[g], mno[g], mxo[g], bins[3 + t])
            assert bad == 0 and np.array_equal(idx[g].astype(np.int32), ref), f"indices of dof {g}"
        if max(bins[3:]) > 255:
            with pytest.raises(pa.ParentError):
                eng.bins()
        for rows in (0, 3):  # as many rows per slab as the LDS holds; 3 rows per slab
            eng.set_variant(pa.METHOD_SLABS, rows)
            got = eng.pair_counts(pairs, pa.METHOD_SLABS, bins[3:])
            for (gi, gj), h in zip(pairs, got):
                ti, tj = O.type_of(nD, gi), O.type_of(nD, gj)
                _, ho = O.entropy2d(ti, shifted[gi], mno[gi], mxo[gi], bins[3 + ti],
                                    tj, shifted[gj], mno[gj], mxo[gj], bins[3 + tj], counts=True)
                assert np.array_equal(h.astype(np.int64), ho), f"counts of {(gi, gj)}, rows {rows}"
            eng.pairs(pa.METHOD_AUTO)
            assert eng.stats()["method"] == pa.METHOD_SLABS
            _check(nD, eng.s1d(), eng.s2d(), s1o, s2o, f"bins {bins} rows {rows}")
        eng.set_variant(pa.METHOD_SLABS, -1)


def test_slabs_at_small_bins_equal_plain():
    """The slab kernel forced at 50 bins, 1 and 7 rows a slab (50 and 8
    slabs), against the plain kernel."""
    pa = _pa()
    vals = pa.synth_values(20, 4000, 3)
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.prepare([50] * 6)
        eng.pairs(pa.METHOD_PLAIN)
        ref = eng.s2d()
        for rows in (1, 7, 0):
            eng.set_variant(pa.METHOD_SLABS, rows)
            eng.pairs(pa.METHOD_SLABS)
            assert eng.stats()["method"] == pa.METHOD_SLABS
            assert np.max(np.abs(eng.s2d() - ref)) <= 1e-13, f"rows {rows}"
        eng.set_variant(pa.METHOD_SLABS, -1)


def test_sweep_and_batch_at_wide_bins():
    pa = _pa()
    sets = [[50] * 6, [256] * 6, [220, 30, 300, 210, 40, 260]]
    vals = pa.synth_values(13, 2500, 4)
    nD = 10
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.sweep(sets, pa.METHOD_AUTO)
        eng.sweep(sets, pa.METHOD_AUTO)  # a replay of the captured graph
        for k, b in enumerate(sets):
            eng.select(k)
            s1o, s2o, *_ = O.compute(nD, vals.shape[0], vals.T.astype(np.float64), b)
            _check(nD, eng.s1d(), eng.s2d(), s1o, s2o, f"sweep set {b}")
    shapes = [(9, 1200, 1), (15, 3000, 2), (11, 700, 3)]
    bins = [300] * 6
    with pa.Batch(bins) as bt:
        trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
        for v in trajs:
            bt.add(v)
        bt.run(pa.METHOD_AUTO)
        assert bt.stats()["method"] == pa.METHOD_SLABS
        for k, v in enumerate(trajs):
            nDk = v.shape[1] // 3 - 1
            s1, s2 = bt.results(k)
            s1o, s2o, *_ = O.compute(nDk, v.shape[0], v.T.astype(np.float64), bins)
            _check(nDk, s1, s2, s1o, s2o, f"batch trajectory {k}")
