"""HIP engine vs CPU oracle on the same seeded trajectories (needs an MI355X).

Bars:
  * ranges (min, max, modfit shift) and every bin index: bit-exact;
  * histogram counts of every checked pair, each fill path: bit-exact;
  * S1D: same formula as the reference; only libm's log/sin may differ by an ulp,
    so |dS1D| <= 1e-12;
  * S2D: |dS2D| <= 1e-10 and per-pair MI within 1e-6 relative of the FP64
    oracle with an absolute floor of 1e-11 (MI ~ 0 for independent pairs);
  * MIST from identical S1D/S2D: identical edge list, MI bits and totals.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

MI_REL, MI_ABS, S_ABS = 1e-6, 1e-11, 1e-10


def _pa():
    import parent_amd
    return parent_amd


def _case(atoms, frames, seed):
    pa = _pa()
    vals = pa.synth_values(atoms, frames, seed)
    return vals, atoms - 3


def _oracle_full(vals, nD, bins):
    N = vals.shape[0]
    return O.compute(nD, N, vals.T.astype(np.float64), bins)


def _mi(nD, s1, s2):
    D = 3 * nD + 3
    gi, gj = np.triu_indices(D, 1)
    pos = np.array([O.lib().oracle_par_pos(nD, int(a), int(b)) for a, b in zip(gi, gj)])
    return s1[gi] + s1[gj] - s2[pos]


@pytest.fixture(scope="module")
def engine():
    pa = _pa()
    print("engine: create on device 0", flush=True)
    e = pa.Engine(0)
    yield e
    e.close()


CASES = [
    (12, 1000, 3, [50] * 6),
    (30, 5000, 11, [50] * 6),
    (20, 3000, 5, [7, 13, 29, 5, 40, 17]),
    (25, 4097, 8, [10] * 6),
    (16, 2000, 9, [100, 100, 100, 100, 100, 100]),
]


@pytest.mark.parametrize("atoms,frames,seed,bins", CASES)
def test_prepare_bit_exact(engine, atoms, frames, seed, bins):
    print(f"prepare atoms={atoms} frames={frames} seed={seed} bins={bins}", flush=True)
    vals, nD = _case(atoms, frames, seed)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    s1o, _, mno, mxo, mfo, shifted = _oracle_full(vals, nD, bins)
    mn, mx, mf = engine.ranges()
    assert np.array_equal(mn.view(np.uint64), mno.view(np.uint64))
    assert np.array_equal(mx.view(np.uint64), mxo.view(np.uint64))
    assert np.array_equal(mf.view(np.uint64), mfo.view(np.uint64))
    gb = engine.bins()
    D = 3 * nD + 3
    for g in range(D):
        t = O.type_of(nD, g)
        idx, bad = O.bin_index(shifted[g], mno[g], mxo[g], bins[3 + t])
        assert bad == 0
        assert np.array_equal(gb[g].astype(np.int32), idx), f"bin indices differ for dof {g}"
    s1 = engine.s1d()
    assert np.max(np.abs(s1 - s1o)) <= 1e-12


def test_fp64_trajectory_vs_oracle(engine):
    """An FP64 trajectory (values no FP32 file could hold) through the double
    instantiations of the scans and the pre-bin pass, with 1D bin counts that
    differ from the 2D ones for bonds and angles (a second index set) and
    equal for dihedrals: ranges and every index bit-exact, S1D and S2D within
    tolerance of the FP64 oracle."""
    pa = _pa()
    atoms, frames, bins = 30, 5001, [50, 40, 30, 20, 25, 30]
    print(f"fp64 atoms={atoms} frames={frames} bins={bins}", flush=True)
    vals = pa.synth_values(atoms, frames, 12).astype(np.float64)
    vals += np.linspace(0, 1e-7, vals.shape[1])
    nD = atoms - 3
    engine.set_trajectory(vals)
    engine.prepare(bins)
    s1o, s2o, mno, mxo, mfo, shifted = O.compute(nD, frames, vals.T.copy(), bins)
    mn, mx, mf = engine.ranges()
    assert np.array_equal(mn.view(np.uint64), mno.view(np.uint64))
    assert np.array_equal(mx.view(np.uint64), mxo.view(np.uint64))
    assert np.array_equal(mf.view(np.uint64), mfo.view(np.uint64))
    gb = engine.bins()
    for g in range(3 * nD + 3):
        t = O.type_of(nD, g)
        idx, bad = O.bin_index(shifted[g], mno[g], mxo[g], bins[3 + t])
        assert bad == 0
        assert np.array_equal(gb[g].astype(np.int32), idx), f"2D bin indices differ for dof {g}"
    assert np.max(np.abs(engine.s1d() - s1o)) <= 1e-12
    engine.pairs(pa.METHOD_AUTO)
    assert np.max(np.abs(engine.s2d() - s2o)) <= S_ABS


def test_modfit_shift_of_crafted_and_random_histograms(engine):
    """Dihedrals whose 100-bin modfit histograms have hand-checked answers (ties,
    runs through bins 99 and 0 or across bins 63/64, one occupied or one empty
    bin, no empty bin) and 200 random histograms: the modfit shift, the ranges
    and S1D against the oracle, and the crafted ones against their known bin."""
    import modfit_cases as MC
    rng = np.random.default_rng(21)
    cols = [MC.column(c, O.PI, rng) for _, c, _ in MC.CASES]
    cols += [MC.column(MC.random_counts(rng), O.PI, rng) for _ in range(200)]
    nD, N = len(cols), MC.FRAMES
    D = 3 * nD + 3
    vals = np.empty((N, D), dtype=np.float32)
    vals[:, :nD + 2] = rng.uniform(1.0, 1.6, (N, nD + 2))
    vals[:, nD + 2:2 * nD + 3] = rng.uniform(1.5, 2.5, (N, nD + 1))
    vals[:, 2 * nD + 3:] = np.stack(cols, axis=1)
    bins = [50] * 6
    print(f"modfit histograms nD={nD} frames={N}", flush=True)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    s1o, _, mno, mxo, mfo, _ = O.compute(nD, N, vals.T.astype(np.float64), bins)
    mn, mx, mf = engine.ranges()
    for j, (name, _, expect) in enumerate(MC.CASES):
        assert mf[2 * nD + 3 + j] == 2 * O.PI - (expect + 0.5) * MC.bin_width(O.PI), name
    assert np.array_equal(mf.view(np.uint64), mfo.view(np.uint64))
    assert np.array_equal(mn.view(np.uint64), mno.view(np.uint64))
    assert np.array_equal(mx.view(np.uint64), mxo.view(np.uint64))
    assert np.max(np.abs(engine.s1d() - s1o)) <= 1e-12


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("atoms,frames,seed,bins", CASES[:4])
def test_counts_bit_exact(engine, method, atoms, frames, seed, bins):
    print(f"counts method={method} atoms={atoms} frames={frames}", flush=True)
    vals, nD = _case(atoms, frames, seed)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    D = 3 * nD + 3
    rng = np.random.default_rng(seed)
    pairs = set()
    # every type combination, the first/last dofs and random pairs
    pairs.update([(0, 1), (0, D - 1), (D - 2, D - 1), (1, nD + 2), (nD + 2, 2 * nD + 3), (2 * nD + 3, 2 * nD + 4)])
    while len(pairs) < 40:
        a, b = sorted(rng.choice(D, 2, replace=False))
        pairs.add((int(a), int(b)))
    pairs = sorted(pairs)
    got = engine.pair_counts(pairs, method, bins[3:])
    _, _, mno, mxo, _, shifted = _oracle_full(vals, nD, bins)
    for (gi, gj), h in zip(pairs, got):
        ti, tj = O.type_of(nD, gi), O.type_of(nD, gj)
        _, ho = O.entropy2d(ti, shifted[gi], mno[gi], mxo[gi], bins[3 + ti],
                            tj, shifted[gj], mno[gj], mxo[gj], bins[3 + tj], counts=True)
        assert np.array_equal(h.astype(np.int64), ho), f"counts differ for pair {(gi, gj)}"
        assert h.sum() == frames


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("atoms,frames,seed,bins", CASES)
def test_entropies_vs_oracle(engine, method, atoms, frames, seed, bins):
    print(f"entropies method={method} atoms={atoms} frames={frames} bins={bins}", flush=True)
    vals, nD = _case(atoms, frames, seed)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    engine.pairs(method)
    s1, s2 = engine.s1d(), engine.s2d()
    s1o, s2o, *_ = _oracle_full(vals, nD, bins)
    assert np.max(np.abs(s2 - s2o)) <= S_ABS
    mi, mio = _mi(nD, s1, s2), _mi(nD, s1o, s2o)
    err = np.abs(mi - mio)
    assert np.all(err <= MI_REL * np.abs(mio) + MI_ABS), f"max MI err {err.max():.3e}"
    st = engine.stats()
    print(f"  method used {st['method']} variant {st['kernel_variant']} pairs {st['ms_pairs']:.3f} ms "
          f"max|dS2D| {np.max(np.abs(s2 - s2o)):.2e} max|dMI| {err.max():.2e}", flush=True)


def test_every_atomic_variant_bit_exact(engine):
    """Each tile configuration the autotuner can pick: counts bit-exact on a pair
    sample, S2D and MI within tolerance on all pairs (fast and general fills)."""
    pa = _pa()
    for bins, atoms, frames in (([50] * 6, 26, 4097), ([10, 10, 10, 20, 30, 40], 22, 3000),
                                ([75] * 6, 14, 2500), ([150] * 6, 10, 1500), ([10] * 6, 20, 2000),
                                ([30] * 6, 18, 2000), ([20] * 6, 16, 17000),
                                ([7, 13, 29, 5, 40, 17], 18, 3000), ([35] * 6, 12, 2500), ([59] * 6, 10, 2000),
                                # frames past one iteration of the 1,024-thread fills (16,384
                                # frames): the perm + dot2 loop, its prefetch and the 4-frame tail
                                ([50] * 6, 10, 40000), ([35] * 6, 10, 33001),
                                ([7, 13, 29, 5, 40, 17], 10, 20002)):
        vals, nD = _case(atoms, frames, 13)
        engine.set_trajectory(vals)
        engine.prepare(bins)
        s1o, s2o, mno, mxo, _, shifted = _oracle_full(vals, nD, bins)
        D = 3 * nD + 3
        pairs = [(0, 1), (0, D - 1), (5, 6), (nD + 1, nD + 3), (2 * nD + 2, 2 * nD + 3), (D - 2, D - 1), (3, 2 * nD + 5)]
        ref = []
        for gi, gj in pairs:
            ti, tj = O.type_of(nD, gi), O.type_of(nD, gj)
            ref.append(O.entropy2d(ti, shifted[gi], mno[gi], mxo[gi], bins[3 + ti],
                                   tj, shifted[gj], mno[gj], mxo[gj], bins[3 + tj], counts=True)[1])
        tested = 0
        for v in range(engine.variant_count(pa.METHOD_ATOMIC)):
            engine.set_variant(pa.METHOD_ATOMIC, v)
            engine.pairs(pa.METHOD_ATOMIC)
            st = engine.stats()
            if st["method"] != pa.METHOD_ATOMIC:
                continue  # histograms do not fit this variant's slots
            tested += 1
            print(f"  bins={bins} variant {v}", flush=True)
            for h, ho in zip(engine.pair_counts(pairs, pa.METHOD_ATOMIC, bins[3:]), ref):
                assert np.array_equal(h.astype(np.int64), ho), f"variant {v}"
            s1, s2 = engine.s1d(), engine.s2d()
            assert np.max(np.abs(s2 - s2o)) <= S_ABS, f"variant {v}"
        engine.set_variant(pa.METHOD_ATOMIC, -1)
        assert tested >= 1


@pytest.mark.parametrize("atoms,frames", [(6, 1), (6, 2), (8, 5), (10, 16), (10, 17), (12, 63), (14, 200),
                                           (8, 1025)])
def test_short_trajectories_every_kernel(engine, atoms, frames):
    """Trajectories shorter than one 16-frame chunk, one wave of chunks or one
    iteration of the 1,024-thread fills: every atomic variant that fits, the
    MFMA, plain and slab kernels, against the oracle's S2D and a pair's counts."""
    pa = _pa()
    vals, nD = _case(atoms, frames, 29)
    D = 3 * nD + 3
    for bins in ([50] * 6, [5] * 6, [7, 13, 29, 5, 40, 17]):
        print(f"short: atoms={atoms} frames={frames} bins={bins}", flush=True)
        engine.set_trajectory(vals)
        engine.prepare(bins)
        s1o, s2o, mno, mxo, _, shifted = _oracle_full(vals, nD, bins)
        assert np.max(np.abs(engine.s1d() - s1o)) <= 1e-12
        t0, t1 = O.type_of(nD, 0), O.type_of(nD, D - 1)
        ho = O.entropy2d(t0, shifted[0], mno[0], mxo[0], bins[3 + t0],
                         t1, shifted[D - 1], mno[D - 1], mxo[D - 1], bins[3 + t1], counts=True)[1]
        runs = [(pa.METHOD_PLAIN, -1), (pa.METHOD_MFMA, -1), (pa.METHOD_SLABS, -1)]
        runs += [(pa.METHOD_ATOMIC, v) for v in range(engine.variant_count(pa.METHOD_ATOMIC))]
        for method, v in runs:
            if v >= 0:
                engine.set_variant(pa.METHOD_ATOMIC, v)
            engine.pairs(method)
            st = engine.stats()
            if v >= 0 and st["method"] != pa.METHOD_ATOMIC:
                continue  # histograms do not fit this variant's slots
            assert np.max(np.abs(engine.s2d() - s2o)) <= S_ABS, f"method {method} variant {v}"
            h = engine.pair_counts([(0, D - 1)], st["method"], bins[3:])[0]
            assert np.array_equal(h.astype(np.int64), ho), f"method {method} variant {v}"
        engine.set_variant(pa.METHOD_ATOMIC, -1)


MFMA_CASES = [
    ([5] * 6, 30, 3001),
    ([10] * 6, 26, 4097),
    ([2] * 6, 20, 1000),
    ([64] * 6, 12, 1500),
    ([3, 9, 12, 5, 7, 11], 24, 2500),   # mixed types: 12-row dof slots, per-type cells
    ([20] * 6, 16, 17000),              # frames not a multiple of the 128-frame chunk
    ([40] * 6, 14, 2000),
    ([8] * 6, 28, 2049),                # register-built operands: 8 bins from one staged copy,
    ([9] * 6, 22, 3000),                # 9, 11 and 12 bins with the bins 8-11 rows on the
    ([11] * 6, 20, 1100),               # remapped copy, partial last chunks
    ([12] * 6, 18, 1537),
]


@pytest.mark.parametrize("bins,atoms,frames", MFMA_CASES)
def test_mfma_counts_and_entropies(engine, bins, atoms, frames):
    """One-hot MFMA path: every pair's histogram bit-exact against the plain
    kernel, a pair sample against the oracle, S2D and MI within tolerance."""
    pa = _pa()
    print(f"mfma bins={bins} atoms={atoms} frames={frames}", flush=True)
    vals, nD = _case(atoms, frames, 17)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    engine.pairs(pa.METHOD_MFMA)
    st = engine.stats()
    assert st["method"] == pa.METHOD_MFMA
    s1, s2 = engine.s1d(), engine.s2d()
    s1o, s2o, mno, mxo, _, shifted = _oracle_full(vals, nD, bins)
    assert np.max(np.abs(s2 - s2o)) <= S_ABS
    mi, mio = _mi(nD, s1, s2), _mi(nD, s1o, s2o)
    err = np.abs(mi - mio)
    assert np.all(err <= MI_REL * np.abs(mio) + MI_ABS), f"max MI err {err.max():.3e}"
    D = 3 * nD + 3
    gi, gj = np.triu_indices(D, 1)
    pairs = list(zip(gi.tolist(), gj.tolist()))
    got = engine.pair_counts(pairs, pa.METHOD_MFMA, bins[3:])
    plain = engine.pair_counts(pairs, pa.METHOD_PLAIN, bins[3:])
    for p, h, hp in zip(pairs, got, plain):
        assert np.array_equal(h, hp), f"counts differ for pair {p}"
        assert h.sum() == frames
    for a, b in [(0, 1), (0, D - 1), (nD + 1, 2 * nD + 5), (D - 2, D - 1)]:
        ta, tb = O.type_of(nD, a), O.type_of(nD, b)
        ho = O.entropy2d(ta, shifted[a], mno[a], mxo[a], bins[3 + ta],
                         tb, shifted[b], mno[b], mxo[b], bins[3 + tb], counts=True)[1]
        assert np.array_equal(got[pairs.index((a, b))].astype(np.int64), ho)


FP4_CASES = [c for c in MFMA_CASES if 8 <= max(c[0][3:]) <= 12]


@pytest.mark.parametrize("bins,atoms,frames", FP4_CASES)
def test_mfma_fp4_counts_and_entropies(engine, bins, atoms, frames):
    """MFMA variant 1 (FP4 one-hot operands, FP32 sums, 8 to 12 bins): every
    pair's histogram bit-exact against the plain kernel and S2D within
    tolerance of the oracle."""
    pa = _pa()
    print(f"mfma fp4 bins={bins} atoms={atoms} frames={frames}", flush=True)
    vals, nD = _case(atoms, frames, 19)
    engine.set_trajectory(vals)
    engine.prepare(bins)
    engine.set_variant(pa.METHOD_MFMA, 1)
    try:
        engine.pairs(pa.METHOD_MFMA)
        st = engine.stats()
        assert st["method"] == pa.METHOD_MFMA and st["kernel_variant"] == 1
        s2 = engine.s2d()
        D = 3 * nD + 3
        gi, gj = np.triu_indices(D, 1)
        pairs = list(zip(gi.tolist(), gj.tolist()))
        got = engine.pair_counts(pairs, pa.METHOD_MFMA, bins[3:])
    finally:
        engine.set_variant(pa.METHOD_MFMA, -1)
    plain = engine.pair_counts(pairs, pa.METHOD_PLAIN, bins[3:])
    for p, h, hp in zip(pairs, got, plain):
        assert np.array_equal(h, hp), f"counts differ for pair {p}"
    _, s2o, *_ = _oracle_full(vals, nD, bins)
    assert np.max(np.abs(s2 - s2o)) <= S_ABS


def test_mfma_fp4_outside_its_bins_runs_int8(engine):
    """Variant 1 pinned at 5 bins (two staging units per thread, no FP4
    kernel): the int8 kernel runs instead, with the same counts."""
    pa = _pa()
    vals, nD = _case(10, 1000, 23)
    engine.set_trajectory(vals)
    engine.prepare([5] * 6)
    engine.set_variant(pa.METHOD_MFMA, 1)
    try:
        engine.pairs(pa.METHOD_MFMA)
        st = engine.stats()
    finally:
        engine.set_variant(pa.METHOD_MFMA, -1)
    assert st["method"] == pa.METHOD_MFMA and st["kernel_variant"] == 0
    _, s2o, *_ = _oracle_full(vals, nD, [5] * 6)
    assert np.max(np.abs(engine.s2d() - s2o)) <= S_ABS


def test_mfma_out_of_range_falls_back(engine):
    """Above 64 bins the MFMA request runs on the atomic kernels."""
    pa = _pa()
    vals, nD = _case(10, 1000, 2)
    engine.set_trajectory(vals)
    engine.prepare([100] * 6)
    engine.pairs(pa.METHOD_MFMA)
    assert engine.stats()["method"] == pa.METHOD_ATOMIC


@pytest.mark.parametrize("variant,bins", [(53, 50), (51, 30)])
@pytest.mark.parametrize("frames", [130800, 130801, 261872, 261873])
def test_packed_sums_at_their_limits(engine, variant, bins, frames):
    """The multi-tile kernel adds the 16-bit sub-counters of a cell as packed
    halves, 8 at a time up to 130,800 frames and 4 at a time up to 261,872
    (each sub-counter holds at most N/16 + 16 counts).  A constant trajectory
    puts every frame in one cell: at each limit and one frame past it the
    counts and entropies must match the plain kernel."""
    pa = _pa()
    nD = 3
    D = 3 * nD + 3
    row = np.array([0.12] * (nD + 2) + [1.9] * (nD + 1) + [3.0] * nD, dtype=np.float32)
    vals = np.tile(row, (frames, 1))
    vals[frames // 2, 0] = 0.5  # one frame elsewhere: two occupied cells in the bond pairs
    engine.set_trajectory(vals)
    engine.prepare([bins] * 6)
    engine.pairs(pa.METHOD_PLAIN)
    s2p = engine.s2d()
    engine.set_variant(pa.METHOD_ATOMIC, variant)
    engine.pairs(pa.METHOD_ATOMIC)
    st = engine.stats()
    print(f"packed sums: variant {variant} frames {frames} method {st['method']} variant {st['kernel_variant']}",
          flush=True)
    assert st["method"] == pa.METHOD_ATOMIC and st["kernel_variant"] == variant
    assert np.max(np.abs(engine.s2d() - s2p)) <= S_ABS
    pairs = [(0, 1), (0, D - 1), (D - 2, D - 1)]
    h = engine.pair_counts(pairs, pa.METHOD_ATOMIC, [bins] * 3)
    hp = engine.pair_counts(pairs, pa.METHOD_PLAIN, [bins] * 3)
    engine.set_variant(pa.METHOD_ATOMIC, -1)
    for x, xp in zip(h, hp):
        assert np.array_equal(x, xp)
        assert x.sum() == frames and x.max() >= frames - 1


@pytest.mark.parametrize("variant,bins,slices", [(31, 50, 16), (45, 50, 16), (49, 30, 16), (53, 50, 16),
                                                 (51, 30, 16), (56, 100, 4), (57, 75, 4)])
def test_packed_counters_at_the_frame_limit(engine, variant, bins, slices):
    """The packed sliced histograms keep 16-bit counters of the frames of one
    lane position.  A constant trajectory puts every frame of every pair in
    one cell, so at the largest frame count the packed variants accept
    (1,048,304) one counter reaches 65,519: the counts must still be exact.
    One frame more and the engine must not pick a packed variant.  Variant 31
    pairs rows in a word; 45 (50 bins) and 49 (30 bins) interleave the halves
    by row parity, and 53 and 51 are the multi-tile kernel on that layout; 56
    and 57 keep 4 sub-counters per cell, so their limit is (65535 - 16) * 4."""
    pa = _pa()
    n_max = (65535 - 16) * slices
    nD = 3
    D = 3 * nD + 3
    row = np.array([0.12] * (nD + 2) + [1.9] * (nD + 1) + [3.0] * nD, dtype=np.float32)
    for frames, packed_ok in ((n_max, True), (n_max + 1, False)):
        vals = np.tile(row, (frames, 1))
        engine.set_trajectory(vals)
        engine.prepare([bins] * 6)
        engine.pairs(pa.METHOD_PLAIN)
        s2p = engine.s2d()
        engine.set_variant(pa.METHOD_ATOMIC, variant)
        engine.pairs(pa.METHOD_ATOMIC)
        st = engine.stats()
        engine.set_variant(pa.METHOD_ATOMIC, -1)
        print(f"packed limit: variant {variant} frames {frames} method {st['method']} variant {st['kernel_variant']}", flush=True)
        if not packed_ok:
            assert st["method"] == pa.METHOD_PLAIN  # the forced packed variant does not fit
            continue
        assert st["method"] == pa.METHOD_ATOMIC and st["kernel_variant"] == variant
        assert np.array_equal(engine.s2d().view(np.uint64), s2p.view(np.uint64)) or \
            np.max(np.abs(engine.s2d() - s2p)) <= S_ABS
        engine.set_variant(pa.METHOD_ATOMIC, variant)
        h = engine.pair_counts([(0, 1), (0, D - 1), (D - 2, D - 1)], pa.METHOD_ATOMIC, [bins] * 3)
        engine.set_variant(pa.METHOD_ATOMIC, -1)
        for x in h:
            assert x.sum() == frames and x.max() == frames


@pytest.mark.parametrize("atoms,frames,seed", [(12, 1000, 3), (40, 3000, 21)])
def test_mist_identical(engine, atoms, frames, seed):
    print(f"mist atoms={atoms} frames={frames}", flush=True)
    vals, nD = _case(atoms, frames, seed)
    bins = [50] * 6
    s1o, s2o, *_ = _oracle_full(vals, nD, bins)
    engine.set_entropies(nD, s1o, s2o)
    edges, mi, left, tot = engine.mist()
    eo, mio, lo, to = O.mist(nD, s1o, s2o)
    assert np.array_equal(edges, eo)
    assert np.array_equal(mi.view(np.uint64), mio.view(np.uint64))
    assert np.array_equal(left, lo)
    assert np.array_equal(tot.view(np.uint64), to.view(np.uint64))


def test_mist_ties_follow_reference_scan(engine):
    """Many exactly equal MI values: the tie rule must reproduce the reference's scan order."""
    nD = 6
    D = 3 * nD + 3
    s1 = np.zeros(D)
    rng = np.random.default_rng(1)
    P = D * (D - 1) // 2
    s2 = -rng.integers(0, 3, P).astype(np.float64)  # MI in {0, 1, 2}: heavy ties
    engine.set_entropies(nD, s1, s2)
    edges, mi, left, tot = engine.mist()
    eo, mio, lo, to = O.mist(nD, s1, s2)
    assert np.array_equal(edges, eo)
    assert np.array_equal(mi, mio)
    assert np.array_equal(tot, to)


def test_mist_ties_between_signed_zeros(engine):
    """MI values of exactly +0 and -0 (S1D of mixed signed zeros): the
    reference's '>' treats them as equal, so the scan order decides, and each
    edge's MI keeps the bits of the pair that won."""
    nD = 7
    D = 3 * nD + 3
    rng = np.random.default_rng(4)
    s1 = np.where(rng.integers(0, 2, D) == 1, -0.0, 0.0)
    P = D * (D - 1) // 2
    s2 = -rng.integers(0, 2, P).astype(np.float64)  # MI in {+-0, 1}
    engine.set_entropies(nD, s1, s2)
    edges, mi, left, tot = engine.mist()
    eo, mio, lo, to = O.mist(nD, s1, s2)
    assert np.array_equal(edges, eo)
    assert np.array_equal(mi.view(np.uint64), mio.view(np.uint64))
    assert np.array_equal(tot.view(np.uint64), to.view(np.uint64))


def _mist_both_paths(engine, nD, s1, s2):
    """parent_engine_mist through the spanning-tree path (default) and the
    O(V^2) Prim: (results, path) of each."""
    engine.set_entropies(nD, s1, s2)
    engine.set_mist_path(0)
    r0 = engine.mist()
    p0 = engine.stats()["mist_path"]
    engine.set_mist_path(1)
    r1 = engine.mist()
    p1 = engine.stats()["mist_path"]
    engine.set_mist_path(0)
    return r0, p0, r1, p1


def _same_mist(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint64) if x.dtype == np.float64 else x,
                              np.asarray(y).view(np.uint64) if y.dtype == np.float64 else y)


@pytest.mark.parametrize("nD,seed,scale", [(3, 1, 1.0), (17, 2, 1.0), (97, 3, 0.01), (330, 4, 1.0),
                                           (1497, 5, 1.0)])
def test_mist_tree_path_matches_prim(engine, nD, seed, scale):
    """Continuous entropies (no equal MI): the spanning-tree path runs and gives
    the O(V^2) Prim's edges, MI bits, left counts and totals; both match the
    oracle's reference-order scan up to 300 dofs."""
    D = 3 * nD + 3
    rng = np.random.default_rng(seed)
    s1 = rng.normal(size=D) * scale + 3.0
    s2 = rng.normal(size=D * (D - 1) // 2) * scale + 6.0
    print(f"mist tree path: {D} dofs", flush=True)
    r0, p0, r1, p1 = _mist_both_paths(engine, nD, s1, s2)
    assert p0 == 1 and p1 == 0
    _same_mist(r0, r1)
    if D <= 300:
        _same_mist(r0, O.mist(nD, s1, s2))


def test_mist_tree_path_on_trajectory_entropies(engine):
    """Entropies of a synthetic trajectory (bins 50): tree path == Prim == oracle."""
    vals, nD = _case(40, 3000, 9)
    s1o, s2o, *_ = _oracle_full(vals, nD, [50] * 6)
    r0, p0, r1, p1 = _mist_both_paths(engine, nD, s1o, s2o)
    assert p0 == 1
    _same_mist(r0, r1)
    _same_mist(r0, O.mist(nD, s1o, s2o))


@pytest.mark.parametrize("kind", ["ties", "signed_zeros", "nan", "no_candidate", "rounded"])
def test_mist_tree_path_falls_back(engine, kind):
    """Equal candidates in a Borůvka choice, NaN, MI <= -1e200: the tree path
    gives up and the Prim's result (equal to the oracle's) stands.  Rounded
    entropies (some equal MI values, not always in a choice) may take either
    path; the result is the same."""
    nD = 9
    D = 3 * nD + 3
    P = D * (D - 1) // 2
    rng = np.random.default_rng(11)
    s1 = rng.normal(size=D) + 3.0
    s2 = rng.normal(size=P) + 6.0
    if kind == "ties":
        s1 = np.zeros(D)
        s2 = -rng.integers(0, 3, P).astype(np.float64)
    elif kind == "signed_zeros":
        s1 = np.where(rng.integers(0, 2, D) == 1, -0.0, 0.0)
        s2 = -rng.integers(0, 2, P).astype(np.float64)
    elif kind == "nan":
        s2[5] = np.nan
    elif kind == "no_candidate":
        s2[7] = 1e201  # MI = s1 + s1 - 1e201 < -1e200
    else:
        s1 = np.round(s1, 1)
        s2 = np.round(s2, 1)
    r0, p0, r1, p1 = _mist_both_paths(engine, nD, s1, s2)
    print(f"mist fallback {kind}: path {p0}", flush=True)
    if kind != "rounded":
        assert p0 == 0
    _same_mist(r0, r1)
    _same_mist(r0, O.mist(nD, s1, s2))


def test_drop_in_executables(tmp_path):
    """PARENT_hip.x / get_PAR_MIST_hip.x against the oracle's PARENT.x restatement."""
    pa = _pa()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    bat = str(tmp_path / "chain.bat")
    pa.synth_write_bat(bat, 24, 2500, 17)
    ours, theirs = str(tmp_path / "gpu.par"), str(tmp_path / "oracle.par")
    bins = ["50", "50", "50", "50", "50", "50"]
    print("run PARENT_hip.x", flush=True)
    r = subprocess.run([os.path.join(root, "parent_amd", "bin", "PARENT_hip.x"), bat, ours] + bins,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.rstrip().endswith("PROGRAM FINISHED SUCCESSFULLY.")
    r2 = subprocess.run([O.ORACLE_EXE, bat, theirs] + bins, capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    a, b = open(ours, "rb").read(), open(theirs, "rb").read()
    assert len(a) == len(b)
    nD, N, binsr, s1, s2 = pa.read_par(ours)
    _, _, _, s1o, s2o = pa.read_par(theirs)
    hdr = len(a) - 8 * (s1.size + s2.size)
    assert a[:hdr] == b[:hdr], ".par headers differ"
    assert np.max(np.abs(s1 - s1o)) <= 1e-12 and np.max(np.abs(s2 - s2o)) <= S_ABS
    # MIST on the oracle's .par: same text as the oracle's MIST
    txt = str(tmp_path / "mist.txt")
    print("run get_PAR_MIST_hip.x", flush=True)
    r3 = subprocess.run([os.path.join(root, "parent_amd", "bin", "get_PAR_MIST_hip.x"), theirs, txt],
                        capture_output=True, text=True, timeout=120)
    assert r3.returncode == 0, r3.stderr
    eo, mio, lo, to = O.mist(nD, s1o, s2o)
    ref_txt = str(tmp_path / "mist_oracle.txt")
    pa.write_mist_text(ref_txt, eo, mio, lo, to)
    assert open(txt).read() == open(ref_txt).read()


def test_bad_inputs_raise(engine):
    pa = _pa()
    vals, nD = _case(10, 500, 2)
    bad = vals.copy()
    bad[7, 3] = -0.5  # a negative bond (PARENT.cpp:522-525)
    engine.set_trajectory(bad)
    with pytest.raises(pa.ParentError) as ei:
        engine.prepare([20] * 6)
    assert ei.value.code == 4
    bad = vals.copy()
    bad[3, 3 * nD + 2] = np.float32(2 * np.pi)  # float32(2pi) > 2pi + 5e-9 (modfit histogram)
    engine.set_trajectory(bad)
    with pytest.raises(pa.ParentError) as ei:
        engine.prepare([20] * 6)
    assert ei.value.code == 5
    engine.set_trajectory(vals)
    with pytest.raises(pa.ParentError) as ei:
        engine.prepare([4097] + [20] * 5)  # above the library's 4,096 bins per dof
    assert ei.value.code == 7
    with pytest.raises(pa.ParentError) as ei:
        engine.prepare([0] + [20] * 5)  # PARENT.cpp:211-234: bins < 1
    assert ei.value.code == 1


@pytest.mark.parametrize("bins,expect_method", [(5, 2), (50, 1)])
def test_autotune_table_and_choice(bins, expect_method):
    """The autotuner's table is consistent, AUTO runs the kernel it chose, and
    each method wins where it is several times faster at full size: the one-hot
    MFMA kernel at 5 bins (its 25 x 25-dof tiles need a sample grown to fill the
    GPU, or it is timed on a fraction of a wave and dropped), the LDS-atomic
    kernels at 50 bins.  Counts of the chosen kernel are checked bit-exact on a
    pair sample."""
    pa = _pa()
    atoms, frames = 1500, 20000
    print(f"autotune: {atoms} atoms x {frames} frames, {bins} bins", flush=True)
    vals = pa.synth_values(atoms, frames, 7)
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.prepare([bins] * 6)
        tune = eng.autotune()
        assert len(tune) >= 2
        keys = [(t["method"], t["variant"]) for t in tune]
        assert len(set(keys)) == len(keys)
        for t in tune:
            assert t["method"] in (pa.METHOD_ATOMIC, pa.METHOD_MFMA)
            assert 0 <= t["variant"] < eng.variant_count(t["method"])
            assert t["ms"] > 0 and t["updates"] > 0
            assert t["updates_per_s"] == pytest.approx(t["updates"] / (t["ms"] * 1e-3), rel=1e-9)
        eng.pairs(pa.METHOD_AUTO)
        st = eng.stats()
        assert (st["method"], st["kernel_variant"]) in keys
        assert st["method"] == expect_method, tune
        D = 3 * atoms - 6
        rng = np.random.default_rng(bins)
        pairs = sorted({tuple(int(x) for x in sorted(rng.choice(D, 2, replace=False))) for _ in range(8)})
        idx = eng.bins()
        got = eng.pair_counts(pairs, pa.METHOD_AUTO, [bins] * 3)
        for (i, j), h in zip(pairs, got):
            ref = np.zeros((bins, bins), dtype=np.int64)
            np.add.at(ref, (idx[i].astype(np.int64), idx[j].astype(np.int64)), 1)
            assert np.array_equal(h.astype(np.int64), ref), (i, j)
