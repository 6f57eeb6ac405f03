"""Batched paths on the GPU: the bin-count sweep (one captured hipGraph per set
of bin counts) and the batch of uneven trajectories (one launch over a merged
tile queue).  Bars as in test_gpu_parity.py: every replay of the graph equals
an independent eager engine bit for bit; batch results match the oracle
(|dS2D| <= 1e-10, MI within 1e-6 relative with a 1e-11 floor) and the
single-trajectory engine's S1D bit for bit.
"""
import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

MI_REL, MI_ABS, S_ABS = 1e-6, 1e-11, 1e-10

SWEEP = [5, 10, 20, 30, 40, 50, 75, 100, [7, 13, 29, 5, 40, 17]]


def _pa():
    import parent_amd
    return parent_amd


def _rows(bin_sets):
    return [[b] * 6 if np.ndim(b) == 0 else list(b) for b in bin_sets]


def _mi(nD, s1, s2):
    D = 3 * nD + 3
    gi, gj = np.triu_indices(D, 1)
    pos = np.array([O.lib().oracle_par_pos(nD, int(a), int(b)) for a, b in zip(gi, gj)])
    return s1[gi] + s1[gj] - s2[pos]


def _check_vs_oracle(nD, vals, bins, s1, s2):
    s1o, s2o, *_ = O.compute(nD, vals.shape[0], vals.T.astype(np.float64), bins)
    assert np.max(np.abs(s1 - s1o)) <= 1e-12
    assert np.max(np.abs(s2 - s2o)) <= S_ABS
    mi, mio = _mi(nD, s1, s2), _mi(nD, s1o, s2o)
    err = np.abs(mi - mio)
    assert np.all(err <= MI_REL * np.abs(mio) + MI_ABS), f"max MI err {err.max():.3e}"


def test_sweep_replays_match_eager_runs():
    """The sweep's graph, launched three times with eager kernels (MIST,
    pair_counts) and host syncs between the launches: every launch equals an
    independent engine that prepares and pairs each bin set eagerly.  The MFMA
    request makes the kernel choice deterministic (MFMA up to 64 bins, the
    default atomic configuration above), so both engines run the same kernels."""
    pa = _pa()
    vals = pa.synth_values(22, 4100, 31)
    nD = 22 - 3
    rows = _rows(SWEEP)
    print(f"sweep: {len(rows)} bin sets, eager references", flush=True)
    ref = []
    with pa.Engine(0) as eager:
        eager.set_trajectory(vals)
        for b in rows:
            eager.prepare(b)
            eager.pairs(pa.METHOD_MFMA)
            ref.append((eager.s1d(), eager.s2d(), eager.stats()["method"]))
            edges, mi, left, tot = eager.mist()
            ref[-1] = ref[-1] + (tot,)
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        for launch in range(3):
            print(f"sweep launch {launch}", flush=True)
            eng.sweep(rows, pa.METHOD_MFMA)
            st = eng.stats()
            assert st["sweep_sets"] == len(rows) and st["ms_sweep"] > 0
            for k, (s1r, s2r, _, totr) in enumerate(ref):
                eng.select(k)
                s1, s2 = eng.s1d(), eng.s2d()
                assert np.array_equal(s1.view(np.uint64), s1r.view(np.uint64)), f"set {k} S1D, launch {launch}"
                assert np.array_equal(s2.view(np.uint64), s2r.view(np.uint64)), f"set {k} S2D, launch {launch}"
            # eager work between the launches: MIST on two sets, counts on one
            for k in (0, len(rows) - 1):
                eng.select(k)
                _, _, _, tot = eng.mist()
                assert np.array_equal(tot.view(np.uint64), ref[k][3].view(np.uint64))
            eng.select(1)
            h = eng.pair_counts([(0, 1), (2, 3 * nD + 2)], pa.METHOD_PLAIN, rows[1][3:])
            assert all(x.sum() == vals.shape[0] for x in h)
    methods = [m for _, _, m, _ in ref]
    assert methods[0] == pa.METHOD_MFMA and methods[-2] == pa.METHOD_ATOMIC  # 5 bins / 100 bins


def test_sweep_against_oracle():
    pa = _pa()
    vals = pa.synth_values(14, 3000, 5)
    nD = 11
    rows = _rows([5, 50, [10, 10, 10, 20, 30, 40]])
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.sweep(rows, pa.METHOD_AUTO)
        for k, b in enumerate(rows):
            print(f"sweep vs oracle: bins {b}", flush=True)
            eng.select(k)
            _check_vs_oracle(nD, vals, b, eng.s1d(), eng.s2d())


@pytest.mark.parametrize("bins,method", [([30] * 6, 3), ([10] * 6, 2), ([5] * 6, 2), ([20] * 6, 2),
                                         ([7, 13, 29, 5, 40, 17], 1), ([50] * 6, 0)])
def test_batch_of_uneven_trajectories(bins, method):
    """Trajectories of different dof and frame counts in one launch: each
    trajectory's S1D equals the single engine's bit for bit, S2D and MI
    match the oracle; MIST totals equal the single engine's."""
    pa = _pa()
    shapes = [(10, 1500, 1), (23, 4097, 2), (17, 800, 3), (40, 3000, 4), (6, 5000, 5)]
    trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
    print(f"batch bins={bins} method={method}: {len(trajs)} trajectories", flush=True)
    with pa.Batch(bins) as b, pa.Engine(0) as single:
        for v in trajs:
            b.add(v)
        b.run(method)
        st = b.stats()
        assert st["pair_updates"] == sum((3 * a - 6) * (3 * a - 7) // 2 * n for a, n, _ in shapes)
        for k, v in enumerate(trajs):
            nD = v.shape[1] // 3 - 1
            s1, s2 = b.results(k)
            single.set_trajectory(v)
            single.prepare(bins)
            assert np.array_equal(s1.view(np.uint64), single.s1d().view(np.uint64))
            _check_vs_oracle(nD, v, bins, s1, s2)
            single.pairs(pa.METHOD_PLAIN)
            s2p = single.s2d()
            assert np.max(np.abs(s2 - s2p)) <= S_ABS
            _, _, _, tot = b.mist(k)
            _, _, _, tot1 = single.mist()
            assert np.max(np.abs(tot - tot1)) <= 1e-8


def test_batch_frame_classes():
    """PARENT_METHOD_AUTO on a batch of short (under 18,000 frames at 30 bins)
    and long trajectories: forced into two class launches, forced into one,
    and left to the autotuner, every trajectory's S1D bit for bit and S2D
    and MI against the oracle, and the two-launch S2D within 1e-10 of the
    one-launch S2D."""
    pa = _pa()
    bins = [30] * 6
    shapes = [(12, 20000, 1), (9, 24001, 2), (15, 3000, 3), (8, 900, 4), (20, 6000, 5)]
    trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
    with pa.Batch(bins) as b:
        for v in trajs:
            b.add(v)
        runs = {}
        for mode, launches in ((2, 2), (0, 1), (1, None)):
            print(f"frame classes: mode {mode}", flush=True)
            b.set_classes(mode)
            b.run(pa.METHOD_AUTO)
            st = b.stats()
            if launches is not None:
                assert st["batch_launches"] == launches
            assert st["pair_updates"] == sum((3 * a - 6) * (3 * a - 7) // 2 * n for a, n, _ in shapes)
            runs[mode] = [b.results(k) for k in range(len(trajs))]
        for k, v in enumerate(trajs):
            nD = v.shape[1] // 3 - 1
            (s1a, s2a), (s1b, s2b), (s1c, s2c) = runs[2][k], runs[0][k], runs[1][k]
            assert np.array_equal(s1a.view(np.uint64), s1b.view(np.uint64))
            assert np.array_equal(s1a.view(np.uint64), s1c.view(np.uint64))
            assert np.max(np.abs(s2a - s2b)) <= S_ABS
            _check_vs_oracle(nD, v, bins, s1a, s2a)
            _check_vs_oracle(nD, v, bins, s1c, s2c)


def test_batch_mfma_fp4():
    """The batch instantiation of MFMA variant 1 (FP4) on uneven
    trajectories at 10 bins: S2D against the oracle."""
    pa = _pa()
    bins = [10] * 6
    shapes = [(10, 1500, 1), (23, 4097, 2), (17, 800, 3), (9, 130, 4)]
    trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
    print("batch mfma fp4", flush=True)
    with pa.Batch(bins) as b:
        for v in trajs:
            b.add(v)
        b.set_variant(pa.METHOD_MFMA, 1)
        b.run(pa.METHOD_MFMA)
        st = b.stats()
        assert st["method"] == pa.METHOD_MFMA and st["kernel_variant"] == 1
        for k, v in enumerate(trajs):
            s1, s2 = b.results(k)
            _check_vs_oracle(v.shape[1] // 3 - 1, v, bins, s1, s2)


@pytest.mark.parametrize("method", [1, 2, 3])  # atomic, MFMA, auto
def test_batch_with_short_trajectories(method):
    """Trajectories of 1 to 17 frames beside longer ones in one launch: the
    short ones' tiles hold fewer frames than one 16-frame chunk of a thread."""
    pa = _pa()
    bins = [10] * 6 if method == 2 else [30] * 6
    shapes = [(8, 1, 7), (12, 2, 8), (9, 5, 9), (14, 17, 10), (20, 3000, 11), (6, 16, 12)]
    trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
    print(f"short batch bins={bins} method={method}", flush=True)
    with pa.Batch(bins) as b:
        for v in trajs:
            b.add(v)
        b.run(method)
        for k, v in enumerate(trajs):
            nD = v.shape[1] // 3 - 1
            s1, s2 = b.results(k)
            _check_vs_oracle(nD, v, bins, s1, s2)


@pytest.mark.parametrize("bins,variant", [([50] * 6, 53), ([30] * 6, 51), ([40] * 6, 54), ([75] * 6, 57)])
def test_batch_on_pinned_variants(bins, variant):
    """parent_batch_set_variant pins the merged launch's kernel: the multi-tile
    kernels (one tile per 1,024-thread workgroup at 40-50 bins, two per
    256-thread workgroup at 30, 4 sub-counters at 75) on trajectories of
    different dof and frame counts: S2D within tolerance of the plain kernel
    and of the oracle."""
    pa = _pa()
    shapes = [(10, 1500, 1), (23, 4097, 2), (17, 800, 3), (40, 3000, 4), (6, 5000, 5), (30, 20001, 6)]
    trajs = [pa.synth_values(a, n, s) for a, n, s in shapes]
    print(f"batch on pinned variant {variant}: bins={bins}", flush=True)
    with pa.Batch(bins) as b, pa.Engine(0) as single:
        for v in trajs:
            b.add(v)
        b.set_variant(pa.METHOD_ATOMIC, variant)
        b.run(pa.METHOD_ATOMIC)
        st = b.stats()
        assert st["method"] == pa.METHOD_ATOMIC and st["kernel_variant"] == variant
        for k, v in enumerate(trajs):
            nD = v.shape[1] // 3 - 1
            s1, s2 = b.results(k)
            single.set_trajectory(v)
            single.prepare(bins)
            single.pairs(pa.METHOD_PLAIN)
            assert np.max(np.abs(s2 - single.s2d())) <= S_ABS, f"trajectory {k}"
            _check_vs_oracle(nD, v, bins, s1, s2)


def test_batch_rejects_bad_data():
    pa = _pa()
    v = pa.synth_values(10, 500, 2)
    bad = v.copy()
    bad[7, 3] = -0.5
    with pa.Batch([20] * 6) as b:
        b.add(v)
        with pytest.raises(pa.ParentError) as ei:
            b.add(bad)
        assert ei.value.code == 4
        b.run(pa.METHOD_AUTO)  # the good trajectory still runs
        s1, s2 = b.results(0)
        assert np.all(np.isfinite(s2))


def test_queue_longer_than_one_launch():
    """4.48M one-pair tiles on a 1,024-thread configuration: the queue is
    issued as two launches (grid x block must stay below 2^32 work-items);
    every pair matches the plain kernel."""
    pa = _pa()
    vals = pa.synth_values(1000, 600, 7)  # 2,994 dofs, 4,480,521 pairs
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.prepare([20] * 6)
        eng.pairs(pa.METHOD_PLAIN)
        s2p = eng.s2d()
        eng.set_variant(pa.METHOD_ATOMIC, 19)  # sliced 1 x 1 tiles, 1,024 threads
        print("long queue: variant 19", flush=True)
        eng.pairs(pa.METHOD_ATOMIC)
        assert eng.stats()["method"] == pa.METHOD_ATOMIC and eng.stats()["kernel_variant"] == 19
        s2 = eng.s2d()
        assert np.max(np.abs(s2 - s2p)) <= S_ABS


def test_batch_rerun_after_rejected_longer_trajectory():
    """A rejected add of a trajectory with more frames than any accepted one
    may grow the shared n ln n table; the next run must rebuild the queue
    descriptors that point to it and give the first run's S2D bit for bit."""
    pa = _pa()
    good = [pa.synth_values(12, 900, 3), pa.synth_values(9, 1200, 4)]
    bad = pa.synth_values(10, 20000, 5)
    bad[11, 2] = -0.25  # a negative bond: PARENT_E_DATA
    with pa.Batch([30] * 6) as b:
        for v in good:
            b.add(v)
        b.run(pa.METHOD_ATOMIC)
        first = [b.results(k)[1].copy() for k in range(len(good))]
        print("batch: adding the rejected 20,000-frame trajectory", flush=True)
        with pytest.raises(pa.ParentError) as ei:
            b.add(bad)
        assert ei.value.code == 4
        b.run(pa.METHOD_ATOMIC)
        for k in range(len(good)):
            s2 = b.results(k)[1]
            assert np.array_equal(s2.view(np.uint64), first[k].view(np.uint64)), f"trajectory {k}"


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_batch_with_mixed_bins_in_one_launch(method):
    """Trajectories with different bin counts in one merged-queue launch (the
    kernel holds the batch's largest counts, each tile reads its own): every
    trajectory against the oracle at its own bins and against the single
    engine's S1D bit for bit."""
    pa = _pa()
    items = [((10, 1500, 1), [20] * 6), ((23, 4097, 2), [30] * 6), ((17, 800, 3), [7, 13, 29, 5, 40, 17]),
             ((12, 3000, 4), [5] * 6), ((8, 2000, 5), [12, 9, 11, 10, 8, 12])]
    trajs = [(pa.synth_values(a, n, s), bins) for (a, n, s), bins in items]
    print(f"mixed-bins batch, method {method}", flush=True)
    with pa.Batch([30] * 6) as b, pa.Engine(0) as single:
        for v, bins in trajs:
            b.add(v, bins)
        b.run(method)
        for k, (v, bins) in enumerate(trajs):
            nD = v.shape[1] // 3 - 1
            s1, s2 = b.results(k)
            single.set_trajectory(v)
            single.prepare(bins)
            assert np.array_equal(s1.view(np.uint64), single.s1d().view(np.uint64))
            _check_vs_oracle(nD, v, bins, s1, s2)
