"""Parity on the GPU at the benchmark shapes, against the oracle.

  c1  chain_100x10k_b50     100-atom chain (294 dofs) x 10,000 frames, 50 bins:
                            every S1D / S2D / MI with the plain and the autotuned
                            kernel, and the whole MIST (edges, MI bits, totals)
  c2  protein_1500x100k_b50 1,500-atom chain (4,494 dofs, 10,095,771 pairs) x
                            100,000 frames, 50 bins: every S2D finite, tuned vs
                            plain within 1e-10 on every pair, a seeded 1% pair
                            sample vs the oracle, and the MIST edge list vs the
                            oracle's Prim on the same entropies
  c3  sweep_1500x100k       the c2 trajectory at bins 5, 10, 20, 30, 40, 50, 75,
                            100 through one sweep graph: a seeded pair sample of
                            every set vs the oracle
  c4  batch64_b30           the 64 uneven trajectories of the benchmark (seed 4),
                            30 bins, one merged-queue launch: a seeded pair
                            sample of every trajectory vs the oracle

Bars (the same as every parity test here): S1D within 1e-12, S2D within 1e-10
absolute, MI = S_i + S_j - S_ij within 1e-6 relative of the oracle's with a
1e-11 absolute floor.  The oracle restates PARENT.cpp:503-1479 in FP64.
"""
import os
import sys

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

MI_REL, MI_ABS, S1_ABS, S_ABS = 1e-6, 1e-11, 1e-12, 1e-10
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pa():
    import parent_amd
    return parent_amd


def _threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def par_pos(nD, gi, gj):
    """Vectorised .par body position of pairs gi < gj (util.cpp:47-81,
    PARENT.cpp:1107-1112): the same scheme as oracle_par_pos."""
    gi = np.asarray(gi, dtype=np.int64)
    gj = np.asarray(gj, dtype=np.int64)
    nB, nA, nd = nD + 2, nD + 1, nD
    first = np.array([0, nB, nB + nA])
    t1 = (gi >= nB).astype(np.int64) + (gi >= nB + nA)
    t2 = (gj >= nB).astype(np.int64) + (gj >= nB + nA)
    i, j = gi - first[t1], gj - first[t2]
    oBA = nB * (nB - 1) // 2
    oBD = oBA + nB * nA
    oAA = oBD + nB * nd
    oAD = oAA + nA * (nA - 1) // 2
    oDD = oAD + nA * nd
    rev = lambda n, s, g: (n - s) * (n - s - 1) // 2 + s - g  # noqa: E731
    out = np.empty(gi.shape, dtype=np.int64)
    for a, b, val in ((0, 0, rev(nB, i, j)), (0, 1, oBA + i * nA + j), (0, 2, oBD + i * nd + j),
                      (1, 1, oAA + rev(nA, i, j)), (1, 2, oAD + i * nd + j), (2, 2, oDD + rev(nd, i, j))):
        m = (t1 == a) & (t2 == b)
        out[m] = val[m]
    return out


def _check_sample(nD, N, s1, s2, dofs, cols, bins, label):
    """Every pair among `dofs` (cols: their FP32 values, dof-major) vs the oracle."""
    a, b = np.triu_indices(len(dofs), 1)
    pairs = np.stack([a, b], axis=1).astype(np.int32)
    s1o, s2o = O.compute_subset(nD, N, dofs, cols, bins, pairs, _threads())
    d1 = np.max(np.abs(s1[dofs] - s1o))
    got = s2[par_pos(nD, dofs[a], dofs[b])]
    d2 = np.max(np.abs(got - s2o))
    mi, mio = s1[dofs[a]] + s1[dofs[b]] - got, s1o[a] + s1o[b] - s2o
    err = np.abs(mi - mio)
    print(f"{label}: {len(dofs)} dofs, {len(pairs)} pairs: max|dS1D| {d1:.2e} max|dS2D| {d2:.2e} "
          f"max|dMI| {err.max():.2e}", flush=True)
    assert d1 <= S1_ABS, label
    assert d2 <= S_ABS, label
    assert np.all(err <= MI_REL * np.abs(mio) + MI_ABS), f"{label}: max MI err {err.max():.3e}"


def _sample_dofs(D, n, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(D, min(n, D), replace=False)).astype(np.int32)


def test_c1_chain_100x10k_b50_every_pair_and_mist():
    pa = _pa()
    atoms, N, B = 100, 10_000, 50
    nD = atoms - 3
    vals = pa.synth_values(atoms, N, 1)
    s1o, s2o, *_ = O.compute(nD, N, vals.T.astype(np.float64), [B] * 6, _threads())
    D = 3 * nD + 3
    gi, gj = np.triu_indices(D, 1)
    pos = par_pos(nD, gi, gj)
    mio = s1o[gi] + s1o[gj] - s2o[pos]
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        for m in (pa.METHOD_PLAIN, pa.METHOD_AUTO):
            print(f"c1: method {m}", flush=True)
            eng.prepare([B] * 6)
            eng.pairs(m)
            s1, s2 = eng.s1d(), eng.s2d()
            assert np.max(np.abs(s1 - s1o)) <= S1_ABS
            assert np.max(np.abs(s2 - s2o)) <= S_ABS
            mi = s1[gi] + s1[gj] - s2[pos]
            assert np.all(np.abs(mi - mio) <= MI_REL * np.abs(mio) + MI_ABS)
            edges, mi_e, left, tot = eng.mist()
            eo, mo, lo, to = O.mist(nD, s1, s2)
            assert np.array_equal(edges, eo) and np.array_equal(left, lo)
            assert np.array_equal(mi_e.view(np.uint64), mo.view(np.uint64))
            assert np.max(np.abs(tot - to)) <= 1e-9 * np.max(np.abs(to))
        # the autotuned step's MIST against the oracle's own entropies: same edges
        eo2, *_ = O.mist(nD, s1o, s2o)
        assert np.array_equal(edges, eo2)


@pytest.fixture(scope="module")
def c2_values():
    pa = _pa()
    print("c2: synthesising 100,000 x 4,494", flush=True)
    return pa.synth_values(1500, 100_000, 2, _threads())


def test_c2_protein_1500x100k_b50(c2_values):
    pa = _pa()
    vals = c2_values
    N, B = vals.shape[0], 50
    nD = vals.shape[1] // 3 - 1
    D = 3 * nD + 3
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        eng.prepare([B] * 6)
        eng.pairs(pa.METHOD_PLAIN)
        s2p = eng.s2d()
        eng.pairs(pa.METHOD_AUTO)
        st = eng.stats()
        print(f"c2: tuned method {st['method']} variant {st['kernel_variant']} pairs {st['ms_pairs']:.1f} ms",
              flush=True)
        s1, s2 = eng.s1d(), eng.s2d()
        edges, mi_e, left, tot = eng.mist()
    assert s2.size == 10_095_771
    assert np.all(np.isfinite(s2))
    dmax = np.max(np.abs(s2 - s2p))
    print(f"c2: tuned vs plain on all {s2.size} pairs: max |dS2D| {dmax:.2e}", flush=True)
    assert dmax <= S_ABS
    # 450 dofs: 101,025 pairs, 1.0% of all pairs
    dofs = _sample_dofs(D, 450, 2)
    _check_sample(nD, N, s1, s2, dofs, np.ascontiguousarray(vals[:, dofs].T), [B] * 6, "c2 sample")
    print("c2: oracle MIST on the GPU's entropies", flush=True)
    eo, mo, lo, to = O.mist(nD, s1, s2)
    assert np.array_equal(edges, eo) and np.array_equal(left, lo)
    assert np.array_equal(mi_e.view(np.uint64), mo.view(np.uint64))


def test_c3_sweep_1500x100k(c2_values):
    pa = _pa()
    vals = c2_values
    N = vals.shape[0]
    nD = vals.shape[1] // 3 - 1
    D = 3 * nD + 3
    sweep = [5, 10, 20, 30, 40, 50, 75, 100]
    with pa.Engine(0) as eng:
        eng.set_trajectory(vals)
        print("c3: sweep (first call: prepare, tune, capture, launch)", flush=True)
        eng.sweep(sweep, pa.METHOD_AUTO)
        eng.sweep(sweep, pa.METHOD_AUTO)  # a replay of the captured graph
        print(f"c3: graph {eng.stats()['ms_sweep']:.1f} ms", flush=True)
        methods = []
        for k, b in enumerate(sweep):
            eng.select(k)
            methods.append(eng.stats()["method"])
            s1, s2 = eng.s1d(), eng.s2d()
            assert np.all(np.isfinite(s2))
            dofs = _sample_dofs(D, 120, 30 + k)
            _check_sample(nD, N, s1, s2, dofs, np.ascontiguousarray(vals[:, dofs].T), [b] * 6, f"c3 bins {b}")
    assert methods[0] == pa.METHOD_MFMA and methods[-1] == pa.METHOD_ATOMIC


def test_c4_batch64_b30():
    pa = _pa()
    sys.path.insert(0, ROOT)
    import bench
    shapes = bench.batch_shapes()
    B = bench.BATCH_BINS
    samples = []
    with pa.Batch([B] * 6) as b:
        for k, (a, n) in enumerate(shapes):
            v = pa.synth_values(a, n, 1000 + k, _threads())
            D = v.shape[1]
            dofs = _sample_dofs(D, 40, 400 + k)
            samples.append((a - 3, n, dofs, np.ascontiguousarray(v[:, dofs].T)))
            b.add(v)
            del v
        print(f"c4: {len(shapes)} trajectories added", flush=True)
        b.run(pa.METHOD_AUTO)
        st = b.stats()
        print(f"c4: method {st['method']} variant {st['kernel_variant']} pairs {st['ms_pairs']:.1f} ms", flush=True)
        assert st["pair_updates"] == sum((3 * a - 6) * (3 * a - 7) // 2 * n for a, n in shapes)
        for k, (nD, n, dofs, cols) in enumerate(samples):
            s1, s2 = b.results(k)
            assert np.all(np.isfinite(s2))
            _check_sample(nD, n, s1, s2, dofs, cols, [B] * 6, f"c4 trajectory {k}")
        _, _, _, tot = b.mist(0)
        assert np.all(np.isfinite(tot))
