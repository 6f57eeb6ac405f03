"""Pins the CPU oracle (no GPU needed).

The oracle is the checker of the HIP engine, so it is checked first:
  * against the reference's own outputs (golden .par / MIE / MIST text written
    by the reference programs built from its sources, tests/golden/);
  * against closed forms and brute force where the reference has no vectors:
    uniform dihedral entropy, product-grid MI, duplicated dofs, range rule,
    modfit gap placement, bin centres, and MIST vs Kruskal.
"""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

import modfit_cases as MC
import oracle_api as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BAT = os.path.join(GOLD, "chain12_1000.bat")


def _par_body(path, nD):
    D = 3 * nD + 3
    raw = open(path, "rb").read()
    n = D + D * (D - 1) // 2
    body = np.frombuffer(raw[len(raw) - 8 * n:], dtype=np.float64)
    return raw[:len(raw) - 8 * n], body[:D].copy(), body[D:].copy()


def _fmt(x):
    return "%.12g" % x


@pytest.mark.parametrize("tag,bins", [("b50", [50] * 6), ("mixed", [7, 13, 29, 5, 40, 17]),
                                      ("wide", [210, 256, 300, 500, 256, 210])])
def test_oracle_executable_matches_reference_par(tmp_path, tag, bins):
    out = str(tmp_path / "o.par")
    r = subprocess.run([O.ORACLE_EXE, BAT, out] + [str(b) for b in bins], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.rstrip().endswith("PROGRAM FINISHED SUCCESSFULLY.")
    assert open(out, "rb").read() == open(os.path.join(GOLD, f"chain12_{tag}.par"), "rb").read()


@pytest.mark.parametrize("tag", ["b50", "mixed", "wide"])
def test_oracle_mist_matches_reference_text(tag):
    import parent_amd
    nD = 9
    _, s1, s2 = _par_body(os.path.join(GOLD, f"chain12_{tag}.par"), nD)
    edges, mi, left, tot = O.mist(nD, s1, s2)
    names = ["TOTAL 1D BONDS ENTROPY", "TOTAL 1D ANGLES ENTROPY", "TOTAL 1D DIHEDRALS ENTROPY",
             "TOTAL 2D BONDS-BONDS MUTUAL INFORMATION", "TOTAL 2D BONDS-ANGLES MUTUAL INFORMATION",
             "TOTAL 2D BONDS-DIHEDRALS MUTUAL INFORMATION", "TOTAL 2D ANGLES-ANGLES MUTUAL INFORMATION",
             "TOTAL 2D ANGLES-DIHEDRALS MUTUAL INFORMATION",
             "TOTAL 2D DIHEDRALS-DIHEDRALS MUTUAL INFORMATION", "TOTAL CONFIGURATIONAL ENTROPY"]
    lines = [f"{'bad'[e[0]]} {'bad'[e[1]]} {e[2] + 1} {e[3] + 1}   MI: {_fmt(m)}   left to process: {n}"
             for e, m, n in zip(edges, mi, left)]
    lines += [f"{n} = {_fmt(t)}" for n, t in zip(names, tot)]
    ref = open(os.path.join(GOLD, f"chain12_{tag}_MIST.txt")).read().splitlines()
    assert lines == ref
    del parent_amd


def test_mie_listing_and_totals_from_oracle():
    """get_PAR_MIE.cpp:109-199 listing order, MI via get_mutual and the 16 totals."""
    nD = 9
    nB, nA, D = nD + 2, nD + 1, 3 * nD + 3
    _, s1, s2 = _par_body(os.path.join(GOLD, "chain12_b50.par"), nD)
    ref = open(os.path.join(GOLD, "chain12_b50_MIE.txt")).read().splitlines()
    out, tot = [], {}
    for k in range(nB):
        out.append(f"bond {k + 1}   {_fmt(s1[k])}")
    for k in range(nA):
        out.append(f"angle {k + 1}   {_fmt(s1[nB + k])}")
    for k in range(nD):
        out.append(f"dihedral {k + 1}   {_fmt(s1[nB + nA + k])}")
    cats = [("bond-bond", 0, 0), ("bond-angle", 0, 1), ("bond-dihedral", 0, 2), ("angle-angle", 1, 1),
            ("angle-dihedral", 1, 2), ("dihedral-dihedral", 2, 2)]
    first, cnt = (0, nB, nB + nA), (nB, nA, nD)
    sums = []
    for name, ta, tb in cats:
        se = sm = 0.0
        for i in range(cnt[ta]):
            for j in range(i + 1 if ta == tb else 0, cnt[tb]):
                gi, gj = first[ta] + i, first[tb] + j
                s = s2[O.lib().oracle_par_pos(nD, gi, gj)]
                m = O.lib().oracle_mutual(nD, O._p(s1), O._p(s2), gi, gj)
                out.append(f"{name} {i + 1}  {j + 1}   {_fmt(s)}   {_fmt(m)}")
                se += s
                sm += m
        sums.append((se, sm))
    assert out == ref[:len(out)]
    tB, tA, tD = sum(s1[:nB]), sum(s1[nB:nB + nA]), sum(s1[nB + nA:])
    total = tB + tA + tD
    for _, sm in sums:
        total -= sm
    assert ref[-1] == f"TOTAL CONFIGURATIONAL ENTROPY = {_fmt(total)}"
    assert ref[len(out)] == f"TOTAL 1D BONDS ENTROPY = {_fmt(tB)}"
    del tot


def test_uniform_dihedral_closed_form():
    """Values at the B bin centres, m frames each: S = ln(B*delta) + (B-1)/(2N)."""
    B, m = 20, 7
    vmin, vmax = 0.5, 4.5
    delta = (vmax - vmin) / B
    x = np.repeat(vmin + (np.arange(B) + 0.5) * delta, m)
    N = x.size
    s = O.entropy1d(O.D, x, vmin, vmax, B)
    assert abs(s - (math.log(B * delta) + (B - 1) / (2 * N))) < 1e-12


def test_product_grid_mi_is_herzel_only():
    """N = m*Bi*Bj frames covering every cell equally: plug-in MI = 0, so
    MI = S1 + S2 - S12 = (Bi + Bj - Bi*Bj - 1) / (2N) for a dihedral pair."""
    Bi, Bj, m = 6, 9, 3
    a = (np.arange(Bi) + 0.5) / Bi * 2.0
    b = (np.arange(Bj) + 0.5) / Bj * 3.0
    xa = np.repeat(np.repeat(a, Bj), m)
    xb = np.repeat(np.tile(b, Bi), m)
    N = xa.size
    s1 = O.entropy1d(O.D, xa, 0.0, 2.0, Bi)
    s2 = O.entropy1d(O.D, xb, 0.0, 3.0, Bj)
    s12, h = O.entropy2d(O.D, xa, 0.0, 2.0, Bi, O.D, xb, 0.0, 3.0, Bj, counts=True)
    assert np.all(h == m)
    mi = s1 + s2 - s12
    assert abs(mi - (Bi + Bj - Bi * Bj - 1) / (2 * N)) < 1e-12


def test_duplicated_dof_mi_equals_entropy_terms():
    """y = x: the 2D histogram is diagonal (n_kk = h_k), so
    S12 = -sum (h/N) ln(h/(N delta^2)) + herzel = S1 + ln(delta) and MI = S1 - ln(delta)."""
    rng = np.random.default_rng(4)
    x = rng.uniform(0.3, 2.7, 5000)
    rc, mn, mx, _, xs = O.dof_range(O.D, x.copy())
    assert rc == 0
    B = 25
    s1 = O.entropy1d(O.D, xs, mn, mx, B)
    s12, h = O.entropy2d(O.D, xs, mn, mx, B, O.D, xs, mn, mx, B, counts=True)
    assert np.count_nonzero(h - np.diag(np.diag(h))) == 0
    delta = (mx - mn) / B
    assert abs(s12 - (s1 + math.log(delta))) < 1e-10
    assert abs((s1 + s1 - s12) - (s1 - math.log(delta))) < 1e-10


def test_range_rule_closed_form():
    x = np.array([0.12, 0.13, 0.125])
    rc, mn, mx, mf, _ = O.dof_range(O.B, x.copy())
    assert rc == 0 and mf == 0.0
    assert mn == 0.12 - 0.000000005 and mx == 0.13 + 0.000000005
    # narrow range -> +0.05 on max
    rc, mn, mx, _, _ = O.dof_range(O.B, np.array([0.1, 0.1 + 1e-6]))
    assert mx == (0.1 + 1e-6 + 0.000000005) + 0.05
    # bonds: min clamps at 0; negative bond is an error
    rc, mn, _, _, _ = O.dof_range(O.B, np.array([1e-9, 0.2]))
    assert mn == 0.0
    assert O.dof_range(O.B, np.array([0.1, -0.01]))[0] == 1
    # angles clamp max at pi
    rc, mn, mx, _, _ = O.dof_range(O.A, np.array([1.0, O.PI - 1e-12]))
    assert mx == O.PI


def test_modfit_moves_the_largest_gap_to_the_end():
    """Dihedrals on [0, 2pi) with one empty arc: after the modfit shift the data
    occupy one contiguous arc, so max - min ~ 2pi - gap (within one modfit bin)."""
    rng = np.random.default_rng(7)
    two_pi = 2 * O.PI
    binsize = (two_pi + 0.000000005) / 100
    for start in (0.3, 2.0, 5.5, 6.0):  # 5.5/6.0: the occupied arc wraps through 0
        width = 2.5
        x = np.mod(start + rng.uniform(0, width, 4000), two_pi)
        rc, mn, mx, mf, xs = O.dof_range(O.D, x.copy())
        assert rc == 0
        assert np.all(xs >= 0) and np.all(xs < two_pi)
        assert abs((mx - mn) - width) < 2 * binsize
    # no empty modfit bin: shift puts the first least-populated bin at the end
    x = np.repeat((np.arange(100) + 0.5) * binsize, 5)
    x = np.concatenate([x, np.repeat(x[:5 * 100][::5][np.arange(100) != 37], 1)])
    rc, mn, mx, mf, xs = O.dof_range(O.D, x.copy())
    assert mf == two_pi - (37 + 0.5) * binsize


@pytest.mark.parametrize("name,counts,expect", MC.CASES, ids=[c[0] for c in MC.CASES])
def test_modfit_bin_of_crafted_histograms(name, counts, expect):
    """The oracle's modfit shift on histograms whose bin the reference's two scans
    (PARENT.cpp:573-607) fix by hand: ties, runs through bins 99 and 0, runs at
    the 63/64 boundary, a single occupied or a single empty bin."""
    x = MC.column(counts, O.PI, np.random.default_rng(5))
    rc, mn, mx, mf, xs = O.dof_range(O.D, x.copy())
    assert rc == 0
    assert mf == 2 * O.PI - (expect + 0.5) * MC.bin_width(O.PI)


def test_bin_centres_index_exactly():
    B = 37
    mn, mx = 0.25, 3.0
    delta = (mx - mn) / B
    x = mn + (np.arange(B) + 0.5) * delta
    idx, bad = O.bin_index(x, mn, mx, B)
    assert bad == 0 and np.array_equal(idx, np.arange(B))


def _kruskal(V, w):
    """Maximum spanning tree by Kruskal on a dense weight matrix."""
    parent = list(range(V))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    es = sorted(((w[i, j], i, j) for i in range(V) for j in range(i + 1, V)), reverse=True)
    tree = set()
    for wt, i, j in es:
        ri, rj = find(i), find(j)
        if ri != rj:
            parent[ri] = rj
            tree.add((i, j))
    return tree


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mist_is_the_maximum_spanning_tree(seed):
    nD = 4
    D = 3 * nD + 3
    rng = np.random.default_rng(seed)
    s1 = rng.normal(size=D)
    s2 = rng.normal(size=D * (D - 1) // 2)
    edges, mi, left, tot = O.mist(nD, s1, s2)
    first = (0, nD + 2, 2 * nD + 3)
    got = {tuple(sorted((first[e[0]] + e[2], first[e[1]] + e[3]))) for e in edges}
    w = np.zeros((D, D))
    for i, j in itertools.combinations(range(D), 2):
        w[i, j] = w[j, i] = O.lib().oracle_mutual(nD, O._p(s1), O._p(s2), i, j)
    assert got == _kruskal(D, w)
    assert list(left) == list(range(D - 1, 0, -1))
    assert abs(tot[9] - (s1.sum() - mi.sum())) < 1e-9


UBQ_DIR = os.path.join(GOLD, "ubq")


def test_oracle_reproduces_ubq_sample_totals(tmp_path):
    """End-to-end known answer of the reference (README.md:108-123): the UBQ_UBM2
    test system at 50 bins.  The .bat comes from the library's BAT builder (its
    bytes are pinned to the reference builder's in test_cpu_bat_builder.py);
    every TOTAL line of the reference's sample get_PAR_MIE / get_PAR_MIST
    output is reproduced at 12 digits."""
    import parent_amd as pa
    bat = str(tmp_path / "UBQ.bat")
    assert pa.bat_build(os.path.join(UBQ_DIR, "UBQ_UBM2.top"), os.path.join(UBQ_DIR, "UBQ_UBM2.xtc"), bat) == 801
    par = str(tmp_path / "UBQ.par")
    r = subprocess.run([O.ORACLE_EXE, bat, par] + ["50"] * 6, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr
    nD = 1091
    _, s1, s2 = _par_body(par, nD)
    mie = pa.mie_write(None, nD, s1, s2)
    mie_tail = open(os.path.join(GOLD, "ubq_sample_MIE_tail.txt")).read().splitlines()
    assert [ln.split(" = ")[0] + " = " + _fmt(v) for ln, v in zip(mie_tail, mie)] == mie_tail
    edges, mi, left, tot = O.mist(nD, s1, s2)
    mist_tail = open(os.path.join(GOLD, "ubq_sample_MIST_tail.txt")).read().splitlines()
    assert [ln.split(" = ")[0] + " = " + _fmt(v) for ln, v in zip(mist_tail, tot)] == mist_tail


def test_subset_computation_equals_full():
    """oracle_compute_subset on a dof selection gives the bits of oracle_compute."""
    rng = np.random.default_rng(5)
    nD, N = 7, 1500
    D = 3 * nD + 3
    v = np.empty((D, N))
    v[: nD + 2] = rng.normal(0.13, 0.003, (nD + 2, N))
    v[nD + 2: 2 * nD + 3] = rng.uniform(1.8, 2.1, (nD + 1, N))
    v[2 * nD + 3:] = rng.uniform(0.0, 6.28, (nD, N))
    bins = [20, 30, 40, 25, 35, 45]
    s1, s2, *_ = O.compute(nD, N, v, bins)
    dofs = np.array([0, 3, nD + 2, nD + 5, 2 * nD + 3, D - 1], dtype=np.int32)
    pairs = [(a, b) for a in range(dofs.size) for b in range(a + 1, dofs.size)]
    s1s, s2s = O.compute_subset(nD, N, dofs, v[dofs], bins, pairs)
    assert np.array_equal(s1s, s1[dofs])
    pos = [O.lib().oracle_par_pos(nD, int(dofs[a]), int(dofs[b])) for a, b in pairs]
    assert np.array_equal(s2s, s2[pos])
