"""One test body, two implementations: every test here runs against the CPU
oracle (the CPU suite) and against the HIP engine (marked gpu), through the
common interface of engines.py.  The checks are closed forms that the
reference's formulas fix (PARENT.cpp:644-772 1D, 774-1479 2D; util.cpp:47-81
MI; get_PAR_MIST.cpp:396-766 MIST), so each implementation is pinned on its
own, not only against the other:

  * a dihedral whose values sit one to a bin, m frames each:
    S1D = ln(B * delta) + (B - 1) / (2N)  (delta from the engine's own range);
  * dofs whose frames enumerate every cell of their joint grid m times:
    plug-in MI = 0, so MI = (Bi + Bj - Bi*Bj - 1) / (2N), also between a bond
    and a dihedral (the Jacobian terms cancel in MI);
  * a dihedral duplicated: the joint histogram is diagonal and
    MI = S1D - ln(delta);
  * MIST on distinct random MI: the maximum spanning tree (Kruskal), `left`
    counting down, total = sum S1D - sum of the tree's MI.
"""
import itertools
import math

import numpy as np
import pytest

import engines as E

IMPLS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


@pytest.fixture(params=IMPLS)
def engine(request):
    print(f"engine: {request.param}", flush=True)
    e = E.make(request.param)
    yield e
    e.close()


def _chain(nD, N, rng):
    """Frame-major float32 values of a chain: bonds in [1.0, 1.6], angles in
    [1.5, 2.5], dihedrals in [0, 6.2]."""
    D = 3 * nD + 3
    v = np.empty((N, D), dtype=np.float32)
    v[:, :nD + 2] = rng.uniform(1.0, 1.6, (N, nD + 2))
    v[:, nD + 2:2 * nD + 3] = rng.uniform(1.5, 2.5, (N, nD + 1))
    v[:, 2 * nD + 3:] = rng.uniform(0.0, 6.2, (N, nD))
    return v


def test_dihedral_one_value_per_bin(engine):
    """Values x0 + k h (k < B), m frames each, land one to a bin: S1D =
    ln(B delta) + (B - 1) / (2N) for the dihedral."""
    nD, B, m = 5, 20, 50
    N = B * m
    rng = np.random.default_rng(11)
    v = _chain(nD, N, rng)
    g = 2 * nD + 3 + 2
    v[:, g] = rng.permutation(np.repeat(0.7 + 0.11 * np.arange(B), m))
    r = engine.run(v, [B] * 6)
    delta = (r.vmax[g] - r.vmin[g]) / B
    assert abs(r.s1d[g] - (math.log(B * delta) + (B - 1) / (2 * N))) < 1e-12


def test_product_grid_mi(engine):
    """A bond and two dihedrals whose frames enumerate every (ka, kb, kc) cell
    m times: each pair's plug-in MI is 0 and MI = (Bi + Bj - Bi Bj - 1)/(2N)."""
    nD, Bb, Ba, Bd, m = 4, 6, 7, 9, 2
    N = m * Bb * Bd * Bd
    rng = np.random.default_rng(12)
    v = _chain(nD, N, rng)
    ka, kb, kc = (np.array(t) for t in zip(*itertools.product(range(Bb), range(Bd), range(Bd))))
    order = rng.permutation(N)
    bond, d1, d2 = 0, 2 * nD + 3, 2 * nD + 3 + 3
    v[:, bond] = np.tile(1.0 + 0.05 * ka, m)[order]
    v[:, d1] = np.tile(0.5 + 0.2 * kb, m)[order]
    v[:, d2] = np.tile(1.1 + 0.3 * kc, m)[order]
    r = engine.run(v, [Bb, Ba, Bd, Bb, Ba, Bd])
    for (i, bi), (j, bj) in [((bond, Bb), (d1, Bd)), ((bond, Bb), (d2, Bd)), ((d1, Bd), (d2, Bd))]:
        mi = E.mutual(nD, r, i, j)
        assert abs(mi - (bi + bj - bi * bj - 1) / (2 * N)) < 1e-12, (i, j)


def test_duplicated_dihedral_mi(engine):
    """Dihedral 3 a copy of dihedral 1: the same modfit shift and range, a
    diagonal joint histogram, so MI = S1D - ln(delta)."""
    nD, B, N = 6, 25, 4000
    rng = np.random.default_rng(13)
    v = _chain(nD, N, rng)
    i, j = 2 * nD + 3 + 1, 2 * nD + 3 + 3
    v[:, i] = rng.uniform(0.3, 2.7, N)
    v[:, j] = v[:, i]
    r = engine.run(v, [B] * 6)
    assert r.vmin[i] == r.vmin[j] and r.vmax[i] == r.vmax[j] and r.modfit[i] == r.modfit[j]
    delta = (r.vmax[i] - r.vmin[i]) / B
    assert abs(E.mutual(nD, r, i, j) - (r.s1d[i] - math.log(delta))) < 1e-10


def _kruskal(V, w):
    """Maximum spanning tree by Kruskal on a dense weight matrix (distinct weights)."""
    parent = list(range(V))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    tree = set()
    for _, i, j in sorted(((w[i, j], i, j) for i in range(V) for j in range(i + 1, V)), reverse=True):
        ri, rj = find(i), find(j)
        if ri != rj:
            parent[ri] = rj
            tree.add((i, j))
    return tree


@pytest.mark.parametrize("seed", [3, 4])
def test_mist_is_the_maximum_spanning_tree(engine, seed):
    nD = 5
    D = 3 * nD + 3
    rng = np.random.default_rng(seed)
    s1 = rng.normal(size=D)
    s2 = rng.normal(size=D * (D - 1) // 2)
    edges, mi, left, tot = engine.mist(nD, s1, s2)
    first = (0, nD + 2, 2 * nD + 3)
    got = {tuple(sorted((first[e[0]] + e[2], first[e[1]] + e[3]))) for e in edges}
    r = E.Result(s1, s2, None, None, None)
    w = np.zeros((D, D))
    for i, j in itertools.combinations(range(D), 2):
        w[i, j] = w[j, i] = E.mutual(nD, r, i, j)
    assert got == _kruskal(D, w)
    assert list(left) == list(range(D - 1, 0, -1))
    assert abs(tot[9] - (s1.sum() - mi.sum())) < 1e-9
