"""The reasoning behind the spanning-tree path of MIST, checked on the CPU.

A small Python model of that path -- Borůvka rounds in which every component
must have a strictly largest outgoing candidate, the tree rooted at vertex 0
by its Euler tour and list ranking, then Prim over the tree edges alone with
the reference's order words -- is compared with the oracle's reference-order
Prim (get_PAR_MIST.cpp:441-748) on random entropies: continuous ones, coarse
ones (many equal MI values), rounded ones, and ones with NaN or MI below
-1e200.  Wherever the model's certificate holds, its edges and MI bits must
equal the oracle's; where it does not, the GPU path hands over to its O(V^2)
Prim (the GPU tests check that hand-over).  This checks the claim the GPU
path rests on, not the GPU code."""
import struct

import numpy as np
import pytest

import oracle_api as O


def _ordered(x):
    if not (x > -1e200):
        return 0
    b = 0 if x == 0.0 else struct.unpack("<Q", struct.pack("<d", x))[0]
    return (~b & 0xFFFFFFFFFFFFFFFF) if (b >> 63) else (b | 0x8000000000000000)


def _type(nD, g):
    return 0 if g < nD + 2 else (1 if g < 2 * nD + 3 else 2)


def _first(nD, t):
    return 0 if t == 0 else (nD + 2 if t == 1 else 2 * nD + 3)


def _dense(nD, s1, s2):
    V = 3 * nD + 3
    M = np.zeros((V, V))
    for i in range(V):
        for j in range(i + 1, V):
            M[i, j] = M[j, i] = s1[i] + s1[j] - s2[O.lib().oracle_par_pos(nD, i, j)]
    return M


def _boruvka(M):
    """Tree edges, or None where a choice is not strict (or an MI is no candidate)."""
    V = M.shape[0]
    key = [[_ordered(M[v, u]) for u in range(V)] for v in range(V)]
    if any(key[v][u] == 0 for v in range(V) for u in range(V) if u != v):
        return None
    comp, tree = list(range(V)), []
    while len(tree) < V - 1:
        best = []
        for v in range(V):
            b1 = b2 = 0
            bi = -1
            for u in range(V):
                if u == v or comp[u] == comp[v]:
                    continue
                if key[v][u] > b1:
                    b2, b1, bi = b1, key[v][u], u
                elif key[v][u] > b2:
                    b2 = key[v][u]
            best.append((b1, b2, bi))
        reps = [c for c in range(V) if comp[c] == c]
        cb = {c: max(best[v][0] for v in range(V) if comp[v] == c) for c in reps}
        hook = {}
        for c in reps:
            holders = [v for v in range(V) if comp[v] == c and best[v][0] == cb[c]]
            if len(holders) != 1 or best[holders[0]][1] == cb[c]:
                return None
            hook[c] = (holders[0], best[holders[0]][2])
        for c in reps:
            x, y = hook[c]
            d = comp[y]
            if comp[hook[d][1]] == c and c > d:
                continue  # the pair chose each other: one copy of the edge
            tree.append((x, y))
        # components: connected pieces of the chosen edges
        parent = list(range(V))

        def find(a):
            while parent[a] != a:
                a = parent[a]
            return a

        for x, y in tree:
            parent[find(x)] = find(y)
        roots = {}
        comp = [roots.setdefault(find(v), min(u for u in range(V) if find(u) == find(v))) for v in range(V)]
    return tree


def _root(V, tree):
    """Parents with the tree rooted at 0: Euler tour from the first arc out of 0, list ranking."""
    A = 2 * len(tree)
    src = [0] * A
    for i, (x, y) in enumerate(tree):
        src[2 * i], src[2 * i + 1] = x, y
    out = [[] for _ in range(V)]
    pos = [0] * A
    for a in range(A):
        pos[a] = len(out[src[a]])
        out[src[a]].append(a)
    a0 = out[0][0]
    succ = []
    for a in range(A):
        v = src[a ^ 1]
        s = out[v][(pos[a ^ 1] + 1) % len(out[v])]
        succ.append(-1 if s == a0 else s)
    dist = [0 if s < 0 else 1 for s in succ]
    while any(s >= 0 for s in succ):
        dist = [dist[a] + (dist[succ[a]] if succ[a] >= 0 else 0) for a in range(A)]
        succ = [succ[succ[a]] if succ[a] >= 0 else -1 for a in range(A)]
    parent = [-1] * V
    for a in range(A):
        if dist[a] > dist[a ^ 1]:
            parent[src[a ^ 1]] = src[a]
    return parent


def _tree_prim(nD, M, parent):
    V = M.shape[0]
    live, cand, edges, mi = set(range(1, V)), {}, [], []

    def activate(w, pos):
        for v in live:
            if parent[v] == w:
                order = (_type(nD, w) << 62) | (_type(nD, v) << 60) | (pos << 30) | v
                cand[v] = (_ordered(M[w, v]), order, w)

    activate(0, 0)
    for step in range(V - 1):
        w = max(cand, key=lambda v: (cand[v][0], -cand[v][1]))
        u = cand.pop(w)[2]
        tu, tw = _type(nD, u), _type(nD, w)
        edges.append((tu, tw, u - _first(nD, tu), w - _first(nD, tw)))
        mi.append(M[u, w])
        live.discard(w)
        activate(w, step + 1)
    return edges, mi


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tree_path_model_equals_reference_prim(seed):
    rng = np.random.default_rng(seed)
    held = 0
    for trial in range(16):
        nD = int(rng.integers(1, 7))
        V = 3 * nD + 3
        P = V * (V - 1) // 2
        kind = trial % 4
        s1 = rng.normal(size=V) * 2
        s2 = rng.normal(size=P) * 3
        if kind == 1:
            s1, s2 = np.zeros(V), rng.integers(-3, 4, size=P).astype(float)
        elif kind == 2:
            s1, s2 = np.round(s1, 1), np.round(s2, 1)
        elif kind == 3:
            s2[rng.integers(0, P, 2)] = np.nan if trial % 8 == 3 else 1e201
        M = _dense(nD, s1, s2)
        tree = _boruvka(M)
        if tree is None:
            continue
        held += 1
        edges, mi = _tree_prim(nD, M, _root(V, tree))
        eo, mio, _, _ = O.mist(nD, s1, s2)
        assert [tuple(int(x) for x in r) for r in eo] == edges, f"trial {trial}"
        assert np.array_equal(np.asarray(mi).view(np.uint64), mio.view(np.uint64)), f"trial {trial}"
    assert held >= 4  # the continuous cases always hold the certificate
