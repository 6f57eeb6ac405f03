"""Two implementations of the engine's operations behind one interface, so that
the same test body runs against either (tests/test_engine_swap.py):

  * OracleEngine -- the CPU restatement of the reference (oracle/, through
    oracle_api.py), the checker;
  * HipEngine    -- the product: the HIP library through parent_amd.Engine.

Operations (PARENT.cpp:503-1479 and get_PAR_MIST.cpp:396-766): run a
frame-major trajectory [N][3*nD+3] at six bin counts and return S1D, S2D in
.par order and the per-dof ranges; MIST from given S1D / S2D.
"""
from dataclasses import dataclass

import numpy as np

import oracle_api as O


@dataclass
class Result:
    s1d: np.ndarray
    s2d: np.ndarray
    vmin: np.ndarray
    vmax: np.ndarray
    modfit: np.ndarray


class OracleEngine:
    name = "oracle"

    def run(self, values, bins) -> Result:
        nD = (values.shape[1] - 3) // 3
        s1, s2, mn, mx, mf, _ = O.compute(nD, values.shape[0], values.T.astype(np.float64), bins)
        return Result(s1, s2, mn, mx, mf)

    def mist(self, nD, s1d, s2d):
        return O.mist(nD, s1d, s2d)

    def close(self):
        pass


class HipEngine:
    name = "hip"

    def __init__(self):
        import parent_amd
        self._pa = parent_amd
        self._e = parent_amd.Engine(0)

    def run(self, values, bins) -> Result:
        self._e.set_trajectory(values)
        self._e.prepare(bins)
        self._e.pairs(self._pa.METHOD_AUTO)
        mn, mx, mf = self._e.ranges()
        return Result(self._e.s1d(), self._e.s2d(), mn, mx, mf)

    def mist(self, nD, s1d, s2d):
        self._e.set_entropies(nD, s1d, s2d)
        return self._e.mist()

    def close(self):
        self._e.close()


def make(name):
    return OracleEngine() if name == "oracle" else HipEngine()


def pair_pos(nD, i, j):
    """Position of pair (i, j) in the S2D array (.par order, util.cpp:47-81)."""
    return int(O.lib().oracle_par_pos(nD, int(i), int(j)))


def mutual(nD, r: Result, i, j):
    return r.s1d[i] + r.s1d[j] - r.s2d[pair_pos(nD, i, j)]
