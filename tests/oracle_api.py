"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker the HIP engine is compared with.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
ORACLE_EXE = os.path.join(ROOT, "oracle", "_build", "oracle_parent.x")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
B, A, D = 0, 1, 2
_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        p, i32, dp = C.c_void_p, C.c_int, C.c_double
        sig = {
            "oracle_bat_read": (i32, [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(p)]),
            "oracle_free": (None, [p]),
            "oracle_dof_range": (i32, [i32, i32, p, dp, p, p, p]),
            "oracle_bin_index": (i32, [i32, p, dp, dp, i32, p]),
            "oracle_entropy1d": (dp, [i32, i32, p, dp, dp, i32]),
            "oracle_entropy2d": (dp, [i32, p, dp, dp, i32, i32, p, dp, dp, i32, i32, p]),
            "oracle_par_pos": (C.c_int64, [i32, i32, i32]),
            "oracle_compute": (i32, [i32, i32, p, p, p, p, p, p, p, i32]),
            "oracle_compute_subset": (i32, [i32, i32, i32, p, p, p, i32, p, p, p, i32]),
            "oracle_mutual": (dp, [i32, p, p, i32, i32]),
            "oracle_mist": (i32, [i32, p, p, p, p, p, p]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _o = L
    return _o


def _p(a):
    return C.c_void_p(a.ctypes.data)


PI = float(np.arccos(-1.0))


def type_of(nD, g):
    return B if g < nD + 2 else (A if g < 2 * nD + 3 else D)


def bat_read(path):
    """(nD, N, double_prec, values dof-major [D][N] float64)."""
    nD, N, dp, v = C.c_int(), C.c_int(), C.c_int(), C.c_void_p()
    rc = lib().oracle_bat_read(os.fsencode(path), C.byref(nD), C.byref(N), C.byref(dp), C.byref(v))
    if rc:
        raise RuntimeError(f"oracle_bat_read failed: {rc}")
    Dn = 3 * nD.value + 3
    arr = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_double)), shape=(Dn, N.value)).copy()
    lib().oracle_free(v)
    return nD.value, N.value, dp.value, arr


def dof_range(t, x):
    """Range rule of one dof; x (float64) is shifted in place for dihedrals. -> (rc, min, max, modfit)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    mn, mx, mf = C.c_double(), C.c_double(), C.c_double()
    rc = lib().oracle_dof_range(t, x.size, _p(x), PI, C.byref(mn), C.byref(mx), C.byref(mf))
    return rc, mn.value, mx.value, mf.value, x


def bin_index(x, vmin, vmax, bins):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(x.size, dtype=np.int32)
    bad = lib().oracle_bin_index(x.size, _p(x), vmin, vmax, bins, _p(out))
    return out, bad


def entropy1d(t, x, vmin, vmax, bins):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_entropy1d(t, x.size, _p(x), vmin, vmax, bins)


def entropy2d(t1, x1, mn1, mx1, b1, t2, x2, mn2, mx2, b2, counts=False):
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    h = np.zeros((b1, b2), dtype=np.int64) if counts else None
    s = lib().oracle_entropy2d(t1, _p(x1), mn1, mx1, b1, t2, _p(x2), mn2, mx2, b2, x1.size,
                               _p(h) if counts else None)
    return (s, h) if counts else s


def compute(nD, N, values_dof_major, bins, nthreads=0):
    """Full single-rank computation -> (s1d, s2d, vmin, vmax, modfit, shifted values)."""
    v = np.array(values_dof_major, dtype=np.float64, order="C", copy=True)
    Dn = 3 * nD + 3
    s1 = np.empty(Dn)
    s2 = np.empty(Dn * (Dn - 1) // 2)
    mn, mx, mf = np.empty(Dn), np.empty(Dn), np.empty(Dn)
    b = np.ascontiguousarray(bins, dtype=np.int32)
    rc = lib().oracle_compute(nD, N, _p(v), _p(b), _p(s1), _p(s2), _p(mn), _p(mx), _p(mf), nthreads)
    if rc:
        raise RuntimeError(f"oracle_compute failed: {rc}")
    return s1, s2, mn, mx, mf, v


def compute_subset(nD, N, dofs, cols_dof_major, bins, pairs, nthreads=0):
    """S1D of the selected dofs and S2D of pairs among them -> (s1d_sel, s2d_pairs).
    cols_dof_major: [len(dofs)][N] values; pairs: [n][2] indices into dofs."""
    d = np.ascontiguousarray(dofs, dtype=np.int32)
    v = np.array(cols_dof_major, dtype=np.float64, order="C", copy=True)
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    b = np.ascontiguousarray(bins, dtype=np.int32)
    s1 = np.empty(d.size)
    s2 = np.empty(pr.shape[0])
    rc = lib().oracle_compute_subset(nD, N, d.size, _p(d), _p(v), _p(b), pr.shape[0], _p(pr), _p(s1), _p(s2),
                                     nthreads)
    if rc:
        raise RuntimeError(f"oracle_compute_subset failed: {rc}")
    return s1, s2


def mist(nD, s1d, s2d):
    Dn = 3 * nD + 3
    s1 = np.ascontiguousarray(s1d, dtype=np.float64)
    s2 = np.ascontiguousarray(s2d, dtype=np.float64)
    edges = np.empty((Dn - 1, 4), dtype=np.int32)
    mi = np.empty(Dn - 1)
    left = np.empty(Dn - 1, dtype=np.int32)
    tot = np.empty(10)
    rc = lib().oracle_mist(nD, _p(s1), _p(s2), _p(edges), _p(mi), _p(left), _p(tot))
    if rc:
        raise RuntimeError("oracle_mist failed")
    return edges, mi, left, tot
