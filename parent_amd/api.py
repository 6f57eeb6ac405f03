"""ctypes binding of ``libparent_hip.so`` (C ABI in ``include/parent_hip.h``)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libparent_hip.so")
_lib = None

METHOD_PLAIN, METHOD_ATOMIC, METHOD_MFMA, METHOD_AUTO, METHOD_SLABS = 0, 1, 2, 3, 4
STATUS = {0: "OK", 1: "E_INVALID", 2: "E_NOMEM", 3: "E_IO", 4: "E_DATA", 5: "E_RANGE",
          6: "E_DEVICE", 7: "E_UNSUPPORTED"}
TYPE_CHAR = "bad"


class ParentError(RuntimeError):
    """A C-ABI call returned a nonzero status (``code`` = PARENT_E_*)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def library_path() -> str:
    return _LIB_PATH


def lib() -> C.CDLL:
    """Loads the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ParentError(6, f"{_LIB_PATH} is missing: build it with __graft_entry__.build() "
                             "or `make -C parent_amd`")
    L = C.CDLL(_LIB_PATH)
    p, i32, i64, u64, dp = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_void_p
    sig = {
        "parent_bat_open": (i32, [C.c_char_p, C.POINTER(p)]),
        "parent_bat_info": (i32, [p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "parent_bat_read_frames": (i32, [p, p]),
        "parent_bat_close": (None, [p]),
        "parent_par_write": (i32, [C.c_char_p, p, p, dp, dp]),
        "parent_par_read": (i32, [C.c_char_p, C.POINTER(i32), C.POINTER(i32), p, dp, dp]),
        "parent_last_error": (C.c_char_p, []),
        "parent_engine_create": (i32, [i32, C.POINTER(p)]),
        "parent_engine_destroy": (None, [p]),
        "parent_engine_error": (C.c_char_p, [p]),
        "parent_engine_set_trajectory": (i32, [p, i32, i32, i32, p, i32]),
        "parent_engine_prepare": (i32, [p, p]),
        "parent_engine_pairs": (i32, [p, i32]),
        "parent_engine_read_s1d": (i32, [p, dp]),
        "parent_engine_read_s2d": (i32, [p, dp]),
        "parent_engine_read_ranges": (i32, [p, dp, dp, dp]),
        "parent_engine_read_bins": (i32, [p, p]),
        "parent_engine_read_bins16": (i32, [p, p]),
        "parent_engine_pair_counts": (i32, [p, i32, i32, p, p]),
        "parent_engine_mist": (i32, [p, p, dp, p, dp]),
        "parent_engine_set_entropies": (i32, [p, i32, dp, dp]),
        "parent_engine_mist_noalloc": (i32, [p, i32, p, p, p, p, p, p, p, p]),
        "parent_engine_sync": (i32, [p]),
        "parent_engine_autotune": (i32, [p, p, i32, C.POINTER(i32)]),
        "parent_batch_set_variant": (i32, [p, i32, i32]),
        "parent_batch_set_classes": (i32, [p, i32]),
        "parent_engine_set_variant": (i32, [p, i32, i32]),
        "parent_engine_set_mist_path": (i32, [p, i32]),
        "parent_engine_variant_count": (i32, [p, i32, C.POINTER(i32)]),
        "parent_engine_stats": (i32, [p, p]),
        "parent_engine_sweep": (i32, [p, i32, p, i32]),
        "parent_engine_select": (i32, [p, i32]),
        "parent_batch_create": (i32, [i32, p, C.POINTER(p)]),
        "parent_batch_destroy": (None, [p]),
        "parent_batch_error": (C.c_char_p, [p]),
        "parent_batch_add": (i32, [p, i32, i32, i32, p, i32, C.POINTER(i32)]),
        "parent_batch_add_binned": (i32, [p, p, i32, i32, i32, p, i32, C.POINTER(i32)]),
        "parent_batch_run": (i32, [p, i32]),
        "parent_batch_read": (i32, [p, i32, dp, dp]),
        "parent_batch_mist": (i32, [p, i32, p, dp, p, dp]),
        "parent_batch_stats": (i32, [p, p]),
        "parent_batch_autotune": (i32, [p, p, i32, C.POINTER(i32)]),
        "parent_synth_values": (i32, [i32, i32, u64, p, i32]),
        "parent_synth_write_bat": (i32, [C.c_char_p, i32, i32, u64, i32]),
        "parent_mie_write": (i32, [C.c_char_p, i32, dp, dp, i32, dp]),
        "parent_bat_record_bytes": (i64, [p]),
        "parent_bat_read_raw": (i32, [p, i32, i32, p]),
        "parent_engine_load_bat": (i32, [p, p]),
        "parent_bat_build": (i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, i32, i32, C.POINTER(i32)]),
    }
    del i64
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class _TuneEntry(C.Structure):
    _fields_ = [("method", C.c_int), ("variant", C.c_int), ("tile_rows", C.c_int), ("tile_cols", C.c_int),
                ("block", C.c_int), ("ms", C.c_double), ("updates", C.c_double), ("updates_per_s", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("ms_upload", C.c_double), ("ms_prepare", C.c_double), ("ms_pairs", C.c_double),
                ("ms_mist", C.c_double), ("pair_updates", C.c_double), ("method", C.c_int),
                ("kernel_variant", C.c_int), ("ms_sweep", C.c_double), ("sweep_sets", C.c_int),
                ("mist_path", C.c_int), ("batch_launches", C.c_int)]


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _check_host(rc: int) -> None:
    if rc != 0:
        raise ParentError(rc, lib().parent_last_error().decode())


def dof_counts(n_dihedrals: int):
    """(nB, nA, nD, D) of a trajectory (PARENT.cpp:298-299)."""
    return n_dihedrals + 2, n_dihedrals + 1, n_dihedrals, 3 * n_dihedrals + 3


def par_pos(n_dihedrals: int, gi: int, gj: int) -> int:
    """Offset of pair gi < gj in the .par 2D section (io_binary.cpp:606-631, util.cpp:52-78)."""
    nB, nA, nD, _ = dof_counts(n_dihedrals)
    firsts = (0, nB, nB + nA)

    def typ(g):
        return 0 if g < nB else (1 if g < nB + nA else 2)
    t1, t2 = typ(gi), typ(gj)
    i, j = gi - firsts[t1], gj - firsts[t2]
    oBA = nB * (nB - 1) // 2
    oBD = oBA + nB * nA
    oAA = oBD + nB * nD
    oAD = oAA + nA * (nA - 1) // 2
    oDD = oAD + nA * nD
    if (t1, t2) == (0, 0):
        return (nB - i) * (nB - i - 1) // 2 + i - j
    if (t1, t2) == (0, 1):
        return oBA + i * nA + j
    if (t1, t2) == (0, 2):
        return oBD + i * nD + j
    if (t1, t2) == (1, 1):
        return oAA + (nA - i) * (nA - i - 1) // 2 + i - j
    if (t1, t2) == (1, 2):
        return oAD + i * nD + j
    return oDD + (nD - i) * (nD - i - 1) // 2 + i - j


def get_mutual(n_dihedrals: int, s1d: np.ndarray, s2d: np.ndarray, gi: int, gj: int) -> float:
    """MI of two dofs from .par arrays: (S1D[lo] + S1D[hi]) - S2D (util.cpp:47-81)."""
    lo, hi = min(gi, gj), max(gi, gj)
    return (s1d[lo] + s1d[hi]) - s2d[par_pos(n_dihedrals, lo, hi)]


class Bat:
    """An opened .bat (read_BAT_header, io_binary.cpp:107-200)."""

    def __init__(self, path: str):
        L = lib()
        h = C.c_void_p()
        _check_host(L.parent_bat_open(os.fsencode(path), C.byref(h)))
        self._h = h
        nD, N, dp = C.c_int(), C.c_int(), C.c_int()
        L.parent_bat_info(h, C.byref(nD), C.byref(N), C.byref(dp))
        self.n_dihedrals, self.n_frames, self.double_prec = nD.value, N.value, dp.value
        self.D = 3 * self.n_dihedrals + 3

    def read_frames(self) -> np.ndarray:
        """Frame-major values [N][D] (float32 or float64)."""
        out = np.empty((self.n_frames, self.D), dtype=np.float64 if self.double_prec else np.float32)
        _check_host(lib().parent_bat_read_frames(self._h, _ptr(out)))
        return out

    @property
    def record_bytes(self) -> int:
        """Bytes of one frame record (io_binary.cpp:281-368)."""
        return int(lib().parent_bat_record_bytes(self._h))

    def read_raw(self, first: int, n: int) -> np.ndarray:
        """Frame records [first, first + n) as stored in the file, [n][record_bytes] uint8."""
        out = np.empty((n, self.record_bytes), dtype=np.uint8)
        _check_host(lib().parent_bat_read_raw(self._h, first, n, _ptr(out)))
        return out

    def write_par(self, path: str, bins, s1d: np.ndarray, s2d: np.ndarray) -> None:
        b = np.ascontiguousarray(bins, dtype=np.int32)
        s1 = np.ascontiguousarray(s1d, dtype=np.float64)
        s2 = np.ascontiguousarray(s2d, dtype=np.float64)
        _check_host(lib().parent_par_write(os.fsencode(path), self._h, _ptr(b), _ptr(s1), _ptr(s2)))

    def close(self) -> None:
        if self._h:
            lib().parent_bat_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_par(path: str):
    """(n_dihedrals, n_frames, bins[6], s1d, s2d) of a .par (io_binary.cpp:446-673)."""
    L = lib()
    nD, N = C.c_int(), C.c_int()
    bins = np.zeros(6, dtype=np.int32)
    _check_host(L.parent_par_read(os.fsencode(path), C.byref(nD), C.byref(N), _ptr(bins), None, None))
    _, _, _, D = dof_counts(nD.value)
    s1 = np.empty(D, dtype=np.float64)
    s2 = np.empty(D * (D - 1) // 2, dtype=np.float64)
    _check_host(L.parent_par_read(os.fsencode(path), C.byref(nD), C.byref(N), _ptr(bins), _ptr(s1), _ptr(s2)))
    return nD.value, N.value, bins, s1, s2


def mie_write(path, n_dihedrals: int, s1d: np.ndarray, s2d: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Writes the get_PAR_MIE.x listing of a .par body to `path` ("-" for
    stdout, None for totals only); returns the 16 totals in the order of the
    listing's TOTAL lines (get_PAR_MIE.cpp:109-199)."""
    s1 = np.ascontiguousarray(s1d, dtype=np.float64)
    s2 = np.ascontiguousarray(s2d, dtype=np.float64)
    _, _, _, D = dof_counts(n_dihedrals)
    if s1.size != D or s2.size != D * (D - 1) // 2:
        raise ValueError("s1d / s2d sizes do not match n_dihedrals")
    tot = np.empty(16, dtype=np.float64)
    _check_host(lib().parent_mie_write(os.fsencode(path) if path is not None else None, n_dihedrals,
                                       _ptr(s1), _ptr(s2), nthreads, _ptr(tot)))
    return tot


def bat_build(top: str, xtc: str, bat: str, backbone: str = "CA C N H1 O1", double_prec: int = 1,
              nthreads: int = 0) -> int:
    """BAT_builder.x -t top -x xtc -o bat -bb backbone (bat.cpp:82-176): writes the .bat and
    returns its frame count."""
    n = C.c_int()
    _check_host(lib().parent_bat_build(os.fsencode(top), os.fsencode(xtc), os.fsencode(bat), backbone.encode(),
                                       double_prec, nthreads, C.byref(n)))
    return n.value


def synth_values(atoms: int, n_frames: int, seed: int, nthreads: int = 0) -> np.ndarray:
    """Synthetic chain trajectory, frame-major float32 [N][3*atoms-6]."""
    out = np.empty((n_frames, 3 * atoms - 6), dtype=np.float32)
    _check_host(lib().parent_synth_values(atoms, n_frames, seed, _ptr(out), nthreads))
    return out


def synth_write_bat(path: str, atoms: int, n_frames: int, seed: int, nthreads: int = 0) -> None:
    _check_host(lib().parent_synth_write_bat(os.fsencode(path), atoms, n_frames, seed, nthreads))


class Engine:
    """One GPU engine (one HIP stream); mirrors parent_engine_* of the C ABI."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        _check_host(L.parent_engine_create(device, C.byref(h)))
        self._h = h
        self.n_dihedrals = None
        self.n_frames = None
        self._keep = None

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ParentError(rc, lib().parent_engine_error(self._h).decode())

    @property
    def D(self) -> int:
        return 3 * self.n_dihedrals + 3

    def set_trajectory(self, values, n_dihedrals: int | None = None, device_ptr: int | None = None,
                       n_frames: int | None = None, double_prec: int | None = None) -> None:
        """values: frame-major [N][D] float32/float64 host array, or a device pointer via device_ptr."""
        if device_ptr is not None:
            self._check(lib().parent_engine_set_trajectory(self._h, n_dihedrals, n_frames, double_prec,
                                                           C.c_void_p(device_ptr), 1))
            self.n_dihedrals, self.n_frames = n_dihedrals, n_frames
            return
        v = np.ascontiguousarray(values)
        if v.dtype not in (np.float32, np.float64) or v.ndim != 2 or (v.shape[1] - 3) % 3:
            raise ValueError("values must be [N][3*nD+3] float32 or float64")
        nD = (v.shape[1] - 3) // 3
        self._check(lib().parent_engine_set_trajectory(self._h, nD, v.shape[0], int(v.dtype == np.float64),
                                                       _ptr(v), 0))
        self.n_dihedrals, self.n_frames = nD, v.shape[0]

    def load_bat(self, bat: "Bat") -> None:
        """Streams an opened .bat's frames to the device (parent_engine_load_bat)."""
        self._check(lib().parent_engine_load_bat(self._h, bat._h))
        self.n_dihedrals, self.n_frames = bat.n_dihedrals, bat.n_frames

    def prepare(self, bins) -> None:
        b = np.ascontiguousarray(bins, dtype=np.int32)
        if b.size == 1:
            b = np.full(6, int(b.reshape(-1)[0]), dtype=np.int32)
        self._check(lib().parent_engine_prepare(self._h, _ptr(b)))

    def pairs(self, method: int = METHOD_AUTO) -> None:
        self._check(lib().parent_engine_pairs(self._h, method))

    def sweep(self, bin_sets, method: int = METHOD_AUTO) -> None:
        """Prepare + all pairs for every bin set (rows of 6, or single counts) as one
        captured hipGraph launch; parent_engine_sweep.  select(k) then picks set k."""
        rows = [np.full(6, int(b), dtype=np.int32) if np.ndim(b) == 0 else np.asarray(b, dtype=np.int32)
                for b in bin_sets]
        arr = np.ascontiguousarray(np.stack(rows), dtype=np.int32)
        self._check(lib().parent_engine_sweep(self._h, arr.shape[0], _ptr(arr), method))

    def select(self, k: int) -> None:
        self._check(lib().parent_engine_select(self._h, k))

    def s1d(self) -> np.ndarray:
        out = np.empty(self.D, dtype=np.float64)
        self._check(lib().parent_engine_read_s1d(self._h, _ptr(out)))
        return out

    def s2d(self) -> np.ndarray:
        D = self.D
        out = np.empty(D * (D - 1) // 2, dtype=np.float64)
        self._check(lib().parent_engine_read_s2d(self._h, _ptr(out)))
        return out

    def ranges(self):
        D = self.D
        mn, mx, mf = (np.empty(D, dtype=np.float64) for _ in range(3))
        self._check(lib().parent_engine_read_ranges(self._h, _ptr(mn), _ptr(mx), _ptr(mf)))
        return mn, mx, mf

    def bins(self) -> np.ndarray:
        out = np.empty((self.D, self.n_frames), dtype=np.uint8)
        self._check(lib().parent_engine_read_bins(self._h, _ptr(out)))
        return out

    def bins16(self) -> np.ndarray:
        """The 2D bin indices as uint16 (any bin count)."""
        out = np.empty((self.D, self.n_frames), dtype=np.uint16)
        self._check(lib().parent_engine_read_bins16(self._h, _ptr(out)))
        return out

    def pair_counts(self, pairs, method: int, bins2d) -> list:
        """Histograms of the listed (gi < gj) pairs through `method`'s fill path."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        nB, nA, _, _ = dof_counts(self.n_dihedrals)

        def nb(g):
            return bins2d[0] if g < nB else (bins2d[1] if g < nB + nA else bins2d[2])
        sizes = [nb(a) * nb(b) for a, b in pr]
        out = np.empty(int(sum(sizes)), dtype=np.int32)
        self._check(lib().parent_engine_pair_counts(self._h, method, len(pr), _ptr(pr), _ptr(out)))
        res, o = [], 0
        for (a, b), s in zip(pr, sizes):
            res.append(out[o:o + s].reshape(nb(a), nb(b)))
            o += s
        return res

    def set_entropies(self, n_dihedrals: int, s1d: np.ndarray, s2d: np.ndarray) -> None:
        s1 = np.ascontiguousarray(s1d, dtype=np.float64)
        s2 = np.ascontiguousarray(s2d, dtype=np.float64)
        self._check(lib().parent_engine_set_entropies(self._h, n_dihedrals, _ptr(s1), _ptr(s2)))
        self.n_dihedrals = n_dihedrals

    def mist(self):
        """(edges [V-1,4], mi [V-1], left [V-1], totals [10]) -- get_PAR_MIST.cpp:396-766."""
        V = self.D
        edges = np.empty((V - 1, 4), dtype=np.int32)
        mi = np.empty(V - 1, dtype=np.float64)
        left = np.empty(V - 1, dtype=np.int32)
        tot = np.empty(10, dtype=np.float64)
        self._check(lib().parent_engine_mist(self._h, _ptr(edges), _ptr(mi), _ptr(left), _ptr(tot)))
        return edges, mi, left, tot

    def mist_noalloc(self, n_dihedrals: int, s1d: int, s2d: int, edges: int, mi: int, left: int, totals: int,
                     status: int, stream: int | None = None) -> None:
        """parent_engine_mist_noalloc: every argument is a device address (int);
        enqueued on `stream` (a hipStream_t as int, None: the engine's), not waited for."""
        v = [C.c_void_p(x) for x in (s1d, s2d, edges, mi, left, totals, status)]
        self._check(lib().parent_engine_mist_noalloc(self._h, n_dihedrals, *v,
                                                     C.c_void_p(stream) if stream else None))

    def autotune(self) -> list:
        """Times every pair-kernel configuration on a sample; returns the table."""
        arr = (_TuneEntry * 256)()
        n = C.c_int()
        self._check(lib().parent_engine_autotune(self._h, C.cast(arr, C.c_void_p), 256, C.byref(n)))
        return [{k: getattr(arr[i], k) for k, _ in _TuneEntry._fields_} for i in range(min(n.value, 256))]

    def set_variant(self, method: int, variant: int) -> None:
        """Pins the configuration of `method` (-1: default)."""
        self._check(lib().parent_engine_set_variant(self._h, method, variant))

    def set_mist_path(self, mode: int) -> None:
        """0: the spanning-tree MIST path where it applies (default); 1: always the O(V^2) Prim."""
        self._check(lib().parent_engine_set_mist_path(self._h, mode))

    def variant_count(self, method: int) -> int:
        n = C.c_int()
        self._check(lib().parent_engine_variant_count(self._h, method, C.byref(n)))
        return n.value

    def sync(self) -> None:
        self._check(lib().parent_engine_sync(self._h))

    def stats(self) -> dict:
        s = _Stats()
        self._check(lib().parent_engine_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().parent_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent trajectories with one set of bins, all pairs in one launch
    (parent_batch_*)."""

    def __init__(self, bins, device: int = 0):
        b = np.ascontiguousarray(bins, dtype=np.int32)
        if b.size == 1:
            b = np.full(6, int(b.reshape(-1)[0]), dtype=np.int32)
        h = C.c_void_p()
        _check_host(lib().parent_batch_create(device, _ptr(b), C.byref(h)))
        self._h = h
        self.n_dihedrals = []

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ParentError(rc, lib().parent_batch_error(self._h).decode())

    def add(self, values, bins=None) -> int:
        """Adds a frame-major [N][D] float32/float64 host array, with its own
        bins[6] or the batch's; returns its index."""
        v = np.ascontiguousarray(values)
        if v.dtype not in (np.float32, np.float64) or v.ndim != 2 or (v.shape[1] - 3) % 3:
            raise ValueError("values must be [N][3*nD+3] float32 or float64")
        nD = (v.shape[1] - 3) // 3
        idx = C.c_int()
        if bins is None:
            self._check(lib().parent_batch_add(self._h, nD, v.shape[0], int(v.dtype == np.float64), _ptr(v), 0,
                                               C.byref(idx)))
        else:
            b = np.ascontiguousarray(bins, dtype=np.int32)
            if b.size == 1:
                b = np.full(6, int(b.reshape(-1)[0]), dtype=np.int32)
            self._check(lib().parent_batch_add_binned(self._h, _ptr(b), nD, v.shape[0], int(v.dtype == np.float64),
                                                      _ptr(v), 0, C.byref(idx)))
        self.n_dihedrals.append(nD)
        return idx.value

    def set_variant(self, method: int, variant: int) -> None:
        """Pins the configuration `run(method)` uses for METHOD_ATOMIC,
        METHOD_MFMA (0 int8, 1 FP4) or METHOD_SLABS (-1: default)."""
        self._check(lib().parent_batch_set_variant(self._h, method, variant))

    def set_classes(self, mode: int) -> None:
        """Frame-count classes of `run(METHOD_AUTO)`: 1 split when predicted
        faster (default), 2 always when both classes exist, 0 never."""
        self._check(lib().parent_batch_set_classes(self._h, mode))

    def run(self, method: int = METHOD_AUTO) -> None:
        self._check(lib().parent_batch_run(self._h, method))

    def autotune(self) -> list:
        """The batch autotuner's table (ms = predicted time of the whole queue)."""
        arr = (_TuneEntry * 256)()
        n = C.c_int()
        self._check(lib().parent_batch_autotune(self._h, C.cast(arr, C.c_void_p), 256, C.byref(n)))
        return [{k: getattr(arr[i], k) for k, _ in _TuneEntry._fields_} for i in range(min(n.value, 256))]

    def results(self, k: int):
        """(s1d, s2d) of trajectory k."""
        D = 3 * self.n_dihedrals[k] + 3
        s1 = np.empty(D, dtype=np.float64)
        s2 = np.empty(D * (D - 1) // 2, dtype=np.float64)
        self._check(lib().parent_batch_read(self._h, k, _ptr(s1), _ptr(s2)))
        return s1, s2

    def mist(self, k: int):
        V = 3 * self.n_dihedrals[k] + 3
        edges = np.empty((V - 1, 4), dtype=np.int32)
        mi = np.empty(V - 1, dtype=np.float64)
        left = np.empty(V - 1, dtype=np.int32)
        tot = np.empty(10, dtype=np.float64)
        self._check(lib().parent_batch_mist(self._h, k, _ptr(edges), _ptr(mi), _ptr(left), _ptr(tot)))
        return edges, mi, left, tot

    def stats(self) -> dict:
        s = _Stats()
        self._check(lib().parent_batch_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().parent_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_parent(bat_path: str, par_path: str, bins, method: int = METHOD_AUTO, device: int = 0) -> dict:
    """PARENT.x in.bat out.par b1D a1D d1D b2D a2D d2D (PARENT.cpp:93-1512) on the GPU."""
    with Bat(bat_path) as bat, Engine(device) as eng:
        eng.load_bat(bat)
        eng.prepare(bins)
        eng.pairs(method)
        bat.write_par(par_path, bins, eng.s1d(), eng.s2d())
        return eng.stats()


def write_mist_text(path: str, edges, mi, left, totals) -> None:
    """The text of get_PAR_MIST.x (get_PAR_MIST.cpp:663-741, 754-765), precision 12."""
    def g12(x: float) -> str:
        return _fmt12(x)
    lines = [f"{TYPE_CHAR[e[0]]} {TYPE_CHAR[e[1]]} {e[2] + 1} {e[3] + 1}   MI: {g12(m)}   left to process: {n}"
             for e, m, n in zip(edges, mi, left)]
    names = ["TOTAL 1D BONDS ENTROPY", "TOTAL 1D ANGLES ENTROPY", "TOTAL 1D DIHEDRALS ENTROPY",
             "TOTAL 2D BONDS-BONDS MUTUAL INFORMATION", "TOTAL 2D BONDS-ANGLES MUTUAL INFORMATION",
             "TOTAL 2D BONDS-DIHEDRALS MUTUAL INFORMATION", "TOTAL 2D ANGLES-ANGLES MUTUAL INFORMATION",
             "TOTAL 2D ANGLES-DIHEDRALS MUTUAL INFORMATION", "TOTAL 2D DIHEDRALS-DIHEDRALS MUTUAL INFORMATION",
             "TOTAL CONFIGURATIONAL ENTROPY"]
    lines += [f"{n} = {g12(t)}" for n, t in zip(names, totals)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _fmt12(x: float) -> str:
    """iostream's default formatting at precision(12) (%.12g without a trailing-zero pad)."""
    s = "%.12g" % x
    return s


def run_mist(par_path: str, out_txt: str, device: int = 0):
    """get_PAR_MIST.x in.par out.txt (get_PAR_MIST.cpp:64-791) with the Prim on the GPU."""
    nD, _, _, s1, s2 = read_par(par_path)
    with Engine(device) as eng:
        eng.set_entropies(nD, s1, s2)
        res = eng.mist()
    write_mist_text(out_txt, *res)
    return res
