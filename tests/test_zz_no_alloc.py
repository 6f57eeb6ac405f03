"""parent_engine_mist_noalloc (MIST with its scratch in a static device array,
no device allocation) against parent_engine_mist on the same entropies:
edges, MI bits, left counts and totals identical, on two calls in a row, the
first with 8,991 dofs and the second with 294 on a caller's stream, so stale
scratch of the first call must not reach the second.

The test's own device buffers and stream come from the HIP runtime that the
engine library links (called through ctypes), so one HIP runtime serves the
whole process."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree / streams of the runtime the engine uses."""

    def __init__(self):
        import parent_amd
        parent_amd.lib()  # loads libamdhip64 with the engine library
        self.L = L = C.CDLL("libamdhip64.so.7")
        p = C.c_void_p
        for name, args in (("hipMalloc", [C.POINTER(p), C.c_size_t]), ("hipFree", [p]),
                           ("hipMemcpy", [p, p, C.c_size_t, C.c_int]), ("hipStreamCreate", [C.POINTER(p)]),
                           ("hipStreamDestroy", [p]), ("hipStreamSynchronize", [p]), ("hipSetDevice", [C.c_int])):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        assert L.hipSetDevice(0) == 0
        self.bufs = []

    def alloc(self, nbytes):
        ptr = C.c_void_p()
        assert self.L.hipMalloc(C.byref(ptr), max(int(nbytes), 8)) == 0
        self.bufs.append(ptr)
        return ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(C.c_void_p(d), a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return d

    def filled(self, n, dtype, value):
        return self.upload(np.full(n, value, dtype=dtype))

    def download(self, d, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(d), out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for b in self.bufs:
            self.L.hipFree(b)
        self.bufs = []


def _noalloc(eng, hip, nD, s1, s2, stream):
    V = 3 * nD + 3
    d1, d2 = hip.upload(s1), hip.upload(s2)
    de = hip.filled((V - 1) * 4, np.int32, -7)
    dm = hip.filled(V - 1, np.float64, np.nan)
    dl = hip.filled(V - 1, np.int32, -7)
    dt = hip.filled(10, np.float64, np.nan)
    ds = hip.filled(1, np.int32, -7)
    eng.mist_noalloc(nD, d1, d2, de, dm, dl, dt, ds, stream)
    if stream:
        assert hip.L.hipStreamSynchronize(C.c_void_p(stream)) == 0
    else:
        eng.sync()
    return (hip.download(de, (V - 1) * 4, np.int32).reshape(V - 1, 4), hip.download(dm, V - 1, np.float64),
            hip.download(dl, V - 1, np.int32), hip.download(dt, 10, np.float64),
            int(hip.download(ds, 1, np.int32)[0]))


def test_mist_noalloc_matches_engine_mist():
    import parent_amd as pa
    hip = _Hip()
    st = C.c_void_p()
    assert hip.L.hipStreamCreate(C.byref(st)) == 0
    try:
        with pa.Engine(0) as eng:
            for atoms, frames, seed, own_stream in ((3000, 600, 21, False), (100, 2000, 22, True)):
                nD = atoms - 3
                print(f"noalloc: {3 * nD + 3} dofs", flush=True)
                eng.set_trajectory(pa.synth_values(atoms, frames, seed))
                eng.prepare([20] * 6)
                eng.pairs(pa.METHOD_AUTO)
                s1, s2 = eng.s1d(), eng.s2d()
                e0, m0, l0, t0 = eng.mist()
                e1, m1, l1, t1, status = _noalloc(eng, hip, nD, s1, s2, st.value if own_stream else None)
                assert status == 0
                assert np.array_equal(e1, e0)
                assert np.array_equal(m1.view(np.uint64), m0.view(np.uint64))
                assert np.array_equal(l1, l0)
                assert np.array_equal(t1.view(np.uint64), t0.view(np.uint64))
    finally:
        hip.L.hipStreamDestroy(st)
        hip.close()


def test_mist_noalloc_rejects_too_many_dofs():
    import parent_amd as pa
    with pa.Engine(0) as eng:
        with pytest.raises(pa.ParentError) as ei:
            eng.mist_noalloc(21846, 8, 8, 8, 8, 8, 8, 8)  # 65,541 dofs; no pointer is touched
        assert ei.value.code == 7


def test_mist_noalloc_heavy_ties_memory_keys():
    """16,392 vertices with MI in {0, 1, 2} (heavy ties): the allocation-free
    call (above 16,384 vertices its keys live in the static array, MI gathered
    from S2D) against the engine's MIST (keys in registers, dense MI rows)."""
    import parent_amd as pa
    nD = 5463
    D = 3 * nD + 3
    rng = np.random.default_rng(7)
    s1 = np.zeros(D)
    s2 = -rng.integers(0, 3, D * (D - 1) // 2).astype(np.float64)
    hip = _Hip()
    try:
        with pa.Engine(0) as eng:
            eng.set_entropies(nD, s1, s2)
            e0, m0, l0, t0 = eng.mist()
            e1, m1, l1, t1, status = _noalloc(eng, hip, nD, s1, s2, None)
    finally:
        hip.close()
    assert status == 0
    assert np.array_equal(e1, e0) and np.array_equal(l1, l0)
    assert np.array_equal(m1.view(np.uint64), m0.view(np.uint64))
    assert np.array_equal(t1.view(np.uint64), t0.view(np.uint64))
