"""Streamed .bat read-in on the GPU (parent_engine_load_bat): raw frame
records go through pinned buffers to HBM and are reordered there.  The result
must equal the host reader + parent_engine_set_trajectory path bit for bit:
same ranges, S1D and S2D, over several ~32 MB chunks, FP32 and FP64."""
import os
import struct

import numpy as np
import pytest

import parent_amd as pa

pytestmark = pytest.mark.gpu


def write_bat_fp64(path, values, nD):
    """Double-precision .bat v3 (io_binary.cpp:43-102, 203-278), vectorised."""
    N, D = values.shape
    nB, nA = nD + 2, nD + 1
    order = [0, 1, nB]
    for i in range(nD):
        order += [i + 2, nB + i + 1, nB + nA + i]
    vals = np.zeros((N, 6 + len(order)), dtype="<f8")
    vals[:, 6:] = values[:, order]
    rec = np.zeros((N, 44 + vals.shape[1] * 8), dtype=np.uint8)
    rec[:, 44:] = vals.view(np.uint8).reshape(N, -1)
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", 3, 1, nD, N))
        for _ in range(nD + 3):
            f.write(b"RES\0\0\0\0\0" + struct.pack("<i", 1) + b"A\0\0\0\0\0\0\0" + b"mol".ljust(31, b"\0"))
        for i in range(nD):
            f.write(struct.pack("<6i", i + 1, i + 2, i + 3, i + 4, 0, -1))
        f.write(struct.pack(f"<{nD + 3}f", *([1.0] * (nD + 3))))
        f.write(rec.tobytes())


def _results(eng, bins):
    eng.prepare(bins)
    eng.pairs(pa.METHOD_ATOMIC)
    return eng.ranges(), eng.s1d(), eng.s2d()


@pytest.mark.parametrize("atoms,frames,dp", [(200, 40_000, 0), (120, 40_000, 1), (30, 777, 0)])
def test_streamed_load_equals_host_upload(tmp_path, atoms, frames, dp):
    path = str(tmp_path / "s.bat")
    if dp:
        v = pa.synth_values(atoms, frames, 8).astype(np.float64)
        v += np.linspace(0, 1e-7, v.shape[1])  # values no FP32 file could hold
        write_bat_fp64(path, v, atoms - 3)
    else:
        pa.synth_write_bat(path, atoms, frames, 8)
    bins = [50, 40, 30, 50, 40, 30]
    with pa.Bat(path) as b:
        nchunks = -(-frames // max(1, (32 << 20) // b.record_bytes))
        print(f"stream: {atoms} atoms x {frames} frames dp={dp}, {nchunks} chunks", flush=True)
        host = b.read_frames()
        with pa.Engine(0) as e1, pa.Engine(0) as e2:
            e1.load_bat(b)
            assert e1.stats()["ms_upload"] > 0
            r1 = _results(e1, bins)
            e2.set_trajectory(host, atoms - 3)
            r2 = _results(e2, bins)
    for a, c in zip(r1[0], r2[0]):
        assert np.array_equal(a, c)
    assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])


def test_streamed_load_of_truncated_file_fails_cleanly(tmp_path):
    """A file cut inside the frames: the load stops at the failing read with
    PARENT_E_IO after the earlier chunks' kernels have drained; the engine
    then holds no trajectory and still works for the next load."""
    path = str(tmp_path / "t.bat")
    pa.synth_write_bat(path, 200, 30_000, 3)
    short = str(tmp_path / "short.bat")
    with open(path, "rb") as f, open(short, "wb") as g:
        data = f.read()
        g.write(data[: len(data) - 1000])
    with pa.Engine(0) as e:
        with pa.Bat(short) as b:
            with pytest.raises(pa.ParentError) as ei:
                e.load_bat(b)
            assert ei.value.code == 3
        with pytest.raises(pa.ParentError):
            e.prepare([50] * 6)
        with pa.Bat(path) as b:
            e.load_bat(b)
        e.prepare([50] * 6)
        e.pairs(pa.METHOD_ATOMIC)
        assert np.all(np.isfinite(e.s2d()))
