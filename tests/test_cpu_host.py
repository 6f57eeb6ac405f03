"""Host side of the engine without a GPU: C-ABI exports, .bat/.par codecs,
synthetic data, error paths of the drop-in executables."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle_api as O
import parent_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BAT = os.path.join(GOLD, "chain12_1000.bat")
HEADER = os.path.join(ROOT, "include", "parent_hip.h")


def test_library_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(parent_\w+)\s*\(", open(HEADER).read()))
    assert len(decl) >= 20
    L = pa.lib()
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", pa.library_path()], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (parent_\w+)", out))
    assert decl <= exported


def test_engine_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(pa.ParentError) as ei:
        pa.Engine(0)
    assert ei.value.code == 6


def test_synth_is_deterministic_and_in_range():
    a = pa.synth_values(30, 800, 9, 1)
    b = pa.synth_values(30, 800, 9, 4)
    assert np.array_equal(a, b)
    c = pa.synth_values(30, 800, 10, 4)
    assert not np.array_equal(a, c)
    nB, nA, nD, D = pa.dof_counts(27)
    assert a.shape == (800, D)
    assert np.all(a[:, :nB] > 0)
    assert np.all(a[:, nB:nB + nA] > 0) and np.all(a[:, nB:nB + nA] < np.float32(np.pi))
    dih = a[:, nB + nA:]
    assert np.all(dih >= 0) and np.all(dih <= np.float32(6.283185005187988))
    # 20% of dofs carry a latent factor: some pairs are clearly correlated
    cc = np.corrcoef(a[:, :nB + nA].T)
    assert np.max(np.abs(cc - np.eye(cc.shape[0]))) > 0.1


def test_bat_reader_matches_oracle_reader(tmp_path):
    with pa.Bat(BAT) as b:
        v = b.read_frames()
        assert (b.n_dihedrals, b.n_frames, b.double_prec) == (9, 1000, 0)
    nD, N, dp, vo = O.bat_read(BAT)
    assert np.array_equal(v.T.astype(np.float64), vo)
    # the synthetic writer and the value generator agree
    assert np.array_equal(v, pa.synth_values(12, 1000, 3))


def write_bat_fp64(path, values, nD):
    """A double-precision .bat v3 (io_binary.cpp:43-102, 203-278) for tests."""
    N, D = values.shape
    nB, nA = nD + 2, nD + 1
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", 3, 1, nD, N))
        for i in range(nD + 3):
            f.write(b"RES\0\0\0\0\0" + struct.pack("<i", 1) + b"A\0\0\0\0\0\0\0" + b"mol".ljust(31, b"\0"))
        for i in range(nD):
            f.write(struct.pack("<6i", i + 1, i + 2, i + 3, i + 4, 0, -1))
        f.write(struct.pack(f"<{nD + 3}f", *([1.0] * (nD + 3))))
        for fr in range(N):
            row = values[fr]
            rec = [row[0], row[1], row[nB]]
            for i in range(nD):
                rec += [row[i + 2], row[nB + i + 1], row[nB + nA + i]]
            f.write(struct.pack("<11f", *([0.0] * 11)))
            f.write(np.array([0.0] * 6 + rec, dtype="<f8").tobytes())


def test_fp64_bat_reader(tmp_path):
    v = pa.synth_values(10, 300, 2).astype(np.float64) + 1e-9
    p = str(tmp_path / "d.bat")
    write_bat_fp64(p, v, 7)
    with pa.Bat(p) as b:
        got = b.read_frames()
        assert b.double_prec == 1 and got.dtype == np.float64
    assert np.array_equal(got, v)
    _, _, dp, vo = O.bat_read(p)
    assert dp == 1 and np.array_equal(vo, v.T)


def test_par_writer_header_matches_reference(tmp_path):
    """.par written by the engine's writer from the .bat header: same header bytes
    as the reference PARENT.x output, and the body round-trips."""
    ref = open(os.path.join(GOLD, "chain12_mixed.par"), "rb").read()
    nD, N, bins, s1, s2 = pa.read_par(os.path.join(GOLD, "chain12_mixed.par"))
    assert (nD, N, list(bins)) == (9, 1000, [7, 13, 29, 5, 40, 17])
    out = str(tmp_path / "w.par")
    with pa.Bat(BAT) as b:
        b.write_par(out, bins, s1, s2)
    assert open(out, "rb").read() == ref


def test_par_reader_rejects_corrupt_headers(tmp_path):
    good = open(os.path.join(GOLD, "chain12_b50.par"), "rb").read()

    def bad(offset, value, code):
        p = str(tmp_path / "x.par")
        raw = bytearray(good)
        raw[offset:offset + 4] = struct.pack("<i", value)
        open(p, "wb").write(bytes(raw))
        with pytest.raises(pa.ParentError) as ei:
            pa.read_par(p)
        assert ei.value.code == code
    bad(0, -1, 1)            # version < 0
    bad(4, 2, 1)             # double_prec not 0/1
    bad(12, 0, 1)            # numFrames < 1
    bad(16 + 4 * 3, 0, 1)    # 2D bonds bins < 1
    dih0 = 40 + 12 * 51      # first dihedral record
    bad(dih0, 0, 1)          # dihedral atom id < 1
    bad(dih0 + 16, 2, 1)     # dihedral type not in {-1, 0, 1}
    bad(dih0 + 20, -2, 1)    # parent < -1
    p = str(tmp_path / "t.par")
    open(p, "wb").write(good[:-8])
    with pytest.raises(pa.ParentError) as ei:
        pa.read_par(p)
    assert ei.value.code == 3


def test_bat_reader_rejects_corrupt_and_truncated(tmp_path):
    good = open(BAT, "rb").read()
    for off, val in ((0, -3), (4, 5), (8, -1), (12, 0)):
        raw = bytearray(good)
        raw[off:off + 4] = struct.pack("<i", val)
        p = str(tmp_path / "c.bat")
        open(p, "wb").write(bytes(raw))
        with pytest.raises(pa.ParentError) as ei:
            pa.Bat(p)
        assert ei.value.code == 1
    p = str(tmp_path / "t.bat")
    open(p, "wb").write(good[:-100])
    with pa.Bat(p) as b:
        with pytest.raises(pa.ParentError) as ei:
            b.read_frames()
        assert ei.value.code == 3
    with pytest.raises(pa.ParentError):
        pa.Bat(str(tmp_path / "missing.bat"))


def test_drop_in_cli_errors_exit_1(tmp_path):
    exe = os.path.join(ROOT, "parent_amd", "bin", "PARENT_hip.x")
    cases = [
        [],                                            # wrong argument count
        [BAT, str(tmp_path / "o.par"), "50"],          # wrong argument count
        [BAT, str(tmp_path / "o.txt")] + ["50"] * 6,   # output must be .par
        [BAT.replace(".bat", ".xyz"), str(tmp_path / "o.par")] + ["50"] * 6,
        [BAT, str(tmp_path / "o.par"), "x"] + ["50"] * 5,  # bins not an int
        [str(tmp_path / "missing.bat"), str(tmp_path / "o.par")] + ["50"] * 6,
    ]
    for args in cases:
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (args, r.stdout, r.stderr)
    mist = os.path.join(ROOT, "parent_amd", "bin", "get_PAR_MIST_hip.x")
    assert subprocess.run([mist], capture_output=True, timeout=60).returncode == 1
    assert subprocess.run([mist, str(tmp_path / "none.par"), "x.txt"], capture_output=True, timeout=60).returncode == 1


def test_get_mutual_and_par_pos_match_oracle():
    nD, _, _, s1, s2 = pa.read_par(os.path.join(GOLD, "chain12_b50.par"))
    D = 3 * nD + 3
    seen = set()
    for i in range(D):
        for j in range(i + 1, D):
            p = pa.par_pos(nD, i, j)
            assert p == O.lib().oracle_par_pos(nD, i, j)
            seen.add(p)
            assert pa.get_mutual(nD, s1, s2, i, j) == O.lib().oracle_mutual(nD, O._p(s1), O._p(s2), i, j)
    assert seen == set(range(D * (D - 1) // 2))
