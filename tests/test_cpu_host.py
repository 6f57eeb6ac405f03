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


MIE_EXE = os.path.join(ROOT, "parent_amd", "bin", "get_PAR_MIE_hip.x")


def test_mie_listing_matches_reference_golden(tmp_path):
    """get_PAR_MIE_hip.x on the golden .par prints the bytes the reference's
    get_PAR_MIE.x printed for it (chain12_b50_MIE.txt)."""
    r = subprocess.run([MIE_EXE, os.path.join(GOLD, "chain12_b50.par")], capture_output=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout == open(os.path.join(GOLD, "chain12_b50_MIE.txt"), "rb").read()


def _planted_par(path, atoms, seed, specials):
    """A .par of an `atoms` chain whose S1D/S2D are planted values spanning
    magnitudes, signs, 12-digit rounding ties and (optionally) inf/nan."""
    bat = str(path) + ".bat"
    pa.synth_write_bat(bat, atoms, 1, 1, 1)
    nD = atoms - 3
    _, _, _, D = pa.dof_counts(nD)
    rng = np.random.default_rng(seed)
    P = D * (D - 1) // 2
    def draw(n):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)
        k = rng.integers(0, 8, n)
        v[k == 0] = np.round(v[k == 0], 3)            # short decimals
        v[k == 1] = rng.integers(-10**6, 10**6, (k == 1).sum()) * 1.0
        v[k == 2] = 0.5 + rng.integers(0, 10**11, (k == 2).sum()) * 1e-12 + 5e-13  # 12th-digit ties
        return v
    s1, s2 = draw(D), draw(P)
    if specials:
        s2[:6] = [np.inf, -np.inf, np.nan, -0.0, 5e-324, 1.7976931348623157e308]
        s1[0] = -np.nan
    with pa.Bat(bat) as b:
        b.write_par(str(path), [50] * 6, s1, s2)
    return nD, s1, s2


@pytest.mark.skipif(not os.path.exists(os.path.join(O.REF_DIR, "get_PAR_MIE.x")),
                    reason="reference get_PAR_MIE.x not built (needs /root/reference)")
@pytest.mark.parametrize("atoms,specials", [(40, True), (100, False)])
def test_mie_listing_matches_reference_binary(tmp_path, atoms, specials):
    """Byte comparison with the reference get_PAR_MIE.x (built from its
    sources) on planted values, including inf, nan, -0 and denormals."""
    par = tmp_path / "x.par"
    _planted_par(par, atoms, atoms, specials)
    ref = subprocess.run([os.path.join(O.REF_DIR, "get_PAR_MIE.x"), str(par)], capture_output=True)
    ours = subprocess.run([MIE_EXE, str(par)], capture_output=True, env={**os.environ, "OMP_NUM_THREADS": "3"})
    assert ref.returncode == 0 and ours.returncode == 0, (ref.stderr, ours.stderr)
    assert ours.stdout == ref.stdout


def test_mie_totals_are_sequential_listing_sums(tmp_path):
    """The 16 totals: each TOTAL line's value, and the category sums equal
    sequential FP64 sums over the listed values (order fixed by the listing,
    which the byte comparisons pin)."""
    par = tmp_path / "t.par"
    nD, s1, s2 = _planted_par(par, 30, 3, False)
    out = tmp_path / "t.txt"
    tot = pa.mie_write(str(out), nD, s1, s2, 2)
    lines = open(out).read().splitlines()
    totals = [float(x.split(" = ")[1]) for x in lines[-16:]]
    assert np.allclose(totals, tot, rtol=1e-11, atol=0)
    nB, nA, nd, D = pa.dof_counts(nD)
    assert len(lines) == D + D * (D - 1) // 2 + 16
    acc = 0.0
    for v in s1[:nB]:
        acc += v
    assert acc == tot[0]
    # bond-bond block: entry k of the listing is bb[P-1-k]
    bb = s2[:nB * (nB - 1) // 2][::-1]
    acc_e = acc_m = 0.0
    k = 0
    for i in range(nB - 1):
        for j in range(i + 1, nB):
            acc_e += bb[k]
            acc_m += (s1[i] + s1[j]) - bb[k]
            k += 1
    assert (acc_e, acc_m) == (tot[3], tot[4])
    assert tot[15] == (tot[0] + tot[1] + tot[2] - tot[4] - tot[6] - tot[8] - tot[10] - tot[12] - tot[14])


def test_mie_cli_errors_exit_1(tmp_path):
    r = subprocess.run([MIE_EXE, str(tmp_path / "missing.par")], capture_output=True, text=True)
    assert r.returncode == 1 and "COULD NOT OPEN FILE" in r.stderr
    bad = tmp_path / "bad.par"
    bad.write_bytes(struct.pack("<10i", 4, 2, 3, 10, 50, 50, 50, 50, 50, 50))
    r = subprocess.run([MIE_EXE, str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "ERROR READING HEADER OF FILE" in r.stderr


def test_bat_raw_records_are_the_file_bytes(tmp_path):
    """parent_bat_read_raw hands back the frame records verbatim: the file's
    last N * record_bytes bytes, for any frame window, FP32 and FP64."""
    p64 = str(tmp_path / "d.bat")
    write_bat_fp64(p64, pa.synth_values(10, 70, 2).astype(np.float64), 7)
    for path, rec in [(BAT, 80 + 12 * 9), (p64, 116 + 24 * 7)]:
        data = open(path, "rb").read()
        with pa.Bat(path) as b:
            assert b.record_bytes == rec
            frames = np.frombuffer(data[len(data) - b.n_frames * rec:], dtype=np.uint8).reshape(b.n_frames, rec)
            assert np.array_equal(b.read_raw(0, b.n_frames), frames)
            assert np.array_equal(b.read_raw(13, 29), frames[13:42])
            assert b.read_raw(b.n_frames, 0).shape == (0, rec)
            with pytest.raises(pa.ParentError):
                b.read_raw(b.n_frames - 2, 3)
    short = tmp_path / "short.bat"
    short.write_bytes(open(BAT, "rb").read()[:-50])
    with pa.Bat(str(short)) as b:
        b.read_raw(0, b.n_frames - 1)
        with pytest.raises(pa.ParentError) as ei:
            b.read_raw(0, b.n_frames)
        assert ei.value.code == 3
