"""BAT builder (.top + .xtc -> .bat) on the CPU.

The reference's own test system (tests/golden/ubq: UBQ_UBM2.top and .xtc, 1,094
atoms x 801 frames) goes through parent_bat_build; the output must have the
SHA-256 of the reference BAT_builder.x's output for the same command line
(tests/golden/ubq/bat_sha256.txt, written by make_golden.py): FP64, FP32 and a
second backbone set.  Also: thread count independence, the drop-in command
line, and the .xtc error paths (a truncated last frame, a bad magic number, an
atom count that does not fit the topology)."""
import hashlib
import os
import struct
import subprocess

import pytest

import parent_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UBQ = os.path.join(ROOT, "tests", "golden", "ubq")
TOP = os.path.join(UBQ, "UBQ_UBM2.top")
XTC = os.path.join(UBQ, "UBQ_UBM2.xtc")
EXE = os.path.join(ROOT, "parent_amd", "bin", "BAT_builder_hip.x")
HASH = {ln.split()[0]: (ln.split()[1], int(ln.split()[2]))
        for ln in open(os.path.join(UBQ, "bat_sha256.txt")).read().splitlines()}
NATOMS, ND, NFRAMES = 1094, 1091, 801


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def header_bytes(nd):
    natoms = nd + 3
    return 16 + natoms * (8 + 4 + 8 + 31) + 24 * nd + 4 * natoms


@pytest.mark.parametrize("tag,dp,bb", [("fp64", 1, "CA C N H1 O1"), ("fp32", 0, "CA C N H1 O1"),
                                       ("fp64_bb_CA_N", 1, "CA N")])
def test_bat_bytes_match_reference_builder(tmp_path, tag, dp, bb):
    out = str(tmp_path / "u.bat")
    assert pa.bat_build(TOP, XTC, out, bb, dp, 4) == NFRAMES
    want, size = HASH[tag]
    assert os.path.getsize(out) == size
    assert sha(out) == want


def test_thread_count_does_not_change_the_bytes(tmp_path):
    a, b = str(tmp_path / "a.bat"), str(tmp_path / "b.bat")
    pa.bat_build(TOP, XTC, a, nthreads=1)
    pa.bat_build(TOP, XTC, b, nthreads=7)
    assert sha(a) == sha(b) == HASH["fp64"][0]


def test_bat_header_and_reader_round_trip(tmp_path):
    out = str(tmp_path / "u.bat")
    pa.bat_build(TOP, XTC, out)
    with pa.Bat(out) as b:
        assert (b.n_dihedrals, b.n_frames, b.double_prec) == (ND, NFRAMES, 1)
        v = b.read_frames()
    # bonds > 0, angles in (0, pi], dihedrals and phase angles in [0, 2 pi)
    nB, nA = ND + 2, ND + 1
    assert v.shape == (NFRAMES, 3 * ND + 3)
    assert (v[:, :nB] > 0).all()
    assert (v[:, nB:nB + nA] > 0).all() and (v[:, nB:nB + nA] <= 3.141592653589794).all()
    assert (v[:, nB + nA:] >= 0).all() and (v[:, nB + nA:] < 2 * 3.141592653589794).all()


def test_drop_in_command_line(tmp_path):
    out = str(tmp_path / "cli.bat")
    r = subprocess.run([EXE, "-t", TOP, "-x", XTC, "-o", out, "-bb", "CA C N H1 O1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert sha(out) == HASH["fp64"][0]
    r = subprocess.run([EXE, "-t", TOP, "-x", XTC, "-o", out, "-bb", "CA C N H1 O1", "--single_precision"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and sha(out) == HASH["fp32"][0]
    bad = [[EXE], [EXE, "-t", TOP, "-x", XTC, "-o", str(tmp_path / "x.par"), "-bb", "CA"],
           [EXE, "-b", out, "-o", str(tmp_path / "back.xtc")],
           [EXE, "-t", TOP, "-x", XTC, "-o", out, "-bb", "CA", "--double"]]
    for args in bad:
        assert subprocess.run(args, capture_output=True, timeout=60).returncode == 1, args


def test_truncated_xtc_keeps_the_complete_frames(tmp_path):
    """A file cut inside its last frame: the complete frames are written and
    the header counts them, as the reference's read loop ends at the first
    frame it cannot read (BAT_trajectory.cpp:112-127)."""
    full = str(tmp_path / "full.bat")
    pa.bat_build(TOP, XTC, full)
    cut = tmp_path / "cut.xtc"
    cut.write_bytes(open(XTC, "rb").read()[:-1000])
    out = str(tmp_path / "cut.bat")
    assert pa.bat_build(TOP, str(cut), out) == NFRAMES - 1
    a, b = open(full, "rb").read(), open(out, "rb").read()
    h = header_bytes(ND)
    assert struct.unpack("<i", b[12:16])[0] == NFRAMES - 1
    assert a[:12] == b[:12] and a[16:h] == b[16:h]
    rec = 44 + (9 + 3 * ND) * 8
    assert len(b) == h + (NFRAMES - 1) * rec
    assert a[h:h + (NFRAMES - 1) * rec] == b[h:]


def test_corrupt_xtc_is_an_error(tmp_path):
    data = bytearray(open(XTC, "rb").read())
    bad_magic = tmp_path / "magic.xtc"
    bad_magic.write_bytes(b"\0\0\0\0" + bytes(data[4:]))
    with pytest.raises(pa.ParentError) as ei:
        pa.bat_build(TOP, str(bad_magic), str(tmp_path / "m.bat"))
    assert ei.value.code == 1
    wrong_atoms = tmp_path / "natoms.xtc"
    d2 = bytearray(data)
    d2[4:8] = struct.pack(">i", NATOMS - 1)
    wrong_atoms.write_bytes(bytes(d2))
    with pytest.raises(pa.ParentError) as ei:
        pa.bat_build(TOP, str(wrong_atoms), str(tmp_path / "n.bat"))
    assert ei.value.code == 1 and "DOES NOT FIT THE TOPOLOGY" in str(ei.value)
    with pytest.raises(pa.ParentError) as ei:
        pa.bat_build(TOP, str(tmp_path / "missing.xtc"), str(tmp_path / "x.bat"))
    assert ei.value.code == 3
