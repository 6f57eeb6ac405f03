"""The reference's end-to-end known answer on the GPU: the UBQ_UBM2 test system
(tests/golden/ubq: 1,094 atoms, 1,091 dihedrals, 801 frames) at 50 bins, through
the drop-in programs the reference's run.sh calls:

  BAT_builder_hip.x  .top + .xtc -> .bat (bytes pinned to the reference builder's)
  PARENT_hip.x       .bat -> .par on the GPU (streamed read-in, autotuned pair kernel)
  get_PAR_MIST_hip.x .par -> MIST text (device Prim)
  parent_mie_write   the get_PAR_MIE.x totals

Every TOTAL line of the reference's sample outputs (sample_output_MIE.txt,
sample_output_MIST.txt; README.md:108-123) must agree within 1e-10 relative.
At 12 printed digits the lines are compared as text, and the count that match
exactly is printed.  The CPU oracle matches all 26 of them exactly
(test_cpu_oracle.py).
"""
import os
import subprocess

import pytest

import parent_amd as pa

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
UBQ = os.path.join(GOLD, "ubq")
BIN = os.path.join(ROOT, "parent_amd", "bin")
ND = 1091


def _check(tail_file, values):
    lines = open(os.path.join(GOLD, tail_file)).read().splitlines()
    assert len(lines) == len(values)
    exact = 0
    for ln, v in zip(lines, values):
        name, ref = ln.split(" = ")
        got = "%.12g" % v
        exact += got == ref
        assert abs(v - float(ref)) <= 1e-10 * abs(float(ref)) + 1e-9, (name, got, ref)
    return exact, len(lines)


def test_ubq_sample_totals_through_the_drop_in_programs(tmp_path):
    bat, par, mist = str(tmp_path / "UBQ.bat"), str(tmp_path / "UBQ.par"), str(tmp_path / "UBQ_MIST.txt")
    print("ubq: BAT builder", flush=True)
    r = subprocess.run([os.path.join(BIN, "BAT_builder_hip.x"), "-t", os.path.join(UBQ, "UBQ_UBM2.top"), "-x",
                        os.path.join(UBQ, "UBQ_UBM2.xtc"), "-o", bat, "-bb", "CA C N H1 O1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print("ubq: PARENT_hip.x 50 bins", flush=True)
    r = subprocess.run([os.path.join(BIN, "PARENT_hip.x"), bat, par] + ["50"] * 6, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    assert "PROGRAM FINISHED SUCCESSFULLY." in r.stdout
    nD, N, bins, s1, s2 = pa.read_par(par)
    assert (nD, N, list(bins)) == (ND, 801, [50] * 6)
    mie_exact, n_mie = _check("ubq_sample_MIE_tail.txt", pa.mie_write(None, nD, s1, s2))
    print("ubq: get_PAR_MIST_hip.x", flush=True)
    r = subprocess.run([os.path.join(BIN, "get_PAR_MIST_hip.x"), par, mist], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    text = open(mist).read().splitlines()
    assert len(text) == (3 * ND + 3 - 1) + 10
    tot = [float(ln.split(" = ")[1]) for ln in text[-10:]]
    mist_exact, n_mist = _check("ubq_sample_MIST_tail.txt", tot)
    print(f"ubq: {mie_exact}/{n_mie} MIE and {mist_exact}/{n_mist} MIST total lines identical at 12 digits",
          flush=True)
