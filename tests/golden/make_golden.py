"""Writes the golden fixtures of this directory (run in the build container only).

Inputs are small synthetic chains from parent_amd.synth_write_bat; expected
outputs come from the reference programs compiled from the reference's own
sources by `make -C oracle ref` (oracle/_ref/):

  chain12_1000.bat            12 atoms (30 dofs) x 1000 frames, FP32, seed 3
  chain12_b50.par             PARENT.x chain12_1000.bat ... 50 50 50 50 50 50
  chain12_b50_MIE.txt         get_PAR_MIE.x chain12_b50.par
  chain12_b50_MIST.txt        get_PAR_MIST.x chain12_b50.par
  chain12_mixed.par           PARENT.x ... 7 13 29 5 40 17
  chain12_mixed_MIST.txt      get_PAR_MIST.x chain12_mixed.par
  chain12_wide.par            PARENT.x ... 210 256 300 500 256 210 (bin counts past uint8 indices
                              and past a 2D histogram in LDS)
  chain12_wide_MIST.txt       get_PAR_MIST.x chain12_wide.par
  ubq_sample_MIE_tail.txt     test_system/sample_output/sample_output_MIE.txt (reference data)
  ubq_sample_MIST_tail.txt    test_system/sample_output/sample_output_MIST.txt (reference data)
  ubq/UBQ_UBM2.top, .xtc      test_system/UBQ_UBM2.top / .xtc (reference data: the inputs of its
                              end-to-end test)
  ubq/bat_sha256.txt          SHA-256 of BAT_builder.x -t UBQ_UBM2.top -x UBQ_UBM2.xtc -o UBQ.bat
                              -bb "CA C N H1 O1" (FP64), of the same with --single_precision, and
                              with -bb "CA N"

Usage: python tests/golden/make_golden.py [--ubq-only]
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
REFERENCE = "/root/reference"
sys.path.insert(0, ROOT)


def run(args, mpi=True, **kw):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    if mpi:  # MPICH's libmpi.so.12 lives in /opt/conda/lib
        env["LD_LIBRARY_PATH"] = "/opt/conda/lib"
    r = subprocess.run(args, capture_output=True, text=True, env=env, **kw)
    if r.returncode != 0:
        raise SystemExit(f"{args[0]} failed: {r.stderr[-500:]}")
    return r.stdout


def ubq():
    """UBQ inputs and the hashes of the reference builder's .bat outputs."""
    d = os.path.join(HERE, "ubq")
    os.makedirs(d, exist_ok=True)
    for ext in ("top", "xtc"):
        shutil.copyfile(os.path.join(REFERENCE, "test_system", f"UBQ_UBM2.{ext}"), os.path.join(d, f"UBQ_UBM2.{ext}"))
    lines = []
    with tempfile.TemporaryDirectory() as td:
        for tag, extra, bb in (("fp64", [], "CA C N H1 O1"), ("fp32", ["--single_precision"], "CA C N H1 O1"),
                               ("fp64_bb_CA_N", [], "CA N")):
            bat = os.path.join(td, f"{tag}.bat")
            run([os.path.join(REF, "BAT_builder.x"), "-t", os.path.join(d, "UBQ_UBM2.top"), "-x",
                 os.path.join(d, "UBQ_UBM2.xtc"), "-o", bat, "-bb", bb] + extra, mpi=False, cwd=td)
            lines.append(f"{tag} {hashlib.sha256(open(bat, 'rb').read()).hexdigest()} {os.path.getsize(bat)}")
    with open(os.path.join(d, "bat_sha256.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    if "--ubq-only" in sys.argv:
        ubq()
        return
    import parent_amd

    bat = os.path.join(HERE, "chain12_1000.bat")
    parent_amd.synth_write_bat(bat, 12, 1000, 3)
    for tag, bins in (("b50", [50] * 6), ("mixed", [7, 13, 29, 5, 40, 17]), ("wide", [210, 256, 300, 500, 256, 210])):
        par = os.path.join(HERE, f"chain12_{tag}.par")
        run([os.path.join(REF, "PARENT.x"), bat, par] + [str(b) for b in bins], cwd=HERE)
        with tempfile.TemporaryDirectory() as td:
            txt = os.path.join(td, "mist.txt")
            run([os.path.join(REF, "get_PAR_MIST.x"), par, txt], cwd=td)
            shutil.copy(txt, os.path.join(HERE, f"chain12_{tag}_MIST.txt"))
        if tag == "b50":
            with open(os.path.join(HERE, f"chain12_{tag}_MIE.txt"), "w") as f:
                f.write(run([os.path.join(REF, "get_PAR_MIE.x"), par], mpi=False))
    for name in ("MIE", "MIST"):
        shutil.copy(os.path.join(REFERENCE, "test_system", "sample_output", f"sample_output_{name}.txt"),
                    os.path.join(HERE, f"ubq_sample_{name}_tail.txt"))
    ubq()


if __name__ == "__main__":
    main()
