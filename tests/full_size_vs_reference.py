"""Full-size check against the reference program itself (not collected by
pytest: it runs minutes of CPU work).

Writes the default bench workload (1,500-atom chain, 4,494 dofs x 100,000
frames, FP32, seed 2) as a .bat file, runs PARENT_hip.x on the GPU and the
reference PARENT.x built -O3 from its sources (oracle/_ref/PARENT_O3.x, 16
OpenMP threads) on the CPU at 50 bins, and compares the two .par files: every
S1D and S2D value and the 16 MIE totals.

    python3 -u tests/full_size_vs_reference.py [workdir]

The workdir (default /tmp) receives the 1.8 GB .bat and both .par files.
Exits non-zero when a value differs by more than the build's tolerances
(|dS| <= 1e-10, MIE totals within 1e-10 relative).
"""
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import parent_amd as pa  # noqa: E402


def main():
    wd = sys.argv[1] if len(sys.argv) > 1 else "/tmp"
    bat, pg, pc = (os.path.join(wd, n) for n in ("c2.bat", "c2_gpu.par", "c2_cpu.par"))
    t = time.time()
    pa.synth_write_bat(bat, 1500, 100000, 2)
    print(f"bat written in {time.time() - t:.1f} s, {os.path.getsize(bat) / 1e9:.2f} GB", flush=True)
    runs = (("PARENT_hip.x", os.path.join(ROOT, "parent_amd", "bin", "PARENT_hip.x"), pg, None),
            ("PARENT_O3.x", os.path.join(ROOT, "oracle", "_ref", "PARENT_O3.x"), pc,
             dict(os.environ, OMP_NUM_THREADS="16",
                  LD_LIBRARY_PATH="/opt/conda/lib:" + os.environ.get("LD_LIBRARY_PATH", ""))))
    for name, exe, par, env in runs:
        t = time.time()
        r = subprocess.run([exe, bat, par] + ["50"] * 6, capture_output=True, text=True, timeout=1800, env=env)
        calc = [ln for ln in r.stdout.splitlines() if ln.startswith(("Calculation time", "Total execution"))]
        print(f"{name}: rc {r.returncode}, wall {time.time() - t:.1f} s; " + "; ".join(calc), flush=True)
        if r.returncode != 0:
            print(r.stdout[-2000:], r.stderr[-2000:])
            return 1
    g, c = pa.read_par(pg), pa.read_par(pc)
    assert g[0] == c[0] and g[1] == c[1] and list(g[2]) == list(c[2]), (g[:3], c[:3])
    d1, d2 = np.abs(g[3] - c[3]), np.abs(g[4] - c[4])
    print(f"S1D: max |d| {d1.max():.3e}, bit-identical {int((g[3] == c[3]).sum())} of {g[3].size}")
    print(f"S2D: max |d| {d2.max():.3e}, bit-identical {int((g[4] == c[4]).sum())} of {g[4].size}")
    tg, tc = pa.mie_write(None, g[0], g[3], g[4]), pa.mie_write(None, c[0], c[3], c[4])
    rel = np.max(np.abs(tg - tc) / np.maximum(np.abs(tc), 1e-300))
    print(f"MIE totals: max relative |d| {rel:.3e}; total entropy GPU {tg[-1]:.12g}, CPU {tc[-1]:.12g}")
    ok = d1.max() <= 1e-10 and d2.max() <= 1e-10 and rel <= 1e-10
    print("PASS" if ok else "FAIL", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
