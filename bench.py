"""Benchmark of the all-pairs entropy / MI hot path on one MI355X.

A step is one pass of the hot path over the resident trajectory: per-dof
ranges + modfit + pre-binning + 1D entropies, then all D(D-1)/2 pair
histograms with their entropy epilogue (the reference's "Calculation time"
span, PARENT.cpp:498-1489).  The trajectory (synthetic, float32, frame-major)
is uploaded to HBM before the timed region.

Workloads (--workload):
  protein_1500x100k_b50   default: 1,500-atom chain, 4,494 dofs (10,095,771 pairs)
                          x 100,000 frames, 50 bins for every type
  protein_1500x100k_bB    same at B bins (listed: 5, 10, 20, 30, 75, 100; any B from 1 to 4,096)
  chain_100x10k_b50       100-atom chain, 294 dofs x 10,000 frames (the small case)
  large_5000x200k_b30     5,000-atom chain, 14,994 dofs (112,402,521 pairs) x 200,000 frames, 30 bins
  sweep_1500x100k         the default trajectory at B in {5,10,20,30,40,50,75,100}: one step = one
                          launch of the captured graph holding every set's prepare and pair kernels
  batch64_b30             64 chains, atoms log-uniform in [50, 2000], frames log-uniform in
                          [5k, 200k] (seed 4), 30 bins: one step = the pre-binning of every
                          trajectory (timed on the device) + one pair launch over the merged queue
  readin_1500x100k        read-in of the default trajectory as a 1.8 GB FP32 .bat file (page-cached):
                          one step = parent_engine_load_bat (file -> pinned buffers -> HBM ->
                          frame-major values); reported in bytes/s beside the host path
                          (parent_bat_read_frames + parent_engine_set_trajectory)

Prints one JSON line.  Extra keys: "plain" (the one-pair-per-workgroup
kernel on the same workload), "phases", "mist", "roofline" and the CPU
baseline.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (atoms, frames, bins)
    "protein_1500x100k_b50": (1500, 100_000, 50),
    "protein_1500x100k_b5": (1500, 100_000, 5),
    "protein_1500x100k_b10": (1500, 100_000, 10),
    "protein_1500x100k_b20": (1500, 100_000, 20),
    "protein_1500x100k_b30": (1500, 100_000, 30),
    "protein_1500x100k_b75": (1500, 100_000, 75),
    "protein_1500x100k_b100": (1500, 100_000, 100),
    "chain_100x10k_b50": (100, 10_000, 50),
    "large_5000x200k_b30": (5000, 200_000, 30),
}
SWEEP_BINS = [5, 10, 20, 30, 40, 50, 75, 100]
BATCH_N, BATCH_BINS, BATCH_SEED = 64, 30, 4
INT8_PEAK_OPS = 5.0e15  # MI355X dense int8 MFMA
FP4_PEAK_OPS = 10.1e15  # MI355X dense FP4 MFMA
HBM_PEAK = 8.0e12
LDS_PEAK_FALLBACK = {"conflict_free": 9.5e12, "random_2500": 5.2e12}  # lds_peak.x on MI355X
# HBM bytes per pass of the pair kernel, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled
# by the elementwise-kernel calibration (FETCH x 2.00, WRITE x 1.00):
# profiles/pairs_multi_protein_1500x100k_b50.md, tiles in bands of 32 row dofs
# (k_pairs_multi<1024, 1275, 2, 1>, variant 53: 32.84 GB read + 0.80 GB written;
# k_pairs_sliced<1, 1, 1024, 1275, 2, true>, variant 45: 32.09 GB + 0.80 GB; the autotuner picks
# one of the two at 50 bins).
# Keyed by (workload, method, variant): the figure is used only when the step ran that same kernel.
PROFILED_TRAFFIC = {("protein_1500x100k_b50", 1, 53): (32.84e9 + 0.80e9,
                                                      "profiles/pairs_multi_protein_1500x100k_b50.md"),
                    ("protein_1500x100k_b50", 1, 45): (32.09e9 + 0.80e9,
                                                      "profiles/pairs_multi_protein_1500x100k_b50.md")}
METRIC = "pair-frame updates/s, all-pairs MI (4,491 dofs,100k fr,50 bins): tuned vs plain"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lds_peak():
    """LDS atomic rate of this GPU (parent_amd/bin/lds_peak.x), atomics/s."""
    exe = os.path.join(ROOT, "parent_amd", "bin", "lds_peak.x")
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        d = json.loads(r.stdout.strip().splitlines()[-1])
        return {"conflict_free": d["conflict_free"]["atomics_per_s"],
                "random_2500": d["random_2500"]["atomics_per_s"], "source": "lds_peak.x (this run)"}
    except Exception as ex:  # noqa: BLE001
        log(f"lds_peak.x unavailable ({ex}); using the recorded MI355X figures")
        return dict(LDS_PEAK_FALLBACK, source="recorded lds_peak.x figures")


def cpu_host():
    """The measuring host: lscpu model, sockets, cores, the affinity mask and
    the OpenMP thread count the CPU legs use (the process's CPU share: every
    CPU of its affinity mask, capped by OMP_NUM_THREADS when the host sets
    it for this process)."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.lower().replace("(s)", "s").replace(" ", "_")] = v
    except Exception as ex:  # noqa: BLE001
        info["lscpu_error"] = str(ex)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = info["affinity_cpus"]
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
        info["omp_num_threads_env"] = int(env)
    info["threads_used"] = threads
    return info


def _native_oracle(td):
    """The oracle restatement built -O3 -march=native for this host (the
    image has g++), or None."""
    exe = os.path.join(td, "oracle_parent_native.x")
    src = [os.path.join(ROOT, "oracle", "oracle_main.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp")]
    try:
        r = subprocess.run(["g++", "-O3", "-march=native", "-fopenmp", "-std=c++17", "-o", exe] + src,
                           capture_output=True, text=True, timeout=300)
        return exe if r.returncode == 0 else None
    except Exception:  # noqa: BLE001
        return None


def _time_cpu(exe, atoms_sample, frames, bins, nthreads, td):
    """`exe` (a PARENT.x-compatible program) on a synthetic chain of
    `atoms_sample` atoms: its 'Calculation time' line (PARENT.cpp:498-1489)
    and the pair-frame updates it covers, or None."""
    import parent_amd
    bat = os.path.join(td, f"sample_{atoms_sample}x{frames}.bat")
    if not os.path.exists(bat):
        parent_amd.synth_write_bat(bat, atoms_sample, frames, 12345)
    env = dict(os.environ, OMP_NUM_THREADS=str(nthreads))
    try:
        r = subprocess.run([exe, bat, os.path.join(td, "o.par")] + [str(bins)] * 6,
                           capture_output=True, text=True, timeout=600, env=env)
    except subprocess.TimeoutExpired:
        return None
    calc = [ln for ln in r.stdout.splitlines() if ln.startswith("Calculation time:")]
    if r.returncode != 0 or not calc:
        log(f"cpu run of {os.path.basename(exe)} failed: rc={r.returncode} {r.stderr[-300:]}")
        return None
    D = 3 * atoms_sample - 6
    return float(calc[-1].split(":")[1]), D * (D - 1) / 2 * frames


def cpu_baselines(atoms_sample, shipped_atoms, frames, bins):
    """The CPU path timed on this host: the oracle's restatement of PARENT.x
    (oracle_parent.x: the reference's per-row OpenMP schedule and FP64
    on-the-fly binning, PARENT.cpp:774-1479) built (1) -O3 -march=native for
    this host and (2) with the reference's shipped flags (-fopenmp -Wall, no
    -O: Makefile:9,50), each on a bounded sample of the workload.  Returns
    (cpu_baseline, cpu_shipped_flags); either may be None."""
    host = cpu_host()
    nthreads = host["threads_used"]
    log(f"cpu host: {host}")
    out = []
    with tempfile.TemporaryDirectory() as td:
        native = _native_oracle(td)
        o3 = os.path.join(ROOT, "oracle", "_build", "oracle_parent_O3.x")
        shipped = os.path.join(ROOT, "oracle", "_build", "oracle_parent_shipped.x")
        legs = ((native or o3, atoms_sample,
                 "-O3 -march=native -fopenmp (built on this host)" if native else "-O3 -march=x86-64-v3 -fopenmp"),
                (shipped, shipped_atoms, "the reference's shipped flags: -fopenmp -Wall, no -O (i.e. -O0)"))
        for exe, atoms, flags in legs:
            if atoms <= 0 or not exe or not os.path.exists(exe):
                out.append(None)
                continue
            t = _time_cpu(exe, atoms, frames, bins, nthreads, td)
            if t is None:
                out.append(None)
                continue
            secs, upd = t
            D = 3 * atoms - 6
            log(f"cpu {flags}: {secs:.2f} s on {D} dofs x {frames} frames -> {upd / secs:.3e} updates/s")
            out.append({"value": upd / secs, "unit": "pair-frame updates/s", "cores": nthreads, "kind": "port",
                        "sample": f"{atoms}-atom chain ({D} dofs, {D * (D - 1) // 2} pairs) x {frames} frames, "
                                  f"{bins} bins, same distributions; 'Calculation time' span (PARENT.cpp:498-1489) "
                                  f"= {secs:.2f} s; oracle restatement of PARENT.x (reference OpenMP schedule) "
                                  f"built {flags}; {nthreads} OpenMP threads",
                        "host": host})
    return out[0], out[1]


def roofline_of(pa, method, variant, bins, updates, kernel_ms, workload, D, P, frames):
    """Roofline of the dominant kernel (the pair kernel).  On the LDS-atomic
    paths one pair-frame update is one ds_add_u32, so the bound is the LDS
    atomic rate (measured: no published figure exists); the one-hot int8 MFMA
    path needs 2 B^2 int8 ops per update against the int8 peak."""
    rate = updates / (kernel_ms / 1e3)
    mfma_ceiling = INT8_PEAK_OPS / (2.0 * bins * bins)
    if method == pa.METHOD_MFMA:
        # variant 1 multiplies FP4 operands: its peak is the dense FP4 rate
        peak, kind = (FP4_PEAK_OPS, "fp4") if variant == 1 else (INT8_PEAK_OPS, "int8")
        r = {"bound": "mfma", "achieved": 2.0 * bins * bins * rate / 1e12, "peak": peak / 1e12,
             "unit": "TFLOP/s", "frac": 2.0 * bins * bins * rate / peak, "traffic": None,
             "algorithmic_work_per_launch": f"{2.0 * bins * bins * updates:.4e} {kind} ops (2 B^2 per update)"}
    else:
        peak = lds_peak()
        r = {"bound": "lds_atomic", "achieved": rate / 1e9, "peak": peak["conflict_free"] / 1e9,
             "unit": "Gatomic/s", "frac": rate / peak["conflict_free"], "traffic": None,
             "frac_of_random_scatter_peak": rate / peak["random_2500"],
             "peak_source": peak["source"] + "; conflict-free ds_add_u32, 16 lanes/clk/CU",
             "algorithmic_work_per_launch": f"{updates:.4e} pair-frame updates (1 LDS atomic each)"}
    traffic, source = PROFILED_TRAFFIC.get((workload, method, variant), (None, None))
    r["traffic"] = traffic
    if traffic is not None:
        r["traffic_source"] = f"{source} (method {method} variant {variant}, the kernel this step ran)"
        r["hbm"] = {"achieved": traffic / (kernel_ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": traffic / (kernel_ms / 1e3) / HBM_PEAK,
                    "algorithmic_bytes": D * (-(-frames // 4096) * 4096) + P * 8}
    r.update({"kernel": "pairs", "kernel_ms": kernel_ms, "mfma_one_hot_ceiling_updates_per_s": mfma_ceiling})
    return r


def emit(args, value, ms_step, config, extra, metric=METRIC, unit="pair-frame updates/s"):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def method_of(pa, name):
    return {"auto": pa.METHOD_AUTO, "plain": pa.METHOD_PLAIN, "atomic": pa.METHOD_ATOMIC,
            "mfma": pa.METHOD_MFMA}[name]


def run_single(args, pa, nthreads):
    atoms, frames, bins = WORKLOADS[args.workload]
    D = 3 * atoms - 6
    P = D * (D - 1) // 2
    updates = P * frames
    method = method_of(pa, args.method)
    t0 = time.time()
    vals = pa.synth_values(atoms, frames, 2, nthreads)
    log(f"synthetic trajectory {vals.shape} in {time.time() - t0:.1f} s")
    eng = pa.Engine(0)
    eng.set_trajectory(vals)
    del vals
    b6 = [bins] * 6
    if args.variant >= 0:
        eng.set_variant(pa.METHOD_MFMA if args.method == "mfma" else pa.METHOD_ATOMIC, args.variant)

    def step(m):
        eng.prepare(b6)
        eng.pairs(m)

    tune = None
    if args.tune_report:
        eng.prepare(b6)
        tune = eng.autotune()
        for t in tune:
            log(f"tune: method {t['method']} variant {t['variant']} tile {t['tile_rows']}x{t['tile_cols']} "
                f"block {t['block']}: {t['ms']:.3f} ms, {t['updates_per_s']:.3e} updates/s")
    first_ms = None
    for w in range(args.warmup):
        t = time.perf_counter()
        step(method)
        eng.sync()
        if w == 0:  # the first AUTO step also runs the autotuner
            first_ms = 1e3 * (time.perf_counter() - t)
            log(f"first step (with tuning under auto): {first_ms:.1f} ms")
    eng.sync()
    times, prep_ms, pair_ms = [], [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        step(method)
        eng.sync()
        times.append(time.perf_counter() - t)
        st = eng.stats()
        prep_ms.append(st["ms_prepare"])
        pair_ms.append(st["ms_pairs"])
        log(f"step: {times[-1] * 1e3:.1f} ms (prepare {st['ms_prepare']:.2f} ms, pairs {st['ms_pairs']:.2f} ms, "
            f"method {st['method']} variant {st['kernel_variant']})")
    st = eng.stats()
    ms_step = 1e3 * sum(times) / len(times)
    pair_avg = sum(pair_ms) / len(pair_ms)
    kernel_rate = updates / (pair_avg / 1e3)

    mist = None
    if not args.no_mist:
        _, _, _, tot = eng.mist()
        mist = {"ms": eng.stats()["ms_mist"], "vertices": D, "total_configurational_entropy": float(tot[-1])}
        log(f"mist: {mist['ms']:.1f} ms, total {tot[-1]:.6f}")

    plain = None
    if args.plain_steps > 0 and method != pa.METHOD_PLAIN:
        step(pa.METHOD_PLAIN)
        pt = []
        for _ in range(args.plain_steps):
            t = time.perf_counter()
            step(pa.METHOD_PLAIN)
            eng.sync()
            pt.append(time.perf_counter() - t)
            log(f"plain step: {pt[-1] * 1e3:.1f} ms (pairs {eng.stats()['ms_pairs']:.2f} ms)")
        pms = 1e3 * sum(pt) / len(pt)
        plain = {"value": updates / (pms / 1e3), "ms_per_step": pms,
                 "speedup": (updates / (ms_step / 1e3)) / (updates / (pms / 1e3))}

    roofline = roofline_of(pa, st["method"], st["kernel_variant"], bins, updates, pair_avg, args.workload, D, P, frames)
    cpu, shipped = (None, None) if args.no_cpu else cpu_baselines(args.cpu_sample_atoms, args.cpu_shipped_atoms,
                                                                   frames, bins)
    extra = {"phases": {"prepare_ms": sum(prep_ms) / len(prep_ms), "pairs_ms": pair_avg,
                        "pairs_kernel_updates_per_s": kernel_rate, "first_step_ms": first_ms},
             "plain": plain, "mist": mist, "roofline": roofline, "cpu_baseline": cpu,
             "cpu_shipped_flags": shipped}
    if tune is not None:
        extra["tuning"] = tune
    eng.close()
    emit(args, updates / (ms_step / 1e3), ms_step,
         {"workload": args.workload, "atoms": atoms, "dofs": D, "pairs": P, "frames": frames, "bins": bins,
          "method": st["method"], "variant": st["kernel_variant"]}, extra)


def run_sweep(args, pa, nthreads):
    atoms, frames = 1500, 100_000
    D = 3 * atoms - 6
    P = D * (D - 1) // 2
    updates = P * frames * len(SWEEP_BINS)
    method = method_of(pa, args.method)
    vals = pa.synth_values(atoms, frames, 2, nthreads)
    eng = pa.Engine(0)
    eng.set_trajectory(vals)
    del vals
    t0 = time.time()
    eng.sweep(SWEEP_BINS, method)  # first call: per-set prepare, tuning, tile lists, capture + one launch
    log(f"sweep planned and captured in {time.time() - t0:.1f} s")
    for _ in range(args.warmup):
        eng.sweep(SWEEP_BINS, method)
    times, graph_ms = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        eng.sweep(SWEEP_BINS, method)
        times.append(time.perf_counter() - t)
        graph_ms.append(eng.stats()["ms_sweep"])
        log(f"sweep step: {times[-1] * 1e3:.1f} ms (graph {graph_ms[-1]:.1f} ms)")
    ms_step = 1e3 * sum(times) / len(times)
    per_bin = []
    for k, b in enumerate(SWEEP_BINS):
        eng.select(k)
        st = eng.stats()
        s1 = eng.s1d()
        entry = {"bins": b, "method": st["method"], "variant": st["kernel_variant"], "sum_s1d": float(s1.sum())}
        if not args.no_mist:
            _, mi, _, tot = eng.mist()
            entry.update({"mist_ms": eng.stats()["ms_mist"], "mist_total_entropy": float(tot[-1]),
                          "mist_sum_mi": float(mi.sum())})
        per_bin.append(entry)
        log(f"bins {b}: method {entry['method']} variant {entry['variant']} "
            + (f"MIST total {entry['mist_total_entropy']:.6f}" if not args.no_mist else ""))
    eng.close()
    cpu = None if args.no_cpu else cpu_baselines(args.cpu_sample_atoms, 0, frames, 50)[0]
    if cpu:
        cpu["sample"] += "; 50 bins stands for the sweep (CPU cost per update is flat in B)"
    emit(args, updates / (ms_step / 1e3), ms_step,
         {"workload": "sweep_1500x100k", "atoms": atoms, "dofs": D, "pairs": P, "frames": frames,
          "bins": SWEEP_BINS, "method": args.method},
         {"phases": {"graph_ms": sum(graph_ms) / len(graph_ms)}, "per_bin_count": per_bin, "cpu_baseline": cpu,
          "roofline": None})


def batch_shapes():
    """64 (atoms, frames) pairs: log-uniform atoms in [50, 2000] and frames in
    [5k, 200k], from a fixed splitmix64 stream (seed 4)."""
    s = BATCH_SEED

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return ((z ^ (z >> 31)) >> 11) / float(1 << 53)
    out = []
    for _ in range(BATCH_N):
        a = int(round(math.exp(math.log(50) + nxt() * (math.log(2000) - math.log(50)))))
        n = int(round(math.exp(math.log(5000) + nxt() * (math.log(200_000) - math.log(5000)))))
        out.append((a, n))
    return out


def run_batch(args, pa, nthreads):
    shapes = batch_shapes()
    bins = BATCH_BINS
    method = method_of(pa, args.method)
    updates = sum((3 * a - 6) * (3 * a - 7) // 2 * n for a, n in shapes)
    t0 = time.time()
    b = pa.Batch([bins] * 6)
    prep_ms = 0.0
    for k, (a, n) in enumerate(shapes):
        v = pa.synth_values(a, n, 1000 + k, nthreads)
        b.add(v)
        prep_ms += b.stats()["ms_prepare"]
        del v
    log(f"{len(shapes)} trajectories synthesised, uploaded and pre-binned in {time.time() - t0:.1f} s "
        f"(device pre-binning {prep_ms:.1f} ms); {updates:.4e} pair-frame updates")
    tune = None
    if args.tune_report:
        tune = b.autotune()
        for t in tune:
            log(f"batch tune: method {t['method']} variant {t['variant']} tile {t['tile_rows']}x{t['tile_cols']} "
                f"block {t['block']}: predicted {t['ms']:.1f} ms")
    for _ in range(args.warmup):
        b.run(method)
    run_ms = []
    for _ in range(args.steps):
        b.run(method)
        st = b.stats()
        run_ms.append(st["ms_pairs"])
        log(f"batch pairs: {run_ms[-1]:.1f} ms in {st['batch_launches']} launch(es) (method {st['method']} "
            f"variant {st['kernel_variant']} for the most updates)")
    st = b.stats()
    pair_avg = sum(run_ms) / len(run_ms)
    ms_step = prep_ms + pair_avg
    mist = []
    if not args.no_mist:
        for k in range(len(shapes)):
            _, _, _, tot = b.mist(k)
            mist.append(float(tot[-1]))
    b.close()
    roofline = roofline_of(pa, st["method"], st["kernel_variant"], bins, updates, pair_avg, "batch64_b30", 0, 0, 0)
    cpu = None if args.no_cpu else cpu_baselines(args.cpu_sample_atoms, 0, 100_000, bins)[0]
    emit(args, updates / (ms_step / 1e3), ms_step,
         {"workload": "batch64_b30", "trajectories": len(shapes), "bins": bins, "updates": updates,
          "atoms_frames": shapes, "method": st["method"], "variant": st["kernel_variant"],
          "launches": st["batch_launches"]},
         {"phases": {"prepare_ms_total": prep_ms, "pairs_ms": pair_avg,
                     "pairs_kernel_updates_per_s": updates / (pair_avg / 1e3)},
          "mist_totals": mist, "roofline": roofline, "cpu_baseline": cpu, "tuning": tune})


def run_readin(args, pa, nthreads):
    atoms, frames = 1500, 100_000
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "protein.bat")
        t0 = time.time()
        pa.synth_write_bat(path, atoms, frames, 2, nthreads)
        size = os.path.getsize(path)
        log(f"wrote {path}: {size / 1e9:.2f} GB in {time.time() - t0:.1f} s")
        eng = pa.Engine(0)

        def streamed():
            with pa.Bat(path) as b:
                eng.load_bat(b)
            eng.sync()

        def host():
            with pa.Bat(path) as b:
                v = b.read_frames()
            eng.set_trajectory(v)
            eng.sync()
        for _ in range(args.warmup):
            streamed()
            host()
        st_ms, host_ms = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            streamed()
            st_ms.append(1e3 * (time.perf_counter() - t))
            t = time.perf_counter()
            host()
            host_ms.append(1e3 * (time.perf_counter() - t))
            log(f"read-in: streamed {st_ms[-1]:.1f} ms, host path {host_ms[-1]:.1f} ms")
        # the streamed values feed the same prepare as the host-path values
        eng.prepare([50] * 6)
        eng.close()
    ms = sum(st_ms) / len(st_ms)
    hms = sum(host_ms) / len(host_ms)
    emit(args, size / (ms / 1e3), ms,
         {"workload": "readin_1500x100k", "atoms": atoms, "dofs": 3 * atoms - 6, "frames": frames,
          "file_bytes": size, "precision": "fp32"},
         {"host_path": {"ms": hms, "bytes_per_s": size / (hms / 1e3),
                        "what": "parent_bat_read_frames (host de-interleave) + set_trajectory (pageable H2D)"},
          "speedup_vs_host_path": hms / ms, "roofline": None, "cpu_baseline": None},
         metric=".bat read-in (file -> frame-major values in HBM), bytes/s", unit="B/s")


def workload_name(name):
    """A listed workload, or the default trajectory at another bin count
    (protein_1500x100k_bB, added to WORKLOADS)."""
    if name in WORKLOADS or name in ("sweep_1500x100k", "batch64_b30", "readin_1500x100k"):
        return name
    if name.startswith("protein_1500x100k_b") and name[19:].isdigit() and 1 <= int(name[19:]) <= 4096:
        WORKLOADS[name] = (1500, 100_000, int(name[19:]))
        return name
    raise argparse.ArgumentTypeError(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="protein_1500x100k_b50", type=workload_name,
                    help="one of " + ", ".join(sorted(WORKLOADS) + ["sweep_1500x100k", "batch64_b30",
                                                                   "readin_1500x100k"])
                    + "; protein_1500x100k_bB also takes any other bin count B from 1 to 4,096")
    ap.add_argument("--method", default="auto", choices=["auto", "plain", "atomic", "mfma"])
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain kernel (0: skip)")
    ap.add_argument("--cpu-sample-atoms", type=int, default=500)
    ap.add_argument("--cpu-shipped-atoms", type=int, default=200,
                    help="sample for the restatement built with the reference's shipped flags (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mist", action="store_true")
    ap.add_argument("--tune-report", action="store_true", help="print the autotuner's table")
    ap.add_argument("--variant", type=int, default=-1,
                    help="pin the atomic method's configuration (with --method mfma: 0 int8, 1 FP4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU sharding is a design note, not a runnable path.", file=sys.stderr)
        sys.exit(2)
    import parent_amd as pa
    nthreads = cpu_host()["threads_used"]  # synthetic-data generation threads
    if args.workload == "sweep_1500x100k":
        run_sweep(args, pa, nthreads)
    elif args.workload == "batch64_b30":
        run_batch(args, pa, nthreads)
    elif args.workload == "readin_1500x100k":
        run_readin(args, pa, nthreads)
    else:
        run_single(args, pa, nthreads)


if __name__ == "__main__":
    main()
