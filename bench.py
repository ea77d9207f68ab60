"""Benchmark: Q1 Poisson residual + Jacobian on a 256^3 hex8 cube (precision form).

Workload: 3D scalar Poisson on the unit cube, n^3 hex8 cells (default n=256:
16,777,216 elements, 16,974,593 DOFs, 454,756,609 CRS nonzeros), 2x2x2 Gauss,
source f = 1, k = exp(g) with g ~ N(0, 0.5^2) iid per element, u ~ U[-1,1],
Dirichlet u = 0 on x = 0.  One step = one residual + Jacobian assembly of the
whole mesh (volume kernel + Dirichlet rows) with inputs resident in HBM.

The timed form integrates elements in FP32 (geometry and the scatter
accumulation in FP64); the plain FP64 GPU path runs beside it and the max
relative error of FP32 vs FP64 is reported for R and J.  Prints one JSON line.

After the headline, the default run times the same workload on jittered cells
(the general kernel) and appends compact summaries of the other workloads under
"other_workloads" (Q2 hex27 elasticity 64^3, a batch of 512 Poisson meshes,
transient elasticity served at 20 requests/s, transient hex8 elasticity 128^3);
--no-extra skips them, --workload runs one of them alone with its full line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_ELEM = 408        # algorithmic bytes per Q1 element (residual + Jacobian)
HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
SEED_K, SEED_U = 1001, 2002


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n=128, reps=10, n1=48, reps1=10):
    """Oracle (reference algorithm: cached worksets, forward-AD Jacobian, row-search
    scatter) on host cores, on an n^3 sample of the same workload with all threads
    (at most 16), and on an n1^3 sample with one thread."""
    import numpy as np

    import oracle as O
    from tianxin_amd import synth

    cores = min(16, os.cpu_count() or 1)

    def run(nn, threads, rr):
        conn, ev, _ = O.brick_mesh(O.HEX8, nn, nn, nn)
        _, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        X = synth.brick_coords(nn, nn, nn)
        k = synth.lognormal(SEED_K, nn ** 3, 0.5)
        x = synth.uniform(SEED_U, owned.size, -1, 1)
        A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, k, rp, col, src=(1.0,), workset_size=1000, nthreads=threads)
        for _ in range(2):
            A.evaluate(x, mode=O.MODE_FAD)
        ts = []
        for _ in range(rr):
            t0 = time.perf_counter()
            A.evaluate(x, mode=O.MODE_FAD)
            ts.append(time.perf_counter() - t0)
        return nn ** 3 / float(np.median(ts))

    v = run(n, cores, reps)
    v1 = run(n1, 1, reps1)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": v, "unit": "elements/s", "cores": cores, "kind": "port",
            "host": {"cpu_model": model, "logical_cpus": os.cpu_count(), "threads_used": cores},
            "sample": f"Q1 Poisson {n}^3 hex8 ({n ** 3} elements), FP64, forward-AD Jacobian, workset 1000, "
                      f"median of {reps} after 2 warm-ups, {cores} OpenMP threads",
            "single_core": {"value": v1, "unit": "elements/s", "cores": 1,
                            "sample": f"{n1}^3 hex8, median of {reps1} after 2 warm-ups, 1 thread"}}


COPY_BYTES = 1 << 30
# the unit cube's cells are parallelepipeds: the one-metric hex8 kind (17); jittered cells run kind 1
FP32_KERNEL = "k_q1_hex8_diffusion<17, float, false, false>"   # FP32 integration, steady
FP64_KERNEL = "k_q1_hex8_diffusion<17, double, false, true>"   # FP64, steady, beta = 1
FP32_KERNEL_GENERAL = "k_q1_hex8_diffusion<1, float, false, false>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "q1_hex8_diffusion_fp32_256_traffic.json")


def _csvs(d, suffix):
    import glob
    out = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not out:
        raise SystemExit(f"no *{suffix} under {d}")
    return out


def _rows(d, suffix):
    import csv
    rows = []
    for fn in _csvs(d, suffix):
        with open(fn) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _match(name, kernel):
    """Demangled-name match of a kernel instance: 'k_q1_hex8_ring<1, float, false>'."""
    base, targs = kernel.split("<", 1)
    targs = [t.strip() for t in targs.rstrip(">").split(",")]
    if base not in name:
        return False
    inside = name.split(base, 1)[1]
    return all(t in inside.split("(")[0] for t in targs) and inside.split("(")[0].replace(" ", "").startswith(
        "<" + ",".join(targs))


def _counter_per_dispatch(d, counter, pred, last=None):
    """Mean per-dispatch value of `counter` over dispatches whose kernel name satisfies
    pred (only the `last` such dispatches when given)."""
    per = {}
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter or not pred(r.get("Kernel_Name", "")):
            continue
        key = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} samples for the requested kernel in {d}")
    keys = sorted(per)[-last:] if last else sorted(per)
    return sum(per[k] for k in keys) / len(keys), len(keys)


PROF_STEPS, PROF_WARMUP = 20, 3   # the stats pass: bench.py --steps 20 --warmup 3


def _trace_durations(d, kernel):
    """Durations (us) of the dispatches of `kernel`, in launch order, from a kernel trace."""
    rows = [r for r in _rows(d, "kernel_trace.csv") if _match(r.get("Kernel_Name", ""), kernel)]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows]


def summarize_profiles(stats_dir, fetch_dir, write_dir):
    """Writes profiles/q1_hex8_diffusion_fp32_256_kernel_stats.md and the traffic JSON.

    The kernel times come from the kernel trace of the stats pass.  bench.py launches each
    form's volume kernel in this order: PROF_WARMUP untimed steps, PROF_STEPS timed steps
    (volume kernel + Dirichlet rows), one untimed and PROF_STEPS timed volume-only steps,
    and one last full step; the averages below are over the timed dispatches only.
    FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by what the same
    counter reads for a 1 GiB elementwise kernel (dst = 2 src: COPY_BYTES read,
    COPY_BYTES written)."""
    import numpy as np
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    stats = _rows(stats_dir, "kernel_stats.csv")
    K, W = PROF_STEPS, PROF_WARMUP
    cmd_stats = ("rocprofv3 --kernel-trace --stats --output-format csv -d <out>/prof_stats -- "
                 f"python3 bench.py --steps {K} --warmup {W} --no-cpu --no-extra")
    lines = ["# Kernel statistics: Q1 Poisson hex8 256^3 (bench.py workload)", "",
             "Command (from the repository root, on one MI355X):", "", "    " + cmd_stats, ""]
    timed = {}
    for label, kern in (("FP32 form", FP32_KERNEL), ("FP64 path", FP64_KERNEL)):
        d = _trace_durations(stats_dir, kern)
        if len(d) != W + 2 * K + 2:
            raise SystemExit(f"{kern}: {len(d)} dispatches in the trace, expected {W + 2 * K + 2}")
        with_bc, alone = np.array(d[W:W + K]), np.array(d[W + K + 1:W + 2 * K + 1])
        timed[label] = (kern, with_bc, alone, d)
    dirich = [t for r in _rows(stats_dir, "kernel_trace.csv") if "k_dirichlet" in r.get("Kernel_Name", "")
              for t in [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3]]
    nel = 256 ** 3
    lines += ["Per-dispatch durations of the volume kernel from the kernel trace (us), timed dispatches only:", "",
              "| form | kernel | timed steps (with Dirichlet pass) mean / min / max | "
              "volume-only timed steps mean / min / max | HBM fraction at 408 B/element (volume-only mean) |",
              "|---|---|---|---|---:|"]
    for label, (kern, wb, al, _) in timed.items():
        lines.append(f"| {label} | `{kern}` | {wb.mean():.1f} / {wb.min():.1f} / {wb.max():.1f} | "
                     f"{al.mean():.1f} / {al.min():.1f} / {al.max():.1f} | "
                     f"{BYTES_PER_ELEM * nel / (al.mean() / 1e6) / HBM_PEAK:.3f} |")
    if dirich:
        lines += ["", f"`k_dirichlet` (the Dirichlet-row pass of each full step): {len(dirich)} calls, "
                      f"mean {np.mean(dirich):.1f} us"]
    lines += ["", "All dispatches of the run (setup and checks included), from the stats file:", "",
              "| kernel | calls | average us | min us | max us | share % |", "|---|---:|---:|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:12]:
        lines.append(f"| `{r['Name'][:110]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    # torch.mul(src, 2, out=dst), the last three dispatches of the run: reads and writes COPY_BYTES
    copy_pred = lambda n: "MulFunctor<double>" in n
    kf, nf = _counter_per_dispatch(fetch_dir, "FETCH_SIZE", lambda n: _match(n, FP32_KERNEL))
    kw, nw = _counter_per_dispatch(write_dir, "WRITE_SIZE", lambda n: _match(n, FP32_KERNEL))
    cf, _ = _counter_per_dispatch(fetch_dir, "FETCH_SIZE", copy_pred, last=3)
    cw, _ = _counter_per_dispatch(write_dir, "WRITE_SIZE", copy_pred, last=3)
    ff, fw = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    fetch_b, write_b = kf * 1024.0 * ff, kw * 1024.0 * fw
    al = timed["FP32 form"][2]
    tr = {"kernel": FP32_KERNEL, "workload": "Q1 Poisson hex8 256^3, FP32 integration",
          "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw, "dispatches": [nf, nw],
          "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_factor": ff, "write_factor": fw,
          "fetch_bytes": fetch_b, "write_bytes": write_b, "bytes_per_launch": fetch_b + write_b,
          "algorithmic_bytes": BYTES_PER_ELEM * nel, "kernel_us_volume_only_timed_mean": float(al.mean()),
          "commands": [f"rocprofv3 --pmc {c} --output-format csv -d <out>/pmc_{c.lower()} -- "
                       "python3 bench.py --steps 2 --warmup 1 --no-cpu --no-extra --copy-calibration"
                       for c in ("FETCH_SIZE", "WRITE_SIZE")]}
    json.dump(tr, open(TRAFFIC_FILE, "w"), indent=1)
    lines += ["", "## HBM traffic of the FP32 kernel (rocprofv3 --pmc, one pass per counter)", ""]
    lines += ["    " + c for c in tr["commands"]]
    lines += ["", f"- `{FP32_KERNEL}` average over {nf} / {nw} dispatches: FETCH_SIZE {kf:.0f} KiB, "
                  f"WRITE_SIZE {kw:.0f} KiB",
              f"- 1 GiB elementwise calibration kernel (dst = 2 src): FETCH_SIZE {cf:.0f} KiB, WRITE_SIZE {cw:.0f} KiB -> scale factors "
              f"{ff:.3f} (fetch), {fw:.3f} (write)",
              f"- traffic per launch: {fetch_b / 1e9:.3f} GB read + {write_b / 1e9:.3f} GB written = "
              f"{(fetch_b + write_b) / 1e9:.3f} GB; algorithmic {BYTES_PER_ELEM * nel / 1e9:.3f} GB "
              f"({BYTES_PER_ELEM} B/element x {nel} elements)"]
    out = os.path.join(ROOT, "profiles", "q1_hex8_diffusion_fp32_256_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out, TRAFFIC_FILE)


UTIL_A = ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY "
          "SQ_ACTIVE_INST_VALU SQ_INSTS_VALU GRBM_GUI_ACTIVE")
UTIL_B = ("SQ_INSTS_LDS SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS "
          "SQ_ACTIVE_INST_LDS SQ_INST_LEVEL_LDS GRBM_GUI_ACTIVE")


def summarize_counters(dir_a, dir_b):
    """Writes profiles/q1_hex8_diffusion_fp32_256_counters.md from two --pmc passes (UTIL_A,
    UTIL_B) of the headline workload: issue counters of the FP32 and FP64 volume kernels."""
    kernels = (FP32_KERNEL, FP64_KERNEL, FP32_KERNEL_GENERAL)
    vals = {}
    for d in (dir_a, dir_b):   # GRBM_GUI_ACTIVE is in both passes, whose dispatch ids overlap: keyed per pass
        for r in _rows(d, "counter_collection.csv"):
            for k in kernels:
                if _match(r.get("Kernel_Name", ""), k):
                    key = (k, r["Counter_Name"])
                    did = (d, r["Dispatch_Id"])
                    vals.setdefault(key, {}).setdefault(did, 0.0)
                    vals[key][did] += float(r["Counter_Value"])

    def avg(k, c):
        v = vals.get((k, c), {})
        return sum(v.values()) / len(v) if v else float("nan")
    present = [k for k in kernels if (k, "SQ_INSTS_VALU") in vals]
    cmd = "rocprofv3 --pmc {} --output-format csv -d <out>/{} -- python3 bench.py --steps 2 --warmup 1 --no-cpu --no-extra"
    lines = ["# Issue counters: Q1 Poisson hex8 256^3 (bench.py workload)", "",
             "Two `--pmc` passes over the bench command, from the repository root on one MI355X:", "",
             "    " + cmd.format(UTIL_A, "util_a"), "    " + cmd.format(UTIL_B, "util_b"), "",
             "Values are per dispatch, averaged over each kernel's dispatches. GRBM_GUI_ACTIVE is summed over",
             "the 8 XCDs, so the shader clock is GRBM_GUI_ACTIVE / 8 / duration. A wave64 VALU instruction",
             "occupies a SIMD for 4 cycles; VALU busy = 4 x SQ_INSTS_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8).",
             "The SQ_WAIT_* and SQ_ACTIVE_* rows are ratios to SQ_WAVE_CYCLES. The unit cube's cells are",
             "parallelepipeds (one metric per cell: kind 17); the general kind (1) appears when the run",
             "includes jittered cells.", "",
             "| | " + " | ".join(f"`{k}`" for k in present) + " |", "|---|" + "---:|" * len(present)]

    def row(label, f):
        lines.append(f"| {label} | " + " | ".join(f(k) for k in present) + " |")
    row("SQ_WAVES", lambda k: f"{avg(k, 'SQ_WAVES'):.0f}")
    row("SQ_INSTS_VALU", lambda k: f"{avg(k, 'SQ_INSTS_VALU'):.4g}")
    row("VALU busy (of SIMD cycles)",
        lambda k: f"{4 * avg(k, 'SQ_INSTS_VALU') / (1024 * avg(k, 'GRBM_GUI_ACTIVE') / 8):.3f}")
    row("VALU wave-instructions per 64 elements (element and owner waves)",
        lambda k: f"{avg(k, 'SQ_INSTS_VALU') / 256 ** 3 * 64:.0f}")
    for c in ("SQ_ACTIVE_INST_ANY", "SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_WAIT_INST_LDS"):
        row(f"{c} / SQ_WAVE_CYCLES", lambda k, c=c: f"{avg(k, c) / avg(k, 'SQ_WAVE_CYCLES'):.3f}")
    row("SQ_INSTS_LDS", lambda k: f"{avg(k, 'SQ_INSTS_LDS'):.4g}")
    row("SQ_LDS_BANK_CONFLICT (cycles)", lambda k: f"{avg(k, 'SQ_LDS_BANK_CONFLICT'):.3g}")
    row("SQ_INSTS_VMEM_RD / SQ_INSTS_VMEM_WR", lambda k: f"{avg(k, 'SQ_INSTS_VMEM_RD'):.3g} / {avg(k, 'SQ_INSTS_VMEM_WR'):.3g}")
    row("SQ_INSTS_SALU", lambda k: f"{avg(k, 'SQ_INSTS_SALU'):.4g}")
    out = os.path.join(ROOT, "profiles", "q1_hex8_diffusion_fp32_256_counters.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out)


def measured_hbm_peaks(gib=2, reps=10):
    """Achievable HBM bandwidth on this GPU (GB/s): a device-to-device copy of `gib` GiB
    (read + write bytes counted) and a fill of the same size (write bytes), each the
    median of `reps` event-timed runs after a warm-up."""
    import numpy as np
    import torch
    n = gib * (1 << 30) // 8
    src = torch.ones(n, device="cuda", dtype=torch.float64)
    dst = torch.empty_like(src)

    def med(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / 1e3)
        return float(np.median(ts))
    t_copy = med(lambda: dst.copy_(src))
    t_fill = med(lambda: dst.fill_(0.5))
    del src, dst
    torch.cuda.empty_cache()
    return 2 * 8 * n / t_copy / 1e9, 8 * n / t_fill / 1e9


def log_uniform_sizes(seed, count, lo, hi):
    from tianxin_amd import synth
    import numpy as np
    u = synth.uniform(seed, count, np.log(lo), np.log(hi))
    return np.rint(np.exp(u)).astype(int)


def bench_batch(args):
    """Many meshes in one persistent launch: 512 meshes, n^3 cells with n log-uniform in
    [8, 64], hex8 or tet4 (6-tet split) 50/50 by seed, Poisson with lognormal k
    (sigma 0.5), u ~ U[-1, 1], u = 0 on x = 0.  Times the autotuned FP32 batch (graph
    replay), the FP64 batch and the plain FP64 path (one txa_assemble per mesh)."""
    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    nm = args.meshes
    sizes = log_uniform_sizes(31, nm, 8, 64)
    kinds = np.where(synth.uniform(32, nm, 0, 1) < 0.5, txa.HEX8, txa.TET4)
    eng = txa.Engine(0)
    t0 = time.time()
    ms, per = [], []
    for i in range(nm):
        n = int(sizes[i])
        m = eng.brick(int(kinds[i]), n, n, n, synth.brick_coords(n, n, n))
        nd = m.build_dofmap(1)
        nz = m.build_graph()
        pl = m.point_lids()[:, 0]
        bc = torch.as_tensor(np.unique(pl[np.arange(pl.size) % (n + 1) == 0]).astype(np.int32), device="cuda")
        k = torch.as_tensor(synth.lognormal(1000 + i, m.nelem, 0.5), device="cuda")
        x = torch.as_tensor(synth.uniform(5000 + i, nd, -1, 1), device="cuda")
        bufs = {p: (torch.empty(nd, device="cuda", dtype=torch.float64),
                    torch.empty(nz, device="cuda", dtype=torch.float64)) for p in (txa.FP32, txa.FP64)}
        ms.append(m)
        per.append(dict(coef=k, x=x, bc=bc, bufs=bufs))
    nel = sum(m.nelem for m in ms)
    cells = int(np.sum(sizes.astype(np.int64) ** 3))
    # compulsory HBM bytes of one batch call: CRS values and f written once, x and k read
    cbytes = sum(8 * m.nnz + 16 * m.ndofs + 8 * m.nelem for m in ms)
    log(f"setup {time.time() - t0:.1f}s: {nm} meshes, {cells} cells, {nel} elements")

    def kw(i, prec):
        f, v = per[i]["bufs"][prec]
        return dict(physics=txa.DIFFUSION, coef=per[i]["coef"], x=per[i]["x"], f=f, vals=v, src=(1.0,),
                    bc_lids=per[i]["bc"])

    def time_it(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    out = {}
    for prec, name in ((txa.FP32, "fp32_batch"), (txa.FP64, "fp64_batch")):
        b = txa.Batch(eng, ms, precision=prec)
        # the argument blocks are checked and packed once, as a time loop over the same
        # device arrays would; each timed step is one batch call (graph replay)
        args_list = b.pack([kw(i, prec) for i in range(nm)])
        tuned_ms = b.tune(args_list, reps=3)
        info = b.info()
        t = time_it(lambda: b.assemble(args_list, use_graph=True), args.steps, 2)
        out[name] = {"ms_per_step": t, "elements_per_s": nel / (t / 1e3), "cells_per_s": cells / (t / 1e3),
                     "hbm_frac": cbytes / (t / 1e3) / HBM_PEAK,
                     "tuner_best_ms": tuned_ms,
                     "slab": info["slab"], "grid": info["grid"], "work_items": info["items"],
                     "tile": f"{info['tile_x']}x{256 // info['tile_x']}", "launches": info["launches"]}
        log(f"{name}: {t:.3f} ms (tile {info['tile_x']}, slab {info['slab']}, grid {info['grid']})")
        b.destroy()

    def plain():
        for i, m in enumerate(ms):
            m.assemble(precision=txa.FP64, **kw(i, txa.FP64))
    t = time_it(plain, max(2, args.steps // 4), 1)
    out["fp64_per_mesh"] = {"ms_per_step": t, "elements_per_s": nel / (t / 1e3), "cells_per_s": cells / (t / 1e3)}
    log(f"fp64 one launch per mesh: {t:.3f} ms")
    # FP32 batch vs FP64, per mesh
    worst = {"R": 0.0, "J": 0.0, "R_entrywise_over_1e-12_max": 0.0, "J_entrywise_over_1e-12_max": 0.0}

    def ent(a, b, tau=1e-12):
        mask = b.abs() > tau * b.abs().max()
        return float(((a - b).abs()[mask] / b.abs()[mask]).max())
    for i in range(nm):
        (f32, v32), (f64, v64) = per[i]["bufs"][txa.FP32], per[i]["bufs"][txa.FP64]
        worst["R"] = max(worst["R"], float((f32 - f64).abs().max() / f64.abs().max()))
        worst["J"] = max(worst["J"], float((v32 - v64).abs().max() / v64.abs().max()))
        worst["R_entrywise_over_1e-12_max"] = max(worst["R_entrywise_over_1e-12_max"], ent(f32, f64))
        worst["J_entrywise_over_1e-12_max"] = max(worst["J_entrywise_over_1e-12_max"], ent(v32, v64))
    value = out["fp32_batch"]["elements_per_s"]
    line = {"metric": "elements/s (residual+Jacobian), batched", "value": value, "unit": "elements/s",
            "n_gpus": 1, "steps": args.steps, "ms_per_step": out["fp32_batch"]["ms_per_step"],
            "higher_is_better": True, "dtype": "f32 element integration, f64 geometry and scatter",
            "data": "synthetic",
            "config": {"workload": f"{nm} Poisson meshes, n^3 cells with n log-uniform in [8, 64], hex8/tet4 50/50",
                       "cells": cells, "elements": nel, "hex8_meshes": int(np.sum(kinds == txa.HEX8))},
            "forms": out, "max_rel_err_fp32_vs_fp64_per_mesh": worst, "compulsory_bytes": cbytes,
            "speedup_vs_fp64_per_mesh": value / out["fp64_per_mesh"]["elements_per_s"]}
    for m in ms:
        m.destroy()
    return line


Q2_FLOP_PER_ELEM = 2 * 81 * 162 * 81 + 2 * 162 * 6 * 81   # dense B^T (W B), B in R^{162 x 81}
Q2_BYTES_PER_ELEM = 37.5e3   # compulsory at 64^3: CRS values written once (8 nnz) plus f and x, per element


def bench_q2(args):
    """Q2 hex27 linear elasticity, 3 DOFs/node, 81x81 element matrices, 3x3x3 Gauss:
    n^3 cells (default 64), E lognormal (sigma 0.3) per element, nu = 0.3, interior
    vertices jittered by U[-0.1h, 0.1h], u ~ U[-1, 1], u = 0 on x = 0.  Times the
    split-BF16 MFMA form, the FP32 form and the plain FP64 path; reports their errors
    against FP64 and a full-size rigid-motion check (R(rigid u) = 0)."""
    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    n = args.n if args.n != 256 else 64
    eng = txa.Engine(0)
    t0 = time.time()
    X = synth.brick_coords(n, n, n, jitter=0.1, seed=61)
    m = eng.brick(txa.HEX27, n, n, n, X)
    nd = m.build_dofmap(3)
    nnz = m.build_graph()
    pl = m.point_lids()
    px = 2 * n + 1
    bc = torch.as_tensor(np.unique(pl[np.arange(pl.shape[0]) % px == 0].ravel()).astype(np.int32), device="cuda")
    E = torch.as_tensor(synth.lognormal(62, m.nelem, 0.3), device="cuda")
    x = torch.as_tensor(synth.uniform(63, nd, -1, 1), device="cuda")
    nel = m.nelem
    compulsory = 8.0 * nnz + 16.0 * nd   # CRS values and f written once, x read once
    log(f"setup {time.time() - t0:.1f}s: {nel} elements, {nd} dofs, {nnz} nnz")
    out, res = {}, {}
    for prec, name in ((txa.SPLIT_BF16, "split_bf16"), (txa.FP32, "fp32"), (txa.FP64, "fp64")):
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)

        def step():
            m.assemble(physics=txa.ELASTICITY, precision=prec, coef=E, x=x, f=f, vals=v, nu=0.3, bc_lids=bc)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        ev[0].record()
        for i in range(args.steps):
            step()
            ev[i + 1].record()
        torch.cuda.synchronize()
        per = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])
        ms = float(per.mean())
        out[name] = {"ms_per_step": ms, "elements_per_s": nel / (ms / 1e3),
                     "dense_tflops": Q2_FLOP_PER_ELEM * nel / (ms / 1e3) / 1e12,
                     "hbm_frac": compulsory / (ms / 1e3) / HBM_PEAK,
                     "p50": float(np.percentile(per, 50)), "p99": float(np.percentile(per, 99))}
        res[name] = (f, v)
        log(f"{name}: {ms:.3f} ms/step")
    f64, v64 = res["fp64"]

    def rel(a, b):
        return float((a - b).abs().max() / b.abs().max())

    def ent(a, b, tau=1e-3):
        mask = b.abs() > tau * b.abs().max()
        return float(((a - b).abs()[mask] / b.abs()[mask]).max())
    errs = {k: {"R_normwise": rel(res[k][0], f64), "J_normwise": rel(res[k][1], v64),
                "R_entrywise_over_1e-12_max": ent(res[k][0], f64, 1e-12),
                "J_entrywise_over_1e-12_max": ent(res[k][1], v64, 1e-12),
                "R_entrywise_over_1e-3_max": ent(res[k][0], f64), "J_entrywise_over_1e-3_max": ent(res[k][1], v64)}
            for k in ("split_bf16", "fp32")}
    del res
    # size-independent property: a rigid motion is stress free (no load, no Dirichlet rows)
    P = synth.q2_positions(X, n, n, n)
    u = np.array([1.0, -2.0, 0.5]) + np.cross(np.array([0.2, -0.3, 0.1]), P)
    xr = np.zeros(nd)
    for k in range(3):
        xr[pl[:, k]] = u[:, k]
    xr = torch.as_tensor(xr, device="cuda")
    rigid = {}
    for prec, name in ((txa.SPLIT_BF16, "split_bf16"), (txa.FP64, "fp64")):
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.ELASTICITY, precision=prec, coef=E, x=xr, f=f, vals=v, nu=0.3)
        torch.cuda.synchronize()
        rigid[name] = float(f.abs().max() / (v.abs().max() * xr.abs().max()))
        del f, v
    best = out["split_bf16"]
    line = {"metric": "elements/s (residual+Jacobian), Q2 elasticity", "value": best["elements_per_s"],
            "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["ms_per_step"], "higher_is_better": True,
            "dtype": "split-bf16 MFMA (fp32 accumulate) element integration, f64 geometry and scatter",
            "data": "synthetic",
            "config": {"workload": f"Q2 hex27 linear elasticity {n}^3, 3 dofs/node, jitter 0.1h interior vertices",
                       "elements": nel, "dofs": nd, "nnz": nnz, "quadrature": "3x3x3 Gauss", "E": "lognormal sigma=0.3",
                       "nu": 0.3, "dirichlet": "u=0 on x=0"},
            "forms": out, "max_rel_err_vs_fp64": errs, "rigid_motion_residual_rel": rigid,
            "speedup_split_vs_fp64": best["elements_per_s"] / out["fp64"]["elements_per_s"]}
    m.destroy()
    return line


def bench_serve(args):
    """Serving: Newmark steps of transient hex8 elasticity (beta 0.25, gamma 0.5,
    Rayleigh a 0.01, b 0.001, dt 1e-3) on a pool of resident meshes with n^3 cells,
    n log-uniform in [16, 128].  For each served precision (FP64, then FP32): phase 1,
    Poisson arrivals at --qps, p50/p99 submit-to-completion latency; phase 2, the same
    requests queued at once (throughput at saturation).  Beside them, the plain FP64
    path: the same requests one txa_assemble after another on one stream.  Each
    request carries a seeded state U[-1e-3, 1e-3] generated on the device.  The
    served results of the first 8 requests are checked against the plain FP64 path
    (normwise and entrywise over |ref| > 1e-12 max|ref|)."""
    import threading

    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    head = {"fp64": txa.FP64, "fp32": txa.FP32}[args.serve_precision]
    eng = txa.Engine(0)
    sizes = log_uniform_sizes(41, args.pool, 16, 128)
    t0 = time.time()
    pool = []
    for i, n in enumerate(sizes):
        n = int(n)
        m = eng.brick(txa.HEX8, n, n, n, synth.brick_coords(n, n, n))
        m.build_dofmap(3)
        m.build_graph()
        pl = m.point_lids()
        bc = torch.as_tensor(np.unique(pl[np.arange(pl.shape[0]) % (n + 1) == 0].ravel()).astype(np.int32),
                             device="cuda")
        E = torch.as_tensor(synth.lognormal(2000 + i, m.nelem, 0.3), device="cuda")
        pool.append((m, E, bc))
    log(f"setup {time.time() - t0:.1f}s: pool n = {list(map(int, sizes))}")
    pick = np.minimum((synth.uniform(43, args.requests, 0, 1) * len(pool)).astype(int), len(pool) - 1)
    gaps = -np.log(np.maximum(synth.uniform(44, args.requests, 0, 1), 1e-12)) / args.qps
    states = []
    g = torch.Generator(device="cuda")
    for r in range(args.requests):
        m = pool[pick[r]][0]
        g.manual_seed(7000 + r)
        states.append([torch.rand(m.ndofs, generator=g, device="cuda", dtype=torch.float64) * 2e-3 - 1e-3
                       for _ in range(3)])
    torch.cuda.synchronize()
    nel = sum(int(pool[i][0].nelem) for i in pick)
    kw = dict(physics=txa.ELASTICITY, dt=1e-3, newmark=(0.25, 0.5), nu=0.3, rho=1.0, rayleigh=(0.01, 0.001))

    def serve(prec, arrivals):
        srv = txa.Server(eng, [p_[0] for p_ in pool], hbm_budget=int(args.budget_gb * 2 ** 30), slots=args.slots)
        rids = [None] * args.requests
        ready = threading.Semaphore(0)

        def collect():
            for r in range(args.requests):
                ready.acquire()
                st = txa.QUEUED
                while st != txa.DONE:
                    st, _, _ = srv.poll(rids[r], 60.0)
                srv.release(rids[r])

        th = threading.Thread(target=collect)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        for r in range(args.requests):
            if arrivals:
                t_next += gaps[r]
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            m, E, bc = pool[pick[r]]
            u, v, a = states[r]
            rids[r] = srv.submit(m, coef=E, u=u, v=v, a=a, bc_lids=bc, precision=prec, **kw)
            ready.release()
        th.join()
        wall = time.perf_counter() - t_start
        st = srv.stats()
        srv.destroy()
        return st, wall

    # plain FP64 path: the same requests back to back on one stream
    outs = {}
    for i, (m, E, bc) in enumerate(pool):
        outs[i] = (torch.empty(m.ndofs, device="cuda", dtype=torch.float64),
                   torch.empty(m.nnz, device="cuda", dtype=torch.float64))

    def plain(r):
        m, E, bc = pool[pick[r]]
        u, v, a = states[r]
        f, vals = outs[pick[r]]
        m.assemble(physics=txa.ELASTICITY, precision=txa.FP64, coef=E, x=u, xdot=v, xddot=a, f=f, vals=vals,
                   alpha=0.5 / (0.25 * 1e-3), beta=1.0, gamma2=1.0 / (0.25 * 1e-6), transient=True, nu=0.3, rho=1.0,
                   rayleigh=(0.01, 0.001), bc_lids=bc)
    # warm-up: one call per resident mesh, so the context's workspace pool has grown to the
    # largest request before the timed pass (the server's arena is allocated up front too)
    for i in sorted(set(int(k) for k in pick[:args.requests])):
        plain(int(np.nonzero(pick == i)[0][0]))
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for r in range(args.requests):
        plain(r)
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - tp

    def ent(a, b, tau):
        mk = np.abs(b) > tau * np.abs(b).max()
        return float((np.abs(a - b)[mk] / np.abs(b)[mk]).max())

    def errors(prec, sample=8):
        """Served results of the first `sample` requests against the plain FP64 path."""
        worst = dict.fromkeys(("R_normwise", "J_normwise", "R_entrywise_over_1e-12_max", "J_entrywise_over_1e-12_max",
                               "R_entrywise_over_1e-3_max", "J_entrywise_over_1e-3_max"), 0.0)
        srv = txa.Server(eng, [p_[0] for p_ in pool], hbm_budget=int(args.budget_gb * 2 ** 30), slots=args.slots)
        try:
            for r in range(min(sample, args.requests)):
                m, E, bc = pool[pick[r]]
                u, v, a = states[r]
                plain(r)
                torch.cuda.synchronize()
                fr, vr = (t_.cpu().numpy() for t_ in outs[pick[r]])
                rid = srv.submit(m, coef=E, u=u, v=v, a=a, bc_lids=bc, precision=prec, **kw)
                fs, vs = srv.result(rid, m)
                srv.release(rid)
                for key, x, y in (("R", fs, fr), ("J", vs, vr)):
                    worst[key + "_normwise"] = max(worst[key + "_normwise"], float(np.abs(x - y).max() / np.abs(y).max()))
                    for tau, tag in ((1e-12, "1e-12"), (1e-3, "1e-3")):
                        k_ = f"{key}_entrywise_over_{tag}_max"
                        worst[k_] = max(worst[k_], ent(x, y, tau))
        finally:
            srv.destroy()
        return worst

    forms = {}
    for name, prec in (("fp64", txa.FP64), ("fp32", txa.FP32)):
        lat, _ = serve(prec, True)
        sat, wall_sat = serve(prec, False)
        forms[name] = {"latency_ms": {"p50": lat["p50_ms"], "p99": lat["p99_ms"], "mean": lat["mean_ms"],
                                      "max": lat["max_ms"], "at_qps": args.qps,
                                      "requests_per_s": lat["requests_per_s"], "peak_inflight": lat["peak_inflight"]},
                       "saturated": {"elements_per_s": nel / wall_sat, "wall_s": wall_sat, "p50_ms": sat["p50_ms"],
                                     "p99_ms": sat["p99_ms"], "peak_inflight": sat["peak_inflight"],
                                     "peak_arena_bytes": sat["peak_bytes"]},
                       "speedup_vs_plain_fp64": wall_plain / wall_sat,
                       "max_rel_err_vs_plain_fp64": errors(prec)}
        log(f"served {name}: p50 {lat['p50_ms']:.2f} ms p99 {lat['p99_ms']:.2f} ms; saturated "
            f"{nel / wall_sat:.3g} elements/s")
    hn = "fp64" if head == txa.FP64 else "fp32"
    h = forms[hn]
    line = {"metric": "serving: p50/p99 latency and elements/s (transient elasticity residual+Jacobian)",
            "value": h["saturated"]["elements_per_s"], "unit": "elements/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f64" if head == txa.FP64 else "f32 element integration, f64 geometry and scatter",
            "data": "synthetic",
            "config": {"workload": "hex8 transient elasticity (Newmark), n^3 cells, n log-uniform in [16, 128]",
                       "pool_n": list(map(int, sizes)), "requests": args.requests, "qps": args.qps,
                       "slots": args.slots, "hbm_budget_gb": args.budget_gb, "dt": 1e-3, "served_precision": hn},
            "latency_ms": h["latency_ms"], "saturated": h["saturated"], "forms": forms,
            "plain_fp64_sequential": {"elements_per_s": nel / wall_plain, "wall_s": wall_plain},
            "speedup_vs_plain_fp64": h["speedup_vs_plain_fp64"], "elements_served": nel}
    for m, _, _ in pool:
        m.destroy()
    return line


Q2_KERNELS = {"split_bf16": "k_elastic<27, 27, 256, 2>", "fp32": "k_elastic<27, 27, 256, 1>",
              "fp64": "k_elastic<27, 27, 256, 0>"}
# second phase (node-centric gather) of each form: FP32 element blocks, FP64 in the FP64 form
Q2_GATHER = {"split_bf16": "k_elastic_gather<27, float>", "fp32": "k_elastic_gather<27, float>",
             "fp64": "k_elastic_gather<27, double>"}


def bench_upload(args):
    """Uploaded-mesh path (txa_mesh_upload, the connectivity a ConnManager hands over):
    an n^3 hex8 brick (default 128) uploaded as unstructured connectivity, FP64 (and the
    FP32 form beside it), steady Poisson (k lognormal sigma 0.5 per element, f = 1,
    u ~ U[-1, 1], u = 0 on x = 0).  Times the jittered cells (vertices moved by U[-0.1h, 0.1h]) in brick
    order and in a seeded random order, and unjittered (parallelepiped) cells in brick order.  Checks: in brick order
    the upload numbers LIDs as the lattice mesh does, so f and J must equal the lattice
    FP64 path's to rounding; Poisson rows off the Dirichlet set sum to zero."""
    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    n = args.n if args.n != 256 else 128
    eng = txa.Engine(0)
    X = synth.brick_coords(n, n, n, jitter=0.1, seed=3)
    k = torch.as_tensor(synth.lognormal(5, n ** 3, 0.5), device="cuda")
    out, checks = {}, {}
    ref = None
    X_cube = synth.brick_coords(n, n, n)
    for name, seed, XX in (("brick_order", None, X), ("shuffled", 11, X), ("parallelepiped_cells_brick_order", None, X_cube)):
        t0 = time.time()
        conn = synth.brick_hex8_conn(n, n, n, shuffle_seed=seed)
        m = eng.upload(txa.HEX8, conn, XX)
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        owned, _ = m.dofmap()
        lid_of_gid = np.empty(nd, np.int64)
        lid_of_gid[owned] = np.arange(nd)
        vx = (np.arange((n + 1) ** 2) * (n + 1)).astype(np.int64)     # vertices with i = 0 (GID = vertex id)
        bc = torch.as_tensor(np.sort(lid_of_gid[vx]).astype(np.int32), device="cuda")
        x = torch.as_tensor(synth.uniform(6, nd, -1, 1), device="cuda")
        if seed is not None:
            kk = torch.as_tensor(synth.lognormal(5, n ** 3, 0.5)[np.argsort(synth.splitmix64(seed, n ** 3), kind="stable")],
                                 device="cuda")
        else:
            kk = k
        log(f"{name}: setup {time.time() - t0:.1f}s, {m.nelem} elements, {nd} dofs, {nnz} nnz, patches {m.patches()}")

        def timed(precision, f, v):
            def step():
                m.assemble(physics=txa.DIFFUSION, precision=precision, coef=kk, x=x, f=f, vals=v, src=(1.0,),
                           bc_lids=bc)
            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
            ev[0].record()
            for i in range(args.steps):
                step()
                ev[i + 1].record()
            torch.cuda.synchronize()
            return np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])

        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        per = timed(txa.FP64, f, v)
        ms = float(per.mean())
        out[name] = {"ms_per_step": ms, "elements_per_s": m.nelem / (ms / 1e3),
                     "p50": float(np.percentile(per, 50)), "p99": float(np.percentile(per, 99)),
                     "patches": m.patches()}
        log(f"{name}: {ms:.3f} ms/step")
        # the FP32 form (FP32 element matrices, FP64 geometry and residual) against this FP64 result
        f32 = torch.empty_like(f)
        v32 = torch.empty_like(v)
        per32 = timed(txa.FP32, f32, v32)
        band = f.abs() > 1e-12 * f.abs().max()
        out[name]["fp32"] = {"ms_per_step": float(per32.mean()),
                             "R_entrywise_over_1e-12_max": float(((f32 - f).abs()[band] / f.abs()[band]).max()),
                             "J_normwise": float((v32 - v).abs().max() / v.abs().max())}
        log(f"{name}: FP32 form {float(per32.mean()):.3f} ms/step")
        del f32, v32
        # Poisson rows off the Dirichlet set sum to zero
        rp, _ = m.graph(with_cols=False)
        rows = torch.repeat_interleave(torch.arange(nd, device="cuda"), torch.as_tensor(np.diff(rp), device="cuda"))
        rs = torch.zeros(nd, device="cuda", dtype=torch.float64).index_add_(0, rows, v)
        keep = torch.ones(nd, dtype=torch.bool, device="cuda")
        keep[bc.long()] = False
        checks[name] = {"row_sum_rel_max": float(rs[keep].abs().max() / v.abs().max())}
        if name == "brick_order":
            ref = (f.clone(), v.clone(), x, bc)
        m.destroy()
        del f, v
    # the lattice FP64 path on the same brick, same LIDs and graph
    ml = eng.brick(txa.HEX8, n, n, n, X)
    ml.build_dofmap(1)
    ml.build_graph()
    fu, vu, x, bc = ref
    fl = torch.empty_like(fu)
    vl = torch.empty_like(vu)
    ml.assemble(physics=txa.DIFFUSION, precision=txa.FP64, coef=k, x=x, f=fl, vals=vl, src=(1.0,), bc_lids=bc)
    torch.cuda.synchronize()
    checks["brick_order_vs_lattice_fp64"] = {"R_normwise": float((fu - fl).abs().max() / fl.abs().max()),
                                            "J_normwise": float((vu - vl).abs().max() / vl.abs().max())}
    ml.destroy()
    best = out["brick_order"]
    line = {"metric": "elements/s (residual+Jacobian), uploaded mesh", "value": best["elements_per_s"],
            "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["ms_per_step"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Q1 Poisson hex8 {n}^3 brick uploaded as unstructured connectivity",
                       "elements": n ** 3, "quadrature": "2x2x2 Gauss", "k": "lognormal sigma=0.5",
                       "dirichlet": "u=0 on x=0"},
            "orders": out, "checks": checks}
    return line


def bench_elastic(args):
    """Transient hex8 linear elasticity, the served physics (Newmark seeds alpha 2000, gamma2 4e6,
    Rayleigh a 0.01, b 0.001, nu 0.3, E lognormal sigma 0.3 per cell, state U[-1e-3, 1e-3],
    u = 0 on x = 0) on one n^3 brick (default 128), FP64: parallelepiped cells (the unit cube:
    the node-centric kernel, no workspace), jittered cells (U[-0.1h, 0.1h]: the element kernel
    and the node gather), and the jittered brick uploaded as connectivity (the element kernel and
    the general node gather; cells in brick order and shuffled).  Compulsory bytes: the CRS values, f and the Dirichlet rows written once,
    x, xdot, xddot, coordinates, E and the row pointers read once."""
    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    n = args.n if args.n != 256 else 128
    eng = txa.Engine(0)
    out = {}
    for name, X in (("parallelepiped_cells", synth.brick_coords(n, n, n)),
                    ("jittered_cells", synth.brick_coords(n, n, n, jitter=0.1, seed=21))):
        m = eng.brick(txa.HEX8, n, n, n, X)
        nd = m.build_dofmap(3)
        nnz = m.build_graph()
        pl = m.point_lids()
        bc = torch.as_tensor(np.unique(pl[np.arange(pl.shape[0]) % (n + 1) == 0].ravel()).astype(np.int32),
                             device="cuda")
        E = torch.as_tensor(synth.lognormal(22, m.nelem, 0.3), device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(23)
        u, v, a = (torch.rand(nd, generator=g, device="cuda", dtype=torch.float64) * 2e-3 - 1e-3 for _ in range(3))
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        vals = torch.empty(nnz, device="cuda", dtype=torch.float64)

        def step():
            m.assemble(physics=txa.ELASTICITY, precision=txa.FP64, coef=E, x=u, xdot=v, xddot=a, f=f, vals=vals,
                       alpha=2000.0, beta=1.0, gamma2=4e6, transient=True, nu=0.3, rho=1.0, rayleigh=(0.01, 0.001),
                       bc_lids=bc)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        ev[0].record()
        for i in range(args.steps):
            step()
            ev[i + 1].record()
        torch.cuda.synchronize()
        per = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])
        ms = float(per.mean())
        compulsory = 8.0 * nnz + 8.0 * nd * 4 + 24.0 * m.nverts + 8.0 * m.nelem + 8.0 * nd
        out[name] = {"ms_per_step": ms, "elements_per_s": m.nelem / (ms / 1e3), "mesh_affine": m.affine,
                     "p50": float(np.percentile(per, 50)), "p99": float(np.percentile(per, 99)),
                     "compulsory_bytes": compulsory, "hbm_frac": compulsory / (ms / 1e3) / HBM_PEAK}
        log(f"{name}: {ms:.3f} ms/step")
        if name == "jittered_cells":
            ref = (f.clone(), vals.clone(), u, v, a, bc, E, X)
        m.destroy()
        del f, vals
    # the jittered brick uploaded as unstructured connectivity (txa_mesh_upload): the element
    # kernel and the general node gather, cells in brick order (same LIDs as the brick: checked
    # against the lattice result) and in a seeded random order
    fr, vr, u, v, a, bc, E, X = ref
    up = {}
    for order, seed in (("brick_order", None), ("shuffled", 24)):
        conn = synth.brick_hex8_conn(n, n, n, shuffle_seed=seed)
        perm = np.argsort(synth.splitmix64(seed, n ** 3), kind="stable") if seed is not None else None
        m = eng.upload(txa.HEX8, conn, X)
        nd = m.build_dofmap(3)
        nnz = m.build_graph()
        owned, _ = m.dofmap()
        lid_of_gid = np.empty(nd, np.int64)
        lid_of_gid[owned] = np.arange(nd)
        vx = (np.arange((n + 1) ** 2) * (n + 1)).astype(np.int64)   # vertices with i = 0 (GID = 3 vertex + field)
        ubc = torch.as_tensor(np.sort(lid_of_gid[(3 * vx[:, None] + np.arange(3)).ravel()]).astype(np.int32),
                              device="cuda")
        Eu = E if perm is None else E[torch.as_tensor(perm, device="cuda")]
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        vals = torch.empty(nnz, device="cuda", dtype=torch.float64)
        uu, vv, aa = (t if perm is None else torch.rand(nd, device="cuda", dtype=torch.float64) * 2e-3 - 1e-3
                      for t in (u, v, a))

        def ustep():
            m.assemble(physics=txa.ELASTICITY, precision=txa.FP64, coef=Eu, x=uu, xdot=vv, xddot=aa, f=f,
                       vals=vals, alpha=2000.0, beta=1.0, gamma2=4e6, transient=True, nu=0.3, rho=1.0,
                       rayleigh=(0.01, 0.001), bc_lids=ubc)
        for _ in range(args.warmup):
            ustep()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        ev[0].record()
        for i in range(args.steps):
            ustep()
            ev[i + 1].record()
        torch.cuda.synchronize()
        per = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])
        up[order] = {"ms_per_step": float(per.mean()), "elements_per_s": m.nelem / (float(per.mean()) / 1e3)}
        if perm is None:
            up[order]["vs_brick_normwise"] = {"R": float((f - fr).abs().max() / fr.abs().max()),
                                              "J": float((vals - vr).abs().max() / vr.abs().max())}
        log(f"uploaded jittered cells, {order}: {per.mean():.3f} ms/step")
        m.destroy()
        del f, vals
    out["jittered_cells_uploaded"] = up
    best = out["parallelepiped_cells"]
    return {"metric": "elements/s (residual+Jacobian), transient hex8 elasticity", "value": best["elements_per_s"],
            "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["ms_per_step"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"hex8 transient elasticity {n}^3 (Newmark, Rayleigh), FP64",
                       "elements": n ** 3, "E": "lognormal sigma=0.3", "nu": 0.3, "dirichlet": "u=0 on x=0"},
            "cells": out, "roofline": {"bound": "hbm", "frac": best["hbm_frac"], "peak": HBM_PEAK / 1e9,
                                       "unit": "GB/s"}}


def summarize_q2(stats_dir, pmc_dir):
    """Writes profiles/q2_hex27_elastic_64_kernel_stats.md from a kernel-stats pass and
    one --pmc pass (SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE, SQ_INSTS_VALU,
    SQ_INSTS_VALU_MFMA_MOPS_BF16, SQ_INSTS_VALU_MFMA_MOPS_F64, SQ_INSTS_LDS, SQ_WAVES) of
    `bench.py --workload q2`."""
    stats = _rows(stats_dir, "kernel_stats.csv")
    nel = 64 ** 3
    pmc_list = ("SQ_VALU_MFMA_BUSY_CYCLES GRBM_GUI_ACTIVE SQ_INSTS_VALU SQ_INSTS_VALU_MFMA_MOPS_BF16 "
                "SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_LDS SQ_WAVES")
    lines = ["# Kernel statistics: Q2 hex27 elasticity 64^3 (bench.py --workload q2)", "",
             "Commands (from the repository root, on one MI355X):", "",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <out>/prof_q2 -- "
             "python3 bench.py --workload q2 --steps 3 --warmup 1",
             f"    rocprofv3 --pmc {pmc_list} --output-format csv -d <out>/pmc_q2 -- "
             "python3 bench.py --workload q2 --steps 1 --warmup 0", "",
             "Each form is two kernels: the element kernel (geometry, Gram matrix, element block and",
             "residual to a workspace) and the node-centric gather (sums the blocks around each node into",
             "its CRS rows, writes each row once), each launched once per z slab (8 slabs of 8 element",
             "layers): the elements on one stream, each slab's node planes gathered on another once its",
             "elements are done, so slab k+1's element kernel runs beside slab k's gather.  From the",
             "kernel trace, per step: the span from the first element dispatch's start to the last",
             "gather dispatch's end, and the summed durations of each kernel's dispatches (their sum",
             "exceeds the span by the overlap).  Rates are over the span; the HBM fraction counts the",
             "compulsory bytes (CRS values and f written once, x read: "
             f"{Q2_BYTES_PER_ELEM / 1e3:.1f} KB per element) against 8 TB/s.", "",
             "| form | step span ms | element kernel ms (summed) | gather ms (summed) | overlap ms | elements/s | "
             "dense TFLOP/s | HBM fraction |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    trace = sorted(_rows(stats_dir, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    for name, k in Q2_KERNELS.items():
        # the dispatches of this form's element kernel, grouped into steps: a step's element
        # slabs are issued back to back, and the next step starts after its last gather
        el = [r for r in trace if _match(r["Kernel_Name"], k)]
        if not el:
            continue
        spans, se, sg = [], [], []
        i = 0
        while i < len(el):
            t0 = int(el[i]["Start_Timestamp"])
            j = i
            while j + 1 < len(el) and int(el[j + 1]["Start_Timestamp"]) - int(el[j]["End_Timestamp"]) < 2_000_000 \
                    and j + 1 - i < 8:
                j += 1
            t_el_end = int(el[j]["End_Timestamp"])
            nxt = int(el[j + 1]["Start_Timestamp"]) if j + 1 < len(el) else None
            g = [r for r in trace if _match(r["Kernel_Name"], Q2_GATHER[name]) and int(r["Start_Timestamp"]) >= t0
                 and (nxt is None or int(r["Start_Timestamp"]) < nxt)]
            g = g[:j + 1 - i]   # this step's gathers (one per slab)
            if g:
                t1 = max(t_el_end, max(int(r["End_Timestamp"]) for r in g))
                spans.append((t1 - t0) / 1e6)
                se.append(sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in el[i:j + 1]) / 1e6)
                sg.append(sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in g) / 1e6)
            i = j + 1
        if not spans:
            continue
        keep = slice(1, None) if len(spans) > 1 else slice(None)   # the first step is a warm-up
        ms, me, mg = (sum(v[keep]) / len(v[keep]) for v in (spans, se, sg))
        lines.append(f"| {name} | {ms:.3f} | {me:.3f} | {mg:.3f} | {me + mg - ms:.3f} | {nel / (ms / 1e3):.3g} | "
                     f"{Q2_FLOP_PER_ELEM * nel / (ms / 1e3) / 1e12:.1f} | "
                     f"{Q2_BYTES_PER_ELEM * nel / (ms / 1e3) / HBM_PEAK:.3f} |")
    lines += ["", "Other kernels of the run (setup included):", "", "| kernel | calls | average ms |", "|---|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:8]:
        if "k_elastic" not in r["Name"]:
            lines.append(f"| `{r['Name'][:80]}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} |")
    per = {}
    for r in _rows(pmc_dir, "counter_collection.csv"):
        for name, k in Q2_KERNELS.items():
            if _match(r.get("Kernel_Name", ""), k):
                d = per.setdefault(name, {})
                key = (int(r["Dispatch_Id"]), r["Counter_Name"])
                d[key] = d.get(key, 0.0) + float(r["Counter_Value"])
    lines += ["", "## Element-kernel counters (one dispatch each)", "",
              "| form | MFMA busy (of all SIMD cycles) | VALU wave-instructions / element | "
              "MFMA bf16 MOPs / element | MFMA f64 MOPs / element | LDS wave-instructions / element |",
              "|---|---:|---:|---:|---:|---:|"]
    for name, d in per.items():
        tot = {}
        for (disp, c), v in d.items():
            tot.setdefault(c, {})[disp] = v

        def avg(c):
            vals = list(tot.get(c, {}).values())
            return sum(vals) / len(vals) if vals else float("nan")
        simds = 256 * 4
        util = avg("SQ_VALU_MFMA_BUSY_CYCLES") / (avg("GRBM_GUI_ACTIVE") * simds) if avg("GRBM_GUI_ACTIVE") else 0
        lines.append(f"| {name} | {100 * util:.1f}% | {avg('SQ_INSTS_VALU') / nel:.0f} | "
                     f"{avg('SQ_INSTS_VALU_MFMA_MOPS_BF16') / nel:.3g} | {avg('SQ_INSTS_VALU_MFMA_MOPS_F64') / nel:.3g} | "
                     f"{avg('SQ_INSTS_LDS') / nel:.0f} |")
    lines += ["", "The split-BF16 form runs six v_mfma_f32_16x16x32_bf16 products per 16x16 tile (21",
              "upper-triangle tiles, one K = 32 step over the 27 points), the FP64 form one",
              "v_mfma_f64_16x16x4f64 and the FP32 form one v_mfma_f32_16x16x4f32 per tile and k-step (21",
              "tiles, 7 k-steps; the FP32 MOPs are counted by neither column).  The element kernels run one",
              "element per workgroup through barrier-separated",
              "phases (loads, geometry, gradients, Gram matrix, pair blocks, block store), so per-element",
              "latency, not the MFMA rate, sets their time; the gather moves the workspace blocks in and",
              "the CRS values out."]
    out = os.path.join(ROOT, "profiles", "q2_hex27_elastic_64_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out)


BATCH_PROF_STEPS = 5   # steps of the batch kernel-trace pass


def summarize_batch(stats_dir, bench_json):
    """Writes profiles/batch_512_kernel_stats.md from a kernel-trace pass of
    `bench.py --workload batch` and the line that workload printed."""
    import numpy as np
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    trace = _rows(stats_dir, "kernel_trace.csv")
    f32, f64 = line["forms"]["fp32_batch"], line["forms"]["fp64_batch"]
    lines = ["# Kernel statistics: 512-mesh batch (bench.py --workload batch)", "",
             "Commands (from the repository root, on one MI355X):", "",
             "    python3 bench.py --workload batch --steps 20",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <out>/prof_batch -- "
             f"python3 bench.py --workload batch --steps {BATCH_PROF_STEPS}", "",
             "Workload: " + line["config"]["workload"] + f" ({line['config']['cells']} cells, "
             f"{line['config']['elements']} elements, tets counted singly); Poisson, k lognormal (sigma 0.5),",
             "u = 0 on x = 0 per mesh; a step is one batch call (graph replay: the persistent kernel(s), then",
             "the Dirichlet pass), its argument blocks packed once.", "",
             "| form | ms per step (HIP events) | elements/s | HBM fraction (compulsory bytes) | tile | slab | grid | launches |",
             "|---|---:|---:|---:|---|---:|---:|---:|"]
    for name, fm in (("FP32 batch", f32), ("FP64 batch", f64)):
        lines.append(f"| {name} | {fm['ms_per_step']:.3f} | {fm['elements_per_s']:.3g} | {fm['hbm_frac']:.3f} | "
                     f"{fm['tile']} | {fm['slab']} | {fm['grid']} | {fm['launches']} |")
    lines += ["", f"Compulsory bytes per call: {line['compulsory_bytes'] / 1e9:.2f} GB (CRS values and f written once, "
              "x and k read).  One `txa_assemble` per mesh (FP64): "
              f"{line['forms']['fp64_per_mesh']['ms_per_step']:.2f} ms.", "",
              "Per-dispatch durations from the kernel trace (us), all dispatches of the run (the autotuner's",
              "sweeps over tile shape, slab height and grid included):", "",
              "| kernel | calls | mean | min | max |", "|---|---:|---:|---:|---:|"]
    names = sorted({r["Kernel_Name"] for r in trace if "k_q1_batch" in r["Kernel_Name"] or
                    "k_dirichlet_batch" in r["Kernel_Name"]})
    for n in names:
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace if r["Kernel_Name"] == n]
        short = n.replace("txa::(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        lines.append(f"| `{short}` | {len(d)} | {np.mean(d):.1f} | {np.min(d):.1f} | {np.max(d):.1f} |")
    # the timed FP32 graph replays: the last `steps` dispatches of the chosen shape's float kernels
    tx = int(f32["tile"].split("x")[0])
    steps = BATCH_PROF_STEPS
    lines += ["", f"Timed FP32 steps (the last {steps} dispatches of the {f32['tile']} float kernels):", ""]
    for n in names:
        if f"<{tx}, " not in n or "float" not in n:
            continue
        rs = sorted((r for r in trace if r["Kernel_Name"] == n), key=lambda r: int(r["Start_Timestamp"]))[-steps:]
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rs]
        short = n.replace("txa::(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        lines.append(f"- `{short}`: mean {np.mean(d):.1f} us over {len(d)} dispatches")
    out = os.path.join(ROOT, "profiles", "batch_512_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out)


def summarize_elastic(stats_dir, pmc_dir, bench_json):
    """Writes profiles/elastic_hex8_128_kernel_stats.md from a kernel-trace pass and one --pmc
    pass of `bench.py --workload elastic` and the line that workload printed."""
    import numpy as np
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    trace = _rows(stats_dir, "kernel_trace.csv")
    pmc_list = "SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_WAVES GRBM_GUI_ACTIVE"
    lines = ["# Kernel statistics: transient hex8 elasticity 128^3, FP64 (bench.py --workload elastic)", "",
             "Commands (from the repository root, on one MI355X):", "",
             "    python3 bench.py --workload elastic --steps 10 --warmup 2",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <out>/prof_el -- "
             "python3 bench.py --workload elastic --steps 5 --warmup 1",
             f"    rocprofv3 --pmc {pmc_list} --output-format csv -d <out>/pmc_el -- "
             "python3 bench.py --workload elastic --steps 2 --warmup 0", "",
             "Workload: " + line["config"]["workload"] + ", E lognormal (sigma 0.3) per cell, Rayleigh damping,",
             "Newmark seeds (alpha 2000, gamma2 4e6), u = 0 on x = 0; a step is one assembly with the",
             "Dirichlet rows.  Parallelepiped cells (the unit cube) take the node-centric kernel",
             "`k_elastic8_affine` (closed-form blocks, no workspace); jittered cells take the element",
             "kernel `k_elastic<8, 8, 64, 0>` and the gather `k_elastic_gather<8, double>` in 8 z slabs on two",
             "streams (8 dispatches of each per step, the gather of slab k beside the element kernel of slab",
             "k+1); the uploaded mesh runs the element kernel once and `k_elastic_gather_u` once per step.", "",
             "| cells | ms per step (HIP events) | p50 / p99 ms | elements/s | HBM fraction (compulsory bytes) |",
             "|---|---:|---:|---:|---:|"]
    for name, c in line["cells"].items():
        if "p50" not in c:
            continue
        lines.append(f"| {name.replace('_', ' ')} | {c['ms_per_step']:.3f} | {c['p50']:.3f} / {c['p99']:.3f} | "
                     f"{c['elements_per_s']:.3g} | {c['hbm_frac']:.3f} |")
    for order, c in line["cells"].get("jittered_cells_uploaded", {}).items():
        lines.append(f"| jittered cells uploaded (`txa_mesh_upload`), {order.replace('_', ' ')} | "
                     f"{c['ms_per_step']:.3f} | | {c['elements_per_s']:.3g} | |")
    up = line["cells"].get("jittered_cells_uploaded", {}).get("brick_order", {}).get("vs_brick_normwise")
    if up:
        lines += ["", "The uploaded mesh (element kernel, then `k_elastic_gather_u`: each node's element instances",
                  "from its list, slot maps per instance) in brick order against the brick's lattice gather: "
                  f"R {up['R']:.1e}, J {up['J']:.1e} normwise."]
    lines += ["", "Per-dispatch durations from the kernel trace (us), all dispatches of the run:", "",
              "| kernel | calls | mean | min | max |", "|---|---:|---:|---:|---:|"]
    for k in ("k_elastic8_affine", "k_elastic<8, 8, 64, 0>", "k_elastic_gather<8, double>",
              "k_elastic_gather_u<8, double>", "k_dirichlet"):
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace
             if (k in r["Kernel_Name"] if "<" not in k else _match(r["Kernel_Name"], k))]
        if d:
            lines.append(f"| `{k}` | {len(d)} | {np.mean(d):.1f} | {np.min(d):.1f} | {np.max(d):.1f} |")
    tot = {}
    for r in _rows(pmc_dir, "counter_collection.csv"):
        if "k_elastic8_affine" in r.get("Kernel_Name", ""):
            tot.setdefault(r["Counter_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
            tot[r["Counter_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    if tot:
        avg = {c: sum(v.values()) / len(v) for c, v in tot.items()}
        nodes = 129 ** 3
        lines += ["", "## `k_elastic8_affine` counters (per dispatch, per lattice node)", "",
                  f"- VALU wave-instructions per node: {avg.get('SQ_INSTS_VALU', 0) / nodes:.0f}; LDS: "
                  f"{avg.get('SQ_INSTS_LDS', 0) / nodes:.0f}; SALU: {avg.get('SQ_INSTS_SALU', 0) / nodes:.0f}; "
                  f"vector memory reads: {avg.get('SQ_INSTS_VMEM_RD', 0) / nodes:.0f}",
                  f"- waves per dispatch: {avg.get('SQ_WAVES', 0):.0f}; GRBM_GUI_ACTIVE (summed over the 8 XCDs): "
                  f"{avg.get('GRBM_GUI_ACTIVE', 0):.3g}",
                  "- one wave per node: 8 cells x 8 vertices on the 64 lanes, a 3x3 block per lane from the",
                  "  cell's scaled cofactors and |det J| (a per-mesh table, `k_cell_geom`): Chat S Chat^T, the",
                  "  Newmark/Rayleigh combination, eight cell-ordered LDS adds, rows stored non-temporally.",
                  "  Timing probes (work skipped) put the LDS adds and the state loads' dependence on the",
                  "  LIDs at 1% each and the row stores at about 0.3 ms: the VALU work and its dependences",
                  "  (four waves per SIMD at 125 VGPRs) set the rest."]
    out = os.path.join(ROOT, "profiles", "elastic_hex8_128_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out)


def summarize_upload(stats_dir, fetch_dir, write_dir, util_dir, bench_json):
    """Writes profiles/upload_hex8_128_kernel_stats.md from a kernel-trace pass, the FETCH_SIZE,
    WRITE_SIZE and issue-counter passes of `bench.py --workload upload`, and the line it printed."""
    import numpy as np
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    trace = sorted(_rows(stats_dir, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    n = int(round(line["config"]["elements"] ** (1 / 3)))
    rows_, nnz = (n + 1) ** 3, (3 * n + 1) ** 3   # a vertex row couples 2 or 3 vertices along each axis
    ne = n ** 3
    alg = 8 * nnz + 8 * rows_ + 8 * rows_ + 24 * rows_ + 8 * ne + 32 * ne + 8 * rows_
    util = ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY "
            "SQ_ACTIVE_INST_VALU SQ_INSTS_VALU GRBM_GUI_ACTIVE")
    cmd = "    rocprofv3 --pmc {} --output-format csv -d <out>/{} -o up -- python3 bench.py --workload upload " \
          "--steps 2 --warmup 0 --no-cpu --no-extra"
    o = line["orders"]
    pt = o["brick_order"]["patches"]
    lines = [f"# Kernel statistics: uploaded hex8 Poisson brick {n}^3, node patches (bench.py --workload upload)", "",
             "Commands (from the repository root, on one MI355X):", "",
             "    python3 bench.py --workload upload --steps 20 --warmup 3",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <out>/prof_up -o up -- python3 bench.py "
             "--workload upload --steps 5 --warmup 1 --no-cpu --no-extra",
             cmd.format("FETCH_SIZE", "pmc_f"), cmd.format("WRITE_SIZE", "pmc_w"), cmd.format(util, "pmc_u"), "",
             f"Workload: {line['config']['workload']} ({ne} elements, {rows_} rows, {nnz} nnz), steady Poisson,",
             "k lognormal sigma 0.5 per element, u = 0 on x = 0; jittered cells (vertices moved by",
             "U[-0.1h, 0.1h]) in brick order, then in a seeded random order (the LIDs, numbered by first",
             "touch, follow the element order), then unjittered cells in brick order (every element a",
             "parallelepiped: the one-metric element form, decided once per mesh).  The assembly is",
             f"one launch of `k_patch` (one workgroup per patch of {pt['rows_per_patch']} rows along the Morton "
             "curve) and the", "Dirichlet rows.", "",
             f"Patches: {pt['patches']}, element instances {pt['instances']} ({pt['instances'] / ne:.2f} per element: "
             f"an element is integrated once per patch it touches), {pt['colours']} colours, rows of at most "
             f"{pt['longest_row']} entries.", "",
             "| order | form | ms per step (HIP events) | p50 / p99 ms | elements/s |", "|---|---|---:|---:|---:|"]
    for name in [k for k in ("brick_order", "shuffled", "parallelepiped_cells_brick_order") if k in o]:
        c = o[name]
        lines.append(f"| {name.replace('_', ' ')} | FP64 | {c['ms_per_step']:.3f} | {c['p50']:.3f} / {c['p99']:.3f} | "
                     f"{c['elements_per_s']:.3g} |")
        lines.append(f"| {name.replace('_', ' ')} | FP32 | {c['fp32']['ms_per_step']:.3f} | | "
                     f"{ne / c['fp32']['ms_per_step'] * 1e3:.3g} |")
    lines += ["", "Per-dispatch durations from the kernel trace (us); each case runs the FP64 form (1 + 5",
              "dispatches), then the FP32 form (1 + 5):", "",
              "| kernel | order | calls | mean | min | max |", "|---|---|---:|---:|---:|---:|"]
    for k in ("k_patch<1, double, false, true, false>", "k_patch<1, float, false, true, false>"):
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace if _match(r["Kernel_Name"], k)]
        h = len(d) // 2
        for name, part in (("brick order", d[:h]), ("shuffled", d[h:])):
            if part:
                lines.append(f"| `{k}` | {name} | {len(part)} | {np.mean(part):.1f} | {np.min(part):.1f} | "
                             f"{np.max(part):.1f} |")
    for k in ("k_patch<1, double, false, true, true>", "k_patch<1, float, false, true, true>"):
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace if _match(r["Kernel_Name"], k)]
        if d:
            lines.append(f"| `{k}` | parallelepiped cells, brick order | {len(d)} | {np.mean(d):.1f} | "
                         f"{np.min(d):.1f} | {np.max(d):.1f} |")
    d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace if "k_state_by_pos" in r["Kernel_Name"]]
    if d:
        lines.append(f"| `k_state_by_pos` (the state in curve order, before each assembly) | shuffled | {len(d)} | "
                     f"{np.mean(d):.1f} | {np.min(d):.1f} | {np.max(d):.1f} |")
    lines += ["", "The brick in its own order reads the state by LID (its rows touch 0.19 lines of 128 bytes",
              "per row along the curve); with shuffled elements the LIDs scatter (0.41 lines per row) and",
              "the state is gathered into curve order first (`txa_mesh_patches` reports the choice)."]
    kern = "k_patch<1, double, false, true, false>"

    def per_dispatch(d, counter):
        tot = {}
        for r in _rows(d, "counter_collection.csv"):
            if r["Counter_Name"] == counter and _match(r["Kernel_Name"], kern):
                tot[int(r["Dispatch_Id"])] = tot.get(int(r["Dispatch_Id"]), 0.0) + float(r["Counter_Value"])
        return [tot[k] for k in sorted(tot)]
    f_, w_ = per_dispatch(fetch_dir, "FETCH_SIZE"), per_dispatch(write_dir, "WRITE_SIZE")
    brick_us = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace if _match(r["Kernel_Name"], kern)]
    brick_us = float(np.mean(brick_us[1:len(brick_us) // 2])) if len(brick_us) >= 4 else float("nan")
    if f_ and w_:
        h = len(f_) // 2
        rd = [2.0 * v * 1024 / 1e9 for v in f_]   # KiB, times the copy-kernel calibration factors
        wr = [1.0 * v * 1024 / 1e9 for v in w_]
        lines += ["", f"## HBM traffic of `{kern}` per dispatch", "",
                  "FETCH_SIZE and WRITE_SIZE (KiB) scaled by the factors a copy kernel of known size gives on",
                  "this GPU (fetch 2.000, write 1.000; `q1_hex8_diffusion_fp32_256_kernel_stats.md`):", "",
                  "| order | HBM read GB | HBM written GB |", "|---|---:|---:|",
                  f"| brick order | {np.mean(rd[:h]):.3f} | {np.mean(wr[:h]):.3f} |",
                  f"| shuffled | {np.mean(rd[h:]):.3f} | {np.mean(wr[h:]):.3f} |", "",
                  f"Algorithmic bytes per launch: {alg / 1e9:.3f} GB (values 8 B x nnz, f, x, vertex coordinates",
                  "24 B, rowptr per row; coefficient and 8 LIDs per element).  At the brick-order kernel mean of",
                  f"{brick_us:.0f} us (warm dispatches) that is {alg / (brick_us * 1e-6) / 1e12:.2f} TB/s, "
                  f"{alg / (brick_us * 1e-6) / 8e12:.3f} of 8 TB/s; the actual traffic "
                  f"({np.mean(rd[:h]) + np.mean(wr[:h]):.2f} GB) is about "
                  f"{(np.mean(rd[:h]) + np.mean(wr[:h])) / (alg / 1e9):.1f}x the algorithmic bytes (the instance "
                  "records, 96 B per", "instance, and the row position of each vertex record)."]
    u = {c: per_dispatch(util_dir, c) for c in ("SQ_INSTS_VALU", "GRBM_GUI_ACTIVE", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY",
                                               "SQ_ACTIVE_INST_ANY", "SQ_WAVES")}
    if all(u.values()):
        h = len(u["SQ_INSTS_VALU"]) // 2
        m = {c: float(np.mean(v[:h])) for c, v in u.items()}
        lines += ["", f"## Issue counters of `{kern}`, brick order (per dispatch)", "",
                  f"- waves {m['SQ_WAVES']:.0f}; VALU wave-instructions {m['SQ_INSTS_VALU']:.4g} "
                  f"({m['SQ_INSTS_VALU'] / pt['instances'] * 64:.0f} per instance);",
                  f"- VALU busy 4 x SQ_INSTS_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8) = "
                  f"{4 * m['SQ_INSTS_VALU'] / (1024 * m['GRBM_GUI_ACTIVE'] / 8):.3f};",
                  f"- SQ_ACTIVE_INST_ANY / SQ_WAVE_CYCLES = {m['SQ_ACTIVE_INST_ANY'] / m['SQ_WAVE_CYCLES']:.3f}, "
                  f"SQ_WAIT_ANY / SQ_WAVE_CYCLES = {m['SQ_WAIT_ANY'] / m['SQ_WAVE_CYCLES']:.3f}.", "",
                  "Two waves per SIMD (the FP64 hex8 element routine holds about 210 VGPRs; two 61 KB patches",
                  "per CU) leave the waits of the gathers, the colour rounds' barriers and the row stores",
                  "mostly exposed: the kernel is bound by latency and issue, not by HBM."]
    out = os.path.join(ROOT, "profiles", "upload_hex8_128_kernel_stats.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(out)


def other_workloads(args):
    """Compact summaries of the Q2, batch, serving, hex8 elasticity and uploaded-mesh workloads at
    their own default sizes, appended to the headline line (each run after the headline's buffers
    are freed; a failure is recorded in place of its summary and leaves the headline intact)."""
    import copy
    import torch

    def pick(d, keys):
        return {k: d[k] for k in keys if k in d}

    out = {}
    for name, fn, over in (("q2_elasticity_64", bench_q2, dict(n=64, steps=5, warmup=1)),
                           ("batch_512_meshes", bench_batch, dict(meshes=512, steps=5)),
                           ("serve_hex8_newmark", bench_serve, dict(requests=48, pool=8)),
                           ("elastic_hex8_128", bench_elastic, dict(n=128, steps=5, warmup=1)),
                           ("upload_hex8_128", bench_upload, dict(n=128, steps=10, warmup=2))):
        a = copy.copy(args)
        for k, v in over.items():
            setattr(a, k, v)
        t0 = time.time()
        try:
            line = fn(a)
        except Exception as e:   # keep the headline: report what failed
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name}: failed: {e}")
            continue
        finally:
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
        s = pick(line, ("metric", "value", "unit", "ms_per_step", "dtype"))
        s["config"] = line["config"]["workload"]
        if fn is bench_q2:
            f = line["forms"]
            s["forms_ms"] = {k: v["ms_per_step"] for k, v in f.items()}
            s["max_rel_err_vs_fp64"] = line["max_rel_err_vs_fp64"]
            s["rigid_motion_residual_rel"] = line["rigid_motion_residual_rel"]
            s["roofline"] = {"bound": "hbm", "frac": f["split_bf16"]["hbm_frac"],
                             "bytes_per_element": Q2_BYTES_PER_ELEM,
                             "dense_bf16_tflops": f["split_bf16"]["dense_tflops"],
                             "mfma_frac": f["split_bf16"]["dense_tflops"] / 2500.0}
        elif fn is bench_batch:
            f = line["forms"]
            s["forms_ms"] = {k: v["ms_per_step"] for k, v in f.items()}
            s["cells_per_s"] = f["fp32_batch"]["cells_per_s"]
            s["roofline"] = {"bound": "hbm", "frac": f["fp32_batch"]["hbm_frac"],
                             "compulsory_bytes": line["compulsory_bytes"]}
            s["fp32_batch_shape"] = {k: f["fp32_batch"][k] for k in ("tile", "slab", "grid", "launches")}
            s["max_rel_err_fp32_vs_fp64_per_mesh"] = line["max_rel_err_fp32_vs_fp64_per_mesh"]
            s["speedup_vs_fp64_per_mesh"] = line["speedup_vs_fp64_per_mesh"]
        elif fn is bench_upload:
            s["orders"] = {k: {"ms_per_step": v["ms_per_step"], "fp32_ms_per_step": v["fp32"]["ms_per_step"],
                               "patches": v["patches"]} for k, v in line["orders"].items()}
            s["checks"] = line["checks"]
        elif fn is bench_elastic:
            s["cells"] = {k: ({"ms_per_step": v["ms_per_step"], "hbm_frac": v["hbm_frac"]} if "hbm_frac" in v
                              else {o: c["ms_per_step"] for o, c in v.items()})
                          for k, v in line["cells"].items()}
            s["roofline"] = line["roofline"]
        else:
            s["plain_fp64_elements_per_s"] = line["plain_fp64_sequential"]["elements_per_s"]
            s["forms"] = {k: {"p50_ms": v["latency_ms"]["p50"], "p99_ms": v["latency_ms"]["p99"],
                              "saturated_elements_per_s": v["saturated"]["elements_per_s"],
                              "speedup_vs_plain_fp64": v["speedup_vs_plain_fp64"],
                              "max_rel_err_vs_plain_fp64": v["max_rel_err_vs_plain_fp64"]}
                          for k, v in line["forms"].items()}
            s["latency_ms"] = pick(line["latency_ms"], ("p50", "p99", "at_qps"))
        s["wall_s"] = time.time() - t0
        out[name] = s
        log(f"{name}: {s.get('ms_per_step', s.get('latency_ms'))} ({s['wall_s']:.1f}s)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only: skip the Q2 / batch / serving summaries appended to the default line")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the measurement, run a 1 GiB -> 1 GiB elementwise kernel (dst = 2 src) three "
                         "times: the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--summarize-profiles", nargs=3, metavar=("STATS_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="CPU only: turn rocprofv3 csv output (kernel stats, FETCH_SIZE pass, WRITE_SIZE pass) "
                         "into the summaries under profiles/")
    ap.add_argument("--workload", choices=("single", "batch", "serve", "q2", "upload", "elastic"), default="single",
                    help="single: the 256^3 Q1 line (default); batch / serve / q2: the other forms; "
                         "upload: a 128^3 hex8 brick through txa_mesh_upload; elastic: 128^3 transient hex8 "
                         "elasticity")
    ap.add_argument("--meshes", type=int, default=512)
    ap.add_argument("--pool", type=int, default=8)
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--qps", type=float, default=20.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-precision", choices=("fp64", "fp32"), default="fp64",
                    help="the served form whose numbers head the serve line (both are measured)")
    ap.add_argument("--summarize-q2", nargs=2, metavar=("STATS_DIR", "PMC_DIR"),
                    help="CPU only: summarise rocprofv3 csv output of --workload q2 into profiles/")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("UTIL_A_DIR", "UTIL_B_DIR"),
                    help="CPU only: summarise the two issue-counter passes of the headline into profiles/")
    ap.add_argument("--summarize-elastic", nargs=3, metavar=("STATS_DIR", "PMC_DIR", "BENCH_JSON"),
                    help="CPU only: summarise rocprofv3 csv output of --workload elastic into profiles/")
    ap.add_argument("--summarize-batch", nargs=2, metavar=("STATS_DIR", "BENCH_JSON"),
                    help="CPU only: summarise rocprofv3 csv output of --workload batch into profiles/")
    ap.add_argument("--summarize-upload", nargs=5,
                    metavar=("STATS_DIR", "FETCH_DIR", "WRITE_DIR", "UTIL_DIR", "BENCH_JSON"),
                    help="CPU only: summarise rocprofv3 csv output of --workload upload into profiles/")
    args = ap.parse_args()
    if args.summarize_upload:
        summarize_upload(*args.summarize_upload)
        return
    if args.summarize_batch:
        summarize_batch(*args.summarize_batch)
        return
    if args.summarize_q2:
        summarize_q2(*args.summarize_q2)
        return
    if args.summarize_elastic:
        summarize_elastic(*args.summarize_elastic)
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters)
        return
    if args.summarize_profiles:
        summarize_profiles(*args.summarize_profiles)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)
    other = {"batch": bench_batch, "serve": bench_serve, "q2": bench_q2, "upload": bench_upload,
             "elastic": bench_elastic}
    if args.workload in other:
        print(json.dumps(other[args.workload](args)), flush=True)
        return

    import numpy as np
    import torch

    import tianxin_amd as txa
    from tianxin_amd import synth

    n = args.n
    nel = n ** 3
    t0 = time.time()
    X = synth.brick_coords(n, n, n)
    eng = txa.Engine(0)
    m = eng.brick(txa.HEX8, n, n, n, X)
    ndofs = m.build_dofmap(1)
    nnz = m.build_graph()
    pl = m.point_lids()[:, 0]
    ii = np.arange(pl.size) % (n + 1)
    bc = torch.as_tensor(pl[ii == 0].astype(np.int32), device="cuda")
    k = torch.as_tensor(synth.lognormal(SEED_K, nel, 0.5), device="cuda")
    x = torch.as_tensor(synth.uniform(SEED_U, ndofs, -1, 1), device="cuda")
    f64 = torch.empty(ndofs, device="cuda", dtype=torch.float64)
    v64 = torch.empty(nnz, device="cuda", dtype=torch.float64)
    f32 = torch.empty_like(f64)
    v32 = torch.empty_like(v64)
    torch.cuda.synchronize()
    log(f"setup {time.time() - t0:.1f}s: {nel} elements, {ndofs} dofs, {nnz} nnz")

    def step(prec, f, v, with_bc=True):
        m.assemble(physics=txa.DIFFUSION, precision=prec, coef=k, x=x, f=f, vals=v, src=(1.0,),
                   bc_lids=bc if with_bc else None)

    def timed(prec, f, v, steps, warmup, with_bc=True):
        for _ in range(warmup):
            step(prec, f, v, with_bc)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        t_wall = time.perf_counter()
        ev[0].record()
        for i in range(steps):
            step(prec, f, v, with_bc)
            ev[i + 1].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        per = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)])
        return wall, per

    # headline: FP32 element integration (timed region = exactly `steps` steps)
    wall32, per32 = timed(txa.FP32, f32, v32, args.steps, args.warmup)
    log(f"fp32 form: {per32.mean():.3f} ms/step")
    # the plain FP64 path beside it
    wall64, per64 = timed(txa.FP64, f64, v64, args.steps, args.warmup)
    log(f"fp64 path: {per64.mean():.3f} ms/step")
    # volume kernel alone (no Dirichlet pass) for the roofline figure
    _, kper32 = timed(txa.FP32, f32, v32, args.steps, 1, with_bc=False)
    _, kper64 = timed(txa.FP64, f64, v64, args.steps, 1, with_bc=False)
    step(txa.FP32, f32, v32)
    step(txa.FP64, f64, v64)
    torch.cuda.synchronize()

    def rel(a, b):
        return float((a - b).abs().max() / b.abs().max())

    def ent(a, b, tau=1e-3):
        mask = b.abs() > tau * b.abs().max()
        return float(((a - b).abs()[mask] / b.abs()[mask]).max())

    err = {"R_normwise": rel(f32, f64), "J_normwise": rel(v32, v64),
           # entrywise over |ref| > 1e-12 max|ref| (the stated measure), and over 1e-3 max|ref|
           "R_entrywise_over_1e-12_max": ent(f32, f64, 1e-12), "J_entrywise_over_1e-12_max": ent(v32, v64, 1e-12),
           "R_entrywise_over_1e-3_max": ent(f32, f64), "J_entrywise_over_1e-3_max": ent(v32, v64)}
    # size-independent property: each interior Jacobian row sums to zero (k grad.grad form)
    rp = m.graph(with_cols=False)[0]
    rows = torch.repeat_interleave(torch.arange(ndofs, device="cuda"),
                                   torch.as_tensor(np.diff(rp), device="cuda"))
    rs = torch.zeros(ndofs, device="cuda", dtype=torch.float64).index_add_(0, rows, v64)
    ra = torch.zeros(ndofs, device="cuda", dtype=torch.float64).index_add_(0, rows, v64.abs())
    isbc = torch.zeros(ndofs, device="cuda", dtype=torch.bool)
    isbc[bc.long()] = True
    zero_rows = int((ra == 0).sum())
    rowsum = float((rs.abs() / ra.clamp_min(1e-300))[~isbc].max())
    del rows, rs, ra

    ms = float(per32.mean())
    value = nel / (ms / 1e3)
    kt32 = float(kper32.mean())
    general = None
    if not args.no_extra:
        # the same workload on jittered cells (interior vertices moved by U[-0.1h, 0.1h]): the
        # general kernel, one Jacobian per integration point
        Xj = synth.brick_coords(n, n, n, jitter=0.1, seed=7)
        mj = eng.brick(txa.HEX8, n, n, n, Xj)
        mj.build_dofmap(1)
        mj.build_graph()
        bcj = torch.as_tensor(mj.point_lids()[:, 0][ii == 0].astype(np.int32), device="cuda")

        def step_j(prec, f, v, with_bc=True):
            mj.assemble(physics=txa.DIFFUSION, precision=prec, coef=k, x=x, f=f, vals=v, src=(1.0,),
                        bc_lids=bcj if with_bc else None)
        gen = {}
        for name, prec, f_, v_ in (("fp32", txa.FP32, f32, v32), ("fp64", txa.FP64, f64, v64)):
            for _ in range(args.warmup):
                step_j(prec, f_, v_)
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
            ev[0].record()
            for i in range(args.steps):
                step_j(prec, f_, v_)
                ev[i + 1].record()
            torch.cuda.synchronize()
            per = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])
            gen[name] = {"ms_per_step": float(per.mean()), "elements_per_s": nel / (float(per.mean()) / 1e3),
                         "frac_whole_step": nel / (float(per.mean()) / 1e3) * BYTES_PER_ELEM / HBM_PEAK}
        step_j(txa.FP32, f32, v32)
        step_j(txa.FP64, f64, v64)
        torch.cuda.synchronize()
        gen["max_rel_err_fp32_vs_fp64"] = {
            "R_normwise": rel(f32, f64), "J_normwise": rel(v32, v64),
            "R_entrywise_over_1e-12_max": ent(f32, f64, 1e-12), "J_entrywise_over_1e-12_max": ent(v32, v64, 1e-12),
            "J_entrywise_over_1e-2_max": ent(v32, v64, 1e-2)}
        general = {"workload": f"Q1 Poisson hex8 {n}^3, interior vertices jittered by U[-0.1h, 0.1h] "
                               "(general kernel: one Jacobian per integration point)",
                   "kernel": FP32_KERNEL_GENERAL, "mesh_affine": mj.affine, **gen}
        log(f"jittered cells: fp32 {gen['fp32']['ms_per_step']:.3f} ms, fp64 {gen['fp64']['ms_per_step']:.3f} ms")
        mj.destroy()
        del bcj
    copy_gbs, fill_gbs = measured_hbm_peaks()
    achieved = BYTES_PER_ELEM * nel / (kt32 / 1e3) / 1e9
    traffic = None
    if n == 256 and os.path.exists(TRAFFIC_FILE):
        traffic = json.load(open(TRAFFIC_FILE)).get("bytes_per_launch")
    out = {
        "metric": "elements/s (residual+Jacobian) vs plain FP64 at stated error; p50/p99 latency",
        "value": value, "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 element integration, f64 geometry and scatter", "data": "synthetic",
        "config": {"workload": f"Q1 Poisson hex8 {n}^3 unit cube, precision form (FP32 integration)",
                   "elements": nel, "dofs": ndofs, "nnz": nnz, "quadrature": "2x2x2 Gauss",
                   "k": "lognormal sigma=0.5", "dirichlet": "u=0 on x=0"},
        "latency_ms": {"p50": float(np.percentile(per32, 50)), "p99": float(np.percentile(per32, 99))},
        "fp64_path": {"elements_per_s": nel / (float(per64.mean()) / 1e3), "ms_per_step": float(per64.mean()),
                      "p50": float(np.percentile(per64, 50)), "p99": float(np.percentile(per64, 99)),
                      "kernel_ms": float(kper64.mean()),
                      "roofline_frac": BYTES_PER_ELEM * nel / (float(kper64.mean()) / 1e3) / HBM_PEAK},
        "max_rel_err_vs_fp64": err,
        "check": {"interior_row_sum_rel_max": rowsum, "rows_all_zero": zero_rows},
        "mesh_affine": m.affine,
        "jittered_cells": general,
        "wall_s_timed": wall32,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": traffic,
                     "kernel": FP32_KERNEL, "kernel_ms": kt32,
                     # the same bytes over the whole step (volume kernel + Dirichlet rows): value x 408 B
                     "frac_whole_step": value * BYTES_PER_ELEM / HBM_PEAK,
                     "bytes_per_element": BYTES_PER_ELEM,
                     # measured on this GPU: device-to-device copy (read + write bytes) and fill (write bytes)
                     "measured_copy_GBs": copy_gbs, "measured_fill_GBs": fill_gbs,
                     "traffic_GBs": traffic / (kt32 / 1e3) / 1e9 if traffic else None,
                     "frac_of_measured_copy": achieved / copy_gbs},
    }
    if args.copy_calibration:
        del rp
        src = torch.empty(COPY_BYTES // 8, device="cuda", dtype=torch.float64).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        log("copy calibration done")
        del src, dst
    if not args.no_extra:
        m.destroy()
        del f64, v64, f32, v32, x, k, bc, isbc
        torch.cuda.empty_cache()
        out["other_workloads"] = other_workloads(args)
    if not args.no_cpu:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        log(f"cpu baseline {time.time() - t0:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
