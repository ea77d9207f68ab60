"""Benchmark: Q1 Poisson residual + Jacobian on a 256^3 hex8 cube (precision form).

Workload: 3D scalar Poisson on the unit cube, n^3 hex8 cells (default n=256:
16,777,216 elements, 16,974,593 DOFs, 454,756,609 CRS nonzeros), 2x2x2 Gauss,
source f = 1, k = exp(g) with g ~ N(0, 0.5^2) iid per element, u ~ U[-1,1],
Dirichlet u = 0 on x = 0.  One step = one residual + Jacobian assembly of the
whole mesh (volume kernel + Dirichlet rows) with inputs resident in HBM.

The timed form integrates elements in FP32 (geometry and the scatter
accumulation in FP64); the plain FP64 GPU path runs beside it and the max
relative error of FP32 vs FP64 is reported for R and J.  Prints one JSON line.
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


def cpu_baseline(n=64, reps=10):
    """Oracle (reference algorithm: cached worksets, forward-AD Jacobian, row-search
    scatter) on host cores, on an n^3 sample of the same workload."""
    import numpy as np

    import oracle as O
    from tianxin_amd import synth

    cores = min(16, os.cpu_count() or 1)
    conn, ev, _ = O.brick_mesh(O.HEX8, n, n, n)
    _, owned, el = O.dofmap(conn, 1)
    rp, col = O.graph(el, owned.size)
    X = synth.brick_coords(n, n, n)
    k = synth.lognormal(SEED_K, n ** 3, 0.5)
    x = synth.uniform(SEED_U, owned.size, -1, 1)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, k, rp, col, src=(1.0,), workset_size=1000, nthreads=cores)
    for _ in range(2):
        A.evaluate(x, mode=O.MODE_FAD)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        A.evaluate(x, mode=O.MODE_FAD)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return {"value": n ** 3 / med, "unit": "elements/s", "cores": cores, "kind": "port",
            "sample": f"Q1 Poisson {n}^3 hex8 ({n ** 3} elements), FP64, forward-AD Jacobian, workset 1000, "
                      f"median of {reps} after 2 warm-ups, {cores} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

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

    def ent(a, b, tau=1e-12):
        mask = b.abs() > tau * b.abs().max()
        return float(((a - b).abs()[mask] / b.abs()[mask]).max())

    err = {"R_normwise": rel(f32, f64), "J_normwise": rel(v32, v64),
           "R_entrywise": ent(f32, f64), "J_entrywise": ent(v32, v64)}
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
    achieved = BYTES_PER_ELEM * nel / (kt32 / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "q1_hex8_diffusion_fp32_256_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
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
        "wall_s_timed": wall32,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": traffic,
                     "kernel": "k_q1_hex8_diffusion<float>", "kernel_ms": kt32,
                     "bytes_per_element": BYTES_PER_ELEM},
    }
    if not args.no_cpu:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        log(f"cpu baseline {time.time() - t0:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
