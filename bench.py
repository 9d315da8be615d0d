#!/usr/bin/env python
"""Throughput of the SOD2D compressible SEM right-hand side + Runge-Kutta stage
on one MI355X: GDOF/s per RK stage on the periodic Taylor-Green vortex,
p = 3 (HEX64), 64^3 elements (192^3 = 7,077,888 unique nodes), FP64.

One bench step = one full time step of the reference's explicit integrator
(rk_4_ls_main, SSP(5,3) 2N low-storage, reference sequencing: 4 fresh RHS
evaluations per step) including the per-step entropy-viscosity work (EOS,
entropy, entropy residual, sensor, Sutherland).  All inputs are device resident
when the timed region starts.

  value = 5 * N_unique * rhs_per_step / t_step / 1e9   [GDOF/s per RK stage]

The JSON line also carries:
  roofline     -- the fused RK stage (element kernel + gather/update kernel):
                  algorithmic bytes of one stage B = 128 N + (11 s + 4) EN (s = 8)
                  over the summed average durations of the two kernels, measured
                  with HIP events on the launch stream over up to 3 steps run
                  right after the timed ones (the timed steps carry no events);
                  HBM traffic per stage from the committed rocprofv3 PMC summary
  cpu_baseline -- the CPU restatement of updateF (oracle, OpenMP, all host cores)
                  on a bounded sample at 16^3 elements, same GDOF/s definition.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--precision 8|4]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "GDOF/s per RK stage (TGV p=3, 64^3 elems) + %HBM roofline; surrogate/plain"
HBM_PEAK_GBS = 8000.0


def stage_bytes(N: int, EN: int, s_op: int, s_state: int = 8) -> int:
    """Algorithmic bytes of one RK stage: state in, base state in, state out,
    invMl (5 s_state N each, 8 N), and He + gpvol + mu_e (11 s_op EN) + connec (4 EN)."""
    return 15 * s_state * N + 8 * N + (11 * s_op + 4) * EN


def cpu_baseline(budget_s: float = 15.0) -> dict:
    """updateF of the oracle (CPU restatement, FP64) on all host cores, 16^3 p=3."""
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import oracle as O
    from sod2d_amd import case

    p, n = 3, 16
    g = case.build_case(p, n, periodic=True, lengths=(2 * math.pi,) * 3)
    t = O.tgv_params()
    ic = O.tgv_ic(g.mesh.coord, t)
    run = O.TGVOracle(p, g, ic, t)
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "64")))
    T = run.T

    def rhs():
        lw = run.lpoin_w
        O.eos_stage(lw, t.gamma_gas, t.Rgas, run.rho[1], run.q[1], run.E[1], run.u[1],
                    run.e_int[1], run.pr[1], run.aux_h, run.Tem[1])
        O.sutherland(run.Tem[1], run.mu_factor, run.mu_fluid)
        O.full_convec_ijk(T, run.connec, g.He, g.gpvol, run.u[1], run.q[1], run.rho[1],
                          run.pr[1], run.aux_h, run.Rmass, run.Rmom, run.Rener, True, -1.0,
                          nthreads=cores)
        O.full_diffusion_ijk(T, run.connec, g.He, g.gpvol, t.Cp, t.Prt, run.rho[1], run.rho[1],
                             run.u[1], run.Tem[1], run.mu_fluid, run.mu_e, run.mu_sgs, run.Rmass,
                             run.Rmom, run.Rener, False, -1.0, nthreads=cores)
        O.lumped_solver_scal(lw, g.invMl, run.Rmass)
        O.lumped_solver_scal(lw, g.invMl, run.Rener)
        O.lumped_solver_vect(lw, g.invMl, run.Rmom)
        O.rk_ls_update(lw, 0.0, 1.0, 0.0, run.rho[0], run.q[0], run.E[0], run.rho[1], run.q[1],
                       run.E[1], run.Rmass, run.Rmom, run.Rener)

    rhs()
    t0 = time.perf_counter()
    rhs()
    t1 = time.perf_counter() - t0
    reps = max(1, min(200, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(reps):
        rhs()
    dt = (time.perf_counter() - t0) / reps
    N = g.mesh.npoin_w
    return {"value": round(5 * N / dt / 1e9, 6), "unit": "GDOF/s per RK stage", "cores": cores,
            "kind": "port",
            "sample": f"updateF + stage combine of the CPU oracle, TGV p=3 16^3 elements "
                      f"({N} unique nodes), FP64, {reps} RHS evaluations, "
                      f"{dt * 1e3:.2f} ms each, OpenMP element loop on {cores} threads"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="elements per direction")
    ap.add_argument("--porder", type=int, default=3)
    ap.add_argument("--precision", type=int, default=8, choices=(8, 4),
                    help="operator precision: 8 = FP64 throughout, 4 = FP32 operators")
    ap.add_argument("--clean", action="store_true",
                    help="fresh RHS at every stage instead of the reference's sequencing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run 3 plain FP64 copy kernels (y = 1.0 * x) of "
                         "MIB MiB each: the reference launches for calibrating HBM byte counters")
    a = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus} GPUs, "
              f"WORLD_SIZE={world}); the single-GPU figure is the whole-job value.",
              file=sys.stderr)
        return 2

    import torch
    sys.path.insert(0, ROOT)
    from sod2d_amd.solver import TGVRun

    torch.cuda.set_device(0)
    t_setup = time.perf_counter()
    run = TGVRun(porder=a.porder, nelem_dir=a.n, op_precision=a.precision, parity=not a.clean,
                 dt_mode="fixed")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    for _ in range(a.warmup):
        run.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per-kernel breakdown from further steps bracketed by HIP events on the
    # library's stream (kept out of the timed steps)
    prof_steps = max(1, min(a.steps, 3))
    run.ctx.prof_reset()
    run.ctx.prof_enable(True)
    for _ in range(prof_steps):
        run.step()
    torch.cuda.synchronize()
    prof = {k: run.ctx.prof_get(k) for k in range(4)}
    run.ctx.prof_enable(False)
    diag = run.diagnostics()
    if a.copy_probe > 0:
        n = a.copy_probe * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y

    t_step = (t1 - t0) / a.steps
    N, EN = run.npoin_w, run.nelem * run.nnode
    nrhs = run.rhs_per_step
    value = 5 * N * nrhs / t_step / 1e9
    k_ms, k_n = prof[0]
    g_ms, g_n = prof[1]
    t_elem = k_ms / max(k_n, 1)
    t_gath = g_ms / max(g_n, 1)
    t_stage = t_elem + t_gath
    B = stage_bytes(N, EN, 8 if a.precision == 8 else 4)
    achieved = B / (t_stage * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_rk_stage_tgv{a.n}_p{a.porder}_"
                      f"{'fp64' if a.precision == 8 else 'fp32op'}.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("bytes_per_stage")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == 8 else "f32-operators/f64-state",
        "data": "synthetic",
        "config": {
            "workload": f"TGV p={a.porder} {a.n}^3 elements, periodic, Re=1600 Ma=0.1, "
                        f"SSP(5,3)-2N rk_4_ls_main {'reference' if not a.clean else 'fresh'}"
                        f" sequencing, smart_visc_spectral sensor, fixed dt",
            "porder": a.porder, "elements": run.nelem, "unique_nodes": N,
            "rhs_per_step": nrhs, "operator_precision": a.precision, "dt": run.dt,
            "setup_s": round(t_setup, 2),
        },
        "roofline": {
            "kernel": "fused RK stage: k_rhs (element) + k_gather_update (node)",
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_stage": B,
            "ms_element_kernel": round(t_elem, 5), "ms_gather_update": round(t_gath, 5),
            "stage_only_gdofs": round(5 * N / (t_stage * 1e-3) / 1e9, 4),
        },
        "per_step_ms": {"element_kernels": round(k_ms / prof_steps, 4),
                        "gather_update": round(g_ms / prof_steps, 4),
                        "sensor": round(prof[2][0] / prof_steps, 4),
                        "other": round(prof[3][0] / prof_steps, 4)},
        "diagnostics": {k: diag[k] for k in ("time", "EK", "eps_T", "maxmach")},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
