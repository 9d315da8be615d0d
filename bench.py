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
  ensemble     -- 32 uneven TGV instances (n^3 elements, n ~ U{8..32}, Ma and Re
                  varied) joined into one mesh: GDOF/s and ms per step of HIP-graph
                  replayed steps, the autotuner's launch shape, the feature dump.
  surrogate    -- the learned entropy-viscosity sensor (per-element MLP, BF16 on
                  the matrix cores, plain sensor as fallback and judge every 10
                  steps) against the plain sensor on two fresh runs of the same
                  TGV: ms/step of each and their ratio, the sensor's own ms/step,
                  the kinetic-energy and dissipation-rate errors, mu_e error and
                  fallback count on the judge steps, training cost.
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

    def rhs(cores=cores):
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

    def timed(nc, budget):
        rhs(nc)
        t0 = time.perf_counter()
        rhs(nc)
        t1 = time.perf_counter() - t0
        reps = max(1, min(200, int(budget / max(t1, 1e-6))))
        t0 = time.perf_counter()
        for _ in range(reps):
            rhs(nc)
        return (time.perf_counter() - t0) / reps, reps

    dt, reps = timed(cores, budget_s)
    # the same RHS on one core (a few evaluations)
    dt1, reps1 = timed(1, budget_s / 3)
    N = g.mesh.npoin_w
    return {"value": round(5 * N / dt / 1e9, 6), "unit": "GDOF/s per RK stage", "cores": cores,
            "kind": "port",
            "sample": f"updateF + stage combine of the CPU oracle, TGV p=3 16^3 elements "
                      f"({N} unique nodes), FP64, {reps} RHS evaluations, "
                      f"{dt * 1e3:.2f} ms each, OpenMP element loop on {cores} threads",
            "single_core": {"value": round(5 * N / dt1 / 1e9, 6), "cores": 1,
                            "sample": f"the same, {reps1} RHS evaluations on 1 thread, "
                                      f"{dt1 * 1e3:.2f} ms each"}}


# jitter and seed of the isolated-operator box per order: the workload's
# +-10 % of the local GLL spacing at p=3 (seed 2 keeps every Jacobian
# positive at 128^3; seeds 1 and 20261202 fold one element), +-5 % at p=4,
# where +-10 % folds an element for each of seeds 2..15
OP_JITTER = {3: (0.10, 2), 4: (0.05, 20261202)}


def run_operators(a) -> dict:
    """The isolated-operator workload: full_convec_ijk, full_diffusion_ijk and
    smart_visc_spectral once each on an n^3-element structured box (2,097,152
    hexahedra at n=128), node positions jittered per axis by U(-j, j) x the
    local GLL spacing (OP_JITTER: the workload's +-10 % at p=3; +-5 % at p=4,
    where +-10 % folds an element for every seed tried), fields from
    case.operator_bench_fields; median of 20
    calls after 3 warm-ups.  Algorithmic bytes per call (N nodes, EN
    element-nodes, s operator bytes): convection (4+10s)EN + 9sN read and 5sN
    written; diffusion (4+12s)EN + 7sN read and 5sN written (outputs
    initialised, not accumulated); sensor 7sN + (4+2s)EN (FP64)."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    from sod2d_amd import case
    from sod2d_amd import device as D
    from sod2d_amd.device import to_dev

    p, n = a.porder, a.op_n
    torch.cuda.set_device(0)
    t_setup = time.perf_counter()
    T = case.basis_tables(p)
    jit, seed = OP_JITTER[p]
    m = case.box_mesh(p, n, periodic=False, jitter=jit, seed=seed)
    He, gpvol = case.elem_jacobian(p, m.connec_orig, m.coord)
    helem, helem_l = case.char_length(p, m.connec_orig, m.coord, spectral=(p == 3))
    if helem_l is None:
        helem_l = np.asfortranarray(np.repeat(helem[:, None], T.nnode, axis=1))
    del m.coord
    f = case.operator_bench_fields(m.npoin, m.nelem, T.nnode, seed=20261202 + p)
    dt = torch.float64 if a.precision == 8 else torch.float32
    dv = lambda x: to_dev(x).to(dt)  # noqa: E731
    dcon = to_dev(m.connec)
    tabs = [to_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    dHe, dgv, ddl = dv(He), dv(gpvol), dv(T.dlxigp_ip)
    fl = {k: dv(f[k]) for k in ("u", "q", "rho", "pr", "h", "Tem", "mu_fluid", "mu_e", "mu_sgs")}
    del He, gpvol
    rng = np.random.default_rng(20261202)
    Reta = to_dev(rng.normal(0, 1.0, m.npoin) * 10.0 ** rng.uniform(-2, 4, m.npoin))
    eta = to_dev(rng.normal(0, 1.0, m.npoin))
    rho64, u64, T64 = (to_dev(f[k]) for k in ("rho", "u", "Tem"))
    dhl, dhk, dlw = to_dev(helem_l), to_dev(helem), to_dev(m.lpoin_w)
    NE, NP, nn = m.nelem, m.npoin, T.nnode
    R = [torch.zeros(k, dtype=dt, device="cuda") for k in (NP, 3 * NP, NP)]
    mu = torch.zeros(NE * nn, dtype=torch.float64, device="cuda")
    ml = torch.zeros_like(mu)
    prm = D.visc_params(287.0, flag_normalise_entropy=1)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    def convec():
        D.full_convec_ijk(NE, NP, dcon, None, None, dHe, dgv, ddl, None, *tabs,
                          *[fl[k] for k in ("u", "q", "rho", "pr", "h")], *R, True, -1.0)

    def diffusion():
        D.full_diffusion_ijk(NE, NP, dcon, None, dHe, dgv, ddl, *tabs, 1004.0, 0.71, fl["rho"],
                             fl["rho"], fl["u"], fl["Tem"], fl["mu_fluid"], fl["mu_e"],
                             fl["mu_sgs"], None, *R, True, -1.0)

    def sensor():
        D.smart_visc_spectral(NE, NP, m.npoin_w, dcon, dlw, Reta, None, None, None, None, None,
                              1.4, rho64, u64, None, T64, eta, dhl, dhk, None, mu, tabs[0], None,
                              None, None, ml, prm)

    s_op = 8 if a.precision == 8 else 4
    EN, N = NE * nn, NP
    ops = {"full_convec_ijk": (convec, (4 + 10 * s_op) * EN + 9 * s_op * N + 5 * s_op * N),
           "full_diffusion_ijk": (diffusion, (4 + 12 * s_op) * EN + 7 * s_op * N + 5 * s_op * N),
           "smart_visc_spectral": (sensor, 7 * 8 * N + (4 + 2 * 8) * EN)}
    res, total = {}, 0.0
    for name, (fn, B) in ops.items():
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(20):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        med = float(np.median(ts))
        total += med
        res[name] = {"ms": round(med * 1e3, 4), "algorithmic_bytes": int(B),
                     "GB_s": round(B / med / 1e9, 1), "frac": round(B / med / 1e9 / HBM_PEAK_GBS, 4)}
    dom = max(res, key=lambda k: res[k]["ms"])
    return {
        "metric": "isolated operators: ms per call of full_convec_ijk + full_diffusion_ijk + "
                  "smart_visc_spectral",
        "value": round(total * 1e3, 4), "unit": "ms", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64" if a.precision == 8 else "f32 (sensor f64)", "data": "synthetic",
        "config": {"workload": f"structured {n}^3-element box, p={p}, jitter "
                               f"{round(100 * jit)}% of local GLL spacing, seed {seed}",
                   "porder": p, "elements": NE,
                   "nodes": NP, "element_nodes": EN, "setup_s": round(t_setup, 1)},
        "per_operator": res,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": res[dom]["GB_s"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": res[dom]["frac"],
                     "traffic": None},
    }


def run_surrogate(a, steps: int = 20, window: int = 40) -> dict:
    """The learned sensor beside the plain one on the bench's TGV: train the
    MLP (standard training set, 1500 Adam steps), then two fresh runs, plain
    smart_visc_spectral and the surrogate (plain fallback, judged every 10
    steps), each timed over ``steps`` steps after 3 warmup steps, then both
    advanced ``window`` more steps with the kinetic energy read every step."""
    import numpy as np
    import torch
    from sod2d_amd.solver import TGVRun
    from sod2d_amd.surrogate import trained_surrogate

    from sod2d_amd.device import Context

    ctx = Context(torch.cuda.current_device())
    t1 = time.perf_counter()
    net, hist, nsamp = trained_surrogate(ctx, steps=1500)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    runs = {}
    t0 = time.perf_counter()
    for kind in ("spectral", "surrogate"):
        runs[kind] = TGVRun(porder=a.porder, nelem_dir=a.n, op_precision=a.precision,
                            parity=not a.clean, dt_mode="fixed", sensor_path=kind,
                            surrogate=net if kind == "surrogate" else None, judge_every=10,
                            ctx=ctx)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    ms, sensor_ms = {}, {}
    for kind, r in runs.items():
        for _ in range(3):
            r.step()
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(steps):
            r.step()
        torch.cuda.synchronize()
        ms[kind] = (time.perf_counter() - ta) / steps * 1e3
        # the sensor's share from 10 further steps bracketed by HIP events
        ctx.prof_reset()
        ctx.prof_enable(True)
        for _ in range(10):
            r.step()
        torch.cuda.synchronize()
        sensor_ms[kind] = ctx.prof_get(2)[0] / 10
        ctx.prof_enable(False)
    ek = {k: [] for k in runs}
    for _ in range(window):
        for kind, r in runs.items():
            r.step()
            ek[kind].append(r.diagnostics()["EK"])
    ekp, eks = np.array(ek["spectral"]), np.array(ek["surrogate"])
    dt = runs["spectral"].dt
    dp, ds = -np.diff(ekp) / dt, -np.diff(eks) / dt
    judged = [j for j in runs["surrogate"].judge_log if j[0] >= 10]
    out = {
        "plain_ms_per_step": round(ms["spectral"], 4),
        "surrogate_ms_per_step": round(ms["surrogate"], 4),
        "ratio": round(ms["surrogate"] / ms["spectral"], 4),
        "sensor_ms_per_step": {"plain": round(sensor_ms["spectral"], 4),
                               "surrogate": round(sensor_ms["surrogate"], 4)},
        "ke_dissipation_rel_err": float(abs(ds[-10:].mean() - dp[-10:].mean())
                                        / abs(dp[-10:].mean())),
        "ke_max_rel_err": float(np.abs(eks - ekp).max() / np.abs(ekp).max()),
        "mu_e_rel_err_judged": float(np.mean([j[1] for j in judged])) if judged else None,
        "fallback_elements_judged": float(np.mean([j[2] for j in judged])) if judged else None,
        "elements": runs["spectral"].nelem,
        "judge_every": 10,
        "steps_compared": 3 + steps + 10 + window,
        "train": {"samples": nsamp, "adam_steps": 1500, "batch": 4096, "s": round(t_train, 2),
                  "loss_first": hist[0], "loss_last": hist[-1]},
        "setup_s": round(t_setup, 2),
    }
    del runs
    return out


def run_ensemble(a, count: int = 32, steps: int = 10) -> dict:
    """The batched ensemble: ``count`` periodic TGV instances, p=3, with n^3
    elements, n ~ U{8..32}, Ma ~ U(0.1, 0.5), Re ~ U(400, 3200) (seeded),
    joined into one mesh and advanced at the smallest instance dt; one element
    launch and one gather per RK stage cover every instance.  After the first
    step the autotuner picks the stage's launch shape, two steps are captured
    in a HIP graph, and ``steps`` steps (rounded up to even) are timed as graph
    replays.  The features and mu_e targets of every element are dumped once.
    GDOF/s per RK stage as for the single run, over all instances' nodes."""
    import torch
    from sod2d_amd import autotune
    from sod2d_amd.ensemble import EnsembleRun, config4_cases

    cases = config4_cases(count)
    t0 = time.perf_counter()
    ens = EnsembleRun(cases)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    ens.step()  # the first step builds the library's tables
    tune = autotune.tune_stage(ens, reps=3)
    ens.step()
    graph = ens.capture()
    graph.replay()
    torch.cuda.synchronize()
    nrep = (steps + 1) // 2
    t0 = time.perf_counter()
    for _ in range(nrep):
        graph.replay()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / (2 * nrep)
    graph.release()
    t1 = time.perf_counter()
    dump = ens.dump_features()
    torch.cuda.synchronize()
    t_dump = time.perf_counter() - t1
    N = ens.npoin_w
    ns = [c.n for c in cases]
    ch = tune["choice"]
    return {
        "value": round(5 * N * ens.rhs_per_step / t_step / 1e9, 4),
        "unit": "GDOF/s",
        "ms_per_step": round(t_step * 1e3, 4),
        "steps": 2 * nrep,
        "instances": count,
        "elements": ens.nelem,
        "unique_nodes": N,
        "n_range": [min(ns), max(ns)],
        "mach_range": [round(min(c.M for c in cases), 3), round(max(c.M for c in cases), 3)],
        "re_range": [round(min(c.Re for c in cases), 1), round(max(c.Re for c in cases), 1)],
        "dt": ens.dt,
        "graph": True,
        "autotune": {"elem_blocks_per_cu": ch["elem_blocks_per_cu"],
                     "geo_rows16": ch["geo_rows16"], "gather_threads": ch["gather_threads"],
                     "ms_stage": round(ch["ms_stage"], 5),
                     "ms_stage_default": round(next(
                         r["ms_stage"] for r in tune["table"]
                         if (r["elem_blocks_per_cu"], r["geo_rows16"], r["gather_threads"])
                         == (16, 1, 256)), 5),
                     "candidates": len(tune["table"])},
        "feature_dump": {"elements": int(dump["X"].shape[0]), "features": int(dump["X"].shape[1]),
                         "s": round(t_dump, 3)},
        "setup_s": round(t_setup, 2),
    }


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
    ap.add_argument("--lps", action="store_true",
                    help="the reference's default TGV path instead of the plain method: LPS "
                         "stabilisation and the live sensor path (smart_visc_spectral_imex + "
                         "Ducros limiter), through the drop-in operators (FP64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-sensor comparison (surrogate vs plain ms/step)")
    ap.add_argument("--graph", action="store_true",
                    help="time the steps as replays of a HIP graph of two captured steps "
                         "(--steps rounded up to even)")
    ap.add_argument("--workload", choices=("tgv", "operators", "ensemble"), default="tgv",
                    help="tgv: the TGV time step (default); operators: the isolated "
                         "full_convec_ijk / full_diffusion_ijk / smart_visc_spectral calls; "
                         "ensemble: 32 uneven TGV instances joined into one mesh")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble summary in the default line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape of the fused stage instead of the "
                         "autotuner's (sod2d_amd/autotune.py, run after the warm-up steps)")
    ap.add_argument("--stage-config", default=None, metavar="B,G,T",
                    help="use this launch shape of the fused stage (element blocks per CU, "
                         "16-byte geometry rows 0/1, gather threads) instead of autotuning, "
                         "e.g. to profile the shape a tuned run chose; with --lps the element "
                         "kernel rounds B down to whole rounds of the blocks a CU holds")
    ap.add_argument("--op-n", type=int, default=128,
                    help="elements per direction of the operator workload's box")
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
    if a.workload == "operators":
        print(json.dumps(run_operators(a)), flush=True)
        return 0
    if a.workload == "ensemble":
        sys.path.insert(0, ROOT)
        e = run_ensemble(a, steps=max(a.steps, 2))
        out = {"metric": "GDOF/s per RK stage, batched ensemble of uneven TGV instances",
               "value": e["value"], "unit": "GDOF/s", "n_gpus": 1, "steps": e["steps"],
               "warmup": 3, "ms_per_step": e["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": "32 periodic TGV instances p=3, n^3 elements n~U{8..32}, "
                                      "Ma~U(0.1,0.5), Re~U(400,3200), joined into one mesh, "
                                      "common fixed dt, HIP-graph replayed steps"},
               "ensemble": e}
        print(json.dumps(out), flush=True)
        return 0

    import torch
    sys.path.insert(0, ROOT)
    from sod2d_amd.solver import TGVRun

    torch.cuda.set_device(0)
    t_setup = time.perf_counter()
    if a.lps and a.precision != 8:
        print("bench.py: --lps runs with FP64 operators only", file=sys.stderr)
        return 2
    run = TGVRun(porder=a.porder, nelem_dir=a.n, op_precision=a.precision, parity=not a.clean,
                 dt_mode="fixed", lps=a.lps, sensor_path="live" if a.lps else "spectral")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    for _ in range(max(a.warmup, 1 if (a.graph or not a.no_autotune) else 0)):
        run.step()
    torch.cuda.synchronize()
    tune = None
    if a.stage_config:
        from sod2d_amd import autotune
        cfg = tuple(int(x) for x in a.stage_config.split(","))
        autotune.set_config(run.ctx, cfg)
        tune = {"choice": {"elem_blocks_per_cu": cfg[0], "geo_rows16": cfg[1],
                           "gather_threads": cfg[2]}, "given": True}
    elif not a.no_autotune and not a.lps:
        from sod2d_amd import autotune
        tune = autotune.tune_stage(run, reps=3)
        run.step()
        torch.cuda.synchronize()
    if a.graph:
        graph = run.capture()
        a.steps += a.steps % 2
        graph.replay()  # the first replay uploads the graph
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    if a.graph:
        for _ in range(a.steps // 2):
            graph.replay()
    else:
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
    prof = {k: run.ctx.prof_get(k) for k in range(5)}
    run.ctx.prof_enable(False)
    diag = run.diagnostics()
    # the same run with the adaptive time step (CFL dt reduced and read back
    # every step, stage 1 then evaluated at the start of each step)
    adaptive = None
    if not a.lps and not a.graph:
        run.dt_mode = "adaptive"
        run.step()
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(5):
            run.step()
        torch.cuda.synchronize()
        adaptive = {"ms_per_step": round((time.perf_counter() - ta) / 5 * 1e3, 4), "steps": 5,
                    "dt": run.dt}
        run.dt_mode = "fixed"
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
    if a.lps:
        # one element kernel (convection and the stabilised viscous operator,
        # two waves per element at p=3) and one gather/update per stage;
        # full_proj_ijk (its element kernel and projection gather) is timed apart
        t_elem = k_ms / (prof_steps * nrhs)
        t_gath = g_ms / (prof_steps * nrhs)
    else:
        t_elem = k_ms / max(k_n, 1)
        t_gath = g_ms / max(g_n, 1)
    t_stage = t_elem + t_gath
    B = stage_bytes(N, EN, 8 if a.precision == 8 else 4)
    if a.lps:
        # the LPS stage also reads each node's projection record (96 bytes) and
        # each element's tau_stab
        B += 96 * N + 8 * run.nelem
    achieved = B / (t_stage * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_rk_stage_tgv{a.n}_p{a.porder}_"
                      f"{'fp64' if a.precision == 8 else 'fp32op'}.json")
    if os.path.exists(tf) and not a.lps:
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
            "graph": bool(a.graph),
            "workload": f"TGV p={a.porder} {a.n}^3 elements, periodic, Re=1600 Ma=0.1, "
                        f"SSP(5,3)-2N rk_4_ls_main {'reference' if not a.clean else 'fresh'}"
                        f" sequencing, "
                        + ("LPS stabilisation, live sensor path (smart_visc_spectral_imex + "
                           "Ducros), drop-in operators" if a.lps else
                           "smart_visc_spectral sensor") + ", fixed dt",
            "porder": a.porder, "elements": run.nelem, "unique_nodes": N,
            "rhs_per_step": nrhs, "operator_precision": a.precision, "dt": run.dt,
            "setup_s": round(t_setup, 2),
            "stage_launch": (None if tune is None else {
                k: tune["choice"][k] for k in ("elem_blocks_per_cu", "geo_rows16",
                                               "gather_threads")} | ({"given": True}
                if tune.get("given") else {
                "autotuned": True, "ms_stage_tuned": round(tune["choice"]["ms_stage"], 5),
                "ms_stage_default": round(next(
                    r["ms_stage"] for r in tune["table"]
                    if (r["elem_blocks_per_cu"], r["geo_rows16"], r["gather_threads"])
                    == (16, 1, 256)), 5)})),
        },
        "roofline": {
            "kernel": ("LPS stage (sodhip_rk_stage_lps_f64): k_rhs_split in LPS mode "
                       "(convection and full_diff_stab_ijk) + k_gather_update, per stage; "
                       "full_proj_ijk (ms_projection) excluded" if a.lps else
                       "fused RK stage: k_rhs_split (element) + k_gather_update (node)"),
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_stage": B,
            "ms_element_kernel": round(t_elem, 5), "ms_gather_update": round(t_gath, 5),
            "stage_only_gdofs": round(5 * N / (t_stage * 1e-3) / 1e9, 4),
        } | ({"ms_projection": round(prof[4][0] / (prof_steps * nrhs), 5)} if a.lps else {}),
        "per_step_ms": {"element_kernels": round(k_ms / prof_steps, 4),
                        "gather_update": round(g_ms / prof_steps, 4),
                        "sensor": round(prof[2][0] / prof_steps, 4),
                        "other": round(prof[3][0] / prof_steps, 4)}
        | ({"projection": round(prof[4][0] / prof_steps, 4)} if a.lps else {}),
        "diagnostics": {k: diag[k] for k in ("time", "EK", "eps_T", "maxmach")},
        "adaptive_dt": adaptive,
    }
    if not a.no_surrogate and not a.lps:
        del run
        torch.cuda.empty_cache()
        out["surrogate"] = run_surrogate(a)
    if not a.no_ensemble and not a.lps and a.precision == 8:
        torch.cuda.empty_cache()
        out["ensemble"] = run_ensemble(a)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
