"""Launch-shape autotuner of the fused Runge-Kutta stage.

The stage's element kernel walks the elements grid-stride, so its grid (blocks
per CU) sets how many elements each workgroup takes; its geometry moves
through LDS in 16-byte or 4-byte lane rows (two LDS layouts of the same
records); the gather/update runs in blocks of 64, 128 or 256 nodes.  None of
these changes a result.  tune_stage() times every candidate shape on the
run's own mesh and state, keeps the fastest in the run's context
(sodhip_set_stage_config) and returns the whole table, which is the record
of its choice.  The run is not advanced: the timed stages write scratch
copies of the state.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json

import torch

from ._lib import check, lib
from .device import _ptr

BLOCKS_PER_CU = (4, 8, 12, 16, 24, 32)
GEO_ROWS16 = (1, 0)
GATHER_THREADS = (128, 256)


def candidates():
    return list(itertools.product(BLOCKS_PER_CU, GEO_ROWS16, GATHER_THREADS))


def get_config(ctx) -> tuple:
    b, g, t = C.c_int(), C.c_int(), C.c_int()
    check(lib().sodhip_get_stage_config(ctx.h, C.byref(b), C.byref(g), C.byref(t)),
          "get_stage_config")
    return b.value, g.value, t.value


def set_config(ctx, cfg) -> None:
    check(lib().sodhip_set_stage_config(ctx.h, *[int(x) for x in cfg]), "set_stage_config")


def _stage_call(run, y):
    """One fused stage (the coefficients of stage 2) from the run's level 1
    into the scratch state y."""
    L, G = run.L, run.G
    check(lib().sodhip_rk_stage_f64(
        run.ctx.h, run.p, run.nelem, run.npoin, run.npoin_w, _ptr(run.connec), _ptr(run.lpoin_w),
        run.op_precision, _ptr(run.He_op), _ptr(run.gpvol_op), _ptr(run.dl_op),
        *[_ptr(x) for x in run.tabs], _ptr(run.mu_e_op), _ptr(run.invMl), C.byref(run.sp), run.dt,
        L[2, 0], L[2, 1], G[2, 1], _ptr(run.rho[0]), _ptr(run.q[0]), _ptr(run.E[0]),
        _ptr(y[0]), _ptr(y[1]), _ptr(y[2]), None, None, None), "rk_stage")


def tune_stage(run, reps: int = 5, cands=None, path: str | None = None) -> dict:
    """Times each candidate (blocks per CU, 16-byte geometry rows, gather
    threads) over ``reps`` stages after one warm-up, by HIP events on the
    library's stream; sets the fastest on run.ctx and returns
    {"choice": ..., "table": [...]} (written as JSON to ``path`` if given)."""
    ctx = run.ctx
    cands = cands or candidates()
    y = [run.rho[1].clone(), run.q[1].clone(), run.E[1].clone()]
    table = []
    prof_was = False
    for cfg in cands:
        set_config(ctx, cfg)
        _stage_call(run, y)
        torch.cuda.synchronize()
        ctx.prof_reset()
        ctx.prof_enable(True)
        for _ in range(reps):
            _stage_call(run, y)
        torch.cuda.synchronize()
        te, ne = ctx.prof_get(0)
        tg, ng = ctx.prof_get(1)
        ctx.prof_enable(prof_was)
        table.append({"elem_blocks_per_cu": cfg[0], "geo_rows16": cfg[1],
                      "gather_threads": cfg[2], "ms_element": te / max(ne, 1),
                      "ms_gather": tg / max(ng, 1), "ms_stage": te / max(ne, 1) + tg / max(ng, 1)})
    best = min(table, key=lambda r: r["ms_stage"])
    set_config(ctx, (best["elem_blocks_per_cu"], best["geo_rows16"], best["gather_threads"]))
    out = {"choice": best, "table": table, "reps": reps, "elements": run.nelem,
           "working_nodes": run.npoin_w}
    if path:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return out
