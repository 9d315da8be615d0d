"""The ensemble of joined TGV instances (sod2d_amd/ensemble.py): uneven
periodic boxes with their own Mach and Reynolds numbers run as one mesh, one
launch per kernel for all instances.

* Every instance matches its run alone at the ensemble's common dt, step by
  step: rho, rho*u, E within 1e-12 normwise relative and mu_e within 1e-10
  (the bars of the single-mesh time loop; the join computes each element and
  node as the single mesh does, so the fields are expected to be equal).
* The autotuner's launch shapes do not change a result, and steps replayed
  from a HIP graph give the eager steps' bits after every replay.
* The feature dump of an instance equals the features of its run alone."""
import numpy as np
import pytest
import torch

from helpers import relerr

pytestmark = pytest.mark.gpu

CASES = [(4, 0.1, 1600.0), (6, 0.3, 800.0), (5, 0.45, 2800.0), (3, 0.2, 500.0)]


def _singles(ens, ctx):
    from sod2d_amd.solver import TGVParams, TGVRun
    runs = []
    for c in ens.cases:
        r = TGVRun(3, c.n, params=TGVParams(M=c.M, Re=c.Re), dt_mode="fixed", ctx=ctx)
        r.dt = ens.dt
        runs.append(r)
    return runs


def test_ensemble_matches_single_runs():
    from sod2d_amd.ensemble import EnsembleRun
    ens = EnsembleRun(CASES)
    assert ens.dt == min(ens.instance_dts) and len(set(ens.instance_dts)) > 1
    singles = _singles(ens, ens.ctx)
    worst = 0.0
    for step in range(1, 7):
        ens.step()
        for i, r in enumerate(singles):
            r.step()
            a, b = ens.instance_state(i), r.state()
            w = r.lpoin_w.cpu().numpy() - 1
            for k in ("rho", "q", "E"):
                e = relerr(a[k][w], b[k][w])
                worst = max(worst, e)
                assert e < 1e-12, (step, i, k, e)
            assert relerr(a["mu_e"], b["mu_e"]) < 1e-10, (step, i)
    print("ensemble vs single runs, worst conserved-field error over 6 steps:", worst)


def test_ensemble_autotune_and_graph():
    from sod2d_amd import autotune
    from sod2d_amd.ensemble import EnsembleRun
    eager = EnsembleRun(CASES)
    tuned = EnsembleRun(CASES, ctx=eager.ctx.__class__(0))
    for r in (eager, tuned):
        r.step()
    rec = autotune.tune_stage(tuned, reps=2)
    ch = rec["choice"]
    assert len(rec["table"]) == len(autotune.candidates())
    assert autotune.get_config(tuned.ctx) == (ch["elem_blocks_per_cu"], ch["geo_rows16"],
                                              ch["gather_threads"])
    # a shape other than the default on the eager run
    autotune.set_config(eager.ctx, (4, 0, 128))
    g = tuned.capture()
    try:
        for rep in range(2):
            g.replay()
            torch.cuda.synchronize()
            eager.step()
            eager.step()
            a, b = tuned.state(), eager.state()
            for k in ("rho", "q", "E", "mu_e"):
                assert np.array_equal(a[k], b[k]), (rep, k)
    finally:
        g.release()


def test_ensemble_feature_dump(tmp_path):
    from sod2d_amd.ensemble import EnsembleRun
    from sod2d_amd.surrogate import NF, sensor_samples
    ens = EnsembleRun(CASES)
    singles = _singles(ens, ens.ctx)
    for _ in range(3):
        ens.step()
        for r in singles:
            r.step()
    d = ens.dump_features(str(tmp_path / "dump.npz"))
    assert d["X"].shape == (ens.nelem, NF) and np.all(np.isfinite(d["X"]))
    assert np.all((d["t"] >= 0) & (d["t"] <= 1))
    f = np.load(tmp_path / "dump.npz")
    assert np.array_equal(f["X"], d["X"]) and np.array_equal(f["inst"], d["inst"])
    t = ens.tables
    for i, r in enumerate(singles):
        X, tt = sensor_samples(r)
        e0, e1 = int(t.elem_off[i]), int(t.elem_off[i + 1])
        assert np.array_equal(d["X"][e0:e1], X[:, :NF].cpu().numpy()), i
        assert np.array_equal(d["t"][e0:e1], tt.cpu().numpy()), i
