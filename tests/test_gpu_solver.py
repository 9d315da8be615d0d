"""Parity of the device time loop (rk_4_ls_main sequence with the fused stage)
with the oracle's restatement, step by step, on small periodic TGV meshes.

Bars (from the FP64 target): conserved fields rho, rho*u, E within 1e-12
normwise relative per step; the sensor output mu_e within 1e-10 (it carries a
min/max limiter, so rounding can flip a branch on isolated points only through
ties); dt and the diagnostics EK and eps within 1e-10.  The FP32-operator mode
(FP64 state, gather, mass inverse and combine) is checked against the FP64
oracle with per-field bounds stated below."""
import numpy as np
import pytest

from helpers import make_case, relerr

pytestmark = pytest.mark.gpu


def _setup(oracle, p, n, stages=5, parity=True, op_precision=8, dt_mode="adaptive", M=0.1,
           sensor_path="spectral"):
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    tp = TGVParams(M=M)
    to = oracle.tgv_params(M=M)
    ic = oracle.tgv_ic(g.mesh.coord, to)
    dev = TGVRun(p, geometry=g, params=tp, initial=ic, stages=stages, parity=parity,
                 op_precision=op_precision, dt_mode=dt_mode, sensor_path=sensor_path)
    orc = oracle.TGVOracle(p, g, ic, to, stages=stages, parity=parity, dt_mode=dt_mode,
                           cfl=(tp.cfl_conv, tp.cfl_diff), sensor_path=sensor_path)
    return dev, orc


@pytest.mark.parametrize("p,n,stages,parity", [(3, 4, 5, True), (3, 3, 4, False), (2, 4, 3, True)])
def test_time_loop_fp64(oracle, p, n, stages, parity):
    dev, orc = _setup(oracle, p, n, stages, parity)
    assert abs(dev.dt - orc.dt) <= 1e-12 * orc.dt
    for step in range(6):
        dev.step()
        orc.step()
        s = dev.state()
        w = orc.lpoin_w - 1
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, step
        assert abs(dev.dt - orc.dt) <= 1e-10 * orc.dt, step
    dd, do = dev.diagnostics(), orc.diagnostics()
    for k in ("EK", "eps_S", "eps_D", "eps_T", "maxmach"):
        assert abs(dd[k] - do[k]) <= 1e-10 * abs(do[k]) + 1e-300, (k, dd[k], do[k])


@pytest.mark.parametrize("p,n", [(3, 4), (2, 4)])
def test_time_loop_live_sensor(oracle, p, n):
    """The reference's live path: smart_visc_spectral_imex, divergence,
    vorticity and the Ducros limit after every step (time_integ_ls.f90:
    455-463), same bars as the plain sensor path."""
    dev, orc = _setup(oracle, p, n, sensor_path="live")
    for step in range(6):
        dev.step()
        orc.step()
        s = dev.state()
        w = orc.lpoin_w - 1
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, step
        assert abs(dev.dt - orc.dt) <= 1e-10 * orc.dt, step
    # the limiter acts: mu_e differs from the unlimited imex sensor's output
    assert np.any(orc.mu_e != orc.mue_l)


def test_time_loop_fp32_operators(oracle):
    """FP32 operator arithmetic with FP64 state: per-field error vs the FP64
    oracle after 6 steps at Ma=0.1 (bounds: rho 2e-7, rho*u 1e-5, E 2e-7)."""
    dev, orc = _setup(oracle, 3, 4, op_precision=4)
    for _ in range(6):
        dev.step()
        orc.step()
    s = dev.state()
    w = orc.lpoin_w - 1
    e = {k: relerr(s[k][w], getattr(orc, k)[1][w]) for k in ("rho", "q", "E")}
    assert e["rho"] < 2e-7 and e["q"] < 1e-5 and e["E"] < 2e-7, e


@pytest.mark.parametrize("op_precision", [8, 4])
def test_time_loop_fixed_dt_next_stage(oracle, op_precision):
    """At fixed dt the last stage of a step also stores the next step's stage 1
    (sodhip_rk_stage_next_f64): the same per-step bars against the oracle, and
    the same bits as the separate stage-1 update."""
    dev, orc = _setup(oracle, 3, 4, op_precision=op_precision, dt_mode="fixed")
    ref, _ = _setup(oracle, 3, 4, op_precision=op_precision, dt_mode="fixed")
    ref.fuse_next_stage = False  # stage 1 evaluated at the start of each step
    for step in range(6):
        dev.step()
        ref.step()
        orc.step()
        s, r = dev.state(), ref.state()
        for k in ("rho", "q", "E", "mu_e"):
            assert np.array_equal(s[k], r[k]), (step, k)
        if op_precision == 8:
            w = orc.lpoin_w - 1
            assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
            assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
            assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step


@pytest.mark.parametrize("sensor_path", ["spectral", "live"])
def test_fused_end_of_step_same_bits(oracle, sensor_path):
    """sodhip_eos_entropy_step_f64 (EOS, entropy and Sutherland in one node
    pass) gives the bits of the separate sodhip_eos_step_f64,
    sodhip_thermo_entropy_f64 and sodhip_sutherland_f64 calls, and
    sodhip_entropy_convec_residual_f64 those of the scalar convection, lumped
    solve and entropy residual calls: over five adaptive-dt steps every array
    they write, the state, mu_e and dt agree bitwise."""
    dev, _ = _setup(oracle, 3, 4, sensor_path=sensor_path)
    ref, _ = _setup(oracle, 3, 4, sensor_path=sensor_path)
    ref.fuse_end_state = False
    for step in range(5):
        dev.step()
        ref.step()
        assert dev.dt == ref.dt, step
        for k in ("u", "e_int", "pr", "Tem"):
            assert np.array_equal(getattr(dev, k)[1].cpu().numpy(),
                                  getattr(ref, k)[1].cpu().numpy()), (step, k)
        for k in ("csound", "machno", "eta", "mu_fluid", "mu_e", "Reta", "auxReta"):
            assert np.array_equal(getattr(dev, k).cpu().numpy(),
                                  getattr(ref, k).cpu().numpy()), (step, k)
        s, r = dev.state(), ref.state()
        for k in ("rho", "q", "E"):
            assert np.array_equal(s[k], r[k]), (step, k)


@pytest.mark.parametrize("lps,n", [(False, 4), (True, 4), (False, 8)])
def test_graph_captured_steps(oracle, lps, n):
    """Two steps captured in a HIP graph and replayed give the same bits as the
    same steps launched eagerly, on every replay (with eager work and
    synchronisations between the replays).  The diagnostics called between the
    replays need more reduction space than the captured sensor at 8^3 elements
    (2*nelem + 2*blocks doubles): they must not move a buffer the graph holds."""
    import torch

    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(3, n, periodic=True, lengths=(2 * np.pi,) * 3)
    to = oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    kw = dict(geometry=g, params=TGVParams(), initial=ic, dt_mode="fixed", lps=lps,
              sensor_path="live" if lps else "spectral")
    eager, cap = TGVRun(3, **kw), TGVRun(3, **kw)
    eager.step()
    cap.step()
    graph = cap.capture()
    for rep in range(4):
        eager.step()
        eager.step()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        a, b = eager.state(), cap.state()
        for k in ("rho", "q", "E", "mu_e"):
            assert np.array_equal(a[k], b[k]), (rep, k)
        assert cap.istep == eager.istep and cap.time == eager.time
        da, db = eager.diagnostics(), cap.diagnostics()
        assert da == db, rep


@pytest.mark.parametrize("p,n", [(3, 4), (2, 4)])
def test_time_loop_rk4(oracle, p, n):
    """The classical RK4 integrator (rk_4_main, sodhip_rk4_stage_f64 per
    stage) against the oracle's restatement, with the same entropy viscosity
    held in both: conserved fields at 1e-12 per step."""
    import torch
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    tp, to = TGVParams(), oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    dev = TGVRun(p, geometry=g, params=tp, initial=ic, integrator="rk4", dt_mode="fixed")
    orc = oracle.TGVOracle(p, g, ic, to, integrator="rk4", dt_mode="fixed",
                           cfl=(tp.cfl_conv, tp.cfl_diff))
    rng = np.random.default_rng(5)
    orc.mu_e[...] = rng.uniform(0.0, 2e-3, orc.mu_e.shape)
    dev.mu_e.copy_(torch.from_numpy(orc.mu_e.ravel(order="F")))
    assert abs(dev.dt - orc.dt) <= 1e-12 * orc.dt
    w = orc.lpoin_w - 1
    for step in range(5):
        dev.step()
        orc.step()
        s = dev.state()
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step
    # the state moved: RK4 is not the identity
    assert relerr(s["q"][w], ic["q"][w]) > 1e-6


@pytest.mark.parametrize("n", [4, 3])
def test_time_loop_brick_local_assembly(oracle, monkeypatch, n):
    """The opt-in brick-local assembly of the fused p=3 FP64 stage
    (SODHIP_STAGE_CLUSTERS=1 at context creation): element clusters summed
    in LDS, interior nodes finished in the element kernel, surface nodes by
    the gather of their partial records.  Same per-step bars against the
    oracle as the node-ordered gather; the state after each step agrees with
    that path's to rounding (interior nodes add the same terms in the same
    order, surface nodes reassociate)."""
    from sod2d_amd.device import Context
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(3, n, periodic=True, lengths=(2 * np.pi,) * 3)
    tp, to = TGVParams(), oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    monkeypatch.setenv("SODHIP_STAGE_CLUSTERS", "1")
    ctx_cl = Context(0)
    monkeypatch.delenv("SODHIP_STAGE_CLUSTERS")
    ctx_plain = Context(0)
    dev = TGVRun(3, geometry=g, params=tp, initial=ic, ctx=ctx_cl, dt_mode="fixed")
    ref = TGVRun(3, geometry=g, params=tp, initial=ic, ctx=ctx_plain, dt_mode="fixed")
    orc = oracle.TGVOracle(3, g, ic, to, dt_mode="fixed", cfl=(tp.cfl_conv, tp.cfl_diff))
    w = orc.lpoin_w - 1
    for step in range(4):
        dev.step()
        ref.step()
        orc.step()
        s, r = dev.state(), ref.state()
        for k in ("rho", "q", "E"):
            assert relerr(s[k][w], getattr(orc, k)[1][w]) < 1e-12, (k, step)
            assert relerr(s[k][w], r[k][w]) < 1e-14, (k, step)
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, step
