"""Parity of the LPS stabilisation entry points (elem_stab.f90:207-438,
441-586, 589-616) with the CPU oracle through the C-ABI, and of the device
time loop with LPS on (the reference's default viscous path, flag_lps_stab)
against the oracle's time loop.

Bars: FP64 normwise relative error 1e-13 per operator output (the device sums
in another order and contracts products into FMAs); tau_stab 1e-15; the time
loop as the other time-loop tests: rho, rho*u, E within 1e-12 per step, mu_e
within 1e-10 over 6 steps, dt within 1e-10."""
import numpy as np
import pytest
import torch

from helpers import make_case, relerr

pytestmark = pytest.mark.gpu


def _dev(a):
    from sod2d_amd.device import to_dev
    return to_dev(a)


def _fields(m, seed):
    r = np.random.default_rng(seed)
    n = m.npoin
    rho = r.uniform(0.9, 1.1, n)
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    T = r.uniform(0.9, 1.1, n)
    c = r.uniform(1.0, 1.2, n)
    return rho, u, T, c


CASES = [(3, 4, True, 0.0), (3, (3, 2, 4), False, 0.1), (2, 3, True, 0.0), (4, 2, False, 0.05)]


@pytest.mark.parametrize("p,n,periodic,jitter", CASES)
def test_comp_tau_and_projection(oracle, p, n, periodic, jitter):
    from sod2d_amd import device as D
    g = make_case(p, n, periodic=periodic, jitter=jitter)
    m, T = g.mesh, g.tables
    OT = oracle.Tables(p)
    rho, u, Tem, c = _fields(m, 10 + p)
    otau = oracle.comp_tau(p, m.connec, c, u, g.helem)
    dtau = torch.zeros(m.nelem, dtype=torch.float64, device="cuda")
    D.comp_tau(m.nelem, m.npoin, _dev(m.connec), _dev(c), _dev(u), _dev(g.helem), 0.1, dtau)
    oP = oracle.full_proj_ijk(OT, m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, Tem, g.invMl)
    dP = [torch.full((3 * m.npoin,), 7.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    D.full_proj_ijk(m.nelem, m.npoin, m.npoin_w, _dev(m.connec), _dev(m.lpoin_w), None,
                    _dev(g.He), _dev(g.gpvol), _dev(T.dlxigp_ip), *tabs, 1004.0, 0.71, _dev(rho),
                    _dev(u), _dev(Tem), None, _dev(g.invMl), *dP)
    torch.cuda.synchronize()
    assert relerr(dtau.cpu().numpy(), otau) <= 1e-15
    for a, b in zip(dP, oP):
        assert relerr(a.cpu().numpy(), b.ravel(order="F")) < 1e-13


@pytest.mark.parametrize("p,n,periodic,jitter", CASES)
@pytest.mark.parametrize("initialize", [True, False])
def test_full_diff_stab(oracle, p, n, periodic, jitter, initialize):
    from sod2d_amd import device as D
    g = make_case(p, n, periodic=periodic, jitter=jitter)
    m, T = g.mesh, g.tables
    OT = oracle.Tables(p)
    rho, u, Tem, c = _fields(m, 20 + p)
    r = np.random.default_rng(30 + p)
    nn = OT.nnode
    rho_n = r.uniform(0.9, 1.1, m.npoin)
    mu_f = r.uniform(1e-3, 2e-3, m.npoin)
    mu_e = np.asfortranarray(r.uniform(0, 1e-2, (m.nelem, nn)))
    mu_sgs = np.asfortranarray(r.uniform(0, 1e-3, (m.nelem, nn)))
    tau = r.uniform(0.05, 0.2, m.nelem)
    P = oracle.full_proj_ijk(OT, m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, Tem, g.invMl)
    R0 = [r.normal(size=m.npoin), np.asfortranarray(r.normal(size=(m.npoin, 3))),
          r.normal(size=m.npoin)]
    oR = [x.copy(order="F") for x in R0]
    oracle.full_diff_stab_ijk(OT, m.connec, g.He, g.gpvol, 1004.0, 0.71, rho_n, rho, u, Tem, mu_f,
                              mu_e, mu_sgs, P, tau, *oR, initialize, -1.0)
    dR = [_dev(x.copy(order="F")) for x in R0]
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    D.full_diff_stab_ijk(m.nelem, m.npoin, _dev(m.connec), None, _dev(g.He), _dev(g.gpvol),
                         _dev(T.dlxigp_ip), *tabs, 1004.0, 0.71, _dev(rho_n), _dev(rho), _dev(u),
                         _dev(Tem), _dev(mu_f), _dev(mu_e), _dev(mu_sgs), None,
                         *[_dev(x) for x in P], _dev(tau), *dR, initialze=initialize, fact=-1.0)
    torch.cuda.synchronize()
    for a, b in zip(dR, oR):
        assert relerr(a.cpu().numpy(), b.ravel(order="F")) < 1e-13
    # two launches give the same bits
    dR2 = [_dev(x.copy(order="F")) for x in R0]
    D.full_diff_stab_ijk(m.nelem, m.npoin, _dev(m.connec), None, _dev(g.He), _dev(g.gpvol),
                         _dev(T.dlxigp_ip), *tabs, 1004.0, 0.71, _dev(rho_n), _dev(rho), _dev(u),
                         _dev(Tem), _dev(mu_f), _dev(mu_e), _dev(mu_sgs), None,
                         *[_dev(x) for x in P], _dev(tau), *dR2, initialze=initialize, fact=-1.0)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(dR, dR2))


def _loop(oracle, p, n, steps, M=0.1, high_mach=False, lengths=2 * np.pi):
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(p, n, periodic=True, lengths=(lengths,) * 3)
    tp = TGVParams(M=M)
    to = oracle.tgv_params(M=M)
    ic = oracle.tgv_ic(g.mesh.coord, to)
    dev = TGVRun(p, geometry=g, params=tp, initial=ic, sensor_path="live", lps=True,
                 high_mach=high_mach)
    orc = oracle.TGVOracle(p, g, ic, to, cfl=(tp.cfl_conv, tp.cfl_diff), sensor_path="live",
                           lps=True, high_mach=high_mach)
    assert abs(dev.dt - orc.dt) <= 1e-12 * orc.dt
    for step in range(steps):
        dev.step()
        orc.step()
        s = dev.state()
        w = orc.lpoin_w - 1
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, step
        assert abs(dev.dt - orc.dt) <= 1e-10 * orc.dt, step
    return dev, orc


@pytest.mark.parametrize("p,n", [(3, 4), (2, 4), (4, 2), (2, 3)])
def test_time_loop_lps(oracle, p, n):
    """The reference's default stage: live sensor path plus LPS.  At p=2 on 3^3
    elements nelem*nnode is odd, so the projection records do not follow the
    convection records at an even offset."""
    dev, orc = _loop(oracle, p, n, 6)
    assert np.abs(orc.tau_stab).min() > 0.0
    dd, do = dev.diagnostics(), orc.diagnostics()
    for k in ("EK", "eps_S", "eps_D", "eps_T", "maxmach"):
        assert abs(dd[k] - do[k]) <= 1e-10 * abs(do[k]) + 1e-300, (k, dd[k], do[k])


def test_time_loop_lps_high_mach(oracle):
    """LPS with the sensor's high-Mach normalisation (flag_high_mach: norm =
    min rho, ce = ce_comp) at Ma = 1.25."""
    dev, orc = _loop(oracle, 3, 3, 6, M=1.25, high_mach=True)
    assert dev.vp.ce == orc.prm.ce == 1.0


def test_time_loop_lps_fixed_dt_next_stage(oracle):
    """At fixed dt the LPS stage also stores the next step's stage 1: same bits
    as the separate stage-1 update, and the oracle's per-step bars."""
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(3, 4, periodic=True, lengths=(2 * np.pi,) * 3)
    tp = TGVParams()
    to = oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    kw = dict(geometry=g, params=tp, initial=ic, sensor_path="live", lps=True, dt_mode="fixed")
    dev = TGVRun(3, **kw)
    ref = TGVRun(3, **kw)
    ref.fuse_next_stage = False
    orc = oracle.TGVOracle(3, g, ic, to, cfl=(tp.cfl_conv, tp.cfl_diff), sensor_path="live",
                           lps=True, dt_mode="fixed")
    for step in range(5):
        dev.step()
        ref.step()
        orc.step()
        s, r = dev.state(), ref.state()
        for k in ("rho", "q", "E", "mu_e"):
            assert np.array_equal(s[k], r[k]), (step, k)
        w = orc.lpoin_w - 1
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step


@pytest.mark.parametrize("p,n", [(3, 3), (2, 3)])
def test_fused_stage_projection_arrays_optional(oracle, p, n):
    """sodhip_rk_stage_lps_f64 keeps the projections inside the call when the
    Proj arrays are NULL (its stabilised operator reads them as node records):
    the stage then gives the same bits as with the arrays, and the arrays
    hold the drop-in full_proj_ijk of the stage's EOS output, bit for bit
    (the drop-in is checked against the oracle in the projection test)."""
    import ctypes as C

    from sod2d_amd._lib import check, lib
    from sod2d_amd.device import _ptr
    from sod2d_amd.solver import C_LPS_COMP, TGVParams, TGVRun
    g = make_case(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    tp = TGVParams()
    ic = oracle.tgv_ic(g.mesh.coord, oracle.tgv_params())
    run = TGVRun(p, geometry=g, params=tp, initial=ic, sensor_path="live", lps=True)
    run.step()
    m, N = g.mesh, run.npoin
    outs = []
    for keep in (False, True):
        y = [run.rho[1].clone(), run.q[1].clone(), run.E[1].clone()]
        prim = [torch.zeros(k * N, dtype=torch.float64, device="cuda") for k in (3, 1, 1, 1, 1, 1)]
        proj = [torch.zeros(3 * N, dtype=torch.float64, device="cuda") for _ in range(5)]
        check(lib().sodhip_rk_stage_lps_f64(
            run.ctx.h, p, run.nelem, N, run.npoin_w, _ptr(run.connec), _ptr(run.lpoin_w),
            _ptr(run.He), _ptr(run.gpvol), _ptr(run.dl), *[_ptr(x) for x in run.tabs],
            _ptr(run.mu_e), _ptr(run.invMl), _ptr(run.tau_stab), C.byref(run.sp), C_LPS_COMP,
            run.dt, run.L[2, 0], run.L[2, 1], run.G[2, 1], _ptr(run.rho[0]), _ptr(run.q[0]),
            _ptr(run.E[0]), *[_ptr(x) for x in y], *[_ptr(x) for x in prim],
            *[_ptr(x) if keep else None for x in proj], None, None, None, 0.0, None, None, None),
            "rk_stage_lps")
        torch.cuda.synchronize()
        outs.append((y, prim, proj))
    (y0, _, _), (y1, prim, proj) = outs
    for a, b in zip(y0, y1):
        assert torch.equal(a, b)
    # the arrays: the drop-in full_proj_ijk (pinned against the oracle above)
    # of the stage's EOS output, bit for bit
    from sod2d_amd import device as D
    ref = [torch.full((3 * N,), 7.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    D.full_proj_ijk(m.nelem, N, m.npoin_w, run.connec, run.lpoin_w, None, run.He, run.gpvol,
                    run.dl, *run.tabs, 1004.0, 0.71, run.rho[1], prim[0], prim[4], None,
                    run.invMl, *ref)
    torch.cuda.synchronize()
    for a, b in zip(proj, ref):
        assert torch.equal(a, b)


def test_lps_stage_geometry_row_layouts_agree(oracle):
    """The p=3 LPS stage moves He and gpvol through LDS in 16-byte or 4-byte
    lane rows (sodhip_set_stage_config); both give the same LDS image, so
    the same bits, step after step."""
    from sod2d_amd import autotune
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(3, 4, periodic=True, lengths=(2 * np.pi,) * 3)
    tp = TGVParams()
    ic = oracle.tgv_ic(g.mesh.coord, oracle.tgv_params())
    runs = []
    for geo16 in (1, 0):
        run = TGVRun(3, geometry=g, params=tp, initial=ic, sensor_path="live", lps=True,
                     dt_mode="fixed")
        autotune.set_config(run.ctx, (8, geo16, 128))
        for _ in range(3):
            run.step()
        runs.append(run.state())
        autotune.set_config(run.ctx, (16, 1, 256))
    for k in ("rho", "q", "E", "mu_e"):
        assert np.array_equal(runs[0][k], runs[1][k]), k
