"""Parity of the live-path sensor entry points with the CPU oracle through the
C-ABI: smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227),
eval_divergence (mod_operators.f90:485-570), compute_vorticity
(mod_operators.f90:684-775) and ducros_visc_limit (mod_entropy_viscosity.f90:
277-314).

Bar: FP64 normwise relative error per output <= 1e-13 against the oracle (the
device sums the line derivatives in another order and contracts products into
FMAs, so agreement is to rounding)."""
import numpy as np
import pytest
import torch

from helpers import fz, make_case, relerr
from sod2d_amd import case

pytestmark = pytest.mark.gpu


def _dev(a):
    from sod2d_amd.device import to_dev
    return to_dev(a)


def _sensor_inputs(m, seed):
    r = np.random.default_rng(seed)
    n = m.npoin
    rho = r.uniform(0.9, 1.1, n)
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    c = r.uniform(1.0, 1.2, n)
    eta = r.normal(0.3, 1.0, n)
    Reta = r.normal(0, 1.0, n) * 10.0 ** r.uniform(-4, 2, n)
    return rho, u, c, eta, Reta


@pytest.mark.parametrize("p,normalise,high_mach", [(3, 1, 0), (3, 0, 0), (3, 1, 1), (2, 1, 0),
                                                   (4, 1, 0)])
def test_smart_visc_spectral_imex(oracle, p, normalise, high_mach):
    from sod2d_amd import device as D
    g = make_case(p, 3, periodic=(p != 4), jitter=0.0 if p != 4 else 0.05)
    m = g.mesh
    nn = (p + 1) ** 3
    rho, u, c, eta, Reta = _sensor_inputs(m, 40 + p)
    mu0 = np.asfortranarray(np.random.default_rng(9).uniform(0, 1e-3, (m.nelem, nn)))
    prm_o = oracle.visc_params(1.0, ce=0.1, ce_comp=0.6, cmax=0.5, cglob=0.9,
                               flag_normalise_entropy=normalise)
    omu, oml = oracle.smart_visc_spectral_imex(oracle.Tables(p), m.connec, m.lpoin_w, Reta, rho, u,
                                               c, eta, g.helem, prm_o, mu0.copy(order="F"),
                                               flag_high_mach=high_mach)
    prm = D.visc_params(1.0, ce=0.1, cmax=0.5, cglob=0.9, flag_normalise_entropy=normalise)
    dmu = _dev(mu0.copy(order="F"))
    dml = torch.zeros_like(dmu)
    D.smart_visc_spectral_imex(m.nelem, m.npoin, m.npoin_w, _dev(m.connec), _dev(m.lpoin_w),
                               _dev(Reta), None, None, None, None, None, 1.4, _dev(rho), _dev(u),
                               _dev(c), None, _dev(eta), None, _dev(g.helem), None, dmu, None,
                               None, None, None, dml, prm, ce_comp=0.6,
                               flag_high_mach=high_mach)
    torch.cuda.synchronize()
    assert np.array_equal(dml.cpu().numpy(), mu0.ravel(order="F"))
    assert relerr(dmu.cpu().numpy(), omu.ravel(order="F")) < 1e-13
    assert prm.ce == prm_o.ce == (0.6 if (normalise and high_mach) else 0.1)
    # both branches of min(Ve, beta_e) occur
    conn = m.connec - 1
    umag = np.sqrt((u ** 2).sum(1))
    betae = ((rho[conn] * g.helem[:, None]) * (0.5 / p) * (umag[conn] + c[conn])).max(1)
    capped = np.isclose(omu, 0.9 * betae[:, None], rtol=1e-14, atol=0)
    assert 0 < capped.mean() < 1


@pytest.mark.parametrize("p,n,periodic,jitter", [(3, 4, True, 0.0), (3, (3, 2, 4), False, 0.1),
                                                 (2, 3, True, 0.0), (4, 2, False, 0.08)])
def test_divergence_and_vorticity(oracle, p, n, periodic, jitter):
    from sod2d_amd import device as D
    g = make_case(p, n, periodic=periodic, jitter=jitter)
    m, T = g.mesh, g.tables
    OT = oracle.Tables(p)
    u = np.asfortranarray(np.random.default_rng(50 + p).normal(0, 0.3, (m.npoin, 3)))
    lelpn = case.elem_per_node(m.connec, m.npoin)
    odiv = oracle.eval_divergence(OT, m.connec, g.He, g.gpvol, u)
    ocurl = oracle.compute_vorticity(OT, m.connec, m.lpoin_w, lelpn, g.He, u)
    args = dict(He=_dev(g.He), gpvol=_dev(g.gpvol), dl=_dev(T.dlxigp_ip))
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    dc, dlw, dle, du = _dev(m.connec), _dev(m.lpoin_w), _dev(lelpn), _dev(u)
    ddiv = torch.full((m.npoin,), 7.0, dtype=torch.float64, device="cuda")
    dcurl = torch.full((3 * m.npoin,), 7.0, dtype=torch.float64, device="cuda")
    D.eval_divergence(m.nelem, m.npoin, dc, args["He"], args["gpvol"], args["dl"], *tabs, du, ddiv)
    D.compute_vorticity(m.nelem, m.npoin, m.npoin_w, dlw, dc, dle, args["He"], args["dl"], *tabs,
                        du, dcurl)
    torch.cuda.synchronize()
    assert relerr(ddiv.cpu().numpy(), odiv) < 1e-13
    assert relerr(dcurl.cpu().numpy(), ocurl.ravel(order="F")) < 1e-13
    # the fused pass gives the same bits as the two separate calls
    d2 = torch.zeros_like(ddiv)
    c2 = torch.zeros_like(dcurl)
    D.div_curl(m.nelem, m.npoin, m.npoin_w, dlw, dc, dle, args["He"], args["gpvol"], args["dl"],
               *tabs, du, d2, c2)
    torch.cuda.synchronize()
    assert torch.equal(d2, ddiv) and torch.equal(c2, dcurl)


def test_ducros_visc_limit(oracle):
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 3, periodic=True)
    m = g.mesh
    r = np.random.default_rng(61)
    n, nn = m.npoin, 64
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    divU = r.normal(0, 1.0, n) * 10.0 ** r.uniform(-3, 1, n)
    vort = np.asfortranarray(r.normal(0, 1.0, (n, 3)))
    Ml = r.uniform(0.5, 2.0, n)
    mu0 = np.asfortranarray(r.uniform(0.0, 1e-3, (m.nelem, nn)))
    omu = oracle.ducros_visc_limit(oracle.Tables(p), m.connec, u, divU, vort, Ml, g.helem_l,
                                   mu0.copy(order="F"), 0.1)
    dmu = _dev(mu0.copy(order="F"))
    D.ducros_visc_limit(m.nelem, n, m.npoin_w, _dev(m.connec), _dev(m.lpoin_w), None, None, None,
                        None, None, None, _dev(u), _dev(divU), _dev(vort), None, _dev(g.helem_l),
                        None, _dev(Ml), dmu, None, None, None, None, None, 0.1)
    torch.cuda.synchronize()
    assert relerr(dmu.cpu().numpy(), omu.ravel(order="F")) < 1e-13
