"""Parity of the HIP drop-in operators with the CPU oracle, through the C-ABI.

Bar: FP64 normwise relative error per output field <= 1e-13 against the oracle
(the HIP path evaluates the same split forms through exact algebraic identities
and a different summation order, so agreement is to rounding, not bitwise);
FP32 entry points <= 2e-5 against the FP64 oracle.  Assembly is a node-ordered
gather, so repeated calls are bitwise identical.
"""
import numpy as np
import pytest
import torch

from helpers import fz, make_case, random_state, relerr

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float64):
    from sod2d_amd.device import to_dev
    t = to_dev(a)
    return t.to(dtype) if t.dtype.is_floating_point else t


CASES = [(3, 4, True, 0.0), (3, (3, 2, 4), False, 0.1), (2, 3, True, 0.0), (4, 2, False, 0.08)]


@pytest.mark.parametrize("p,n,periodic,jitter", CASES)
def test_full_convec_ijk_f64(oracle, p, n, periodic, jitter):
    from sod2d_amd import device as D
    g = make_case(p, n, periodic=periodic, jitter=jitter)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=p)
    OT = oracle.Tables(p)
    Rm, Rq, Re = oracle.full_convec_ijk(OT, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"],
                                        s["pr"], s["h"], fact=-1.0)
    dRm, dRq, dRe = (torch.zeros(k, dtype=torch.float64, device="cuda")
                     for k in (m.npoin, 3 * m.npoin, m.npoin))
    D.full_convec_ijk(m.nelem, m.npoin, _dev(m.connec), None, None, _dev(g.He), _dev(g.gpvol),
                      _dev(T.dlxigp_ip), None, _dev(T.invAtoIJK), _dev(T.gmshAtoI),
                      _dev(T.gmshAtoJ), _dev(T.gmshAtoK), _dev(s["u"]), _dev(s["q"]),
                      _dev(s["rho"]), _dev(s["pr"]), _dev(s["h"]), dRm, dRq, dRe, True, -1.0)
    torch.cuda.synchronize()
    assert relerr(dRm.cpu().numpy(), Rm) < 1e-13
    assert relerr(dRq.cpu().numpy(), Rq.ravel(order="F")) < 1e-13
    assert relerr(dRe.cpu().numpy(), Re) < 1e-13


def test_full_convec_accumulate_and_saved_fact(oracle):
    """initialze=.false. accumulates; an omitted fact reuses the last one given."""
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 3, periodic=True)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=11)
    OT = oracle.Tables(p)
    Rm, Rq, Re = oracle.full_convec_ijk(OT, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"],
                                        s["pr"], s["h"], fact=-1.0)
    Rm, Rq, Re = oracle.full_convec_ijk(OT, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"],
                                        s["pr"], s["h"], Rm, Rq, Re, initialize=False, fact=-1.0)
    args = [_dev(x) for x in (m.connec, g.He, g.gpvol, T.dlxigp_ip)]
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    flds = [_dev(s[k]) for k in ("u", "q", "rho", "pr", "h")]
    dRm, dRq, dRe = (torch.zeros(k, dtype=torch.float64, device="cuda")
                     for k in (m.npoin, 3 * m.npoin, m.npoin))
    D.full_convec_ijk(m.nelem, m.npoin, args[0], None, None, args[1], args[2], args[3], None,
                      *tabs, *flds, dRm, dRq, dRe, True, -1.0)
    D.full_convec_ijk(m.nelem, m.npoin, args[0], None, None, args[1], args[2], args[3], None,
                      *tabs, *flds, dRm, dRq, dRe, False)  # fact omitted -> -1 kept
    torch.cuda.synchronize()
    assert relerr(dRm.cpu().numpy(), Rm) < 1e-13
    assert relerr(dRq.cpu().numpy(), Rq.ravel(order="F")) < 1e-13
    assert relerr(dRe.cpu().numpy(), Re) < 1e-13


@pytest.mark.parametrize("p,n,periodic,jitter", CASES)
def test_full_diffusion_ijk_f64(oracle, p, n, periodic, jitter):
    from sod2d_amd import device as D
    g = make_case(p, n, periodic=periodic, jitter=jitter)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=3 + p)
    OT = oracle.Tables(p)
    Cp, Pr = 1004.0, 0.71
    # rho_n differs from rho to pin which density each term uses
    rho_n = s["rho"] * 1.01
    Rm, Rq, Re = oracle.full_diffusion_ijk(OT, m.connec, g.He, g.gpvol, Cp, Pr, rho_n, s["rho"],
                                           s["u"], s["Tem"], s["mu_fluid"], s["mu_e"], s["mu_sgs"],
                                           fact=-1.0)
    dRm, dRq, dRe = (torch.zeros(k, dtype=torch.float64, device="cuda")
                     for k in (m.npoin, 3 * m.npoin, m.npoin))
    D.full_diffusion_ijk(m.nelem, m.npoin, _dev(m.connec), None, _dev(g.He), _dev(g.gpvol),
                         _dev(T.dlxigp_ip), _dev(T.invAtoIJK), _dev(T.gmshAtoI), _dev(T.gmshAtoJ),
                         _dev(T.gmshAtoK), Cp, Pr, _dev(rho_n), _dev(s["rho"]), _dev(s["u"]),
                         _dev(s["Tem"]), _dev(s["mu_fluid"]), _dev(s["mu_e"]), _dev(s["mu_sgs"]),
                         None, dRm, dRq, dRe, True, -1.0)
    torch.cuda.synchronize()
    assert relerr(dRm.cpu().numpy(), Rm) < 1e-13
    assert relerr(dRq.cpu().numpy(), Rq.ravel(order="F")) < 1e-13
    assert relerr(dRe.cpu().numpy(), Re) < 1e-13


def test_generic_scalar_convec_f64(oracle):
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, (2, 3, 2), periodic=False, jitter=0.1)
    m, T = g.mesh, g.tables
    r = np.random.default_rng(5)
    eta = r.normal(size=m.npoin)
    u = np.asfortranarray(r.normal(size=(m.npoin, 3)))
    f = np.asfortranarray(u * eta[:, None])
    OT = oracle.Tables(p)
    R = oracle.generic_scalar_convec_ijk(OT, m.connec, g.He, g.gpvol, f, eta, u)
    dR = torch.zeros(m.npoin, dtype=torch.float64, device="cuda")
    D.generic_scalar_convec_ijk(m.nelem, m.npoin, _dev(m.connec), None, None, _dev(g.He),
                                _dev(g.gpvol), _dev(T.dlxigp_ip), None, _dev(T.invAtoIJK),
                                _dev(T.gmshAtoI), _dev(T.gmshAtoJ), _dev(T.gmshAtoK), _dev(f),
                                _dev(eta), _dev(u), dR)
    torch.cuda.synchronize()
    assert relerr(dR.cpu().numpy(), R) < 1e-13
    # q = NULL: the flux u * eta formed in the kernel gives the same bits
    dR2 = torch.zeros_like(dR)
    D.generic_scalar_convec_ijk(m.nelem, m.npoin, _dev(m.connec), None, None, _dev(g.He),
                                _dev(g.gpvol), _dev(T.dlxigp_ip), None, _dev(T.invAtoIJK),
                                _dev(T.gmshAtoI), _dev(T.gmshAtoJ), _dev(T.gmshAtoK), None,
                                _dev(eta), _dev(u), dR2)
    torch.cuda.synchronize()
    assert torch.equal(dR, dR2)


def test_f32_entry_points(oracle):
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 3, periodic=True)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=21)
    OT = oracle.Tables(p)
    Rm, Rq, Re = oracle.full_convec_ijk(OT, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"],
                                        s["pr"], s["h"], fact=-1.0)
    f32 = torch.float32
    dRm, dRq, dRe = (torch.zeros(k, dtype=f32, device="cuda")
                     for k in (m.npoin, 3 * m.npoin, m.npoin))
    D.full_convec_ijk(m.nelem, m.npoin, _dev(m.connec), None, None, _dev(g.He, f32),
                      _dev(g.gpvol, f32), _dev(T.dlxigp_ip, f32), None, _dev(T.invAtoIJK),
                      _dev(T.gmshAtoI), _dev(T.gmshAtoJ), _dev(T.gmshAtoK), _dev(s["u"], f32),
                      _dev(s["q"], f32), _dev(s["rho"], f32), _dev(s["pr"], f32),
                      _dev(s["h"], f32), dRm, dRq, dRe, True, -1.0)
    torch.cuda.synchronize()
    assert relerr(dRm.cpu().numpy(), Rm) < 2e-5
    assert relerr(dRq.cpu().numpy(), Rq.ravel(order="F")) < 2e-5
    assert relerr(dRe.cpu().numpy(), Re) < 2e-5


def test_deterministic_assembly():
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 4, periodic=True)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=9)
    args = [_dev(x) for x in (m.connec, g.He, g.gpvol, T.dlxigp_ip)]
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    flds = [_dev(s[k]) for k in ("u", "q", "rho", "pr", "h")]
    outs = []
    for _ in range(3):
        R = [torch.zeros(k, dtype=torch.float64, device="cuda")
             for k in (m.npoin, 3 * m.npoin, m.npoin)]
        D.full_convec_ijk(m.nelem, m.npoin, args[0], None, None, args[1], args[2], args[3], None,
                          *tabs, *flds, *R, True, -1.0)
        outs.append(torch.cat(R).cpu())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_lumped_and_stage_update(oracle):
    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 3, periodic=True)
    m = g.mesh
    r = np.random.default_rng(2)
    n = m.npoin
    rho1, E1, rho, E, Rm, Re = (r.normal(size=n) for _ in range(6))
    q1, q, Rq = (np.asfortranarray(r.normal(size=(n, 3))) for _ in range(3))
    L, G = oracle.rk_ls_coeffs(5, 1)
    dt = 1e-3
    l1, ls, gm = L[4, 0], L[4, 3], G[4, 3]
    # oracle: lumped solves then the stage combine
    oRm, oRq, oRe = Rm.copy(), Rq.copy(order="F"), Re.copy()
    oracle.lumped_solver_scal(m.lpoin_w, g.invMl, oRm)
    oracle.lumped_solver_scal(m.lpoin_w, g.invMl, oRe)
    oracle.lumped_solver_vect(m.lpoin_w, g.invMl, oRq)
    orho, oq, oE = rho.copy(), q.copy(order="F"), E.copy()
    oracle.rk_ls_update(m.lpoin_w, l1, ls, gm * dt, rho1, q1, E1, orho, oq, oE, oRm, oRq, oRe)
    dv = {k: _dev(v) for k, v in dict(rho1=rho1, q1=q1, E1=E1, rho=rho, q=q, E=E, Rm=Rm, Rq=Rq,
                                      Re=Re).items()}
    D.rk_stage_update(n, m.npoin_w, _dev(m.lpoin_w), _dev(g.invMl), dt, l1, ls, gm, dv["rho1"],
                      dv["q1"], dv["E1"], dv["rho"], dv["q"], dv["E"], dv["Rm"], dv["Rq"], dv["Re"])
    torch.cuda.synchronize()
    assert relerr(dv["rho"].cpu().numpy(), orho) < 1e-15
    assert relerr(dv["q"].cpu().numpy(), oq.ravel(order="F")) < 1e-15
    assert relerr(dv["E"].cpu().numpy(), oE) < 1e-15
    assert relerr(dv["Rq"].cpu().numpy(), oRq.ravel(order="F")) < 1e-15


@pytest.mark.parametrize("p,normalise", [(3, 1), (3, 0), (2, 1), (4, 1)])
def test_smart_visc_spectral_f64(oracle, p, normalise):
    from sod2d_amd import device as D
    g = make_case(p, 3, periodic=(p != 4), jitter=0.0 if p != 4 else 0.05)
    m, T = g.mesh, g.tables
    r = np.random.default_rng(13 + p)
    n, ne, nn = m.npoin, m.nelem, T.nnode
    rho = r.uniform(0.9, 1.1, n)
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    Tem = r.uniform(0.9, 1.1, n)
    eta = r.normal(0, 1.0, n)
    # Reta spanning both branches of min(Ve, beta_e)
    Reta = r.normal(0, 1.0, n) * 10.0 ** r.uniform(-2, 4, n)
    gamma, Rgas = 1.4, 287.0
    prm_o = oracle.visc_params(Rgas, flag_normalise_entropy=normalise)
    mu_e = fz((ne, nn))
    mu_e, mue_l = oracle.smart_visc_spectral(oracle.Tables(p), m.connec, m.lpoin_w, Reta, gamma, rho,
                                             u, Tem, eta, g.helem_l, g.helem, prm_o, mu_e)
    dmu = torch.zeros(ne * nn, dtype=torch.float64, device="cuda")
    dml = torch.zeros(ne * nn, dtype=torch.float64, device="cuda")
    prm = D.visc_params(Rgas, flag_normalise_entropy=normalise)
    D.smart_visc_spectral(ne, n, m.npoin_w, _dev(m.connec), _dev(m.lpoin_w), _dev(Reta), None,
                          None, None, None, None, gamma, _dev(rho), _dev(u), None, _dev(Tem),
                          _dev(eta), _dev(g.helem_l), _dev(g.helem), None, dmu,
                          _dev(T.invAtoIJK), None, None, None, dml, prm)
    torch.cuda.synchronize()
    assert relerr(dml.cpu().numpy(), mue_l.ravel(order="F")) < 1e-13
    assert relerr(dmu.cpu().numpy(), mu_e.ravel(order="F")) < 1e-13
    # both limiter branches are exercised
    lim = mue_l.ravel(order="F")
    assert 0 < np.mean(lim == lim.max()) < 1


def test_mesh_cache_follows_freed_connectivity(oracle):
    """A connectivity freed and replaced by a different mesh of the same size
    (here the same elements in reversed order, which torch's caching allocator
    puts at the same address) is not matched to the first mesh's cached
    incidence tables."""
    import gc

    from sod2d_amd import device as D
    p = 3
    g = make_case(p, 3, periodic=True)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=5)
    OT = oracle.Tables(p)
    tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    flds = [_dev(s[k]) for k in ("u", "q", "rho", "pr", "h")]
    dl = _dev(T.dlxigp_ip)

    def run(connec, He, gpvol):
        c = _dev(connec)
        out = [torch.zeros(k, dtype=torch.float64, device="cuda")
               for k in (m.npoin, 3 * m.npoin, m.npoin)]
        D.full_convec_ijk(m.nelem, m.npoin, c, None, None, _dev(He), _dev(gpvol), dl, None,
                          *tabs, *flds, *out, True, -1.0)
        torch.cuda.synchronize()
        return c.data_ptr(), [x.cpu().numpy() for x in out]

    ptr_a, _ = run(m.connec, g.He, g.gpvol)
    gc.collect()
    rev = np.asfortranarray(m.connec[::-1, :])
    He_r = np.asfortranarray(g.He[..., ::-1])
    gv_r = np.asfortranarray(g.gpvol[..., ::-1])
    ptr_b, (dRm, dRq, dRe) = run(rev, He_r, gv_r)
    Rm, Rq, Re = oracle.full_convec_ijk(OT, rev, He_r, gv_r, s["u"], s["q"], s["rho"], s["pr"],
                                        s["h"], fact=-1.0)
    print(f"connec addresses: first {ptr_a:#x}, second {ptr_b:#x}")
    assert relerr(dRm, Rm) < 1e-13
    assert relerr(dRq, Rq.ravel(order="F")) < 1e-13
    assert relerr(dRe, Re) < 1e-13
