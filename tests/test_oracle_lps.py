"""Pins of the oracle's LPS stabilisation (elem_stab.f90:207-438, 441-586,
589-616) against what the mathematics fixes; the reference module needs MPI
and no compiled reference is available for it.

* comp_tau of a uniform state: helem_k (c_species_stab/p) (c + |u|), and
  helem_k (c_species_stab/p) |u| with the sound speed dropped.
* full_proj_ijk of linear rho, T and u on meshes whose elements map
  polynomially (jittered included): the gradients are exact at every
  quadrature point, so the lumped projections equal grad rho, grad T and the
  rows of 2 S - 2/3 div u I at every working node.
* full_diff_stab_ijk:
  - with tau_stab = 0, mu_e = mu_sgs = 0 and a uniform mu_fluid it is
    full_diffusion_ijk (the two differ only in where the entropy-viscosity
    coefficients are applied);
  - for linear fields the LPS fluxes vanish (Proj = grad), so the result does
    not depend on tau_stab;
  - the LPS mass flux is dissipative with the magnitude the projection fixes:
    with fact = +1, rho_n = 1 and no viscosity,
        sum_n rho_n Rmass_n = c_lps tau (sum_g w |grad rho|^2 - sum_n Ml |P|^2),
    where the first sum is sum_n rho_n Rmass_n of full_diffusion_ijk with
    nu_e = 1 and the second is >= 0 smaller (Jensen); the energy flux the same
    with the factor Cp/Pr;
  - on a periodic mesh every residual sums to zero.
A sign error, a dropped scale, a transposed gradient or a wrong line weight
fails one of these.
"""
import numpy as np
import pytest

from helpers import make_case


def _linear(coord, A, b):
    return np.asfortranarray(coord @ A.T + b)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_comp_tau_uniform_state(oracle, p):
    g = make_case(p, 2, periodic=True)
    m = g.mesh
    U = np.array([0.3, -0.4, 1.2])
    u = np.asfortranarray(np.tile(U, (m.npoin, 1)))
    c = np.full(m.npoin, 1.7)
    tau = oracle.comp_tau(p, m.connec, c, u, g.helem)
    want = g.helem * (0.5 / p) * (1.7 + np.linalg.norm(U))
    assert np.allclose(tau, want, rtol=1e-15, atol=0)
    tau = oracle.comp_tau(p, m.connec, c, u, g.helem, drop_c=True)
    assert np.allclose(tau, g.helem * (0.5 / p) * np.linalg.norm(U), rtol=1e-15, atol=0)


@pytest.mark.parametrize("p,n,periodic,jitter", [(3, 3, False, 0.1), (2, 3, False, 0.08),
                                                 (4, 2, False, 0.05), (3, 3, True, 0.0)])
def test_projection_of_linear_fields(oracle, p, n, periodic, jitter):
    g = make_case(p, n, periodic=periodic, jitter=jitter, seed=5)
    m = g.mesh
    r = np.random.default_rng(2)
    A, a, c = r.normal(size=(3, 3)), r.normal(size=3), r.normal(size=3)
    if periodic:  # a periodic box holds constants only
        A[...] = 0.0
        a[...] = 0.0
        c[...] = 0.0
    u = _linear(m.coord, A, r.normal(size=3))
    rho = 1.0 + 0.1 * (m.coord @ a)
    T = 2.0 + m.coord @ c
    P = oracle.full_proj_ijk(oracle.Tables(p), m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, T,
                             g.invMl)
    w = m.lpoin_w - 1
    S = A + A.T - (2.0 / 3.0) * np.trace(A) * np.eye(3)
    scale = 1.0 + np.abs(A).max()
    assert np.abs(P[0][w] - 0.1 * a).max() < 1e-11 * scale
    assert np.abs(P[1][w] - c).max() < 1e-11 * scale
    for i in range(3):
        assert np.abs(P[2 + i][w] - S[i]).max() < 1e-11 * scale


def _fields(m, seed):
    r = np.random.default_rng(seed)
    n = m.npoin
    rho = r.uniform(0.9, 1.1, n)
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    T = r.uniform(0.9, 1.1, n)
    return rho, u, T


@pytest.mark.parametrize("p", [2, 3, 4])
def test_diff_stab_reduces_to_diffusion(oracle, p):
    g = make_case(p, 3, periodic=False, jitter=0.05)
    m = g.mesh
    OT = oracle.Tables(p)
    rho, u, T = _fields(m, 3)
    nn = OT.nnode
    mu_f = np.full(m.npoin, 0.02)
    z = np.zeros((m.nelem, nn), order="F")
    P = oracle.full_proj_ijk(OT, m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, T, g.invMl)
    Rs = oracle.full_diff_stab_ijk(OT, m.connec, g.He, g.gpvol, 1004.0, 0.71, rho, rho, u, T,
                                   mu_f, z, z, P, np.zeros(m.nelem))
    Rd = oracle.full_diffusion_ijk(OT, m.connec, g.He, g.gpvol, 1004.0, 0.71, rho, rho, u, T,
                                   mu_f, z, z)
    for a, b in zip(Rs, Rd):
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_diff_stab_lps_vanishes_for_linear_fields(oracle, p):
    g = make_case(p, 3, periodic=False, jitter=0.08, seed=9)
    m = g.mesh
    OT = oracle.Tables(p)
    r = np.random.default_rng(4)
    A = r.normal(size=(3, 3))
    u = _linear(m.coord, A, r.normal(size=3))
    rho = 1.0 + 0.05 * (m.coord @ r.normal(size=3))
    T = 1.0 + 0.1 * (m.coord @ r.normal(size=3))
    nn = OT.nnode
    mu_f = r.uniform(1e-3, 2e-3, m.npoin)
    mu_e = np.asfortranarray(r.uniform(0, 1e-3, (m.nelem, nn)))
    z = np.zeros((m.nelem, nn), order="F")
    P = oracle.full_proj_ijk(OT, m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, T, g.invMl)
    R0 = oracle.full_diff_stab_ijk(OT, m.connec, g.He, g.gpvol, 1004.0, 0.71, rho, rho, u, T,
                                   mu_f, mu_e, z, P, np.zeros(m.nelem))
    R1 = oracle.full_diff_stab_ijk(OT, m.connec, g.He, g.gpvol, 1004.0, 0.71, rho, rho, u, T,
                                   mu_f, mu_e, z, P, r.uniform(0.5, 2.0, m.nelem))
    for a, b in zip(R1, R0):
        assert np.linalg.norm(a - b) <= 1e-11 * np.linalg.norm(b)


@pytest.mark.parametrize("p,periodic", [(3, True), (2, False), (4, True)])
def test_lps_dissipation_identity(oracle, p, periodic):
    g = make_case(p, 3 if p < 4 else 2, periodic=periodic, jitter=0.0 if periodic else 0.05)
    m = g.mesh
    OT = oracle.Tables(p)
    rho, u, T = _fields(m, 7)
    nn = OT.nnode
    z = np.zeros((m.nelem, nn), order="F")
    zn = np.zeros(m.npoin)
    one = np.ones(m.npoin)
    tau = 0.37
    c_lps, Cp, Pr = oracle.C_LPS_COMP, 1004.0, 0.71
    P = oracle.full_proj_ijk(OT, m.connec, m.lpoin_w, g.He, g.gpvol, rho, u, T, g.invMl)
    Rm, Rq, Re = oracle.full_diff_stab_ijk(OT, m.connec, g.He, g.gpvol, Cp, Pr, one, rho, u, T, zn,
                                           z, z, P, np.full(m.nelem, tau))
    # sum_g w |grad f|^2 through full_diffusion_ijk's density flux with nu_e = 1
    ones = np.ones((m.nelem, nn), order="F")
    w = m.lpoin_w - 1
    Ml = g.Ml

    def grad_sq(f):
        R, _, _ = oracle.full_diffusion_ijk(OT, m.connec, g.He, g.gpvol, Cp, Pr, one, f, u, T, zn,
                                            ones, z)
        return float(np.dot(f, R))

    for f, Rf, Pf, s in ((rho, Rm, P[0], 1.0), (T, Re, P[1], Cp / Pr)):
        proj_sq = float(np.dot(Ml[w], (Pf[w] ** 2).sum(1)))
        gsq = grad_sq(f)
        lhs = float(np.dot(f, Rf))
        assert gsq > proj_sq > 0.0
        want = c_lps * tau * s * (gsq - proj_sq)
        assert lhs > 0.0 and abs(lhs - want) <= 1e-10 * abs(want), (lhs, want)
    if periodic:
        for Rf in (Rm, Rq, Re):
            assert abs(Rf.sum()) <= 1e-12 * np.abs(Rf).sum()
