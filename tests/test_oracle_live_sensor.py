"""Pins of the oracle's live-path sensor pieces against what the mathematics
fixes (the reference modules that hold them need MPI, so no compiled
reference is available for them):

* eval_divergence / compute_vorticity of a linear velocity u = A x + b: on any
  mesh whose elements map polynomially of degree p (jittered included), u is
  in the GLL interpolation space, so div u = tr A and the reference's curl
  eps_ijk du_j/dx_k (leviCivi of CFDSolverBase.f90:1605-1611) are exact at
  every GLL point; hence divU / Ml = tr A and curlU = eps_ijk du_j/dx_k =
  (A_23 - A_32, A_31 - A_13, A_12 - A_21), which is -curl u, at every working
  node.  A transposed gradient, a dropped term or a sign error
  fails these.
* elemPerNode: incidence counts sum to nelem * nnode; interior nodes of a
  structured box have 1, 2, 4 or 8 elements.
* smart_visc_spectral_imex: both branches of min(Ve, beta_e) against hand
  evaluation, the low-Mach and high-Mach normalisations (the latter also
  switches ce to ce_comp), mue_l receiving the incoming mu_e.
* ducros_visc_limit: the scaling factor against a hand evaluation, including
  the no-divergence limit (factor = ducros_min_val) and the pure-dilatation
  limit (phi = 1).
"""
import numpy as np
import pytest

from helpers import make_case


def _linear_u(coord, A, b):
    return np.asfortranarray(coord @ A.T + b)


@pytest.mark.parametrize("p,n,periodic,jitter", [(3, 3, False, 0.1), (2, 4, False, 0.08),
                                                 (4, 2, False, 0.05), (3, 3, True, 0.0)])
def test_divergence_and_vorticity_of_linear_field(oracle, p, n, periodic, jitter):
    g = make_case(p, n, periodic=periodic, jitter=jitter, seed=11)
    m = g.mesh
    r = np.random.default_rng(5)
    A = r.normal(size=(3, 3))
    b = r.normal(size=3)
    if periodic:  # a periodic box holds only the constant part
        A = np.zeros((3, 3))
    u = _linear_u(m.coord, A, b)
    T = oracle.Tables(p)
    divU = oracle.eval_divergence(T, m.connec, g.He, g.gpvol, u)
    Ml = oracle.lumped_mass(p, m.connec, g.gpvol, m.npoin)
    w = m.lpoin_w - 1
    scale = max(1.0, np.abs(A).max())
    assert np.abs(divU[w] / Ml[w] - np.trace(A)).max() < 1e-11 * scale
    lelpn = oracle.elem_per_node(m.connec, m.npoin)
    curl = oracle.compute_vorticity(T, m.connec, m.lpoin_w, lelpn, g.He, u)
    # eps_ijk du_j/dx_k with du_j/dx_k = A[j, k]: minus the curl
    want = np.array([A[1, 2] - A[2, 1], A[2, 0] - A[0, 2], A[0, 1] - A[1, 0]])
    got = curl[w, :]
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1.0
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1.0
    ref = np.einsum("ijk,jk->i", eps, A)
    assert np.allclose(ref, want)
    curl_u = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.allclose(ref, -curl_u)
    assert np.abs(got - ref[None, :]).max() < 1e-11 * scale


def test_elem_per_node(oracle):
    g = make_case(3, 3, periodic=False)
    m = g.mesh
    lelpn = oracle.elem_per_node(m.connec, m.npoin)
    assert lelpn.sum() == m.nelem * 64
    assert set(np.unique(lelpn)) <= {1, 2, 4, 8}
    # periodic: every working node of a structured periodic box touches
    # 1, 2, 4 or 8 elements through the working connectivity
    gp = make_case(3, 3, periodic=True)
    lp = oracle.elem_per_node(gp.mesh.connec, gp.mesh.npoin)
    wp = gp.mesh.lpoin_w - 1
    assert set(np.unique(lp[wp])) <= {1, 2, 4, 8}
    assert lp.sum() == gp.mesh.nelem * 64


def _sensor_inputs(m, seed):
    r = np.random.default_rng(seed)
    n = m.npoin
    rho = r.uniform(0.9, 1.1, n)
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    c = r.uniform(1.0, 1.2, n)
    eta = r.normal(0.3, 1.0, n)
    return rho, u, c, eta


def test_smart_visc_spectral_imex_branches(oracle):
    p = 3
    g = make_case(p, 3, periodic=True)
    m = g.mesh
    T = oracle.Tables(p)
    rho, u, c, eta = _sensor_inputs(m, 21)
    hk = g.helem
    cmax, cglob, ce = 0.5, 0.9, 0.1
    prm = oracle.visc_params(1.0, ce=ce, cmax=cmax, cglob=cglob, flag_normalise_entropy=1)
    ne, nn = m.nelem, 64
    conn = m.connec - 1
    w = m.lpoin_w - 1
    mean = eta[w].sum() / len(w)
    norm = np.abs(eta[w] - mean).max()
    umag = np.sqrt((u ** 2).sum(1))
    betae = ((rho[conn] * hk[:, None]) * (cmax / p) * (umag[conn] + c[conn])).max(1)
    # Reta = 0: mu_e = 0; mue_l = incoming mu_e
    mu0 = np.asfortranarray(np.random.default_rng(1).uniform(0, 1, (ne, nn)))
    mu, ml = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, np.zeros(m.npoin), rho, u, c,
                                             eta, hk, prm, mu0.copy(order="F"))
    assert np.array_equal(ml, mu0)
    assert np.all(mu == 0.0)
    # small Reta: the Ve branch, evaluated by hand
    Reta = 1e-6 * np.random.default_rng(2).normal(size=m.npoin)
    mu, _ = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, Reta, rho, u, c, eta, hk, prm,
                                            np.zeros((ne, nn), order="F"))
    Ve = ce * (rho[conn] * np.abs(Reta[conn]) / norm) * (hk[:, None] / p) ** 2
    assert np.all(Ve < betae[:, None])
    assert np.allclose(mu, cglob * Ve, rtol=1e-14, atol=0)
    # huge Reta: beta_e everywhere
    mu, _ = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, 1e12 * np.ones(m.npoin), rho,
                                            u, c, eta, hk, prm, np.zeros((ne, nn), order="F"))
    assert np.allclose(mu, cglob * betae[:, None] * np.ones((1, nn)), rtol=1e-15, atol=0)
    # flag_drop_c_in_envit drops the sound speed from beta_e
    prm2 = oracle.visc_params(1.0, ce=ce, cmax=cmax, cglob=cglob, flag_drop_c_in_envit=1)
    mu, _ = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, 1e12 * np.ones(m.npoin), rho,
                                            u, c, eta, hk, prm2, np.zeros((ne, nn), order="F"))
    b2 = ((rho[conn] * hk[:, None]) * (cmax / p) * umag[conn]).max(1)
    assert np.allclose(mu, cglob * b2[:, None] * np.ones((1, nn)), rtol=1e-15, atol=0)


def test_smart_visc_spectral_imex_high_mach(oracle):
    p = 3
    g = make_case(p, 3, periodic=True)
    m = g.mesh
    T = oracle.Tables(p)
    rho, u, c, eta = _sensor_inputs(m, 23)
    hk = g.helem
    prm = oracle.visc_params(1.0, ce=0.1, ce_comp=0.7, cmax=0.5, cglob=1.0,
                             flag_normalise_entropy=1)
    Reta = 1e-6 * np.random.default_rng(3).normal(size=m.npoin)
    mu, _ = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, Reta, rho, u, c, eta, hk, prm,
                                            np.zeros((m.nelem, 64), order="F"), flag_high_mach=1)
    assert prm.ce == 0.7
    conn = m.connec - 1
    norm = rho[m.lpoin_w - 1].min()
    Ve = 0.7 * (rho[conn] * np.abs(Reta[conn]) / norm) * (hk[:, None] / p) ** 2
    assert np.allclose(mu, Ve, rtol=1e-14, atol=0)


def test_ducros_factor(oracle):
    p = 2
    g = make_case(p, 2, periodic=False)
    m = g.mesh
    T = oracle.Tables(p)
    r = np.random.default_rng(31)
    n = m.npoin
    u = np.asfortranarray(r.normal(0, 0.3, (n, 3)))
    divU = r.normal(0, 1.0, n)
    vort = np.asfortranarray(r.normal(0, 1.0, (n, 3)))
    Ml = r.uniform(0.5, 2.0, n)
    hl = np.asfortranarray(r.uniform(0.1, 0.2, (m.nelem, 27)))
    mu0 = np.asfortranarray(r.uniform(0.0, 1.0, (m.nelem, 27)))
    dmin = 0.1
    mu = oracle.ducros_visc_limit(T, m.connec, u, divU, vort, Ml, hl, mu0.copy(order="F"), dmin)
    conn = m.connec - 1
    dv = divU[conn] / Ml[conn]
    wsq = (vort ** 2).sum(1)[conn]
    um = np.sqrt((u ** 2).sum(1))[conn]
    eps = np.finfo(np.float64).eps
    phi = np.minimum((4.0 / 3.0) * dv ** 2 / (dv ** 2 + wsq + eps), 1.0)
    phi2 = dv ** 2 / (dv ** 2 + (um / hl) ** 2 + eps)
    want = mu0 * (np.maximum(phi - dmin, 0.0) * phi2 + dmin)
    assert np.allclose(mu, want, rtol=1e-14, atol=0)
    # no dilatation: the factor is ducros_min_val
    mu = oracle.ducros_visc_limit(T, m.connec, u, np.zeros(n), vort, Ml, hl, mu0.copy(order="F"),
                                  dmin)
    assert np.allclose(mu, dmin * mu0, rtol=1e-15, atol=0)
    # pure dilatation, still fluid: phi = 1, phi2 = 1, factor = 1
    mu = oracle.ducros_visc_limit(T, m.connec, np.zeros((n, 3), order="F"), 1e3 * np.ones(n),
                                  np.zeros((n, 3), order="F"), np.ones(n), hl, mu0.copy(order="F"),
                                  dmin)
    assert np.allclose(mu, mu0, rtol=1e-14, atol=0)
