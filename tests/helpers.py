"""Shared test helpers: case construction and random reference-layout fields."""
import numpy as np

from sod2d_amd import case


def fz(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def make_case(p, n, periodic=True, jitter=0.0, lengths=None, seed=7):
    return case.build_case(p, n, periodic=periodic, jitter=jitter, lengths=lengths, seed=seed)


def random_state(npoin, nelem, nnode, seed=1, gamma=1.4, Rgas=1.0):
    """Primitive fields as the operator-bench inputs: rho~U(0.9,1.1), u~N(0,0.3),
    p~U(0.9,1.1)/gamma, E slot = enthalpy, T = p/(rho R)."""
    r = np.random.default_rng(seed)
    rho = r.uniform(0.9, 1.1, npoin)
    u = np.asfortranarray(r.normal(0.0, 0.3, (npoin, 3)))
    pr = r.uniform(0.9, 1.1, npoin) / gamma
    q = np.asfortranarray(rho[:, None] * u)
    h = gamma / (gamma - 1.0) * pr / rho
    Tem = pr / (rho * Rgas)
    mu_fluid = r.uniform(1e-4, 2e-4, npoin)
    mu_e = np.asfortranarray(r.uniform(0.0, 1e-3, (nelem, nnode)))
    mu_sgs = np.asfortranarray(r.uniform(0.0, 1e-4, (nelem, nnode)))
    E = pr / (gamma - 1.0) + 0.5 * rho * (u * u).sum(1)
    return dict(rho=rho, u=u, q=q, pr=pr, h=h, E=E, Tem=Tem, mu_fluid=mu_fluid, mu_e=mu_e,
                mu_sgs=mu_sgs)


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
