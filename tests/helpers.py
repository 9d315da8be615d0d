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
    return case.operator_bench_fields(npoin, nelem, nnode, seed, gamma, Rgas)


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
