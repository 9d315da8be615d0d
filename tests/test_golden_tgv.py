"""The reference's own TGV test data (Tests/TGV_Compressible/*/Reference/
analysis_cube-*.dat, copied to tests/golden/tgv_compressible/): the t = 0 row
of each file, regenerated on the same mesh (periodic cube of side
2*3.14159265, n^3 elements, p = 4) from TGVCompSolver's initial conditions
(TGVCompSolver.f90:154-214) at M = 1.25, through the oracle's diagnostics
(volAvg_EK, visc_dissipationRate, maxMach; mod_analysis.f90:10-168) and, on
the GPU, through the device diagnostics.

The files were written by FP32 builds with 8 decimals, so the bars are
absolute: 1e-7 on EK and maxmach, 1e-8 on the dissipation rates.  Later rows
come from runs with LPS stabilisation and FP32 state and are not pinned here.
"""
import json
import math
import os

import numpy as np
import pytest

from sod2d_amd import case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tgv_compressible")
META = json.load(open(os.path.join(GOLD, "cases.json")))
PI_MESH = 3.14159265  # cube.geo:1


def comp_params(M=1.25, V0=1.0, L=1.0, T0=1.0, Re=1600.0, gamma=1.4):
    """TGVCompSolver_initializeParameters (TGVCompSolver.f90:130-138)."""
    Rgas = V0 * V0 / (gamma * T0 * M * M)
    Cp = gamma * Rgas / (gamma - 1.0)
    rho0 = (1.0 / (gamma * M * M)) / (Rgas * T0)
    mul = rho0 * V0 * L / Re
    P0 = 1.0 / (M * M * gamma)
    mur = 0.000001458 * T0 ** 1.5 / (T0 + 110.40)
    return dict(M=M, V0=V0, L=L, T0=T0, gamma=gamma, Rgas=Rgas, Cp=Cp, rho0=rho0, P0=P0,
                mu_factor=mul / mur)


def comp_ic(coord, c):
    """TGVCompSolver_evalInitialConditions (TGVCompSolver.f90:154-214): the
    coordinates are taken as real(4)."""
    x, y, z = (coord[:, d].astype(np.float32).astype(np.float64) for d in range(3))
    V0, L = c["V0"], c["L"]
    u = np.zeros((coord.shape[0], 3), order="F")
    u[:, 0] = V0 * np.sin(x / L) * np.cos(y / L) * np.cos(z / L)
    u[:, 1] = -V0 * np.cos(x / L) * np.sin(y / L) * np.cos(z / L)
    pr = c["P0"] + (V0 * V0 / 16.0) * (np.cos(2 * x / L) + np.cos(2 * y / L)) * \
        (np.cos(2 * z / L) + 2.0)
    rho = pr / c["Rgas"] / c["T0"]
    g = c["gamma"]
    e_int = pr / (rho * (g - 1.0))
    Tem = pr / (rho * c["Rgas"])
    E = rho * (0.5 * (u * u).sum(1) + e_int)
    q = np.asfortranarray(rho[:, None] * u)
    csound = np.sqrt(g * pr / rho)
    machno = np.sqrt((u * u).sum(1)) / csound
    eta = (rho / (g - 1.0)) * np.log(pr / rho ** g)
    return dict(u=u, q=q, rho=rho, pr=pr, E=E, e_int=e_int, Tem=Tem, csound=csound, eta=eta,
                machno=machno)


def golden_row0(name):
    return np.array([float(v) for v in open(os.path.join(GOLD, name)).readline().split()])


def check_row(d, ref):
    assert ref[0] == 0.0
    assert abs(d["EK"] - ref[1]) < 1e-7, (d["EK"], ref[1])
    assert abs(d["eps_S"] - ref[2]) < 1e-8, (d["eps_S"], ref[2])
    assert abs(d["eps_D"] - ref[3]) < 1e-8
    assert abs(d["eps_T"] - ref[4]) < 1e-8, (d["eps_T"], ref[4])
    assert abs(d["maxmach"] - ref[5]) < 1e-7, (d["maxmach"], ref[5])


CASES = sorted(META["cases"].items())


def test_closed_forms_match_golden():
    """EK(0) = (1/8) (pi_mesh/3.14159)^3 and eps_S(0) = 0.75 nu (the analysis
    normalises by (2*3.14159)^3 while the mesh spans 2*3.14159265)."""
    for name, _ in CASES:
        ref = golden_row0(name)
        assert abs(ref[1] - 0.125 * (PI_MESH / 3.14159) ** 3) < 3e-8
        assert abs(ref[2] - 0.75 / 1600.0) < 1e-8


@pytest.mark.parametrize("name,cfg", CASES)
def test_oracle_t0_row(oracle, name, cfg):
    p, n = META["porder"], cfg["n"]
    g = case.build_case(p, n, periodic=True, lengths=(2 * PI_MESH,) * 3)
    c = comp_params()
    ic = comp_ic(g.mesh.coord, c)
    T = oracle.Tables(p)
    m = g.mesh
    mu_fluid = np.zeros(m.npoin)
    oracle.sutherland(ic["Tem"], np.full(m.npoin, c["mu_factor"]), mu_fluid)
    mu_e = np.zeros((m.nelem, T.nnode), order="F")
    ek = oracle.volavg_ek(T, m.connec, g.gpvol, c["rho0"], ic["rho"], ic["u"])
    eS, eD, eT = oracle.visc_dissipation(T, m.connec, c["rho0"], mu_fluid, mu_e, ic["u"],
                                         g.vol_tot, g.gpvol, g.He)
    mm = oracle.max_mach(m.lpoin_w, ic["machno"])
    check_row(dict(EK=ek, eps_S=eS, eps_D=eD, eps_T=eT, maxmach=mm), golden_row0(name))


@pytest.mark.gpu
@pytest.mark.parametrize("name,cfg", CASES[:2])
def test_device_t0_row(name, cfg):
    from sod2d_amd.solver import TGVParams, TGVRun
    p, n = META["porder"], cfg["n"]
    g = case.build_case(p, n, periodic=True, lengths=(2 * PI_MESH,) * 3)
    c = comp_params()
    tp = TGVParams(M=c["M"], Cp=c["Cp"], gamma_gas=c["gamma"])
    tp.Rgas, tp.rho0, tp.mu_factor = c["Rgas"], c["rho0"], c["mu_factor"]
    run = TGVRun(p, geometry=g, params=tp, initial=comp_ic(g.mesh.coord, c))
    check_row(run.diagnostics(), golden_row0(name))
    assert math.isfinite(run.dt) and run.dt > 0
