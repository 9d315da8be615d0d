"""The Fortran layer (sod2d_amd/fortran/mod_sod2d_hip.f90) keeps the
reference's subroutine interfaces of full_convec_ijk, full_diffusion_ijk,
generic_scalar_convec_ijk, smart_visc_spectral, lumped_solver_scal_opt and the
live path's smart_visc_spectral_imex, eval_divergence, compute_vorticity and
ducros_visc_limit, the LPS routines comp_tau, full_proj_ijk and
full_diff_stab_ijk, and adds the fused stage combine.  The test program tests/fortran/shim_driver.f90
calls them, through the C-ABI, on a case written here; its results are checked
against the oracle: the residuals, the stage combine and the sensor to 1e-13
normwise relative (the device arithmetic is reordered, so not bit-exact).

The real(4) build of the module (-D__RP__=4, the reference's default
USE_RP=4) runs the same driver on real(4) copies of the inputs through the
_f32 entry points; its results are checked against the FP64 oracle with the
bars of FP32 arithmetic on rounded inputs stated in the test (the live-path
and LPS operators are real(8) only)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_case, random_state, relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "sod2d_amd", "fortran")


def _w(d, name, a):
    np.asarray(a).ravel(order="F").tofile(os.path.join(d, name + ".bin"))


def _r(d, name, shape):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=np.float64).reshape(shape, order="F")


@pytest.mark.parametrize("p,n,periodic,jitter,rk", [(3, (3, 2, 2), False, 0.1, 8),
                                                    (2, 3, True, 0.0, 8), (4, 2, True, 0.0, 8),
                                                    (3, (3, 2, 2), False, 0.1, 4)])
def test_fortran_shim(oracle, tmp_path, p, n, periodic, jitter, rk):
    exe = os.path.join(FDIR, f"_build_p{p}" + ("_rp4" if rk == 4 else ""), "shim_driver")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m sod2d_amd.build)")
    g = make_case(p, n, periodic=periodic, jitter=jitter, lengths=(1.0, 1.5, 2.0))
    m, T = g.mesh, oracle.Tables(p)
    npoin, nelem, nn = m.npoin, m.nelem, T.nnode
    Rgas, gamma, Cp, Pr = 287.0, 1.4, 1004.0, 0.71
    s = random_state(npoin, nelem, nn, seed=11, gamma=gamma, Rgas=Rgas)
    r = np.random.default_rng(12)
    Reta = r.normal(0, 1e-3, npoin)
    eta = r.uniform(0.5, 1.5, npoin)
    rho1 = s["rho"] * 1.01
    q1 = np.asfortranarray(s["q"] * 0.99)
    E1 = s["h"] * 1.02
    dt, lam1, lams, gam = 1e-3, 0.3, 0.7, 0.45
    csound = r.uniform(0.9, 1.1, npoin) * 300.0
    prm = oracle.visc_params(Rgas)
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{nelem} {npoin} {m.npoin_w}\n{Cp!r} {Pr!r} {gamma!r}\n")
        f.write(f"{prm.ce!r} {prm.cmax!r} {prm.cglob!r} {Rgas!r} 1\n")
        f.write(f"{dt!r} {lam1!r} {lams!r} {gam!r}\n")
    for name, a in [("connec", m.connec), ("lpoin_w", m.lpoin_w), ("invAtoIJK", T.invAtoIJK),
                    ("gI", T.gI), ("gJ", T.gJ), ("gK", T.gK),
                    ("lelpn", oracle.elem_per_node(m.connec, npoin))]:
        _w(d, name, np.asarray(a, dtype=np.int32))
    for name, a in [("He", g.He), ("gpvol", g.gpvol), ("dlxigp_ip", T.dlxigp_ip), ("u", s["u"]),
                    ("q", s["q"]), ("rho", s["rho"]), ("pr", s["pr"]), ("h", s["h"]),
                    ("Tem", s["Tem"]), ("mu_fluid", s["mu_fluid"]), ("mu_e", s["mu_e"]),
                    ("invMl", g.invMl), ("Reta", Reta), ("eta", eta), ("helem_l", g.helem_l),
                    ("helem", g.helem), ("rho1", rho1), ("q1", q1), ("E1", E1),
                    ("csound", csound), ("Ml", g.Ml)]:
        _w(d, name, np.asarray(a, dtype=np.float64))
    print(f"fortran shim case p={p} n={n} periodic={periodic} rp={rk}", flush=True)
    # bars: FP64 build 1e-13 (residuals, sensor) and 1e-15 (stage combine);
    # real(4) build, inputs rounded to FP32: residuals 2e-4, combine 1e-6,
    # sensor 1e-5 (FP64 arithmetic on FP32 inputs)
    t_res, t_y, t_mu = (1e-13, 1e-15, 1e-13) if rk == 8 else (2e-4, 1e-6, 1e-5)
    res = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "shim_driver: ok" in res.stdout

    # oracle: the same sequence
    Rm, Rq, Re = oracle.full_convec_ijk(T, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"],
                                        s["pr"], s["h"], initialize=True, fact=-1.0)
    oracle.full_diffusion_ijk(T, m.connec, g.He, g.gpvol, Cp, Pr, s["rho"], s["rho"], s["u"],
                              s["Tem"], s["mu_fluid"], s["mu_e"], np.zeros_like(s["mu_e"]), Rm, Rq,
                              Re, initialize=False, fact=-1.0)
    assert relerr(_r(d, "out_Rmass", npoin), Rm) < t_res
    assert relerr(_r(d, "out_Rmom", (npoin, 3)), Rq) < t_res
    assert relerr(_r(d, "out_Rener", npoin), Re) < t_res
    oracle.lumped_solver_scal(m.lpoin_w, g.invMl, Rm)
    oracle.lumped_solver_scal(m.lpoin_w, g.invMl, Re)
    oracle.lumped_solver_vect(m.lpoin_w, g.invMl, Rq)
    rho, q, E = s["rho"].copy(), np.asfortranarray(s["q"].copy()), s["h"].copy()
    oracle.rk_ls_update(m.lpoin_w, lam1, lams, gam * dt, rho1, q1, E1, rho, q, E, Rm, Rq, Re)
    w = m.lpoin_w - 1
    assert relerr(_r(d, "out_Rmass_scaled", npoin)[w], Rm[w]) < t_res
    assert relerr(_r(d, "out_rho", npoin), rho) < t_y
    assert relerr(_r(d, "out_q", (npoin, 3)), q) < t_y
    assert relerr(_r(d, "out_E", npoin), E) < t_y
    Rc = oracle.generic_scalar_convec_ijk(T, m.connec, g.He, g.gpvol, s["q"], eta, s["u"])
    oracle.lumped_solver_scal(m.lpoin_w, g.invMl, Rc)
    assert relerr(_r(d, "out_Rconv", npoin), Rc) < t_res
    mu_e = np.zeros((nelem, nn), order="F")
    oracle.smart_visc_spectral(T, m.connec, m.lpoin_w, Reta, gamma, s["rho"], s["u"], s["Tem"],
                               eta, g.helem_l, g.helem, prm, mu_e)
    assert relerr(_r(d, "out_mu_e", (nelem, nn)), mu_e) < t_mu
    if rk == 4:
        return  # the live-path and LPS operators are real(8) only
    # live path: imex sensor, divergence, vorticity, Ducros limit
    mu_i, _ = oracle.smart_visc_spectral_imex(T, m.connec, m.lpoin_w, Reta, s["rho"], s["u"],
                                              csound, eta, g.helem, oracle.visc_params(Rgas),
                                              np.zeros((nelem, nn), order="F"))
    assert relerr(_r(d, "out_mu_imex", (nelem, nn)), mu_i) < 1e-13
    divU = oracle.eval_divergence(T, m.connec, g.He, g.gpvol, s["u"])
    curl = oracle.compute_vorticity(T, m.connec, m.lpoin_w, oracle.elem_per_node(m.connec, npoin),
                                    g.He, s["u"])
    assert relerr(_r(d, "out_divU", npoin), divU) < 1e-13
    assert relerr(_r(d, "out_curlU", (npoin, 3)), curl) < 1e-13
    oracle.ducros_visc_limit(T, m.connec, s["u"], divU, curl, g.Ml, g.helem_l, mu_i, 0.1)
    assert relerr(_r(d, "out_mu_ducros", (nelem, nn)), mu_i) < 1e-13
    # LPS stabilisation
    tau = oracle.comp_tau(p, m.connec, csound, s["u"], g.helem)
    assert relerr(_r(d, "out_tau", nelem), tau) < 1e-15
    P = oracle.full_proj_ijk(T, m.connec, m.lpoin_w, g.He, g.gpvol, s["rho"], s["u"], s["Tem"],
                             g.invMl)
    assert relerr(_r(d, "out_PMass", (npoin, 3)), P[0]) < 1e-13
    assert relerr(_r(d, "out_PMX", (npoin, 3)), P[2]) < 1e-13
    Rs = oracle.full_diff_stab_ijk(T, m.connec, g.He, g.gpvol, Cp, Pr, s["rho"], s["rho"], s["u"],
                                   s["Tem"], s["mu_fluid"], s["mu_e"], np.zeros_like(s["mu_e"]),
                                   P, tau, initialize=True, fact=-1.0)
    assert relerr(_r(d, "out_Rsm", npoin), Rs[0]) < 1e-13
    assert relerr(_r(d, "out_Rsq", (npoin, 3)), Rs[1]) < 1e-13
    assert relerr(_r(d, "out_Rse", npoin), Rs[2]) < 1e-13
