"""The Fortran host program sod2d_amd/fortran/tgv_host.f90 runs the
reference's time loop (the level copy of CFDSolverBase_evalTimeIteration,
then rk_4_ls_main with updateF, time_integ_ls.f90:202-691) on a periodic TGV
case entirely through mod_sod2d_hip, at fixed dt, and writes each step's
conserved fields and mu_e.  Every step is checked against the oracle's
restatement of the same loop: rho, rho*u and E within 1e-12 normwise relative
on the working nodes, mu_e within 1e-10 (its min/max limiter), the bars of the
Python-hosted time loop.  With ``lps`` the program runs the LPS-stabilised
loop (comp_tau at the start of every step, rk_stage_lps_fused for every
updateF) against the oracle's LPS restatement with the same bars."""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_case, relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "sod2d_amd", "fortran")


def _w(d, name, a, dtype=np.float64):
    np.asarray(a, dtype=dtype).ravel(order="F").tofile(os.path.join(d, name + ".bin"))


def _r(d, name, shape):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=np.float64).reshape(shape, order="F")


@pytest.mark.parametrize("p,n,nsteps,stages,lps", [(3, 4, 6, 5, False), (2, 4, 5, 4, False),
                                                   (3, 4, 4, 5, True), (2, 4, 4, 4, True)])
def test_fortran_host_time_loop(oracle, tmp_path, p, n, nsteps, stages, lps):
    exe = os.path.join(FDIR, f"_build_p{p}", "tgv_host")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m sod2d_amd.build)")
    g = make_case(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    m, T = g.mesh, oracle.Tables(p)
    t = oracle.tgv_params(M=0.1)
    ic = oracle.tgv_ic(m.coord, t)
    orc = oracle.TGVOracle(p, g, ic, t, stages=stages, parity=True, dt_mode="fixed", lps=lps)
    prm = oracle.visc_params(t.Rgas)
    L, G = oracle.rk_ls_coeffs(stages, 1)  # 6 x 6; the leading (S+1) x (S+1) block is used
    L, G = L[:stages + 1, :stages + 1], G[:stages + 1, :stages + 1]
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{m.nelem} {m.npoin} {m.npoin_w} {nsteps} {stages}\n")
        f.write(f"{t.gamma_gas!r} {t.Rgas!r} {t.Cp!r} {t.Prt!r} {t.mu_factor!r} {t.po!r} "
                f"{t.rho0!r}\n")
        f.write(f"{orc.dt!r}\n")
        f.write(f"{prm.ce!r} {prm.cmax!r} {prm.cglob!r} {prm.flag_normalise_entropy}\n")
        f.write(f"{int(lps)}\n")
    for name, a in [("connec", m.connec), ("lpoin_w", m.lpoin_w), ("invAtoIJK", T.invAtoIJK),
                    ("gmshAtoI", T.gI), ("gmshAtoJ", T.gJ),
                    ("gmshAtoK", T.gK)]:
        _w(d, name, a, np.int32)
    for name, a in [("He", g.He), ("gpvol", g.gpvol), ("dlxigp_ip", T.dlxigp_ip),
                    ("invMl", g.invMl), ("helem_l", g.helem_l), ("helem", g.helem),
                    ("rkL", L), ("rkG", G), ("mu_fluid", orc.mu_fluid)]:
        _w(d, name, a)
    for name in ("rho", "q", "E", "u", "pr", "Tem", "e_int", "eta", "csound", "machno"):
        _w(d, name, ic[name])
    print(f"fortran host TGV p={p} n={n} steps={nsteps} stages={stages} lps={lps}", flush=True)
    res = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert f"tgv_host: {nsteps} steps done" in res.stdout
    w = m.lpoin_w - 1
    nn = T.nnode
    for k in range(1, nsteps + 1):
        orc.step()
        rho = _r(d, f"step{k}_rho", m.npoin)
        q = _r(d, f"step{k}_q", (m.npoin, 3))
        E = _r(d, f"step{k}_E", m.npoin)
        mu_e = _r(d, f"step{k}_mu_e", (m.nelem, nn))
        assert relerr(rho[w], orc.rho[1][w]) < 1e-12, k
        assert relerr(q[w], orc.q[1][w]) < 1e-12, k
        assert relerr(E[w], orc.E[1][w]) < 1e-12, k
        assert relerr(mu_e, orc.mu_e) < 1e-10, k
    # the flow moved: the last step differs from the initial state
    assert relerr(q[w], ic["q"][w]) > 1e-8
