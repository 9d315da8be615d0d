"""TGV runs from SOD2D HDF5 files on the GPU.

* A mesh written to the SOD2D layout and loaded back (sod2d_amd.h5mesh.load_case)
  drives the fused time loop to the same bits as the generated box it came from.
* A restart written after a few steps (real(8) fields) and loaded into a new
  run (the state rebuilt as CFDSolverBase_eval_vars_after_load_hdf5_resultsFile
  does) steps like the oracle started from the same loaded state: rho, rho*u
  and E within 1e-12 normwise relative on the working nodes, mu_e within
  1e-10, the bars of the time-loop tests.
"""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


def test_tgv_run_from_mesh_file_matches_generated_box(tmp_path):
    from sod2d_amd import case, h5mesh
    from sod2d_amd.solver import TGVRun

    g = case.build_case(3, 4, periodic=True, lengths=(2 * np.pi,) * 3)
    f = tmp_path / "tgv4.hdf"
    h5mesh.write_mesh(f, g.mesh)
    h = h5mesh.load_case(f)
    print("mesh file TGV p=3 n=4: 3 steps from the file and from the box", flush=True)
    a = TGVRun(3, geometry=g, dt_mode="fixed")
    b = TGVRun(3, geometry=h, dt_mode="fixed")
    assert a.dt == b.dt
    for _ in range(3):
        a.step()
        b.step()
    sa, sb = a.state(), b.state()
    for k in ("rho", "q", "E", "mu_e"):
        assert np.array_equal(sa[k], sb[k]), k


def test_restart_continues_like_the_oracle(oracle, tmp_path):
    from sod2d_amd import case, h5mesh
    from sod2d_amd.solver import TGVParams, TGVRun

    p, n = 3, 4
    g = case.build_case(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    tp = TGVParams()
    a = TGVRun(p, geometry=g, params=tp, dt_mode="fixed")
    print("restart TGV p=3 n=4: 3 steps, save, load, 3 more steps", flush=True)
    for _ in range(3):
        a.step()
    path = tmp_path / h5mesh.restart_file_name(h5mesh.restart_base_name(
        str(tmp_path) + "/", "tgv4", 1), 1)
    h5mesh.save_restart(path, a, rp_vtk=8)
    ic, time, istep = h5mesh.load_restart(path, g, tp)
    assert (time, istep) == (a.time, a.istep)
    w = g.mesh.lpoin_w - 1
    sa = a.state()
    assert np.array_equal(ic["rho"][w], sa["rho"][w])
    assert np.array_equal(ic["E"][w], sa["E"][w])
    assert relerr(ic["q"][w], sa["q"][w]) < 1e-15  # q = rho * (q / rho)
    b = TGVRun(p, geometry=g, params=tp, initial=ic, dt_mode="fixed")
    b.time, b.istep = time, istep
    to = oracle.tgv_params(M=0.1)
    orc = oracle.TGVOracle(p, g, ic, to, parity=True, dt_mode="fixed")
    assert abs(b.dt - orc.dt) <= 1e-14 * orc.dt
    for k in range(3):
        b.step()
        orc.step()
        s = b.state()
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, k
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, k
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, k
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, k
