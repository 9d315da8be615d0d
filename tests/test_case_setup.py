"""Host-side case set-up of the product (C++ in libsod2d_hip.so, no GPU) against
the oracle's restatement and against closed forms."""
import numpy as np
import pytest

from helpers import make_case, relerr
from sod2d_amd import case


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_tables_match_oracle(oracle, p):
    T = case.basis_tables(p)
    R = oracle.Tables(p)
    # integer maps are bit-exact
    assert (T.a2ijk == R.a2ijk).all()
    assert (T.invAtoIJK == R.invAtoIJK).all()
    assert (T.gmshAtoI == R.gI).all() and (T.gmshAtoJ == R.gJ).all() and (T.gmshAtoK == R.gK).all()
    # floating tables to a few ulps (closed-form derivative matrix vs product formula)
    assert np.abs(T.xgp - R.xgp).max() < 4e-16
    assert np.abs(T.wgp - R.wgp).max() < 4e-16
    assert np.abs(T.dlxigp_ip - R.dlxigp_ip).max() < 1e-13 * (p + 1) ** 2
    assert np.abs(T.dNgp - R.dNgp).max() < 1e-13 * (p + 1) ** 2


@pytest.mark.parametrize("p,n,periodic,jitter", [(3, 3, True, 0.0), (3, (2, 3, 4), False, 0.1),
                                                 (4, 2, False, 0.05)])
def test_geometry_matches_oracle(oracle, p, n, periodic, jitter):
    g = make_case(p, n, periodic=periodic, jitter=jitter, lengths=(1.0, 2.0, 3.0))
    R = oracle.Tables(p)
    m = g.mesh
    gpvol, He = oracle.elem_jacobian(p, m.connec_orig, m.coord, R.dNgp, R.wgp)
    assert relerr(g.gpvol, gpvol) < 1e-14
    assert relerr(g.He, He) < 1e-13
    Ml = oracle.lumped_mass(p, m.connec, g.gpvol, m.npoin)
    assert relerr(g.Ml, Ml) < 1e-15
    assert relerr(g.invMl, oracle.inv_ml(Ml, m.lpoin_w)) < 1e-15
    assert relerr(g.helem, oracle.char_length(p, m.connec_orig, m.coord)) < 1e-15
    assert relerr(g.helem_l, oracle.char_length_spectral(p, m.connec_orig, m.coord)) < 1e-15
    # the box volume (GLL quadrature of det J is exact for straight-sided elements)
    if jitter == 0.0:
        assert abs(g.vol_tot - 6.0) < 1e-12


def test_box_mesh_structure():
    p, n = 3, (3, 4, 5)
    m = case.box_mesh(p, n, periodic=True, lengths=(2 * np.pi,) * 3)
    nn = (p + 1) ** 3
    assert m.nelem == 60 and m.npoin == (3 * p + 1) * (4 * p + 1) * (5 * p + 1)
    assert m.npoin_w == 9 * 12 * 15
    # working connectivity references working nodes only, each exactly as often
    # as a periodic structured grid requires (8 elements per vertex node, ...)
    w = np.zeros(m.npoin + 1, bool)
    w[m.lpoin_w] = True
    assert w[m.connec].all()
    counts = np.bincount(m.connec.ravel(), minlength=m.npoin + 1)[m.lpoin_w]
    assert counts.sum() == m.nelem * nn
    assert set(np.unique(counts)) == {1, 2, 4, 8}
    # node a of each element sits at the GLL point xgp(a) of the element
    T = case.basis_tables(p)
    h = 2 * np.pi / np.array(n)
    e = 17
    ex, ey, ez = e % 3, (e // 3) % 4, e // 12
    xe = m.coord[m.connec_orig[e] - 1]
    x0 = np.array([ex, ey, ez]) * h
    assert np.allclose(xe, x0 + (T.xgp + 1) / 2 * h, atol=1e-13)


def test_tgv_ic_matches_oracle(oracle):
    from sod2d_amd.solver import TGVParams, tgv_initial_conditions
    g = make_case(3, 4)
    to, tp = oracle.tgv_params(), TGVParams()
    for k in ("Rgas", "to", "po", "mu_factor"):
        assert getattr(to, k) == getattr(tp, k), k
    a = oracle.tgv_ic(g.mesh.coord, to)
    b = tgv_initial_conditions(g.mesh.coord, tp)
    # the velocity is evaluated in real(4), as the reference: libm and numpy
    # single-precision sin/cos may differ by an ulp
    for k in ("u", "q", "E", "machno"):
        assert relerr(b[k], a[k]) < 2e-7, k
    for k in ("rho", "pr", "e_int", "Tem", "csound", "eta"):
        assert relerr(b[k], a[k]) < 1e-15, k
