"""Host-side case set-up (CPU): element tables, structured box meshes in SOD2D's
partition format, element geometry, lumped mass and element lengths.

All arrays are numpy arrays in the reference's Fortran layout (``order='F'``,
the reference's dimension order) and integer tables hold 1-based values, so
they can be passed unchanged to the reference-shaped operators.  The work is
done by the C++ set-up code of libsod2d_hip.so (no device needed).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def fz(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


@dataclass
class ElementTables:
    porder: int
    nnode: int
    a2ijk: np.ndarray
    invAtoIJK: np.ndarray
    gmshAtoI: np.ndarray
    gmshAtoJ: np.ndarray
    gmshAtoK: np.ndarray
    xgp: np.ndarray
    wgp: np.ndarray
    dNgp: np.ndarray
    dlxigp_ip: np.ndarray

    @property
    def Ngp(self):
        """Shape functions at the GLL points: exactly the identity (nodal basis)."""
        return np.asfortranarray(np.eye(self.nnode))


def basis_tables(porder: int) -> ElementTables:
    n1 = porder + 1
    nn = n1 ** 3
    a2 = np.zeros(nn, np.int32)
    inv = fz((n1, n1, n1), np.int32)
    gI, gJ, gK = (np.zeros(nn, np.int32) for _ in range(3))
    xgp, wgp = fz((nn, 3)), np.zeros(nn)
    dNgp, dl = fz((3, nn, nn)), fz((nn, 3, n1))
    check(lib().sodhip_basis_tables(porder, _p(a2), _p(inv), _p(gI), _p(gJ), _p(gK), _p(xgp),
                                    _p(wgp), _p(dNgp), _p(dl)), "sodhip_basis_tables")
    return ElementTables(porder, nn, a2, inv, gI, gJ, gK, xgp, wgp, dNgp, dl)


@dataclass
class BoxMesh:
    """A structured box in SOD2D partition format (mod_mpi_mesh.f90:15-48)."""
    porder: int
    dims: tuple
    periodic: bool
    nelem: int
    npoin: int
    npoin_w: int
    connec_orig: np.ndarray   # connecParOrig(nelem,nnode)
    connec: np.ndarray        # connecParWork(nelem,nnode)
    lpoin_w: np.ndarray       # workingNodesPar(npoin_w)
    coord: np.ndarray         # coordPar(npoin,3)
    extra: dict = field(default_factory=dict)


def box_mesh(porder: int, n, periodic: bool = True, lengths=None, jitter: float = 0.0,
             seed: int = 20261202) -> BoxMesh:
    if isinstance(n, int):
        n = (n, n, n)
    nx, ny, nz = n
    if lengths is None:
        lengths = (2.0 * math.pi,) * 3
    ne, npo, nw = C.c_int(), C.c_int(), C.c_int()
    check(lib().sodhip_box_mesh_sizes(porder, nx, ny, nz, int(periodic), C.byref(ne), C.byref(npo),
                                      C.byref(nw)), "sodhip_box_mesh_sizes")
    nn = (porder + 1) ** 3
    co = fz((ne.value, nn), np.int32)
    cw = fz((ne.value, nn), np.int32)
    lw = np.zeros(nw.value, np.int32)
    coord = fz((npo.value, 3))
    L = np.asarray(lengths, dtype=np.float64)
    check(lib().sodhip_box_mesh(porder, nx, ny, nz, int(periodic), _p(L), float(jitter),
                                C.c_ulonglong(seed), _p(co), _p(cw), _p(lw), _p(coord)),
          "sodhip_box_mesh")
    return BoxMesh(porder, (nx, ny, nz), periodic, ne.value, npo.value, nw.value, co, cw, lw, coord)


def elem_jacobian(porder: int, connec_orig: np.ndarray, coord: np.ndarray):
    nelem, nn = connec_orig.shape
    He = fz((3, 3, nn, nelem))
    gpvol = fz((1, nn, nelem))
    check(lib().sodhip_elem_jacobian(porder, nelem, coord.shape[0], _p(connec_orig), _p(coord),
                                     _p(He), _p(gpvol)), "sodhip_elem_jacobian")
    return He, gpvol


def lumped_mass(porder: int, connec: np.ndarray, gpvol: np.ndarray, npoin: int,
                lpoin_w: np.ndarray):
    Ml = np.zeros(npoin)
    invMl = np.zeros(npoin)
    check(lib().sodhip_lumped_mass(porder, connec.shape[0], npoin, _p(connec), _p(gpvol),
                                   lpoin_w.shape[0], _p(lpoin_w), _p(Ml), _p(invMl)),
          "sodhip_lumped_mass")
    return Ml, invMl


def elem_per_node(connec: np.ndarray, npoin: int) -> np.ndarray:
    """elemPerNode (mod_geom.f90:68-99): incidences of each node in connec."""
    nelem, nn = connec.shape
    lelpn = np.zeros(npoin, dtype=np.int32)
    check(lib().sodhip_elem_per_node(nelem, nn, npoin, _p(connec), _p(lelpn)),
          "sodhip_elem_per_node")
    return lelpn


def char_length(porder: int, connec_orig: np.ndarray, coord: np.ndarray, spectral: bool = True):
    nelem, nn = connec_orig.shape
    helem = np.zeros(nelem)
    helem_l = fz((nelem, nn)) if spectral else None
    check(lib().sodhip_char_length(porder, nelem, coord.shape[0], _p(connec_orig), _p(coord),
                                   _p(helem), _p(helem_l)), "sodhip_char_length")
    return helem, helem_l


@dataclass
class Geometry:
    tables: ElementTables
    mesh: BoxMesh
    He: np.ndarray
    gpvol: np.ndarray
    Ml: np.ndarray
    invMl: np.ndarray
    helem: np.ndarray
    helem_l: np.ndarray
    vol_tot: float


def build_case(porder: int, n, periodic: bool = True, lengths=None, jitter: float = 0.0,
               seed: int = 20261202, spectral_lengths: bool = True) -> Geometry:
    """Mesh + tables + geometry + mass + lengths, as CFDSolverBase_run's set-up."""
    T = basis_tables(porder)
    m = box_mesh(porder, n, periodic, lengths, jitter, seed)
    He, gpvol = elem_jacobian(porder, m.connec_orig, m.coord)
    Ml, invMl = lumped_mass(porder, m.connec, gpvol, m.npoin, m.lpoin_w)
    helem, helem_l = char_length(porder, m.connec_orig, m.coord, spectral_lengths)
    return Geometry(T, m, He, gpvol, Ml, invMl, helem, helem_l, float(gpvol.sum(dtype=np.float64)))


def operator_bench_fields(npoin: int, nelem: int, nnode: int, seed: int = 1, gamma: float = 1.4,
                          Rgas: float = 1.0) -> dict:
    """Synthetic inputs of the isolated-operator workload, in the reference
    layouts: rho ~ U(0.9, 1.1), u_i ~ N(0, 0.3), p ~ U(0.9, 1.1)/gamma, E
    consistent with them, the E slot of full_convec_ijk as the enthalpy
    h = gamma p / ((gamma - 1) rho) (time_integ_ls.f90:584), T = p/(rho R),
    mu_fluid ~ U(1e-4, 2e-4), mu_e ~ U(0, 1e-3) and mu_sgs ~ U(0, 1e-4) per
    element-GP.  Drawn in that order from one seeded generator."""
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
