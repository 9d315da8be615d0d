"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the cpu_baseline leg of ``bench.py``; the product never imports it.

Arrays are numpy arrays in the reference's Fortran layout (``order='F'``) with the
reference's dimension order; index tables hold the reference's 1-based values.
The functions mirror the C entry points in ``sod2d_oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
    return _LIB


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), a.dtype
    return a.ctypes.data_as(_D)


def _i(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim <= 1), a.dtype
    return a.ctypes.data_as(_I)


def fzeros(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


class ViscParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("ce", "cmax", "cglob", "Rgas", "c_rho", "c_ener", "ce_comp", "factor_comp")] + \
               [("flag_normalise_entropy", C.c_int), ("flag_drop_c_in_envit", C.c_int)]


class TGVParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("Cp", "Prt", "M", "Re", "rho0", "gamma_gas", "Rgas", "to", "po", "mu_factor")]


# ---------------------------------------------------------------- basis
def gll(p):
    xi = np.zeros(p + 1)
    w = np.zeros(p + 1)
    lib().or_gll(p, _d(xi), _d(w))
    return xi, w


def lagrange(p, xi, x):
    out = np.zeros(p + 1)
    lib().or_lagrange(p, _d(np.ascontiguousarray(xi, dtype=np.float64)), C.c_double(x), _d(out))
    return out


def lagrange_deriv(p, xi, x):
    out = np.zeros(p + 1)
    lib().or_lagrange_deriv(p, _d(np.ascontiguousarray(xi, dtype=np.float64)), C.c_double(x),
                            _d(out))
    return out


def gmsh_hex_table(p):
    t = fzeros(((p + 1) ** 3, 3), np.int32)
    lib().or_gmsh_hex_table(p, _i(t))
    return t


def a2ijk(p):
    a = np.zeros((p + 1) ** 3, dtype=np.int32)
    lib().or_a2ijk(p, _i(a))
    return a


def atoijk_inverse(p, a2):
    n1 = p + 1
    inv = fzeros((n1, n1, n1), np.int32)
    gI = np.zeros(n1 ** 3, np.int32)
    gJ = np.zeros(n1 ** 3, np.int32)
    gK = np.zeros(n1 ** 3, np.int32)
    lib().or_atoijk_inverse(p, _i(a2), _i(inv), _i(gI), _i(gJ), _i(gK))
    return inv, gI, gJ, gK


def shape(p, a2):
    ng = (p + 1) ** 3
    xgp = fzeros((ng, 3))
    wgp = np.zeros(ng)
    Ngp = fzeros((ng, ng))
    dNgp = fzeros((3, ng, ng))
    dl = fzeros((ng, 3, p + 1))
    lib().or_shape(p, _i(a2), _d(xgp), _d(wgp), _d(Ngp), _d(dNgp), _d(dl))
    return xgp, wgp, Ngp, dNgp, dl


class Tables:
    """All element tables of order p, built by the oracle."""

    def __init__(self, p):
        self.p = p
        self.nnode = (p + 1) ** 3
        self.a2ijk = a2ijk(p)
        self.invAtoIJK, self.gI, self.gJ, self.gK = atoijk_inverse(p, self.a2ijk)
        self.xgp, self.wgp, self.Ngp, self.dNgp, self.dlxigp_ip = shape(p, self.a2ijk)


# ---------------------------------------------------------------- geometry
def elem_jacobian(p, connec, coord, dNgp, wgp):
    nelem = connec.shape[0]
    ng = (p + 1) ** 3
    gpvol = fzeros((1, ng, nelem))
    He = fzeros((3, 3, ng, nelem))
    lib().or_elem_jacobian(p, nelem, coord.shape[0], _i(connec), _d(coord), _d(dNgp), _d(wgp),
                           _d(gpvol), _d(He))
    return gpvol, He


def lumped_mass(p, connec, gpvol, npoin):
    Ml = np.zeros(npoin)
    lib().or_lumped_mass(p, connec.shape[0], npoin, _i(connec), _d(gpvol), _d(Ml))
    return Ml


def inv_ml(Ml, lpoin_w):
    invMl = np.zeros_like(Ml)
    lib().or_inv_ml(Ml.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(Ml), _d(invMl))
    return invMl


def char_length(p, connec, coord):
    h = np.zeros(connec.shape[0])
    lib().or_char_length(p, connec.shape[0], coord.shape[0], _i(connec), _d(coord), _d(h))
    return h


def char_length_spectral(p, connec, coord):
    h = fzeros(connec.shape)
    lib().or_char_length_spectral(p, connec.shape[0], coord.shape[0], _i(connec), _d(coord),
                                  _d(h))
    return h


# ---------------------------------------------------------------- operators
def full_convec_ijk(T, connec, He, gpvol, u, q, rho, pr, E, Rmass=None, Rmom=None, Rener=None,
                    initialize=True, fact=1.0, nthreads=1):
    npoin = rho.shape[0]
    if Rmass is None:
        Rmass, Rmom, Rener = np.zeros(npoin), fzeros((npoin, 3)), np.zeros(npoin)
    lib().or_full_convec_ijk(T.p, connec.shape[0], npoin, _i(connec), _d(He), _d(gpvol),
                             _d(T.dlxigp_ip), _i(T.invAtoIJK), _i(T.gI), _i(T.gJ), _i(T.gK),
                             _d(u), _d(q), _d(rho), _d(pr), _d(E), _d(Rmass), _d(Rmom), _d(Rener),
                             int(bool(initialize)), C.c_double(fact), int(nthreads))
    return Rmass, Rmom, Rener


def full_diffusion_ijk(T, connec, He, gpvol, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs,
                       Rmass=None, Rmom=None, Rener=None, initialize=True, fact=1.0,
                       c_rho=1.0, c_ener=1.0, nthreads=1):
    npoin = rho.shape[0]
    if Rmass is None:
        Rmass, Rmom, Rener = np.zeros(npoin), fzeros((npoin, 3)), np.zeros(npoin)
    lib().or_full_diffusion_ijk(T.p, connec.shape[0], npoin, _i(connec), _d(He), _d(gpvol),
                                _d(T.dlxigp_ip), _i(T.invAtoIJK), _i(T.gI), _i(T.gJ), _i(T.gK),
                                C.c_double(Cp), C.c_double(Pr), _d(rho_n), _d(rho), _d(u), _d(Tem),
                                _d(mu_fluid), _d(mu_e), _d(mu_sgs), C.c_double(c_rho),
                                C.c_double(c_ener), _d(Rmass), _d(Rmom), _d(Rener),
                                int(bool(initialize)), C.c_double(fact), int(nthreads))
    return Rmass, Rmom, Rener


C_LPS_COMP = float(np.float32(0.1))  # c_lps_comp is a default-real constant
C_SPECIES_STAB = 0.5


def comp_tau(p, connec, csound, u, helem_k, c_species_stab=C_SPECIES_STAB, drop_c=False):
    """comp_tau (elem_stab.f90:589-616): the LPS time scale per element."""
    tau = np.zeros(connec.shape[0])
    lib().or_comp_tau(p, connec.shape[0], csound.shape[0], _i(connec), _d(csound), _d(u),
                      _d(helem_k), C.c_double(c_species_stab), int(bool(drop_c)), _d(tau))
    return tau


def full_proj_ijk(T, connec, lpoin_w, He, gpvol, rho, u, Tem, invMl, nthreads=1):
    """full_proj_ijk (elem_stab.f90:441-586): returns ProjMass, ProjEner, ProjMX,
    ProjMY, ProjMZ, each (npoin,3)."""
    npoin = rho.shape[0]
    P = [fzeros((npoin, 3)) for _ in range(5)]
    lib().or_full_proj_ijk(T.p, connec.shape[0], npoin, lpoin_w.shape[0], _i(connec),
                           _i(lpoin_w), _d(He), _d(gpvol), _d(T.dlxigp_ip), _i(T.invAtoIJK),
                           _i(T.gI), _i(T.gJ), _i(T.gK), _d(rho), _d(u), _d(Tem), _d(invMl),
                           *[_d(x) for x in P], int(nthreads))
    return P


def full_diff_stab_ijk(T, connec, He, gpvol, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs,
                       proj, tau_stab, Rmass=None, Rmom=None, Rener=None, initialize=True,
                       fact=1.0, c_rho=1.0, c_ener=1.0, c_lps=C_LPS_COMP, nthreads=1):
    """full_diff_stab_ijk (elem_stab.f90:207-438); ``proj`` is full_proj_ijk's
    (ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ)."""
    npoin = rho.shape[0]
    if Rmass is None:
        Rmass, Rmom, Rener = np.zeros(npoin), fzeros((npoin, 3)), np.zeros(npoin)
    lib().or_full_diff_stab_ijk(T.p, connec.shape[0], npoin, _i(connec), _d(He), _d(gpvol),
                                _d(T.dlxigp_ip), _i(T.invAtoIJK), _i(T.gI), _i(T.gJ), _i(T.gK),
                                C.c_double(Cp), C.c_double(Pr), _d(rho_n), _d(rho), _d(u),
                                _d(Tem), _d(mu_fluid), _d(mu_e), _d(mu_sgs),
                                *[_d(x) for x in proj], _d(tau_stab), C.c_double(c_rho),
                                C.c_double(c_ener), C.c_double(c_lps), _d(Rmass), _d(Rmom),
                                _d(Rener), int(bool(initialize)), C.c_double(fact),
                                int(nthreads))
    return Rmass, Rmom, Rener


def generic_scalar_convec_ijk(T, connec, He, gpvol, q, eta, u, Rconvec=None, initialize=True,
                              fact=1.0, nthreads=1):
    npoin = eta.shape[0]
    if Rconvec is None:
        Rconvec = np.zeros(npoin)
    lib().or_generic_scalar_convec_ijk(T.p, connec.shape[0], npoin, _i(connec), _d(He), _d(gpvol),
                                       _d(T.dlxigp_ip), _i(T.invAtoIJK), _i(T.gI), _i(T.gJ),
                                       _i(T.gK), _d(q), _d(eta), _d(u), _d(Rconvec),
                                       int(bool(initialize)), C.c_double(fact), int(nthreads))
    return Rconvec


def convert_ijk(p):
    c = np.zeros(p + 3, np.int32)
    lib().or_convert_ijk(p, _i(c))
    return c


def visc_params(Rgas, ce=0.1, cmax=0.5, cglob=1.0, c_rho=1.0, c_ener=1.0, ce_comp=1.0,
                factor_comp=1.0, flag_normalise_entropy=1, flag_drop_c_in_envit=0):
    return ViscParams(ce, cmax, cglob, Rgas, c_rho, c_ener, ce_comp, factor_comp,
                      flag_normalise_entropy, flag_drop_c_in_envit)


def smart_visc_spectral(T, connec, lpoin_w, Reta, gamma_gas, rho, u, Tem, eta, helem, helem_k,
                        prm, mu_e, mue_l=None):
    nelem = connec.shape[0]
    if mue_l is None:
        mue_l = fzeros(mu_e.shape)
    lib().or_smart_visc_spectral(T.p, nelem, rho.shape[0], lpoin_w.shape[0], _i(connec),
                                 _i(lpoin_w), _d(Reta), C.c_double(gamma_gas), _d(rho), _d(u),
                                 _d(Tem), _d(eta), _d(helem), _d(helem_k), _i(T.invAtoIJK),
                                 C.byref(prm), _d(mu_e), _d(mue_l))
    return mu_e, mue_l


def smart_visc_spectral_imex(T, connec, lpoin_w, Reta, rho, u, csound, eta, helem_k, prm, mu_e,
                             mue_l=None, flag_high_mach=0):
    """smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227); mu_e is
    overwritten, mue_l receives its incoming value; with the normalisation and
    flag_high_mach, prm.ce becomes prm.ce_comp."""
    nelem = connec.shape[0]
    if mue_l is None:
        mue_l = fzeros(mu_e.shape)
    lib().or_smart_visc_spectral_imex(T.p, nelem, rho.shape[0], lpoin_w.shape[0], _i(connec),
                                      _i(lpoin_w), _d(Reta), _d(rho), _d(u), _d(csound), _d(eta),
                                      _d(helem_k), C.byref(prm), int(flag_high_mach), _d(mu_e),
                                      _d(mue_l))
    return mu_e, mue_l


def eval_divergence(T, connec, He, gpvol, u, divU=None):
    """eval_divergence (mod_operators.f90:485-570)."""
    npoin = u.shape[0]
    if divU is None:
        divU = np.zeros(npoin)
    lib().or_eval_divergence(T.p, connec.shape[0], npoin, _i(connec), _d(He), _d(gpvol),
                             _d(T.dlxigp_ip), _i(T.invAtoIJK), _i(T.gI), _i(T.gJ), _i(T.gK),
                             _d(u), _d(divU))
    return divU


def elem_per_node(connec, npoin):
    """elemPerNode (mod_geom.f90:68-99): incidences of each node in connec."""
    lelpn = np.zeros(npoin, dtype=np.int32)
    lib().or_elem_per_node(connec.shape[0], connec.shape[1], npoin, _i(connec), _i(lelpn))
    return lelpn


def compute_vorticity(T, connec, lpoin_w, lelpn, He, u, curlU=None):
    """compute_vorticity with makeConti (mod_operators.f90:684-775)."""
    npoin = u.shape[0]
    if curlU is None:
        curlU = fzeros((npoin, 3))
    lib().or_compute_vorticity(T.p, connec.shape[0], npoin, lpoin_w.shape[0], _i(lpoin_w),
                               _i(connec), _i(lelpn), _d(He), _d(T.dlxigp_ip), _i(T.invAtoIJK),
                               _i(T.gI), _i(T.gJ), _i(T.gK), _d(u), _d(curlU))
    return curlU


def ducros_visc_limit(T, connec, u, divU, vorti, Ml, helem, mu_e, ducros_min_val=0.1):
    """ducros_visc_limit (mod_entropy_viscosity.f90:277-314); mu_e scaled in place."""
    lib().or_ducros_visc_limit(T.p, connec.shape[0], u.shape[0], _i(connec), _d(u), _d(divU),
                               _d(vorti), _d(Ml), _d(helem), C.c_double(ducros_min_val), _d(mu_e))
    return mu_e


def thermo_entropy(lpoin_w, gamma_gas, p_inf, rho_inf, rho, pr, u, eta, feta):
    lib().or_thermo_entropy(rho.shape[0], lpoin_w.shape[0], _i(lpoin_w), C.c_double(gamma_gas),
                            C.c_double(p_inf), C.c_double(rho_inf), _d(rho), _d(pr), _d(u),
                            _d(eta), _d(feta))


# ---------------------------------------------------------------- nodal
def lumped_solver_scal(lpoin_w, invMl, R):
    lib().or_lumped_solver_scal(R.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(invMl), _d(R))


def lumped_solver_vect(lpoin_w, invMl, R):
    lib().or_lumped_solver_vect(R.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(invMl), _d(R))


def sutherland(Tem, mu_factor, mu_fluid):
    lib().or_sutherland(Tem.shape[0], _d(Tem), _d(mu_factor), _d(mu_fluid))


def eos_stage(lpoin_w, gamma_gas, Rgas, rho, q, E, u, e_int, pr, aux_h, Tem):
    lib().or_eos_stage(rho.shape[0], lpoin_w.shape[0], _i(lpoin_w), C.c_double(gamma_gas),
                       C.c_double(Rgas), _d(rho), _d(q), _d(E), _d(u), _d(e_int), _d(pr),
                       _d(aux_h), _d(Tem))


def eos_step(lpoin_w, gamma_gas, Rgas, rho, q, E, u, e_int, pr, csound, machno, Tem):
    lib().or_eos_step(rho.shape[0], lpoin_w.shape[0], _i(lpoin_w), C.c_double(gamma_gas),
                      C.c_double(Rgas), _d(rho), _d(q), _d(E), _d(u), _d(e_int), _d(pr),
                      _d(csound), _d(machno), _d(Tem))


def rk_ls_coeffs(stages=5, variant=1):
    L = fzeros((6, 6))
    G = fzeros((6, 6))
    rc = lib().or_rk_ls_coeffs(stages, variant, _d(L), _d(G))
    if rc != 0:
        raise ValueError("unsupported LS-RK scheme")
    return L, G


def rk4_coeffs():
    a, b, c = np.zeros(4), np.zeros(4), np.zeros(4)
    lib().or_rk4_coeffs(_d(a), _d(b), _d(c))
    return a, b, c


def rk_ls_update(lpoin_w, l1, ls, gdt, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener):
    lib().or_rk_ls_update(rho.shape[0], lpoin_w.shape[0], _i(lpoin_w), C.c_double(l1),
                          C.c_double(ls), C.c_double(gdt), _d(rho1), _d(q1), _d(E1), _d(rho),
                          _d(q), _d(E), _d(Rmass), _d(Rmom), _d(Rener))


# ---------------------------------------------------------------- diagnostics
def volavg_ek(T, connec, gpvol, rho0, rho, u):
    f = lib().or_volavg_ek
    f.restype = C.c_double
    return f(T.p, connec.shape[0], rho.shape[0], _i(connec), _d(gpvol), _d(T.Ngp),
             C.c_double(rho0), _d(rho), _d(u))


def visc_dissipation(T, connec, rho0, mu_fluid, mu_e, u, volT, gpvol, He):
    eS, eD, eT = C.c_double(), C.c_double(), C.c_double()
    lib().or_visc_dissipation(T.p, connec.shape[0], u.shape[0], _i(connec), C.c_double(rho0),
                              _d(mu_fluid), _d(mu_e), _d(u), C.c_double(volT), _d(gpvol), _d(He),
                              _d(T.dNgp), C.byref(eS), C.byref(eD), C.byref(eT))
    return eS.value, eD.value, eT.value


def max_mach(lpoin_w, machno):
    f = lib().or_max_mach
    f.restype = C.c_double
    return f(machno.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(machno))


def adapt_dt_cfl(p, connec, helem, u, csound, cfl_conv, cfl_diff, mu_fluid, mu_sgs, mu_e, rho):
    f = lib().or_adapt_dt_cfl
    f.restype = C.c_double
    return f(p, connec.shape[0], u.shape[0], _i(connec), _d(helem), _d(u), _d(csound),
             C.c_double(cfl_conv), C.c_double(cfl_diff), _d(mu_fluid), _d(mu_sgs), _d(mu_e),
             _d(rho))


# ---------------------------------------------------------------- LS-RK sequencing
def ls_rk_step(update_f, combine, L, G, stages, dt, first, parity, on_first=None):
    """The stage sequence of one rk_4_ls_main step (time_integ_ls.f90:270-361),
    on any state: ``update_f()`` stores a fresh residual R of the working level
    (updateF, :489-691) and ``combine(l1, ls, gdt)`` sets, on the working
    level, y = l1*y1 + ls*y + gdt*R (the accumulation loops, :323-331 and
    :350-358); y1 is the level copied from the working level before the step
    (CFDSolverBase.f90:1952-1959), a copy the caller makes.

    The first step (``first``) runs ``on_first()`` (the entropy residual and
    sensor of the first-step block, :270-309) and a fresh updateF (:310-316).
    After it, stage 1 reuses the residual left by the previous step's last
    updateF (reference sequencing, ``parity``); ``parity=False`` evaluates a
    fresh one, which turns the step into the plain SSP scheme of the table.
    Returns the number of updateF calls made."""
    n = 0
    if first:
        if on_first is not None:
            on_first()
        update_f()
        n += 1
    if not parity:
        update_f()
        n += 1
    combine(0.0, 1.0, G[1, 0] * dt)
    for s in range(2, stages + 1):
        update_f()
        n += 1
        combine(L[s, 0], L[s, s - 1], G[s, s - 1] * dt)
    return n


def rk4_step(update_f, node_update, a, b, order=4):
    """The stage sequence of one step of the classical RK4 integrator rk_4_main
    (time_integ.f90:250-470; the reference compiles it out with ``#if 0`` at
    :197 and :561, so it is restated from its loops).  ``update_f(i)`` stores
    the residual R of stage i at that stage's state -- y itself at i = 0, where
    a_1 = 0 and R was zeroed (:234-240, :266-275) -- with the mass inverse
    applied (:372-435); ``node_update(a_next, b_i, first, last)`` is
    `rk4_update`'s accumulation and update.  R has the sign of the operators
    called without ``fact``: y advances by -dt * sum b_i R."""
    for i in range(order):
        update_f(i)
        last = i == order - 1
        node_update(0.0 if last else a[i + 1], b[i], i == 0, last)


def rk4_update(lpoin_w, dt, a_next, b, first, last, y, aux, R, Rsum):
    """The node loops of an rk_4_main stage over the working nodes (the
    reference loops over all npoin; with the working connectivity R is zero
    elsewhere): Rsum = Rsum + b R (time_integ.f90:442-451, zero at the first
    stage, :234-240), then aux = y - dt a_next R (:266-275) or, at the last
    stage, y = y - dt Rsum (:460-468).  y, aux, R, Rsum are tuples of
    (rho, q, E) arrays; q is (npoin, 3)."""
    lw = lpoin_w - 1
    for f in range(3):
        r = R[f][lw]
        s = (0.0 if first else Rsum[f][lw]) + b * r
        Rsum[f][lw] = s
        if last:
            y[f][lw] = y[f][lw] - dt * s
        else:
            aux[f][lw] = y[f][lw] - (dt * a_next) * r


def entropy_residual(lpoin_w, Reta, eta, dt, factor_comp=1.0):
    """The end-of-step entropy residual (time_integ_ls.f90:436-444):
    auxReta = 1.5 Reta(:,2) - 0.5 Reta(:,1) + factor_comp (eta(:,2) - eta(:,1))/dt
    on the working nodes, then Reta(:,1) = Reta(:,2).  Reta(:,2) holds the
    convective residual of the level-1 entropy (this step's start), Reta(:,1)
    the previous step's; eta(:,1), eta(:,2) the entropy before and after the
    step.  Returns auxReta (zero off the working nodes)."""
    lw = lpoin_w - 1
    aux = np.zeros(Reta.shape[0])
    aux[lw] = (1.5 * Reta[lw, 1] - 0.5 * Reta[lw, 0]) + \
        factor_comp * (eta[lw, 1] - eta[lw, 0]) / dt
    Reta[lw, 0] = Reta[lw, 1]
    return aux


# ---------------------------------------------------------------- TGV
def tgv_params(Cp=1004.0, Prt=0.71, M=0.1, Re=1600.0, rho0=1.0, gamma_gas=1.4):
    t = TGVParams(Cp, Prt, M, Re, rho0, gamma_gas, 0, 0, 0, 0)
    lib().or_tgv_params_derive(C.byref(t))
    return t


def tgv_ic(coord, t):
    npoin = coord.shape[0]
    u, q = fzeros((npoin, 3)), fzeros((npoin, 3))
    rho, pr, E, e_int, Tem, csound, eta, machno = (np.zeros(npoin) for _ in range(8))
    lib().or_tgv_ic(npoin, _d(coord), C.byref(t), _d(u), _d(q), _d(rho), _d(pr), _d(E), _d(e_int),
                    _d(Tem), _d(csound), _d(eta), _d(machno))
    return dict(u=u, q=q, rho=rho, pr=pr, E=E, e_int=e_int, Tem=Tem, csound=csound, eta=eta,
                machno=machno)


# ---------------------------------------------------------------- time loop
class TGVOracle:
    """rk_4_ls_main (time_integ_ls.f90:202-487) with updateF (:489-691) and the
    time loop of CFDSolverBase_evalTimeIteration (CFDSolverBase.f90:1929-2164),
    restated on the oracle's primitives for the periodic TGV case.  The
    entropy-viscosity sensor is smart_visc_spectral (``sensor_path`` "spectral")
    or the reference's live path (``sensor_path`` "live"):
    smart_visc_spectral_imex at the first step (time_integ_ls.f90:292) and,
    after each step, smart_visc_spectral_imex, eval_divergence,
    compute_vorticity and ducros_visc_limit (:455-463).  ``high_mach`` selects
    the high-Mach normalisation of smart_visc_spectral_imex (flag_high_mach,
    which TGVCompSolver sets, TGVCompSolver.f90:128) and ``lps`` the LPS
    stabilisation of updateF (flag_lps_stab, :607-611): comp_tau at the start
    of each step (:266-268), then full_proj_ijk and full_diff_stab_ijk in
    place of full_diffusion_ijk.

    ``geo`` provides the mesh and geometry (connec_orig/connec/lpoin_w/coord,
    He, gpvol, invMl, helem, helem_l, vol_tot) and ``ic`` the initial fields;
    ``t`` the TGV constants (an oracle TGVParams); ``cfl`` (conv, diff);
    ``nthreads`` the OpenMP threads of the three element operators (their
    assembly then adds under atomics, so the summation order, and with it the
    last bits, depend on the thread schedule)."""

    def __init__(self, p, geo, ic, t, stages=5, variant=1, parity=True, dt_mode="adaptive",
                 cfl=(0.5, 0.5), sensor=True, sensor_path="spectral", nthreads=1, lps=False,
                 high_mach=False, integrator="ls"):
        self.p = p
        self.integrator = integrator
        self.nthreads = nthreads
        self.lps = lps
        self.high_mach = high_mach
        self.T = Tables(p)
        self.geo = geo
        m = geo.mesh
        self.connec, self.lpoin_w = m.connec, m.lpoin_w
        self.npoin, self.nelem = m.npoin, m.nelem
        nn = self.T.nnode
        self.t = t
        self.L, self.G = rk_ls_coeffs(stages, variant)
        self.S = stages
        self.parity = parity
        self.dt_mode = dt_mode
        self.cfl = cfl
        self.sensor = sensor
        self.sensor_path = sensor_path
        self.lelpn = elem_per_node(m.connec, m.npoin)
        c = lambda a: np.array(a, order="F", copy=True)  # noqa: E731
        self.rho = [c(ic["rho"]), c(ic["rho"])]
        self.q = [c(ic["q"]), c(ic["q"])]
        self.E = [c(ic["E"]), c(ic["E"])]
        self.u = [c(ic["u"]), c(ic["u"])]
        self.pr = [c(ic["pr"]), c(ic["pr"])]
        self.Tem = [c(ic["Tem"]), c(ic["Tem"])]
        self.e_int = [c(ic["e_int"]), c(ic["e_int"])]
        self.eta = fzeros((self.npoin, 2))
        self.eta[:, 0] = ic["eta"]
        self.eta[:, 1] = ic["eta"]
        self.csound = c(ic["csound"])
        self.machno = c(ic["machno"])
        self.mu_e = fzeros((self.nelem, nn))
        self.mue_l = fzeros((self.nelem, nn))
        self.mu_sgs = fzeros((self.nelem, nn))
        self.mu_factor = np.full(self.npoin, t.mu_factor)
        self.mu_fluid = np.zeros(self.npoin)
        self.Rmass, self.Rmom, self.Rener = np.zeros(self.npoin), fzeros((self.npoin, 3)), \
            np.zeros(self.npoin)
        self.Reta = fzeros((self.npoin, 2))
        self.auxReta = np.zeros(self.npoin)
        self.f_eta = fzeros((self.npoin, 3))
        self.aux_h = np.zeros(self.npoin)
        self.prm = visc_params(t.Rgas)
        self.tau_stab = np.zeros(self.nelem)
        self.first = True
        self.time = 0.0
        sutherland(self.Tem[1], self.mu_factor, self.mu_fluid)
        self.dt = self.adapt_dt()

    def adapt_dt(self):
        return adapt_dt_cfl(self.p, self.connec, self.geo.helem, self.u[1], self.csound,
                            self.cfl[0], self.cfl[1], self.mu_fluid, self.mu_sgs, self.mu_e,
                            self.rho[1])

    def update_f(self):
        """updateF (time_integ_ls.f90:489-691)."""
        lw, t, g = self.lpoin_w, self.t, self.geo
        eos_stage(lw, t.gamma_gas, t.Rgas, self.rho[1], self.q[1], self.E[1], self.u[1],
                  self.e_int[1], self.pr[1], self.aux_h, self.Tem[1])
        sutherland(self.Tem[1], self.mu_factor, self.mu_fluid)
        full_convec_ijk(self.T, self.connec, g.He, g.gpvol, self.u[1], self.q[1], self.rho[1],
                        self.pr[1], self.aux_h, self.Rmass, self.Rmom, self.Rener, True, -1.0,
                        nthreads=self.nthreads)
        if self.lps:
            proj = full_proj_ijk(self.T, self.connec, lw, g.He, g.gpvol, self.rho[1], self.u[1],
                                 self.Tem[1], g.invMl, nthreads=self.nthreads)
            full_diff_stab_ijk(self.T, self.connec, g.He, g.gpvol, t.Cp, t.Prt, self.rho[1],
                               self.rho[1], self.u[1], self.Tem[1], self.mu_fluid, self.mu_e,
                               self.mu_sgs, proj, self.tau_stab, self.Rmass, self.Rmom,
                               self.Rener, False, -1.0, nthreads=self.nthreads)
        else:
            full_diffusion_ijk(self.T, self.connec, g.He, g.gpvol, t.Cp, t.Prt, self.rho[1],
                               self.rho[1], self.u[1], self.Tem[1], self.mu_fluid, self.mu_e,
                               self.mu_sgs, self.Rmass, self.Rmom, self.Rener, False, -1.0,
                               nthreads=self.nthreads)
        lumped_solver_scal(lw, g.invMl, self.Rmass)
        lumped_solver_scal(lw, g.invMl, self.Rener)
        lumped_solver_vect(lw, g.invMl, self.Rmom)

    def _scalar_convec(self, col):
        out = np.zeros(self.npoin)
        generic_scalar_convec_ijk(self.T, self.connec, self.geo.He, self.geo.gpvol, self.f_eta,
                                  np.ascontiguousarray(self.eta[:, 0]), self.u[0], out,
                                  nthreads=self.nthreads)
        lumped_solver_scal(self.lpoin_w, self.geo.invMl, out)
        self.Reta[:, col] = out

    def _sensor(self, Reta, rho, u, Tem, eta, ducros=False):
        if self.sensor_path == "spectral":
            smart_visc_spectral(self.T, self.connec, self.lpoin_w, np.ascontiguousarray(Reta),
                                self.t.gamma_gas, rho, u, Tem, np.ascontiguousarray(eta),
                                self.geo.helem_l, self.geo.helem, self.prm, self.mu_e, self.mue_l)
            return
        g = self.geo
        smart_visc_spectral_imex(self.T, self.connec, self.lpoin_w, np.ascontiguousarray(Reta),
                                 rho, u, self.csound, np.ascontiguousarray(eta), g.helem,
                                 self.prm, self.mu_e, self.mue_l,
                                 flag_high_mach=int(self.high_mach))
        if ducros:
            divU = eval_divergence(self.T, self.connec, g.He, g.gpvol, u)
            vorti = compute_vorticity(self.T, self.connec, self.lpoin_w, self.lelpn, g.He, u)
            ducros_visc_limit(self.T, self.connec, u, divU, vorti, g.Ml, g.helem_l, self.mu_e,
                              0.1)

    def update_f_rk4(self, rho, q, E):
        """The RHS of an rk_4_main stage (time_integ.f90:276-435) at the state
        (rho, q, E): EOS on the working nodes, Sutherland, full_diffusion_ijk
        and full_convec_ijk without ``fact`` (aux_fact = 1), their sum and the
        lumped solve, into Rmass/Rmom/Rener."""
        lw, t, g = self.lpoin_w, self.t, self.geo
        u, e_int, pr, Tem = fzeros((self.npoin, 3)), np.zeros(self.npoin), np.zeros(self.npoin), \
            np.zeros(self.npoin)
        eos_stage(lw, t.gamma_gas, t.Rgas, rho, q, E, u, e_int, pr, self.aux_h, Tem)
        sutherland(Tem, self.mu_factor, self.mu_fluid)
        Rd = [np.zeros(self.npoin), fzeros((self.npoin, 3)), np.zeros(self.npoin)]
        full_diffusion_ijk(self.T, self.connec, g.He, g.gpvol, t.Cp, t.Prt, rho, rho, u, Tem,
                           self.mu_fluid, self.mu_e, self.mu_sgs, *Rd, True, 1.0,
                           nthreads=self.nthreads)
        full_convec_ijk(self.T, self.connec, g.He, g.gpvol, u, q, rho, pr, self.aux_h,
                        self.Rmass, self.Rmom, self.Rener, True, 1.0, nthreads=self.nthreads)
        self.Rmass += Rd[0]
        self.Rener += Rd[2]
        self.Rmom += Rd[1]
        lumped_solver_scal(lw, g.invMl, self.Rmass)
        lumped_solver_scal(lw, g.invMl, self.Rener)
        lumped_solver_vect(lw, g.invMl, self.Rmom)

    def step_rk4(self):
        """One step of rk_4_main (time_integ.f90:250-470) for the TGV: the four
        stages with the entropy viscosity held (the step's per-stage entropy
        residual and end-of-step sensor are not restated), then the EOS and
        Sutherland of the new state (:489-533)."""
        a, b, _ = rk4_coeffs()
        y = (self.rho[1], self.q[1], self.E[1])
        aux = tuple(np.array(v, order="F", copy=True) for v in y)
        Rsum = (np.zeros(self.npoin), fzeros((self.npoin, 3)), np.zeros(self.npoin))
        R = (self.Rmass, self.Rmom, self.Rener)

        def update_f(i):
            st = y if i == 0 else aux
            self.update_f_rk4(*st)

        def node_update(a_next, bi, first, last):
            rk4_update(self.lpoin_w, self.dt, a_next, bi, first, last, y, aux, R, Rsum)

        rk4_step(update_f, node_update, a, b)
        t = self.t
        eos_step(self.lpoin_w, t.gamma_gas, t.Rgas, self.rho[1], self.q[1], self.E[1], self.u[1],
                 self.e_int[1], self.pr[1], self.csound, self.machno, self.Tem[1])
        sutherland(self.Tem[1], self.mu_factor, self.mu_fluid)
        self.time += self.dt
        if self.dt_mode == "adaptive":
            self.dt = self.adapt_dt()

    def step(self):
        if self.integrator == "rk4":
            return self.step_rk4()
        for lv in (self.rho, self.q, self.E, self.u, self.pr, self.Tem, self.e_int):
            lv[0][...] = lv[1]
        self.eta[:, 0] = self.eta[:, 1]
        lw = self.lpoin_w - 1
        if self.lps:
            self.tau_stab = comp_tau(self.p, self.connec, self.csound, self.u[1], self.geo.helem)

        def on_first():
            self.f_eta[lw, :] = self.u[0][lw, :] * self.eta[lw, 0][:, None]
            self._scalar_convec(0)
            if self.sensor:
                self._sensor(self.Reta[:, 0], self.rho[0], self.u[0], self.Tem[0], self.eta[:, 0])

        def combine(l1, ls, gdt):
            rk_ls_update(self.lpoin_w, l1, ls, gdt, self.rho[0], self.q[0], self.E[0],
                         self.rho[1], self.q[1], self.E[1], self.Rmass, self.Rmom, self.Rener)

        ls_rk_step(self.update_f, combine, self.L, self.G, self.S, self.dt, self.first,
                   self.parity, on_first)
        self.first = False
        t = self.t
        eos_step(self.lpoin_w, t.gamma_gas, t.Rgas, self.rho[1], self.q[1], self.E[1], self.u[1],
                 self.e_int[1], self.pr[1], self.csound, self.machno, self.Tem[1])
        thermo_entropy(self.lpoin_w, t.gamma_gas, t.po, t.rho0, self.rho[1], self.pr[1],
                       self.u[0], self.eta, self.f_eta)
        self._scalar_convec(1)
        self.auxReta[...] = entropy_residual(self.lpoin_w, self.Reta, self.eta, self.dt)
        if self.sensor:
            self._sensor(self.auxReta, self.rho[1], self.u[1], self.Tem[1], self.eta[:, 1],
                         ducros=True)
        sutherland(self.Tem[1], self.mu_factor, self.mu_fluid)
        self.time += self.dt
        if self.dt_mode == "adaptive":
            self.dt = self.adapt_dt()

    def diagnostics(self):
        ek = volavg_ek(self.T, self.connec, self.geo.gpvol, self.t.rho0, self.rho[1], self.u[1])
        eS, eD, eT = visc_dissipation(self.T, self.connec, self.t.rho0, self.mu_fluid, self.mu_e,
                                      self.u[1], self.geo.vol_tot, self.geo.gpvol, self.geo.He)
        mm = max_mach(self.lpoin_w, self.machno)
        return dict(time=self.time, EK=ek, eps_S=eS, eps_D=eD, eps_T=eT, maxmach=mm)
