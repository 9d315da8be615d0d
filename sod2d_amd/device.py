"""Device context and the reference-named operators on device tensors.

The functions below keep the reference's Fortran argument lists
(elem_convec.f90:227-242, elem_diffu.f90:12-27, mod_entropy_viscosity.f90:14-28,
mod_solver.f90:40-104) and call the C-ABI of libsod2d_hip.so.  A Fortran array
A(d1,...,dk) is a contiguous torch tensor whose memory is in Fortran order: a
flat tensor, or one shaped (dk,...,d1).  Unused reference arguments (Ngp, dNgp,
xgp, coord, wgp, csound, Ml, ...) are accepted and ignored, as the reference
ignores them.

The optional arguments follow the reference: ``initialze`` absent means zero the
outputs first; ``fact`` absent means the last value given to the same operator
(the reference's ``aux_fact`` is initialised in its declaration, hence SAVE'd),
starting at 1.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np
import torch

from ._lib import SodhipError, ViscParams, check, lib


def _ptr(t, dtype=None, numel=None, what="array"):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: expected a torch tensor")
    if t.device.type != "cuda":
        raise SodhipError(f"{what}: tensor must live on the GPU (got {t.device})")
    if not t.is_contiguous():
        raise SodhipError(f"{what}: tensor must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise SodhipError(f"{what}: expected {dtype}, got {t.dtype}")
    if numel is not None and t.numel() < numel:
        raise SodhipError(f"{what}: expected at least {numel} elements, got {t.numel()}")
    return C.c_void_p(t.data_ptr())


class Context:
    """A sodhip device context bound to torch's current stream."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise SodhipError("no GPU available: the sod2d_amd operators run on the device only")
        self.device = device
        h = C.c_void_p()
        check(lib().sodhip_create(device, C.byref(h)), "sodhip_create")
        self.h = h
        self.sync_stream()

    def sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().sodhip_set_stream(self.h, C.c_void_p(s)), "sodhip_set_stream")

    def synchronize(self):
        check(lib().sodhip_synchronize(self.h), "sodhip_synchronize")

    def prof_enable(self, on=True):
        check(lib().sodhip_prof_enable(self.h, int(on)), "sodhip_prof_enable")

    def prof_reset(self):
        check(lib().sodhip_prof_reset(self.h), "sodhip_prof_reset")

    def prof_get(self, kind: int):
        ms, n = C.c_double(), C.c_longlong()
        check(lib().sodhip_prof_get(self.h, kind, C.byref(ms), C.byref(n)), "sodhip_prof_get")
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            lib().sodhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(torch.cuda.current_device())
    return _default


_tracked: dict = {}


def _bind(ctx, *mesh_arrays):
    """The context to use (ctx or the default one).  The library caches tables
    derived from connectivity and working-node arrays by device address; each
    such tensor gets a finaliser that drops those tables when it is freed, so a
    later array at the same address is never matched to them."""
    ctx = ctx or default_context()
    for t in mesh_arrays:
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            continue
        key = (id(ctx), id(t))
        if key in _tracked:
            continue
        ptr = t.data_ptr()
        _tracked[key] = weakref.finalize(t, _forget, weakref.ref(ctx), ptr, key)
    return ctx


def _forget(ctx_ref, ptr, key):
    _tracked.pop(key, None)
    ctx = ctx_ref()
    if ctx is not None and getattr(ctx, "h", None):
        lib().sodhip_forget_mesh(ctx.h, C.c_void_p(ptr))


def _porder(nelem, connec):
    nn = connec.numel() // nelem
    p = round(nn ** (1.0 / 3.0)) - 1
    if (p + 1) ** 3 != nn:
        raise SodhipError(f"connec has {nn} nodes per element: not a hexahedron of order p")
    return p


_saved_fact = {"convec": 1.0, "diffu": 1.0, "scalar": 1.0, "diffstab": 1.0}


def _resolve(name, initialze, fact):
    if fact is not None:
        _saved_fact[name] = float(fact)
    init = 1 if initialze is None else int(bool(initialze))
    return init, _saved_fact[name]


def full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK,
                    gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener,
                    initialze=None, fact=None, ctx: Context | None = None):
    """full_convec_ijk (elem_convec.f90:227-470) on the device."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    init, f = _resolve("convec", initialze, fact)
    dt = rho.dtype
    nn = (p + 1) ** 3
    if dt == torch.float64:
        fn, fc = lib().sodhip_full_convec_ijk_f64, C.c_double
    elif dt == torch.float32:
        fn, fc = lib().sodhip_full_convec_ijk_f32, C.c_float
    else:
        raise SodhipError(f"unsupported dtype {dt}")
    I32 = torch.int32
    rc = fn(ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
            _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(gpvol, dt, nn * nelem, "gpvol"),
            _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
            _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn),
            _ptr(u, dt, 3 * npoin, "u"), _ptr(q, dt, 3 * npoin, "q"), _ptr(rho, dt, npoin, "rho"),
            _ptr(pr, dt, npoin, "pr"), _ptr(E, dt, npoin, "E"), _ptr(Rmass, dt, npoin, "Rmass"),
            _ptr(Rmom, dt, 3 * npoin, "Rmom"), _ptr(Rener, dt, npoin, "Rener"), init, fc(f))
    check(rc, "full_convec_ijk")


def full_diffusion_ijk(nelem, npoin, connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                       gmshAtoJ, gmshAtoK, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, Ml,
                       Rmass, Rmom, Rener, initialze=None, fact=None, c_rho=1.0, c_ener=1.0,
                       ctx: Context | None = None):
    """full_diffusion_ijk (elem_diffu.f90:12-223) on the device."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    init, f = _resolve("diffu", initialze, fact)
    dt = rho.dtype
    nn = (p + 1) ** 3
    if dt == torch.float64:
        fn, fc = lib().sodhip_full_diffusion_ijk_f64, C.c_double
    elif dt == torch.float32:
        fn, fc = lib().sodhip_full_diffusion_ijk_f32, C.c_float
    else:
        raise SodhipError(f"unsupported dtype {dt}")
    I32 = torch.int32
    rc = fn(ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
            _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(gpvol, dt, nn * nelem, "gpvol"),
            _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
            _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn),
            fc(Cp), fc(Pr), _ptr(rho_n, dt, npoin, "rho_n"), _ptr(rho, dt, npoin, "rho"),
            _ptr(u, dt, 3 * npoin, "u"), _ptr(Tem, dt, npoin, "Tem"),
            _ptr(mu_fluid, dt, npoin, "mu_fluid"), _ptr(mu_e, dt, nn * nelem, "mu_e"),
            _ptr(mu_sgs, dt, nn * nelem, "mu_sgs"), fc(c_rho), fc(c_ener),
            _ptr(Rmass, dt, npoin, "Rmass"), _ptr(Rmom, dt, 3 * npoin, "Rmom"),
            _ptr(Rener, dt, npoin, "Rener"), init, fc(f))
    check(rc, "full_diffusion_ijk")


C_LPS_COMP = float(np.float32(0.1))  # c_lps_comp: a default-real constant
C_SPECIES_STAB = 0.5


def comp_tau(nelem, npoin, connec, csound, u, helem_k, dt, tau, c_species_stab=C_SPECIES_STAB,
             flag_drop_c_in_envit=0, ctx: Context | None = None):
    """comp_tau (elem_stab.f90:589-616) on the device (dt is unused there too)."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    F64, I32 = torch.float64, torch.int32
    check(lib().sodhip_comp_tau_f64(ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
                                    _ptr(csound, F64, npoin, "csound"), _ptr(u, F64, 3 * npoin, "u"),
                                    _ptr(helem_k, F64, nelem, "helem_k"), float(c_species_stab),
                                    int(bool(flag_drop_c_in_envit)), _ptr(tau, F64, nelem, "tau")),
          "comp_tau")


def full_proj_ijk(nelem, npoin, npoin_w, connec, lpoin_w, Ngp, He, gpvol, dlxigp_ip, invAtoIJK,
                  gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, Ml, invMl, ProjMass, ProjEner,
                  ProjMX, ProjMY, ProjMZ, ctx: Context | None = None):
    """full_proj_ijk (elem_stab.f90:441-586) on the device."""
    ctx = _bind(ctx, connec, lpoin_w)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    F64, I32 = torch.float64, torch.int32
    out = [_ptr(x, F64, 3 * npoin, nm) for x, nm in
           ((ProjMass, "ProjMass"), (ProjEner, "ProjEner"), (ProjMX, "ProjMX"), (ProjMY, "ProjMY"),
            (ProjMZ, "ProjMZ"))]
    check(lib().sodhip_full_proj_ijk_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(lpoin_w, I32, npoin_w, "lpoin_w"), _ptr(He, F64, 9 * nn * nelem, "He"),
        _ptr(gpvol, F64, nn * nelem, "gpvol"), _ptr(dlxigp_ip, F64, 3 * nn * (p + 1), "dlxigp_ip"),
        _ptr(invAtoIJK, I32, nn), _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn),
        _ptr(gmshAtoK, I32, nn), _ptr(rho, F64, npoin, "rho"), _ptr(u, F64, 3 * npoin, "u"),
        _ptr(Tem, F64, npoin, "Tem"), _ptr(invMl, F64, npoin, "invMl"), *out), "full_proj_ijk")


def full_diff_stab_ijk(nelem, npoin, connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                       gmshAtoJ, gmshAtoK, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, Ml,
                       ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ, tau_stab, Rmass, Rmom, Rener,
                       initialze=None, fact=None, c_rho=1.0, c_ener=1.0, c_lps=C_LPS_COMP,
                       ctx: Context | None = None):
    """full_diff_stab_ijk (elem_stab.f90:207-438) on the device."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    init, f = _resolve("diffstab", initialze, fact)
    nn = (p + 1) ** 3
    F64, I32 = torch.float64, torch.int32
    proj = [_ptr(x, F64, 3 * npoin, nm) for x, nm in
            ((ProjMass, "ProjMass"), (ProjEner, "ProjEner"), (ProjMX, "ProjMX"),
             (ProjMY, "ProjMY"), (ProjMZ, "ProjMZ"))]
    check(lib().sodhip_full_diff_stab_ijk_f64(
        ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(He, F64, 9 * nn * nelem, "He"), _ptr(gpvol, F64, nn * nelem, "gpvol"),
        _ptr(dlxigp_ip, F64, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
        _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn), float(Cp),
        float(Pr), _ptr(rho_n, F64, npoin, "rho_n"), _ptr(rho, F64, npoin, "rho"),
        _ptr(u, F64, 3 * npoin, "u"), _ptr(Tem, F64, npoin, "Tem"),
        _ptr(mu_fluid, F64, npoin, "mu_fluid"), _ptr(mu_e, F64, nn * nelem, "mu_e"),
        _ptr(mu_sgs, F64, nn * nelem, "mu_sgs"), *proj, _ptr(tau_stab, F64, nelem, "tau_stab"),
        float(c_rho), float(c_ener), float(c_lps), _ptr(Rmass, F64, npoin, "Rmass"),
        _ptr(Rmom, F64, 3 * npoin, "Rmom"), _ptr(Rener, F64, npoin, "Rener"), init, float(f)),
        "full_diff_stab_ijk")


def generic_scalar_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp,
                              invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, q, eta, u, Rconvec,
                              initialze=None, fact=None, ctx: Context | None = None):
    """generic_scalar_convec_ijk (elem_convec.f90:472-594) on the device (FP64); q may
    be None: the flux is then u * eta, formed in the kernel."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    init, f = _resolve("scalar", initialze, fact)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    rc = lib().sodhip_generic_scalar_convec_ijk_f64(
        ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(gpvol, dt, nn * nelem, "gpvol"),
        _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
        _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn),
        _ptr(q, dt, 3 * npoin, "q"), _ptr(eta, dt, npoin, "eta"), _ptr(u, dt, 3 * npoin, "u"),
        _ptr(Rconvec, dt, npoin, "Rconvec"), init, C.c_double(f))
    check(rc, "generic_scalar_convec_ijk")


def visc_params(Rgas, ce=0.1, cmax=0.5, cglob=1.0, flag_normalise_entropy=1,
                flag_drop_c_in_envit=0) -> ViscParams:
    """Module globals of mod_numerical_params.f90 read by smart_visc_spectral."""
    return ViscParams(ce, cmax, cglob, Rgas, flag_normalise_entropy, flag_drop_c_in_envit)


def smart_visc_spectral(nelem, npoin, npoin_w, connec, lpoin_w, Reta, Ngp, coord, dNgp, gpvol,
                        wgp, gamma_gas, rho, u, csound, Tem, eta, helem, helem_k, Ml, mu_e,
                        invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mue_l, params: ViscParams,
                        ctx: Context | None = None):
    """smart_visc_spectral (mod_entropy_viscosity.f90:14-111) on the device (FP64)."""
    ctx = _bind(ctx, connec, lpoin_w)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    rc = lib().sodhip_smart_visc_spectral_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(lpoin_w, I32, npoin_w, "lpoin_w"), _ptr(Reta, dt, npoin, "Reta"), float(gamma_gas),
        _ptr(rho, dt, npoin, "rho"), _ptr(u, dt, 3 * npoin, "u"), _ptr(Tem, dt, npoin, "Tem"),
        _ptr(eta, dt, npoin, "eta"), _ptr(helem, dt, nelem * nn, "helem"),
        _ptr(helem_k, dt, nelem, "helem_k"), _ptr(invAtoIJK, I32, nn), C.byref(params),
        _ptr(mu_e, dt, nelem * nn, "mu_e"), _ptr(mue_l, dt, nelem * nn, "mue_l"))
    check(rc, "smart_visc_spectral")


def smart_visc_spectral_imex(nelem, npoin, npoin_w, connec, lpoin_w, Reta, Ngp, coord, dNgp,
                             gpvol, wgp, gamma_gas, rho, u, csound, Tem, eta, helem, helem_k, Ml,
                             mu_e, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mue_l,
                             params: ViscParams, ce_comp: float = 1.0, flag_high_mach: int = 0,
                             ctx: Context | None = None):
    """smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227) on the device
    (FP64), the reference's live-path sensor.  With the normalisation and
    flag_high_mach, params.ce becomes ce_comp (the reference's module global)."""
    ctx = _bind(ctx, connec, lpoin_w)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    rc = lib().sodhip_smart_visc_spectral_imex_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(lpoin_w, I32, npoin_w, "lpoin_w"), _ptr(Reta, dt, npoin, "Reta"),
        _ptr(rho, dt, npoin, "rho"), _ptr(u, dt, 3 * npoin, "u"),
        _ptr(csound, dt, npoin, "csound"), _ptr(eta, dt, npoin, "eta"),
        _ptr(helem_k, dt, nelem, "helem_k"), C.byref(params), C.c_double(ce_comp),
        int(flag_high_mach), _ptr(mu_e, dt, nelem * nn, "mu_e"),
        _ptr(mue_l, dt, nelem * nn, "mue_l"))
    check(rc, "smart_visc_spectral_imex")


def eval_divergence(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                    gmshAtoK, u, divU, ctx: Context | None = None):
    """eval_divergence (mod_operators.f90:485-570) on the device (FP64)."""
    ctx = _bind(ctx, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    check(lib().sodhip_eval_divergence_f64(
        ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(gpvol, dt, nn * nelem, "gpvol"),
        _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
        _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn),
        _ptr(u, dt, 3 * npoin, "u"), _ptr(divU, dt, npoin, "divU")), "eval_divergence")


def compute_vorticity(nelem, npoin, npoin_w, lpoin_w, connec, lelpn, He, dlxigp_ip, invAtoIJK,
                      gmshAtoI, gmshAtoJ, gmshAtoK, u, curlU, ctx: Context | None = None):
    """compute_vorticity with makeConti (mod_operators.f90:684-775) on the device."""
    ctx = _bind(ctx, lpoin_w, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    check(lib().sodhip_compute_vorticity_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(lpoin_w, I32, npoin_w, "lpoin_w"),
        _ptr(connec, I32, nelem * nn, "connec"), _ptr(lelpn, I32, npoin, "lelpn"),
        _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"),
        _ptr(invAtoIJK, I32, nn), _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn),
        _ptr(gmshAtoK, I32, nn), _ptr(u, dt, 3 * npoin, "u"), _ptr(curlU, dt, 3 * npoin, "curlU")),
        "compute_vorticity")


def div_curl(nelem, npoin, npoin_w, lpoin_w, connec, lelpn, He, gpvol, dlxigp_ip, invAtoIJK,
             gmshAtoI, gmshAtoJ, gmshAtoK, u, divU, curlU, ctx: Context | None = None):
    """eval_divergence and compute_vorticity of one u in a single element pass."""
    ctx = _bind(ctx, lpoin_w, connec)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    check(lib().sodhip_div_curl_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(lpoin_w, I32, npoin_w, "lpoin_w"),
        _ptr(connec, I32, nelem * nn, "connec"), _ptr(lelpn, I32, npoin, "lelpn"),
        _ptr(He, dt, 9 * nn * nelem, "He"), _ptr(gpvol, dt, nn * nelem, "gpvol"),
        _ptr(dlxigp_ip, dt, 3 * nn * (p + 1), "dlxigp_ip"), _ptr(invAtoIJK, I32, nn),
        _ptr(gmshAtoI, I32, nn), _ptr(gmshAtoJ, I32, nn), _ptr(gmshAtoK, I32, nn),
        _ptr(u, dt, 3 * npoin, "u"), _ptr(divU, dt, npoin, "divU"),
        _ptr(curlU, dt, 3 * npoin, "curlU")), "div_curl")


def ducros_visc_limit(nelem, npoin, npoin_w, connec, lpoin_w, Ngp, coord, dNgp, gpvol, wgp, rho,
                      u, divU, vorti, csound, helem, helem_k, Ml, mu_e, invAtoIJK, gmshAtoI,
                      gmshAtoJ, gmshAtoK, mue_l, ducros_min_val: float = 0.1,
                      ctx: Context | None = None):
    """ducros_visc_limit (mod_entropy_viscosity.f90:277-314) on the device: mu_e
    scaled in place (rho, csound, helem_k and the other reference arguments do
    not enter the result)."""
    ctx = _bind(ctx, connec, lpoin_w)
    ctx.sync_stream()
    p = _porder(nelem, connec)
    nn = (p + 1) ** 3
    dt, I32 = torch.float64, torch.int32
    check(lib().sodhip_ducros_visc_limit_f64(
        ctx.h, p, nelem, npoin, _ptr(connec, I32, nelem * nn, "connec"),
        _ptr(u, dt, 3 * npoin, "u"), _ptr(divU, dt, npoin, "divU"),
        _ptr(vorti, dt, 3 * npoin, "vorti"), _ptr(Ml, dt, npoin, "Ml"),
        _ptr(helem, dt, nelem * nn, "helem"), C.c_double(ducros_min_val),
        _ptr(mu_e, dt, nelem * nn, "mu_e")), "ducros_visc_limit")


def lumped_solver_scal(npoin, npoin_w, lpoin_w, invMl, R, ctx: Context | None = None):
    ctx = _bind(ctx, lpoin_w)
    ctx.sync_stream()
    check(lib().sodhip_lumped_solver_scal_f64(ctx.h, npoin, npoin_w,
                                              _ptr(lpoin_w, torch.int32, npoin_w),
                                              _ptr(invMl, torch.float64, npoin),
                                              _ptr(R, torch.float64, npoin)), "lumped_solver_scal")


def lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, R, ctx: Context | None = None):
    ctx = _bind(ctx, lpoin_w)
    ctx.sync_stream()
    check(lib().sodhip_lumped_solver_vect_f64(ctx.h, npoin, npoin_w,
                                              _ptr(lpoin_w, torch.int32, npoin_w),
                                              _ptr(invMl, torch.float64, npoin),
                                              _ptr(R, torch.float64, 3 * npoin)),
          "lumped_solver_vect")


def rk_stage_update(npoin, npoin_w, lpoin_w, invMl, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E,
                    Rmass, Rmom, Rener, ctx: Context | None = None):
    """Lumped inverse fused with the Shu-Osher stage combine (time_integ_ls.f90:323-358)."""
    ctx = _bind(ctx, lpoin_w)
    ctx.sync_stream()
    d, I32 = torch.float64, torch.int32
    check(lib().sodhip_rk_stage_update_f64(
        ctx.h, npoin, npoin_w, _ptr(lpoin_w, I32, npoin_w), _ptr(invMl, d, npoin), float(dt),
        float(lam1), float(lams), float(gam), _ptr(rho1, d, npoin), _ptr(q1, d, 3 * npoin),
        _ptr(E1, d, npoin), _ptr(rho, d, npoin), _ptr(q, d, 3 * npoin), _ptr(E, d, npoin),
        _ptr(Rmass, d, npoin), _ptr(Rmom, d, 3 * npoin), _ptr(Rener, d, npoin)),
        "rk_stage_update")


def to_dev(a, device="cuda"):
    """numpy array (any order) -> flat device tensor holding its Fortran-order memory."""
    import numpy as np
    flat = np.asarray(a).ravel(order="F")
    return torch.from_numpy(np.ascontiguousarray(flat)).to(device)


def to_host(t, shape):
    """flat device tensor in Fortran order -> numpy array of the given Fortran shape."""
    return t.detach().cpu().numpy().reshape(shape, order="F")
