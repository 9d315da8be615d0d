"""Time-step sequencer of rk_4_ls_main on the device, for the periodic
Taylor-Green vortex case.

This is the host side of the reference's time loop (CFDSolverBase.f90:1929-2164
calling rk_4_ls_main, time_integ_ls.f90:202-487, whose stages call updateF,
time_integ_ls.f90:489-691), restated over the C-ABI: every array lives on the
device for the whole run and the host only issues launches.  The RHS of a
stage, its lumped solve and the stage combine are one fused C-ABI call
(sodhip_rk_stage_f64).

Modes:
  * ``parity=True`` follows the reference exactly: the first stage of every
    step after the first reuses the residual of the previous step's last RHS
    (time_integ_ls.f90:323-331), so a step performs S-1 fresh RHS evaluations;
  * ``parity=False`` evaluates a fresh RHS at stage 1 (the textbook scheme).
  * ``op_precision`` 8 (FP64 operators) or 4 (FP32 operator arithmetic, FP64
    state, gather, mass inverse, stage combine and reductions).
  * ``dt_mode`` "adaptive" (adapt_dt_cfl after every step, as the reference) or
    "fixed" (the CFL time step of the initial state, held).  At fixed dt with
    the reference sequencing the last stage of a step also stores the next
    step's stage 1 (sodhip_rk_stage_next_f64, the same arithmetic), so that
    stage costs no launch of its own.

The entropy-viscosity sensor (``sensor_path``) is either smart_visc_spectral
(mod_entropy_viscosity.f90:14-111, "spectral") at the points where the
reference's live path calls its _imex variant, or that live path itself
("live"): smart_visc_spectral_imex at the first step (time_integ_ls.f90:292)
and, after every step, smart_visc_spectral_imex, eval_divergence,
compute_vorticity and ducros_visc_limit (time_integ_ls.f90:455-463);
``high_mach`` selects that sensor's high-Mach normalisation (flag_high_mach,
set by the reference's compressible TGV solver).

``lps=True`` turns on the LPS stabilisation of the viscous operator
(flag_lps_stab, the reference default): comp_tau at the start of every step
and, in every updateF, full_proj_ijk and full_diff_stab_ijk in place of
full_diffusion_ijk.  A stage of that path is one C-ABI call,
sodhip_rk_stage_lps_f64, built on the drop-in operators' kernels (convection
and the stabilised viscous terms summed per element, one gather/update).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import case as case_mod
from ._lib import SodhipError, StageParams, ViscParams, check, lib
from .device import C_LPS_COMP, C_SPECIES_STAB, Context, _bind, _ptr

F64, F32, I32 = torch.float64, torch.float32, torch.int32

# classical RK4 tableau of rk_4_main (time_integ.f90:84-87): stage abscissae
# a_i and weights b_i
RK4_A = (0.0, 0.5, 0.5, 1.0)
RK4_B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)


@dataclass
class TGVParams:
    """TGVSolver parameters (TGVSolver.f90:39-152) and their derived constants."""
    Cp: float = 1004.0
    Prt: float = 0.71
    M: float = 0.1
    Re: float = 1600.0
    rho0: float = 1.0
    gamma_gas: float = 1.4
    cfl_conv: float = 0.5
    cfl_diff: float = 0.5

    def __post_init__(self):
        mul = (self.rho0 * 1.0 * 1.0) / self.Re
        self.Rgas = self.Cp * (self.gamma_gas - 1.0) / self.gamma_gas
        self.to = 1.0 * 1.0 / (self.gamma_gas * self.Rgas * self.M * self.M)
        self.po = self.rho0 * self.Rgas * self.to
        mur = 0.000001458 * self.to ** 1.5 / (self.to + 110.40)
        self.mu_factor = mul / mur


def tgv_initial_conditions(coord: np.ndarray, t: TGVParams, uniform_rho: bool = False) -> dict:
    """TGVSolver_evalInitialConditions (TGVSolver.f90:154-221).

    The coordinates are rounded to real(4) and the velocity is evaluated in single
    precision, as the reference; the pressure in double.  ``uniform_rho`` sets
    rho = rho0 (the variant some TGV set-ups use) instead of p/(R T0)."""
    x32 = coord.astype(np.float32)
    x, y, z = x32[:, 0], x32[:, 1], x32[:, 2]
    V0 = np.float32(1.0)
    npoin = coord.shape[0]
    u = np.zeros((npoin, 3), order="F")
    u[:, 0] = (V0 * np.sin(x) * np.cos(y) * np.cos(z)).astype(np.float64)
    u[:, 1] = (-(V0 * np.cos(x) * np.sin(y) * np.cos(z))).astype(np.float64)
    xd, yd, zd = (v.astype(np.float64) for v in (x, y, z))
    pr = t.po + ((1.0 * 1.0 * 1.0) / 16.0) * (np.cos(2.0 * xd) + np.cos(2.0 * yd)) * \
        (np.cos(2.0 * zd) + 2.0)
    rho = np.full(npoin, t.rho0) if uniform_rho else pr / t.Rgas / t.to
    g = t.gamma_gas
    uu = (u * u).sum(1)
    e_int = pr / (rho * (g - 1.0))
    Tem = pr / (rho * t.Rgas)
    E = rho * (0.5 * uu + e_int)
    q = np.asfortranarray(rho[:, None] * u)
    csound = np.sqrt(g * pr / rho)
    eta = (rho / (g - 1.0)) * np.log(pr / rho ** g)
    machno = uu / csound
    return dict(u=u, q=q, rho=rho, pr=pr, E=E, e_int=e_int, Tem=Tem, csound=csound, eta=eta,
                machno=machno)


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()
    return t.to(dtype) if dtype is not None else t


class TGVRun:
    """Device state and time loop of one periodic TGV run."""

    def __init__(self, porder: int = 3, nelem_dir: int = 16, params: TGVParams | None = None,
                 length: float = 2.0 * math.pi, stages: int = 5, variant: int = 1,
                 op_precision: int = 8, parity: bool = True, dt_mode: str = "adaptive",
                 uniform_rho: bool = False, sensor: bool = True, ctx: Context | None = None,
                 geometry=None, initial=None, sensor_path: str = "spectral", lps: bool = False,
                 high_mach: bool = False, integrator: str = "ls", surrogate=None,
                 judge_every: int = 10, instances=None):
        self.t = params or TGVParams()
        # joined instances (sod2d_amd.ensemble): per-instance Sutherland factor
        # and entropy reference state, per-instance entropy normalisation
        self.ens = instances
        if instances is not None and (op_precision != 8 or lps or integrator != "ls" or
                                      sensor_path != "spectral"):
            raise SodhipError("joined instances run the FP64 plain path with the spectral sensor")
        if integrator not in ("ls", "rk4"):
            raise SodhipError(f"unknown integrator {integrator!r}")
        if integrator == "rk4" and lps:
            raise SodhipError("the classical RK4 integrator runs the plain viscous path")
        self.integrator = integrator
        self.ctx = ctx or Context(torch.cuda.current_device())
        self.p = porder
        self.parity = parity
        self.dt_mode = dt_mode
        self.op_precision = op_precision
        self.sensor = sensor
        if sensor_path not in ("spectral", "live", "surrogate"):
            raise SodhipError(f"unknown sensor_path {sensor_path!r}")
        if sensor_path == "surrogate" and (surrogate is None or porder != 3):
            raise SodhipError("the surrogate sensor needs a trained SurrogateMLP and porder 3")
        # the MLP surrogate of smart_visc_spectral; the plain sensor runs
        # instead every judge_every steps (and per element as its fallback)
        self.surrogate = surrogate
        self.judge_every = judge_every
        self.judge_log = []  # (step, relative error of the surrogate's mu_e, fallback elements)
        self.sensor_path = sensor_path
        # LPS stabilisation (flag_lps_stab) runs updateF through the drop-in
        # operators (FP64): convection, full_proj_ijk and full_diff_stab_ijk
        if lps and op_precision != 8:
            raise SodhipError("LPS stabilisation runs with FP64 operators only")
        self.lps = lps
        self.high_mach = high_mach
        g = geometry or case_mod.build_case(porder, nelem_dir, periodic=True,
                                             lengths=(length,) * 3)
        self.geo = g
        m, T = g.mesh, g.tables
        self.nelem, self.npoin, self.npoin_w, self.nnode = m.nelem, m.npoin, m.npoin_w, T.nnode
        self.L, self.G = rk_ls_coeffs(stages, variant)
        self.S = stages
        # static device data
        self.connec = _dev(m.connec)
        self.lpoin_w = _dev(m.lpoin_w)
        _bind(self.ctx, self.connec, self.lpoin_w)
        self.tabs = [_dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
        self.He = _dev(g.He)
        self.gpvol = _dev(g.gpvol)
        self.dl = _dev(T.dlxigp_ip)
        if op_precision == 4:
            self.He_op, self.gpvol_op, self.dl_op = (x.to(F32) for x in (self.He, self.gpvol, self.dl))
        else:
            self.He_op, self.gpvol_op, self.dl_op = self.He, self.gpvol, self.dl
        self.invMl = _dev(g.invMl)
        self.helem = _dev(g.helem)
        self.helem_l = _dev(g.helem_l)
        self.vol_tot = g.vol_tot
        if sensor_path == "surrogate":
            self._sx = torch.empty(m.nelem * 16, dtype=F32, device="cuda")
            self._sbeta = torch.empty(m.nelem, dtype=F64, device="cuda")
            self._sy = torch.empty(m.nelem, dtype=F32, device="cuda")
            self._sflag = torch.zeros(m.nelem, dtype=I32, device="cuda")
            self._snfall = torch.zeros(1, dtype=I32, device="cuda")
        if sensor_path == "live":
            self.Ml = _dev(g.Ml)
            self.lelpn = _dev(case_mod.elem_per_node(m.connec, m.npoin))
        # state
        ic = initial if initial is not None else tgv_initial_conditions(m.coord, self.t, uniform_rho)
        n, ne, nn = self.npoin, self.nelem, self.nnode
        z = lambda k: torch.zeros(k, dtype=F64, device="cuda")  # noqa: E731
        self.rho = [_dev(ic["rho"]), _dev(ic["rho"])]        # levels 1, 2
        self.q = [_dev(ic["q"]), _dev(ic["q"])]
        self.E = [_dev(ic["E"]), _dev(ic["E"])]
        self.u = [_dev(ic["u"]), _dev(ic["u"])]
        self.pr = [_dev(ic["pr"]), _dev(ic["pr"])]
        self.Tem = [_dev(ic["Tem"]), _dev(ic["Tem"])]
        self.e_int = [_dev(ic["e_int"]), _dev(ic["e_int"])]
        self.eta = torch.cat([_dev(ic["eta"]), _dev(ic["eta"])])  # eta(npoin,2)
        self.csound = _dev(ic["csound"])
        self.machno = _dev(ic["machno"])
        self.mu_e = z(ne * nn)
        self.mu_e_op = self.mu_e if op_precision == 8 else self.mu_e.to(F32)
        self.mue_l = z(ne * nn)
        self.mu_fluid = z(n)
        self.Rmass, self.Rmom, self.Rener = z(n), z(3 * n), z(n)
        self.Reta = z(2 * n)
        self.auxReta = z(n)
        if sensor_path == "live":
            self.divU, self.vorti = z(n), z(3 * n)
        if lps:
            # the projections stay inside the stage call (ProjMass, ProjEner,
            # ProjMX, MY, MZ are not kept)
            self.proj = [None] * 5
            self.tau_stab = z(ne)
            self.aux_h = z(n)
        if integrator == "rk4":
            # the stage state (aux_*) and the residual sums of rk_4_main
            self.aux = [z(n), z(3 * n), z(n)]
            self.Rsum = [z(n), z(3 * n), z(n)]
            self.sp4 = StageParams(self.t.gamma_gas, self.t.Rgas, self.t.Cp, self.t.Prt,
                                   self.t.mu_factor, 1.0, 1.0, 1.0,
                                   self.t.po if op_precision == 4 else 0.0)
        self.first = True
        self._next_dt = None  # dt with which the next step's stage 1 was stored
        self.fuse_next_stage = True  # at fixed dt, see the module docstring
        # EOS, entropy and Sutherland of the end of a step in one node pass,
        # and the entropy residual's gather, lumped solve and update in another
        self.fuse_end_state = True
        self.time = 0.0
        self.istep = 0
        # FP32 operators: the uniform reference pressure is removed before the
        # operator arithmetic (only grad p enters the split form)
        self.sp = StageParams(self.t.gamma_gas, self.t.Rgas, self.t.Cp, self.t.Prt,
                              self.t.mu_factor, 1.0, 1.0, -1.0,
                              self.t.po if op_precision == 4 else 0.0)
        if self.ens is not None:
            self.sp.mu_factor_e = self.ens.mu_factor_e.data_ptr()
        self.vp = ViscParams(0.1, 0.5, 1.0, self.t.Rgas, 1, 0)
        self.ctx.sync_stream()
        # evalInitialViscosity + evalInitialDt (CFDSolverBase.f90:1384-1460)
        self._sutherland()
        self.dt = self.adapt_dt()

    # ------------------------------------------------------------ helpers
    def _sutherland(self):
        muf = _ptr(self.ens.mu_factor) if self.ens is not None else None
        check(lib().sodhip_sutherland_f64(self.ctx.h, self.npoin, _ptr(self.Tem[1]), muf,
                                          self.t.mu_factor, _ptr(self.mu_fluid)), "sutherland")

    def adapt_dt(self) -> float:
        dt = C.c_double()
        check(lib().sodhip_adapt_dt_cfl_f64(
            self.ctx.h, self.p, self.nelem, self.npoin, _ptr(self.connec), _ptr(self.helem),
            _ptr(self.u[1]), _ptr(self.csound), self.t.cfl_conv, self.t.cfl_diff,
            _ptr(self.mu_fluid), None, _ptr(self.mu_e), _ptr(self.rho[1]), C.byref(dt)),
            "adapt_dt_cfl")
        return dt.value

    def _scalar_convec(self, out):
        # the flux f_eta = u(:,:,1) * eta(:,1) is formed in the kernel (q = NULL)
        check(lib().sodhip_generic_scalar_convec_ijk_f64(
            self.ctx.h, self.p, self.nelem, self.npoin, _ptr(self.connec), _ptr(self.He),
            _ptr(self.gpvol), _ptr(self.dl), *[_ptr(x) for x in self.tabs], None,
            _ptr(self.eta), _ptr(self.u[0]), _ptr(out), 1, 1.0), "generic_scalar_convec_ijk")
        check(lib().sodhip_lumped_solver_scal_f64(self.ctx.h, self.npoin, self.npoin_w,
                                                  _ptr(self.lpoin_w), _ptr(self.invMl), _ptr(out)),
              "lumped_solver_scal")

    def _surrogate_visc(self, Reta, rho, u, Tem, eta, mu_e):
        h, p, ne, n, nw = self.ctx.h, self.p, self.nelem, self.npoin, self.npoin_w
        check(lib().sodhip_surrogate_visc_spectral_f64(
            h, p, ne, n, nw, _ptr(self.connec), _ptr(self.lpoin_w), _ptr(Reta), self.t.gamma_gas,
            _ptr(rho), _ptr(u), _ptr(Tem), _ptr(eta), _ptr(self.helem_l), _ptr(self.helem),
            _ptr(self.tabs[0]), C.byref(self.vp), C.byref(self.surrogate.struct()), _ptr(mu_e),
            _ptr(self.mue_l), _ptr(self._sx), _ptr(self._sbeta), _ptr(self._sy),
            _ptr(self._sflag), _ptr(self._snfall)), "surrogate_visc_spectral")

    def _smart_visc(self, Reta, rho, u, Tem, eta, ducros=False):
        h, p, ne, n, nw = self.ctx.h, self.p, self.nelem, self.npoin, self.npoin_w
        if self.sensor_path == "surrogate":
            judge = self.judge_every > 0 and self.istep % self.judge_every == 0
            if not judge:
                self._surrogate_visc(Reta, rho, u, Tem, eta, self.mu_e)
                if self.op_precision == 4:
                    self.mu_e_op.copy_(self.mu_e)
                return
            # judge step: the plain sensor sets mu_e; the surrogate's values
            # beside it are logged (relative error, fallback count)
            mu_s = torch.empty_like(self.mu_e)
            self._surrogate_visc(Reta, rho, u, Tem, eta, mu_s)
        if self.ens is not None:
            check(lib().sodhip_smart_visc_spectral_ens_f64(
                h, p, ne, n, nw, _ptr(self.connec), _ptr(self.lpoin_w), C.byref(self.ens.struct),
                _ptr(Reta), self.t.gamma_gas, _ptr(rho), _ptr(u), _ptr(Tem), _ptr(eta),
                _ptr(self.helem_l), _ptr(self.helem), _ptr(self.tabs[0]), C.byref(self.vp),
                _ptr(self.mu_e), _ptr(self.mue_l)), "smart_visc_spectral_ens")
        elif self.sensor_path in ("spectral", "surrogate"):
            check(lib().sodhip_smart_visc_spectral_f64(
                h, p, ne, n, nw, _ptr(self.connec), _ptr(self.lpoin_w), _ptr(Reta),
                self.t.gamma_gas, _ptr(rho), _ptr(u), _ptr(Tem), _ptr(eta), _ptr(self.helem_l),
                _ptr(self.helem), _ptr(self.tabs[0]), C.byref(self.vp), _ptr(self.mu_e),
                _ptr(self.mue_l)), "smart_visc_spectral")
            if self.sensor_path == "surrogate":
                err = float(torch.linalg.norm(mu_s - self.mu_e) /
                            torch.clamp(torch.linalg.norm(self.mu_e), min=1e-300))
                self.judge_log.append((self.istep, err, int(self._snfall.item())))
        else:
            check(lib().sodhip_smart_visc_spectral_imex_f64(
                h, p, ne, n, nw, _ptr(self.connec), _ptr(self.lpoin_w), _ptr(Reta), _ptr(rho),
                _ptr(u), _ptr(self.csound), _ptr(eta), _ptr(self.helem), C.byref(self.vp),
                C.c_double(1.0), int(self.high_mach), _ptr(self.mu_e), _ptr(self.mue_l)),
                "smart_visc_spectral_imex")
            if ducros:
                check(lib().sodhip_div_curl_f64(
                    h, p, ne, n, nw, _ptr(self.lpoin_w), _ptr(self.connec), _ptr(self.lelpn),
                    _ptr(self.He), _ptr(self.gpvol), _ptr(self.dl),
                    *[_ptr(x) for x in self.tabs], _ptr(u), _ptr(self.divU), _ptr(self.vorti)),
                    "div_curl")
                check(lib().sodhip_ducros_visc_limit_f64(
                    h, p, ne, n, _ptr(self.connec), _ptr(u), _ptr(self.divU), _ptr(self.vorti),
                    _ptr(self.Ml), _ptr(self.helem_l), C.c_double(0.1), _ptr(self.mu_e)),
                    "ducros_visc_limit")
        if self.op_precision == 4:
            self.mu_e_op.copy_(self.mu_e)

    def _stage(self, lam1, lams, gam, store_R, next_gam=None):
        if self.lps:
            # updateF with LPS (EOS, Sutherland, full_convec_ijk, full_proj_ijk,
            # full_diff_stab_ijk), lumped solve and stage combine in one call
            R = (_ptr(self.Rmass), _ptr(self.Rmom), _ptr(self.Rener)) if store_R else (None,) * 3
            nxt = ((_ptr(self.rho[0]), _ptr(self.q[0]), _ptr(self.E[0])) if next_gam is not None
                   else (None,) * 3)
            check(lib().sodhip_rk_stage_lps_f64(
                self.ctx.h, self.p, self.nelem, self.npoin, self.npoin_w, _ptr(self.connec),
                _ptr(self.lpoin_w), _ptr(self.He), _ptr(self.gpvol), _ptr(self.dl),
                *[_ptr(x) for x in self.tabs], _ptr(self.mu_e), _ptr(self.invMl),
                _ptr(self.tau_stab), C.byref(self.sp), C_LPS_COMP, self.dt, lam1, lams, gam,
                _ptr(self.rho[0]), _ptr(self.q[0]), _ptr(self.E[0]), _ptr(self.rho[1]),
                _ptr(self.q[1]), _ptr(self.E[1]), _ptr(self.u[1]), _ptr(self.e_int[1]),
                _ptr(self.pr[1]), _ptr(self.aux_h), _ptr(self.Tem[1]), _ptr(self.mu_fluid),
                *[_ptr(x) for x in self.proj], *R, 0.0 if next_gam is None else next_gam, *nxt),
                "rk_stage_lps")
            if next_gam is not None:
                self._next_dt = self.dt
            return
        R = (_ptr(self.Rmass), _ptr(self.Rmom), _ptr(self.Rener)) if store_R else (None,) * 3
        args = (self.ctx.h, self.p, self.nelem, self.npoin, self.npoin_w, _ptr(self.connec),
                _ptr(self.lpoin_w), self.op_precision, _ptr(self.He_op), _ptr(self.gpvol_op),
                _ptr(self.dl_op), *[_ptr(x) for x in self.tabs], _ptr(self.mu_e_op),
                _ptr(self.invMl), C.byref(self.sp), self.dt, lam1, lams, gam, _ptr(self.rho[0]),
                _ptr(self.q[0]), _ptr(self.E[0]), _ptr(self.rho[1]), _ptr(self.q[1]),
                _ptr(self.E[1]), *R)
        if next_gam is None:
            check(lib().sodhip_rk_stage_f64(*args), "rk_stage")
        else:
            # the next step's stage 1 goes to level 1, which becomes level 2 at
            # that step's exchange
            check(lib().sodhip_rk_stage_next_f64(*args, next_gam, _ptr(self.rho[0]),
                                                 _ptr(self.q[0]), _ptr(self.E[0])),
                  "rk_stage_next")
            self._next_dt = self.dt

    # ------------------------------------------------------------ time step
    def step_rk4(self):
        """One step of the classical RK4 integrator rk_4_main (time_integ.f90:
        250-470; compiled out in the reference by #if 0): four fused stages
        (sodhip_rk4_stage_f64; operators without fact, so y advances by
        -dt * sum b_i R), then the EOS of the new state and Sutherland
        (:489-533).  The entropy viscosity mu_e is held: rk_4_main's per-stage
        entropy residual and end-of-step sensor are not part of this path."""
        a, b = RK4_A, RK4_B
        y = (self.rho[1], self.q[1], self.E[1])
        h = self.ctx.h
        for i in range(4):
            last = i == 3
            st = y if i == 0 else self.aux
            check(lib().sodhip_rk4_stage_f64(
                h, self.p, self.nelem, self.npoin, self.npoin_w, _ptr(self.connec),
                _ptr(self.lpoin_w), self.op_precision, _ptr(self.He_op), _ptr(self.gpvol_op),
                _ptr(self.dl_op), *[_ptr(x) for x in self.tabs], _ptr(self.mu_e_op),
                _ptr(self.invMl), C.byref(self.sp4), self.dt, 0.0 if last else a[i + 1], b[i],
                int(i == 0), int(last), *[_ptr(x) for x in y], *[_ptr(x) for x in st],
                *[_ptr(x) for x in self.aux], *[_ptr(x) for x in self.Rsum],
                _ptr(self.Rmass), _ptr(self.Rmom), _ptr(self.Rener)), "rk4_stage")
        t = self.t
        check(lib().sodhip_eos_step_f64(
            h, self.npoin, self.npoin_w, _ptr(self.lpoin_w), t.gamma_gas, t.Rgas,
            _ptr(self.rho[1]), _ptr(self.q[1]), _ptr(self.E[1]), _ptr(self.u[1]),
            _ptr(self.e_int[1]), _ptr(self.pr[1]), _ptr(self.csound), _ptr(self.machno),
            _ptr(self.Tem[1])), "eos_step")
        self._sutherland()
        self.time += self.dt
        self.istep += 1
        if self.dt_mode == "adaptive":
            self.dt = self.adapt_dt()

    def step(self):
        """One time step: evalTimeIteration's level copy, rk_4_ls_main, time += dt,
        then evalDt in adaptive mode."""
        if self.integrator == "rk4":
            return self.step_rk4()
        L, G, S = self.L, self.G, self.S
        # level 2 -> level 1.  The primitive levels are exchanged rather than
        # copied: level 2 of u, pr, Tem and e_int is rewritten by the EOS
        # before anything reads it, and on nodes outside lpoin_w both levels
        # keep their initial values.  With the reference sequencing stage 1
        # reads y^n from level 1 only (lam1 = 1, lams = 0 gives the values
        # of lam1 = 0, lams = 1 on equal levels), so the conserved levels are
        # exchanged too; otherwise stage 1 evaluates its RHS at level 2 and
        # they are copied.
        for lv in (self.rho, self.q, self.E):
            if self.parity:
                lv[0], lv[1] = lv[1], lv[0]
            else:
                lv[0].copy_(lv[1])
        for lv in (self.u, self.pr, self.Tem, self.e_int):
            lv[0], lv[1] = lv[1], lv[0]
        n = self.npoin
        self.eta[:n].copy_(self.eta[n:])
        h = self.ctx.h
        if self.lps:
            # comp_tau at the start of rk_4_ls_main (time_integ_ls.f90:266-268) on
            # the level-1 velocity of this step
            check(lib().sodhip_comp_tau_f64(h, self.p, self.nelem, n, _ptr(self.connec),
                                            _ptr(self.csound), _ptr(self.u[0]),
                                            _ptr(self.helem), C_SPECIES_STAB, 0,
                                            _ptr(self.tau_stab)), "comp_tau")
        if self.first:
            self.first = False
            # Reta(:,1) of f_eta = u(:,:,1)*eta(:,1); sensor; updateF
            self._scalar_convec(self.Reta[:n])
            if self.sensor:
                self._smart_visc(self.Reta[:n], self.rho[0], self.u[0], self.Tem[0], self.eta[:n])
            self._stage(0.0, 1.0, 0.0, True)  # updateF at y^n: keeps R, y unchanged
        if self.parity and self._next_dt is not None and self._next_dt == self.dt:
            pass  # stage 1 was stored by the previous step's last stage
        elif self.parity:
            # stage 1 with the kept residual (time_integ_ls.f90:323-331)
            check(lib().sodhip_rk_stage_update_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), None, self.dt, 1.0, 0.0, G[1, 0],
                _ptr(self.rho[0]), _ptr(self.q[0]), _ptr(self.E[0]), _ptr(self.rho[1]),
                _ptr(self.q[1]), _ptr(self.E[1]), _ptr(self.Rmass), _ptr(self.Rmom),
                _ptr(self.Rener)), "rk_stage_update")
        else:
            self._stage(0.0, 1.0, G[1, 0], False)
        self._next_dt = None
        # at fixed dt the last stage also stores the next step's stage 1
        fuse_next = self.fuse_next_stage and self.parity and self.dt_mode == "fixed"
        for s in range(2, S + 1):
            self._stage(L[s, 0], L[s, s - 1], G[s, s - 1], store_R=(s == S and self.parity),
                        next_gam=G[1, 0] if (s == S and fuse_next) else None)
        self.end_of_step()
        self.time += self.dt
        self.istep += 1
        if self.dt_mode == "adaptive":
            self.dt = self.adapt_dt()

    def end_of_step(self):
        """EOS, entropy, entropy residual, sensor and Sutherland after the stages
        (time_integ_ls.f90:389-473)."""
        h, n, t = self.ctx.h, self.npoin, self.t
        if self.ens is not None:
            # per node: the instance's entropy reference state and Sutherland factor
            check(lib().sodhip_eos_entropy_step_ens_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), t.gamma_gas, t.Rgas,
                C.byref(self.ens.struct), _ptr(self.rho[1]), _ptr(self.q[1]), _ptr(self.E[1]),
                _ptr(self.u[1]), _ptr(self.e_int[1]), _ptr(self.pr[1]), _ptr(self.csound),
                _ptr(self.machno), _ptr(self.Tem[1]), _ptr(self.eta), _ptr(self.mu_fluid)),
                "eos_entropy_step_ens")
        elif self.fuse_end_state:
            # Sutherland moves ahead of the sensor, which does not read mu_fluid
            check(lib().sodhip_eos_entropy_step_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), t.gamma_gas, t.Rgas, t.po, t.rho0,
                t.mu_factor, _ptr(self.rho[1]), _ptr(self.q[1]), _ptr(self.E[1]),
                _ptr(self.u[1]), _ptr(self.e_int[1]), _ptr(self.pr[1]), _ptr(self.csound),
                _ptr(self.machno), _ptr(self.Tem[1]), _ptr(self.eta), _ptr(self.mu_fluid)),
                "eos_entropy_step")
        else:
            check(lib().sodhip_eos_step_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), t.gamma_gas, t.Rgas, _ptr(self.rho[1]),
                _ptr(self.q[1]), _ptr(self.E[1]), _ptr(self.u[1]), _ptr(self.e_int[1]),
                _ptr(self.pr[1]), _ptr(self.csound), _ptr(self.machno), _ptr(self.Tem[1])),
                "eos_step")
            check(lib().sodhip_thermo_entropy_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), t.gamma_gas, t.po, t.rho0,
                _ptr(self.rho[1]), _ptr(self.pr[1]), None, _ptr(self.eta), None),
                "thermo_entropy")
        if self.fuse_end_state or self.ens is not None:
            # Reta(:,2) of f_eta = u(:,:,1)*eta(:,1), its lumped solve and the
            # residual in one element pass and one node pass
            check(lib().sodhip_entropy_convec_residual_f64(
                h, self.p, self.nelem, n, self.npoin_w, _ptr(self.connec), _ptr(self.lpoin_w),
                _ptr(self.He), _ptr(self.gpvol), _ptr(self.dl), *[_ptr(x) for x in self.tabs],
                _ptr(self.u[0]), _ptr(self.eta), _ptr(self.invMl), 1.0, self.dt, _ptr(self.Reta),
                _ptr(self.auxReta)), "entropy_convec_residual")
        else:
            self._scalar_convec(self.Reta[n:])
            check(lib().sodhip_entropy_residual_f64(
                h, n, self.npoin_w, _ptr(self.lpoin_w), 1.0, self.dt, _ptr(self.Reta),
                _ptr(self.eta), _ptr(self.auxReta)), "entropy_residual")
        if self.sensor:
            self._smart_visc(self.auxReta, self.rho[1], self.u[1], self.Tem[1], self.eta[n:],
                             ducros=True)
        if not self.fuse_end_state and self.ens is None:
            self._sutherland()

    # ------------------------------------------------------------ diagnostics
    def diagnostics(self) -> dict:
        """volAvg_EK, visc_dissipationRate and maxMach (mod_analysis.f90:10-168)."""
        h = self.ctx.h
        ek = C.c_double()
        check(lib().sodhip_volavg_ek_f64(h, self.p, self.nelem, self.npoin, _ptr(self.connec),
                                         _ptr(self.gpvol), self.t.rho0, _ptr(self.rho[1]),
                                         _ptr(self.u[1]), C.byref(ek)), "volavg_ek")
        eS, eD, eT = C.c_double(), C.c_double(), C.c_double()
        check(lib().sodhip_visc_dissipation_f64(
            h, self.p, self.nelem, self.npoin, _ptr(self.connec), self.t.rho0, _ptr(self.mu_fluid),
            _ptr(self.mu_e), _ptr(self.u[1]), self.vol_tot, _ptr(self.gpvol), _ptr(self.He),
            _ptr(self.dl), *[_ptr(x) for x in self.tabs], C.byref(eS), C.byref(eD), C.byref(eT)),
            "visc_dissipation")
        mm = C.c_double()
        check(lib().sodhip_max_mach_f64(h, self.npoin, self.npoin_w, _ptr(self.lpoin_w),
                                        _ptr(self.machno), C.byref(mm)), "max_mach")
        return dict(time=self.time, EK=ek.value, eps_S=eS.value, eps_D=eD.value, eps_T=eT.value,
                    maxmach=mm.value)

    def capture(self) -> "StepGraph":
        """Captures two time steps into a HIP graph (torch.cuda.CUDAGraph on the
        library's stream) and returns it; StepGraph.replay() then advances the run
        by two steps with one graph launch.  Two, because the time levels are
        exchanged every step: after two steps every array is back in its place.
        Requires fixed dt (the adaptive time step reads its reduction back to the
        host) and a first step already taken (it builds the library's mesh tables
        and scratch, which allocate).  Nothing in the captured steps allocates,
        synchronises or clears memory with a memset."""
        if self.dt_mode != "fixed":
            raise SodhipError("graph capture needs dt_mode='fixed'")
        if self.first:
            raise SodhipError("take one step before capturing (set-up allocates)")
        saved = (self.time, self.istep)
        self.ctx.prof_enable(False)
        torch.cuda.synchronize()
        # the graph records the library's scratch buffers by address: from here
        # until the graph is released they may not move (a call that would grow
        # them fails instead)
        check(lib().sodhip_pin_scratch(self.ctx.h, 1), "pin_scratch")
        g = torch.cuda.CUDAGraph()
        try:
            with torch.cuda.graph(g):
                self.ctx.sync_stream()  # launches go to the capturing stream
                self.step()
                self.step()
        except BaseException:
            lib().sodhip_pin_scratch(self.ctx.h, 0)
            raise
        self.ctx.sync_stream()
        self.time, self.istep = saved
        return StepGraph(self, g)

    def state(self) -> dict:
        n = self.npoin
        return dict(rho=self.rho[1].cpu().numpy(),
                    q=self.q[1].cpu().numpy().reshape((n, 3), order="F"),
                    E=self.E[1].cpu().numpy(),
                    mu_e=self.mu_e.cpu().numpy().reshape((self.nelem, self.nnode), order="F"))

    @property
    def rhs_per_step(self) -> int:
        """Fresh RHS evaluations per step (after the first step)."""
        return self.S - 1 if self.parity else self.S


class StepGraph:
    """Two captured time steps of a TGVRun (see TGVRun.capture)."""

    def __init__(self, run: TGVRun, graph):
        self.run = run
        self.graph = graph

    def replay(self):
        """Advances the run by two steps."""
        if self.graph is None:
            raise SodhipError("the graph was released")
        self.run.ctx.sync_stream()
        self.graph.replay()
        self.run.time += 2 * self.run.dt
        self.run.istep += 2

    def release(self):
        """Frees the graph and lets the library's scratch buffers move again."""
        if self.graph is not None:
            torch.cuda.synchronize()
            self.graph = None
            if getattr(self.run.ctx, "h", None):
                check(lib().sodhip_pin_scratch(self.run.ctx.h, 0), "pin_scratch")

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def rk_ls_coeffs(stages: int, variant: int = 1):
    """lambda_ij, gamma_ij of init_rk4_ls_solver (time_integ_ls.f90:87-171), 0-based
    (L[i-1, j-1] = lambda_ij).  Constants the reference writes without a d0
    exponent are default-real (FP32) literals and are rounded through float32."""
    L = np.zeros((6, 6))
    G = np.zeros((6, 6))
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    if stages == 3:
        L[1, 0] = 1.0; L[2, 0] = 3.0 / 4.0; L[2, 1] = 1.0 / 4.0; L[3, 0] = 1.0 / 3.0
        L[3, 2] = 2.0 / 3.0
        G[1, 0] = 1.0; G[2, 1] = 0.25; G[3, 2] = 2.0 / 3.0
    elif stages == 4:
        L[1, 0] = 1.0; L[2, 1] = 1.0; L[3, 0] = 2.0 / 3.0; L[3, 2] = 1.0 / 3.0; L[4, 3] = 1.0
        G[1, 0] = 0.5; G[2, 1] = 0.5; G[3, 2] = 1.0 / 6.0; G[4, 3] = 0.5
    elif stages == 5:
        for i in range(1, 6):
            L[i, i - 1] = 1.0
        if variant == 1:
            L[4, 0] = 0.571403511494104; L[4, 3] = 0.428596488505896
            G[1, 0] = 0.443568244942995; G[2, 1] = 0.291111420073766
            G[3, 2] = 0.270612601278217; G[4, 3] = 0.110577759392786
            G[5, 4] = 0.458557505351052
        elif variant == 2:
            L[3, 0] = 0.682342861037239; L[3, 2] = 0.317657138962761
            L[5, 0] = 0.045230974482400; L[5, 4] = 0.954769025517600
            G[1, 0] = f32(0.465388589249323); G[2, 1] = f32(0.465388589249323)
            G[3, 2] = f32(0.124745797313998); G[4, 3] = f32(0.465388589249323)
            G[5, 4] = f32(0.154263303748666)
        elif variant == 3:
            L[4, 0] = f32(0.592032910942121); L[4, 3] = 1.0 - L[4, 0]
            G[1, 0] = f32(0.266541020678955); G[2, 1] = f32(0.548560709048532)
            G[3, 2] = f32(0.289517014154401); G[4, 3] = f32(0.086408328057923)
            G[5, 4] = f32(0.462943578481813)
        elif variant == 4:
            L[3, 0] = f32(0.707858560931430); L[3, 2] = 1.0 - L[3, 0]
            L[5, 0] = f32(0.222853615080669); L[5, 4] = 1.0 - L[5, 0]
            G[1, 0] = f32(0.292845746913355); G[2, 1] = f32(0.339532793976408)
            G[3, 2] = f32(0.200532330324672); G[4, 3] = f32(0.701676169006879)
            G[5, 4] = f32(0.155278812461877)
        else:
            raise SodhipError(f"unknown SSP(5,3) variant {variant}")
    else:
        raise SodhipError(f"no LS-RK scheme with {stages} stages")
    return L, G
