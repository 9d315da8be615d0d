"""An ensemble of periodic TGV runs of uneven sizes, Mach and Reynolds numbers
joined into one disconnected mesh, so that every kernel of a time step covers
all instances in one launch (one element launch and one gather/update per
Runge-Kutta stage).

Each instance is the reference's TGV case (TGVSolver.f90:39-221) on its own
n^3-element periodic box, p=3.  The join concatenates elements, nodes and
working-node lists with offsets; geometry, mass and lengths are per element
or per node and concatenate as they are.  What differs between instances is
carried by sodhip_instances (the instance's working-node segment, each
element's instance, per node the entropy reference pressure and density and
the Sutherland factor) and by the stage's per-element Sutherland factor:

  * Re and Ma set mu_factor (TGVParams), Ma sets the reference pressure po
    of the entropy (Rgas, gamma, Cp, Prt are common);
  * the entropy normalisation of the sensor runs per instance over its own
    working nodes, with the blocking of a single mesh, so it has the bits of
    the instance run alone;
  * one dt for all, the smallest instance dt (each instance then runs at
    CFL <= 0.5): the stage update keeps one scalar.

With the same dt, every instance's fields have the bits of its run alone
(each element and node is computed as in the single mesh, in the same
order).  EnsembleRun is a TGVRun on the joined case; run.capture() records
its step in a HIP graph, sod2d_amd.autotune picks the stage's launch shape,
and dump_features() writes the sensor's features and targets per element.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import case as case_mod
from ._lib import Instances, check, lib
from .device import Context, _ptr
from .solver import TGVParams, TGVRun, tgv_initial_conditions

F64, I32 = torch.float64, torch.int32


@dataclass
class InstanceCase:
    n: int          # elements per direction
    M: float = 0.1  # Mach number
    Re: float = 1600.0


def config4_cases(count: int = 32, seed: int = 4, nmin: int = 8, nmax: int = 32) -> list:
    """Uneven instances: element counts n^3 with n ~ U{nmin..nmax},
    Ma ~ U(0.1, 0.5), Re ~ U(400, 3200), from a seeded generator."""
    r = np.random.default_rng(seed)
    return [InstanceCase(int(r.integers(nmin, nmax + 1)), float(r.uniform(0.1, 0.5)),
                         float(r.uniform(400.0, 3200.0))) for _ in range(count)]


class EnsembleTables:
    """The device arrays of sodhip_instances plus the per-element Sutherland
    factor of the stage, and the instances' offsets on the host."""

    def __init__(self, params, geos):
        counts = [(g.mesh.nelem, g.mesh.npoin, g.mesh.npoin_w) for g in geos]
        self.ninst = len(geos)
        self.elem_off = np.concatenate([[0], np.cumsum([c[0] for c in counts])]).astype(np.int64)
        self.node_off = np.concatenate([[0], np.cumsum([c[1] for c in counts])]).astype(np.int64)
        self.wseg_h = np.concatenate([[0], np.cumsum([c[2] for c in counts])]).astype(np.int32)
        elem_inst = np.repeat(np.arange(self.ninst, dtype=np.int32), [c[0] for c in counts])
        node = lambda v: np.repeat(np.asarray(v, dtype=np.float64), [c[1] for c in counts])  # noqa: E731
        self.wseg = torch.from_numpy(self.wseg_h).cuda()
        self.elem_inst = torch.from_numpy(elem_inst).cuda()
        self.p_inf = torch.from_numpy(node([t.po for t in params])).cuda()
        self.rho_inf = torch.from_numpy(node([t.rho0 for t in params])).cuda()
        self.mu_factor = torch.from_numpy(node([t.mu_factor for t in params])).cuda()
        self.mu_factor_e = torch.from_numpy(
            np.repeat(np.asarray([t.mu_factor for t in params]), [c[0] for c in counts])).cuda()
        self.struct = Instances(self.ninst, self.wseg.data_ptr(), self.elem_inst.data_ptr(),
                                self.p_inf.data_ptr(), self.rho_inf.data_ptr(),
                                self.mu_factor.data_ptr())


def join_geometries(geos) -> case_mod.Geometry:
    """One disconnected mesh from periodic p=3 boxes: connectivities and
    working-node lists shifted by the node offsets, the rest concatenated."""
    T = geos[0].tables
    noff = 0
    co, cw, lw, coord = [], [], [], []
    for g in geos:
        m = g.mesh
        co.append(m.connec_orig + noff)
        cw.append(m.connec + noff)
        lw.append(m.lpoin_w + noff)
        coord.append(m.coord)
        noff += m.npoin
    cat0 = lambda xs: np.asfortranarray(np.concatenate(xs, axis=0))  # noqa: E731
    catl = lambda xs: np.asfortranarray(np.concatenate(xs, axis=-1))  # noqa: E731
    nelem = sum(g.mesh.nelem for g in geos)
    mesh = case_mod.BoxMesh(3, ("joined", len(geos)), True, nelem, noff,
                            sum(g.mesh.npoin_w for g in geos), cat0(co).astype(np.int32),
                            cat0(cw).astype(np.int32), np.concatenate(lw).astype(np.int32),
                            cat0(coord))
    return case_mod.Geometry(T, mesh, catl([g.He for g in geos]), catl([g.gpvol for g in geos]),
                             np.concatenate([g.Ml for g in geos]),
                             np.concatenate([g.invMl for g in geos]),
                             np.concatenate([g.helem for g in geos]),
                             cat0([g.helem_l for g in geos]), sum(g.vol_tot for g in geos))


def instance_dt(geo, t: TGVParams, ctx) -> float:
    """The initial CFL dt of one instance run alone (TGVRun's evalInitialDt)."""
    return TGVRun(3, geometry=geo, params=t, dt_mode="fixed", ctx=ctx).dt


class EnsembleRun(TGVRun):
    """The joined TGV instances as one run (FP64 plain path, spectral sensor,
    the reference sequencing, fixed common dt)."""

    def __init__(self, cases, ctx: Context | None = None, length: float = 2.0 * math.pi):
        self.cases = [c if isinstance(c, InstanceCase) else InstanceCase(*c) for c in cases]
        ctx = ctx or Context(torch.cuda.current_device())
        self.params = [TGVParams(M=c.M, Re=c.Re) for c in self.cases]
        geos = [case_mod.build_case(3, c.n, periodic=True, lengths=(length,) * 3)
                for c in self.cases]
        # the common dt: the smallest instance dt
        self.instance_dts = [instance_dt(g, t, ctx) for g, t in zip(geos, self.params)]
        joined = join_geometries(geos)
        ics = [tgv_initial_conditions(g.mesh.coord, t) for g, t in zip(geos, self.params)]
        ic = {}
        for k in ics[0]:
            ic[k] = np.asfortranarray(np.concatenate([d[k] for d in ics], axis=0))
        tabs = EnsembleTables(self.params, geos)
        self.tables = tabs
        super().__init__(3, geometry=joined, params=TGVParams(), initial=ic, dt_mode="fixed",
                         ctx=ctx, instances=tabs)
        self.dt = min(self.instance_dts)
        del geos

    # ------------------------------------------------------------ per instance
    def instance_state(self, i: int) -> dict:
        """Conserved fields (nodes of instance i) and mu_e (its elements)."""
        t = self.tables
        n0, n1 = int(t.node_off[i]), int(t.node_off[i + 1])
        e0, e1 = int(t.elem_off[i]), int(t.elem_off[i + 1])
        n, np_ = self.npoin, n1 - n0
        q = self.q[1].view(3, n)[:, n0:n1].cpu().numpy().T
        mu = self.mu_e.view(self.nnode, self.nelem)[:, e0:e1].cpu().numpy().T
        return dict(rho=self.rho[1][n0:n1].cpu().numpy(), q=np.asfortranarray(q.reshape(np_, 3)),
                    E=self.E[1][n0:n1].cpu().numpy(), mu_e=np.asfortranarray(mu))

    def dump_features(self, path: str | None = None) -> dict:
        """The sensor's 12 features per element at the current end-of-step
        state (the inputs of the last sensor call), with its target, the
        element's instance and the element mean of mu_e; written to ``path``
        (.npz) when given."""
        from .surrogate import NF, NFP, targets
        ne, n = self.nelem, self.npoin
        X = torch.empty(ne * NFP, dtype=torch.float32, device="cuda")
        beta = torch.empty(ne, dtype=F64, device="cuda")
        check(lib().sodhip_sensor_features_ens_f64(
            self.ctx.h, 3, ne, n, self.npoin_w, _ptr(self.connec), _ptr(self.lpoin_w),
            C.byref(self.tables.struct), _ptr(self.auxReta), self.t.gamma_gas,
            _ptr(self.rho[1]), _ptr(self.u[1]), _ptr(self.Tem[1]), _ptr(self.eta[n:]),
            _ptr(self.helem_l), _ptr(self.helem), _ptr(self.tabs[0]), C.byref(self.vp), _ptr(X),
            _ptr(beta)),
            "sensor_features_ens")
        t = targets(self.mu_e, beta, ne, self.vp.cglob)
        out = dict(X=X.view(ne, NFP)[:, :NF].cpu().numpy(), t=t.cpu().numpy(),
                   beta=beta.cpu().numpy(), inst=self.tables.elem_inst.cpu().numpy(),
                   mu_e_mean=self.mu_e.view(-1, ne).mean(0).cpu().numpy(), step=self.istep)
        if path:
            np.savez(path, **out)
        return out
