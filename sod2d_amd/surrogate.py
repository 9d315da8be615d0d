"""The per-element MLP surrogate of the entropy-viscosity sensor
smart_visc_spectral (mod_entropy_viscosity.f90:14-111) on the device.

SurrogateMLP holds the network 12 -> 64 -> 64 -> 1 (SiLU; BF16 operands on
the matrix cores, FP32 accumulation and master weights) of the C-ABI's
sodhip_mlp and trains it with the library's Adam step.  Its target is, per
element, the mean of the plain sensor's mu_e over the element's nodes divided
by cglob * beta_e (the sensor's cap), a ratio in [0, 1] taken on the log scale
1 + log10(ratio + 1e-10)/10; deployed, the element's viscosity is cglob *
beta_e * (10^(10 (y_e - 1)) - 1e-10) at every node.

Training samples come from the plain sensor itself: at the end of TGV steps
(sensor_samples) and on the isolated-operator workload's random fields on a
jittered box (operator_samples), the features of sodhip_sensor_features_f64
beside the target, repeated with the entropy residual scaled down so that the
samples of coarse meshes also cover finer ones (training_samples).  TGVRun(sensor_path="surrogate", surrogate=...) deploys
the network through sodhip_surrogate_visc_spectral_f64, with the plain sensor
as the per-element fallback and as the judge every ``judge_every`` steps.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from ._lib import MLP, check, lib
from .device import _ptr

F32, F64, I32 = torch.float32, torch.float64, torch.int32
NF, NFP, NH = 12, 16, 64


def _sizes():
    np_, nq, nf = C.c_int(), C.c_int(), C.c_int()
    check(lib().sodhip_mlp_sizes(C.byref(np_), C.byref(nq), C.byref(nf)), "mlp_sizes")
    return np_.value, nq.value, nf.value


class SurrogateMLP:
    """Device parameters, BF16 copy, Adam moments and feature range of the MLP."""

    def __init__(self, ctx, seed: int = 0):
        self.ctx = ctx
        self.np, self.nq, nf = _sizes()
        assert nf == NF
        g = torch.Generator().manual_seed(seed)
        w = torch.zeros(self.np, dtype=F32)
        # uniform initialisation scaled by the fan-in (W1 12x64, W2 64x64, w3 64)
        o = 0
        for fan_in, count in ((NF, NF * NH), (None, NH), (NH, NH * NH), (None, NH), (NH, NH),
                              (None, 1)):
            if fan_in:
                lim = math.sqrt(6.0 / fan_in)
                w[o:o + count] = (torch.rand(count, generator=g) * 2 - 1) * lim
            o += count
        self.w = w.cuda()
        self.wq = torch.zeros(self.nq, dtype=torch.bfloat16, device="cuda")
        self.m = torch.zeros(self.np, dtype=F32, device="cuda")
        self.v = torch.zeros(self.np, dtype=F32, device="cuda")
        self.lo = np.full(NF, -np.inf, dtype=np.float32)
        self.hi = np.full(NF, np.inf, dtype=np.float32)
        self.step = 0
        check(lib().sodhip_mlp_quantize(ctx.h, C.byref(self.struct())), "mlp_quantize")

    def struct(self) -> MLP:
        s = MLP()
        s.w, s.wq = self.w.data_ptr(), self.wq.data_ptr()
        s.m, s.v = self.m.data_ptr(), self.v.data_ptr()
        for i in range(NF):
            s.lo[i], s.hi[i] = float(self.lo[i]), float(self.hi[i])
        return s

    def set_range(self, X: torch.Tensor, margin: float = 0.05):
        """The fallback box: each feature's range over the training data,
        widened by ``margin`` of its span."""
        x = X[:, :NF].float()
        lo, hi = x.min(0).values.cpu().numpy(), x.max(0).values.cpu().numpy()
        span = np.maximum(hi - lo, 1e-6)
        self.lo = (lo - margin * span).astype(np.float32)
        self.hi = (hi + margin * span).astype(np.float32)

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        n = X.shape[0]
        y = torch.empty(n, dtype=F32, device="cuda")
        check(lib().sodhip_mlp_forward(self.ctx.h, C.byref(self.struct()), n,
                                       _ptr(X.contiguous(), F32), _ptr(y)), "mlp_forward")
        return y

    def gradient(self, X: torch.Tensor, t: torch.Tensor, nwaves: int = 64):
        """(y, the gradient of mean (y - t)^2 as the sum of the wave partials)."""
        n = X.shape[0]
        y = torch.empty(n, dtype=F32, device="cuda")
        Z = torch.empty(2 * n * NH, dtype=F32, device="cuda")
        gp = torch.empty(nwaves * self.np, dtype=F32, device="cuda")
        check(lib().sodhip_mlp_gradient(self.ctx.h, C.byref(self.struct()), n, _ptr(X, F32),
                                        _ptr(t, F32), _ptr(y), _ptr(Z), _ptr(gp), nwaves),
              "mlp_gradient")
        return y, gp.view(nwaves, self.np).sum(0)

    def train(self, X: torch.Tensor, t: torch.Tensor, steps: int = 3000, batch: int = 4096,
              lr: float = 2e-3, seed: int = 0, nwaves: int = 64) -> list:
        """Adam on mini-batches drawn with a seeded generator; returns the
        batch loss every 100 steps."""
        X = X.contiguous().float()
        t = t.contiguous().float()
        n = X.shape[0]
        b = min(batch, n)
        y = torch.empty(b, dtype=F32, device="cuda")
        Z = torch.empty(2 * b * NH, dtype=F32, device="cuda")
        gp = torch.empty(nwaves * self.np, dtype=F32, device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(seed)
        hist = []
        st = self.struct()
        for k in range(steps):
            idx = torch.randint(0, n, (b,), device="cuda", generator=gen)
            xb, tb = X[idx].contiguous(), t[idx].contiguous()
            self.step += 1
            lr_k = lr * (0.1 ** (k / max(steps, 1)))  # decays tenfold over the run
            check(lib().sodhip_mlp_train_step(self.ctx.h, C.byref(st), b, _ptr(xb), _ptr(tb),
                                              self.step, lr_k, _ptr(y), _ptr(Z), _ptr(gp),
                                              nwaves), "mlp_train_step")
            if k % 100 == 0 or k == steps - 1:
                hist.append(float(((y - tb) ** 2).mean()))
        return hist

    def state_dict(self) -> dict:
        return dict(w=self.w.cpu().numpy(), lo=self.lo.copy(), hi=self.hi.copy())

    def load(self, d: dict):
        self.w.copy_(torch.from_numpy(np.asarray(d["w"], dtype=np.float32)))
        self.lo = np.asarray(d["lo"], dtype=np.float32)
        self.hi = np.asarray(d["hi"], dtype=np.float32)
        check(lib().sodhip_mlp_quantize(self.ctx.h, C.byref(self.struct())), "mlp_quantize")


# ---------------------------------------------------------------- samples
def features(ctx, p, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma, rho, u, Tem, eta,
             helem_l, helem, inv, vp):
    """sodhip_sensor_features_f64 -> (X (nelem,16) float32, beta (nelem) float64)."""
    X = torch.empty(nelem * NFP, dtype=F32, device="cuda")
    beta = torch.empty(nelem, dtype=F64, device="cuda")
    check(lib().sodhip_sensor_features_f64(
        ctx.h, p, nelem, npoin, npoin_w, _ptr(connec), _ptr(lpoin_w), _ptr(Reta), gamma,
        _ptr(rho), _ptr(u), _ptr(Tem), _ptr(eta), _ptr(helem_l), _ptr(helem), _ptr(inv),
        C.byref(vp), _ptr(X), _ptr(beta)), "sensor_features")
    return X.view(nelem, NFP), beta


def logscale(x):
    """The log scale of the viscosity ratios and of the MLP's output:
    1 + log10(x + 1e-10)/10 maps [0, 1] into [0, 1]."""
    return 1.0 + torch.log10(x + 1e-10) / 10.0


def targets(mu_e: torch.Tensor, beta: torch.Tensor, nelem: int, cglob: float) -> torch.Tensor:
    """The log scale of the mean over the element's nodes of mu_e / (cglob
    beta_e) (mu_e(nelem, nnode) in Fortran order)."""
    mean = mu_e.view(-1, nelem).mean(0)
    ratio = torch.where(beta > 0, mean / (cglob * beta), torch.zeros_like(mean))
    return logscale(torch.clamp(ratio, 0.0, 1.0)).float()


# Scales of the entropy residual in the training samples.  The residual, and
# with it the sensor's local viscosity, falls with the mesh spacing, so a
# sample set taken on coarse meshes and repeated with the residual scaled down
# covers the features of finer meshes.
RETA_SCALES = (1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5)


def plain_samples(ctx, p, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma, rho, u, Tem,
                  eta, helem_l, helem, inv, vp, reta_scales=(1.0,)) -> tuple:
    """(X, t) of the plain sensor smart_visc_spectral on the given inputs, with
    the entropy residual multiplied by each of ``reta_scales`` in turn."""
    Xs, ts = [], []
    mu_e = torch.empty(nelem * 64, dtype=F64, device="cuda")
    mue_l = torch.empty_like(mu_e)
    for sc in reta_scales:
        R = Reta if sc == 1.0 else (Reta * sc).contiguous()
        check(lib().sodhip_smart_visc_spectral_f64(
            ctx.h, p, nelem, npoin, npoin_w, _ptr(connec), _ptr(lpoin_w), _ptr(R), gamma,
            _ptr(rho), _ptr(u), _ptr(Tem), _ptr(eta), _ptr(helem_l), _ptr(helem), _ptr(inv),
            C.byref(vp), _ptr(mu_e), _ptr(mue_l)), "smart_visc_spectral")
        X, beta = features(ctx, p, nelem, npoin, npoin_w, connec, lpoin_w, R, gamma, rho, u, Tem,
                           eta, helem_l, helem, inv, vp)
        Xs.append(X)
        ts.append(targets(mu_e, beta, nelem, vp.cglob))
    return torch.cat(Xs), torch.cat(ts)


def sensor_samples(run, reta_scales=(1.0,)) -> tuple:
    """(X, t) at the current end-of-step state of a TGVRun: the features of
    the sensor's inputs and the plain sensor's target."""
    n = run.npoin
    return plain_samples(run.ctx, run.p, run.nelem, n, run.npoin_w, run.connec, run.lpoin_w,
                         run.auxReta, run.t.gamma_gas, run.rho[1], run.u[1], run.Tem[1],
                         run.eta[n:], run.helem_l, run.helem, run.tabs[0], run.vp, reta_scales)


def operator_samples(ctx, n: int = 8, seeds=(1, 2, 3), jitter: float = 0.05,
                     reta_scales=(1.0,)) -> tuple:
    """(X, t) on the isolated-operator workload's random fields (a jittered
    box, p=3, not periodic): entropy, its convective residual with the lumped
    solve as the sensor's residual, the plain sensor, the features."""
    from . import case as case_mod
    from ._lib import ViscParams
    g = case_mod.build_case(3, n, periodic=False, jitter=jitter, lengths=(2 * np.pi,) * 3)
    m, T = g.mesh, g.tables
    h = ctx.h
    dev = lambda a, dt=None: (torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F")))  # noqa: E731
                              .cuda().to(dt) if dt else torch.from_numpy(
                                  np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda())
    connec, lpoin_w = dev(m.connec), dev(m.lpoin_w)
    tabs = [dev(x) for x in (T.invAtoIJK, T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK)]
    He, gpvol, dl, invMl = dev(g.He), dev(g.gpvol), dev(T.dlxigp_ip), dev(g.invMl)
    helem, helem_l = dev(g.helem), dev(g.helem_l)
    gamma, Rgas = 1.4, 1.0
    vp = ViscParams(0.1, 0.5, 1.0, Rgas, 1, 0)
    Xs, ts = [], []
    for seed in seeds:
        s = case_mod.operator_bench_fields(m.npoin, m.nelem, T.nnode, seed, gamma, Rgas)
        rho, u, pr, Tem = dev(s["rho"]), dev(s["u"]), dev(s["pr"]), dev(s["Tem"])
        eta = torch.zeros(2 * m.npoin, dtype=F64, device="cuda")
        check(lib().sodhip_thermo_entropy_f64(h, m.npoin, m.npoin_w, _ptr(lpoin_w), gamma,
                                              float(s["pr"].mean()), float(s["rho"].mean()),
                                              _ptr(rho), _ptr(pr), None, _ptr(eta), None),
              "thermo_entropy")
        eta1 = eta[m.npoin:].contiguous()
        Reta = torch.zeros(m.npoin, dtype=F64, device="cuda")
        check(lib().sodhip_generic_scalar_convec_ijk_f64(
            h, 3, m.nelem, m.npoin, _ptr(connec), _ptr(He), _ptr(gpvol), _ptr(dl),
            *[_ptr(x) for x in tabs], None, _ptr(eta1), _ptr(u), _ptr(Reta), 1, 1.0),
            "generic_scalar_convec_ijk")
        check(lib().sodhip_lumped_solver_scal_f64(h, m.npoin, m.npoin_w, _ptr(lpoin_w),
                                                  _ptr(invMl), _ptr(Reta)), "lumped_solver_scal")
        X, t = plain_samples(ctx, 3, m.nelem, m.npoin, m.npoin_w, connec, lpoin_w, Reta, gamma,
                             rho, u, Tem, eta1, helem_l, helem, tabs[0], vp, reta_scales)
        Xs.append(X)
        ts.append(t)
    return torch.cat(Xs), torch.cat(ts)


def tgv_samples(ctx, n: int = 24, steps: int = 40, every: int = 4, cases=((0.1, 1600.0),),
                dt_mode: str = "fixed", reta_scales=(1.0,)) -> tuple:
    """(X, t) along plain-sensor TGV runs (p=3, n^3 elements) of the given
    (Mach, Reynolds) cases, every ``every`` steps."""
    from .solver import TGVParams, TGVRun
    Xs, ts = [], []
    for M, Re in cases:
        run = TGVRun(3, n, params=TGVParams(M=M, Re=Re), ctx=ctx, dt_mode=dt_mode)
        for k in range(steps):
            run.step()
            if (k + 1) % every == 0:
                X, t = sensor_samples(run, reta_scales)
                Xs.append(X)
                ts.append(t)
    return torch.cat(Xs), torch.cat(ts)


def training_samples(ctx, reta_scales=RETA_SCALES) -> tuple:
    """The standard training set: the operator workload's random fields on a
    jittered 6^3 box (two seeds) and 12^3 TGV runs at (Ma, Re) = (0.1, 1600)
    and (0.3, 800) every 4 of 40 steps, each with the entropy residual at
    ``reta_scales``."""
    Xa, ta = operator_samples(ctx, n=6, seeds=(1, 2), reta_scales=reta_scales)
    Xb, tb = tgv_samples(ctx, n=12, steps=40, every=4, cases=((0.1, 1600.0), (0.3, 800.0)),
                         reta_scales=reta_scales)
    return torch.cat([Xa, Xb]), torch.cat([ta, tb])


def trained_surrogate(ctx, steps: int = 1500, seed: int = 0, reta_scales=RETA_SCALES):
    """(SurrogateMLP trained on training_samples, its loss history, samples)."""
    X, t = training_samples(ctx, reta_scales)
    net = SurrogateMLP(ctx, seed=seed)
    net.set_range(X)
    hist = net.train(X, t, steps=steps, batch=4096)
    return net, hist, X.shape[0]
