"""The allocation-free stage (sodhip_rk_stage_noalloc_f64) against the
allocating one (sodhip_rk_stage_f64), called through the C-ABI.

Both run the same kernels on the same inputs, so the new state and the kept
residual must agree bit for bit.  Two calls in a row on one GPU, the second on
a smaller mesh, with FP64 and with FP32 operators; plus the two error paths:
a mesh whose records exceed the library's fixed array, and a mesh whose tables
no allocating call has built yet."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(n, op_precision):
    from sod2d_amd.solver import TGVRun
    dev = TGVRun(3, nelem_dir=n, op_precision=op_precision, dt_mode="fixed")
    dev.step()  # a state with entropy viscosity and a kept residual
    return dev


def _args(dev, R, lam1, lams, gam):
    from sod2d_amd.device import _ptr
    return (dev.ctx.h, dev.p, dev.nelem, dev.npoin, dev.npoin_w, _ptr(dev.connec),
            _ptr(dev.lpoin_w), dev.op_precision, _ptr(dev.He_op), _ptr(dev.gpvol_op),
            _ptr(dev.dl_op), *[_ptr(x) for x in dev.tabs], _ptr(dev.mu_e_op), _ptr(dev.invMl),
            C.byref(dev.sp), dev.dt, lam1, lams, gam, _ptr(dev.rho[0]), _ptr(dev.q[0]),
            _ptr(dev.E[0]), _ptr(dev.rho[1]), _ptr(dev.q[1]), _ptr(dev.E[1]),
            *[_ptr(x) for x in R])


def _compare(dev):
    import torch
    from sod2d_amd._lib import check, lib
    L, G = dev.L, dev.G
    coef = (L[3, 0], L[3, 2], G[3, 2])
    y0 = [x.clone() for x in (dev.rho[1], dev.q[1], dev.E[1])]
    outs = []
    for name in ("sodhip_rk_stage_f64", "sodhip_rk_stage_noalloc_f64"):
        for x, v in zip((dev.rho[1], dev.q[1], dev.E[1]), y0):
            x.copy_(v)
        R = [torch.full_like(dev.Rmass, float("nan")), torch.full_like(dev.Rmom, float("nan")),
             torch.full_like(dev.Rener, float("nan"))]
        check(getattr(lib(), name)(*_args(dev, R, *coef)), name)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy().copy() for t in (dev.rho[1], dev.q[1], dev.E[1], *R)])
    w = dev.lpoin_w.cpu().numpy() - 1
    for a, b, what in zip(outs[0], outs[1], ("rho", "q", "E", "Rmass", "Rmom", "Rener")):
        if a.size == 3 * dev.npoin:
            a, b = a.reshape(3, -1)[:, w], b.reshape(3, -1)[:, w]
        else:
            a, b = a[w], b[w]
        assert np.isfinite(a).all(), what
        assert np.array_equal(a, b), (what, dev.nelem, dev.op_precision)


@pytest.mark.parametrize("op_precision", [8, 4])
def test_noalloc_stage_matches_allocating_stage(op_precision):
    first = _run(8, op_precision)
    _compare(first)
    second = _run(4, op_precision)  # a smaller mesh, right after
    _compare(second)


def test_noalloc_stage_rejects_oversized_and_unprepared_meshes():
    import torch
    from sod2d_amd._lib import lib
    from sod2d_amd.device import _ptr
    dev = _run(4, 8)
    R = [torch.zeros_like(dev.Rmass), torch.zeros_like(dev.Rmom), torch.zeros_like(dev.Rener)]
    args = list(_args(dev, R, 0.5, 0.5, 0.1))
    # 27000 elements of p=3 need 27000*64*40 bytes of FP64 records, above 64 MiB;
    # the size is checked before any table is looked up or any kernel launched
    args[2] = 30 ** 3
    assert lib().sodhip_rk_stage_noalloc_f64(*args) != 0
    # a connectivity no allocating call has seen
    other = dev.connec.clone()
    args = list(_args(dev, R, 0.5, 0.5, 0.1))
    args[5] = _ptr(other)
    assert lib().sodhip_rk_stage_noalloc_f64(*args) != 0
