"""The fused time loop on non-affine elements.

Every other time-loop test runs periodic affine boxes, where He and gpvol are
the same at all nodes of an element, so an element kernel that took one
node's geometry for another's would pass them.  Here the nodes of a
non-periodic box are moved by U(-0.1, 0.1) x the local GLL spacing per axis
(every node off the domain boundary), so He and gpvol differ from node to
node, and the device time loop (two-wave element kernel at p=3, its LPS mode,
the single-wave kernel at p=2 and p=4) is compared with the oracle's
restatement step by step, with the bars of the periodic time-loop tests:
rho, rho*u, E within 1e-12 normwise relative, mu_e within 1e-10; the
FP32-operator mode with the bounds of its periodic test."""
import numpy as np
import pytest

from helpers import make_case, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p,n,lps", [(3, 4, False), (3, 3, True), (2, 4, False), (2, 3, True),
                                     (4, 2, False)])
def test_time_loop_on_jittered_box(oracle, p, n, lps):
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(p, n, periodic=False, jitter=0.1, lengths=(2 * np.pi,) * 3, seed=11)
    # the geometry is not affine: He varies over the nodes of an element
    He = g.He.reshape(9, g.tables.nnode, -1, order="F")
    spread = np.abs(He - He.mean(axis=1, keepdims=True)).max()
    assert spread > 1e-3 * np.abs(He).max()
    tp = TGVParams()
    to = oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    kw = dict(sensor_path="live", lps=True) if lps else {}
    dev = TGVRun(p, geometry=g, params=tp, initial=ic, **kw)
    orc = oracle.TGVOracle(p, g, ic, to, cfl=(tp.cfl_conv, tp.cfl_diff), **kw)
    print(f"jittered box p={p} n={n} lps={lps}", flush=True)
    assert abs(dev.dt - orc.dt) <= 1e-12 * orc.dt
    for step in range(4):
        dev.step()
        orc.step()
        s = dev.state()
        w = orc.lpoin_w - 1
        assert relerr(s["rho"][w], orc.rho[1][w]) < 1e-12, step
        assert relerr(s["q"][w], orc.q[1][w]) < 1e-12, step
        assert relerr(s["E"][w], orc.E[1][w]) < 1e-12, step
        assert relerr(s["mu_e"], orc.mu_e) < 1e-10, step
    # the flow moved
    assert relerr(s["q"][w], ic["q"][w]) > 1e-8


def test_time_loop_fp32_operators_on_jittered_box(oracle):
    """The FP32-operator mode of the two-wave kernel (FP64 state, gather, mass
    inverse and combine) on the non-affine box against the FP64 oracle after 4
    steps, with the bounds of the periodic FP32-operator test (rho 2e-7,
    rho*u 1e-5, E 2e-7)."""
    from sod2d_amd.solver import TGVParams, TGVRun
    g = make_case(3, 4, periodic=False, jitter=0.1, lengths=(2 * np.pi,) * 3, seed=11)
    tp = TGVParams()
    to = oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, to)
    dev = TGVRun(3, geometry=g, params=tp, initial=ic, op_precision=4)
    orc = oracle.TGVOracle(3, g, ic, to, cfl=(tp.cfl_conv, tp.cfl_diff))
    for _ in range(4):
        dev.step()
        orc.step()
    s = dev.state()
    w = orc.lpoin_w - 1
    e = {k: relerr(s[k][w], getattr(orc, k)[1][w]) for k in ("rho", "q", "E")}
    print("FP32 operators on the jittered box:", e, flush=True)
    assert e["rho"] < 2e-7 and e["q"] < 1e-5 and e["E"] < 2e-7, e
