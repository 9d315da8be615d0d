"""The MLP surrogate of the entropy-viscosity sensor on the GPU.

The MLP kernels are checked against an FP32 NumPy reference computed in the
test from the same weights, with operands rounded to BF16 where the matrix
cores take them (inputs, weights, hidden activations, back-propagated
deltas); the remaining difference is FP32 summation order, so forward values
agree to 1e-4 and gradients to 2e-3 normwise (a BF16 rounding of a hidden
activation or delta can flip with the last FP32 bit).  The features are checked
against a NumPy restatement of their definition on the plain sensor's own
inputs; the deployed sensor against the plain sensor on fallback elements (bit
for bit) and against its log-scale output elsewhere; a trained surrogate in
the TGV against the plain run (kinetic energy and its dissipation rate within
1 %)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NF, NFP, NH = 12, 16, 64
W1, B1, W2, B2, W3, B3 = 0, 768, 832, 4928, 4992, 5056


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).reshape(a.shape)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[W1:B1].reshape(NF, NH), w[B1:W2], w[W2:B2].reshape(NH, NH), w[B2:W3], w[W3:B3],
            w[B3])


def ref_forward(w, X):
    w1, b1, w2, b2, w3, b3 = unpack(w)
    x = bf16(X[:, :NF]).astype(np.float64)
    z1 = x @ bf16(w1) + b1
    h1 = bf16(silu(z1))
    z2 = h1 @ bf16(w2) + b2
    h2 = bf16(silu(z2))
    y = h2 @ bf16(w3) + b3
    return y, z1, z2


def ref_gradient(w, X, t):
    w1, b1, w2, b2, w3, b3 = unpack(w)
    y, z1, z2 = ref_forward(w, X)
    n = X.shape[0]
    dy = 2.0 * (y - t) / n
    w3b = bf16(w3).astype(np.float64)
    dz2 = dy[:, None] * w3b[None, :] * dsilu(z2)
    dh1 = bf16(dz2) @ bf16(w2).T
    dz1 = dh1 * dsilu(z1)
    g = np.zeros(5057)
    g[W1:B1] = (bf16(X[:, :NF]).T @ bf16(dz1)).ravel()
    g[B1:W2] = dz1.sum(0)
    g[W2:B2] = (bf16(silu(z1)).T @ bf16(dz2)).ravel()
    g[B2:W3] = dz2.sum(0)
    g[W3:B3] = (silu(z2) * dy[:, None]).sum(0)
    g[B3] = dy.sum()
    return y, g


def relerr(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from sod2d_amd.device import Context
    return Context(0)


def _rand_X(n, seed=3):
    r = np.random.default_rng(seed)
    X = np.zeros((n, NFP), dtype=np.float32)
    X[:, :NF] = r.normal(0.0, 1.0, (n, NF))
    return X


@pytest.mark.parametrize("n", [100, 4096])
def test_mlp_forward(ctx, n):
    from sod2d_amd.surrogate import SurrogateMLP
    net = SurrogateMLP(ctx, seed=5)
    X = _rand_X(n)
    y = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    yr, _, _ = ref_forward(net.w.cpu().numpy(), X)
    assert relerr(y, yr) < 1e-4


@pytest.mark.parametrize("n", [77, 2048])
def test_mlp_gradient(ctx, n):
    from sod2d_amd.surrogate import SurrogateMLP
    net = SurrogateMLP(ctx, seed=6)
    X = _rand_X(n, seed=4)
    t = np.random.default_rng(9).uniform(0, 1, n).astype(np.float32)
    y, g = net.gradient(torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda())
    yr, gr = ref_gradient(net.w.cpu().numpy(), X, t)
    g = g.cpu().numpy()
    assert relerr(y.cpu().numpy(), yr) < 1e-4
    for lo, hi in ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3), (B3, B3 + 1)):
        assert relerr(g[lo:hi], gr[lo:hi]) < 2e-3, (lo, hi)


def test_adam_step(ctx):
    """One step of sodhip_mlp_train_step on a full batch equals Adam's first
    step on the gradient of sodhip_mlp_gradient (the update of step 1 is
    lr * g / (|g| + eps), elementwise)."""
    import ctypes as C
    from sod2d_amd._lib import check, lib
    from sod2d_amd.surrogate import SurrogateMLP
    net = SurrogateMLP(ctx, seed=7)
    n = 512
    X = torch.from_numpy(_rand_X(n, seed=8)).cuda()
    t = torch.rand(n, device="cuda")
    _, g = net.gradient(X, t)
    w0 = net.w.cpu().numpy().copy()
    y = torch.empty(n, device="cuda")
    Z = torch.empty(2 * n * NH, device="cuda")
    gp = torch.empty(64 * 5057, device="cuda")
    lr = 1e-3
    check(lib().sodhip_mlp_train_step(ctx.h, C.byref(net.struct()), n,
                                      C.c_void_p(X.data_ptr()), C.c_void_p(t.data_ptr()), 1, lr,
                                      C.c_void_p(y.data_ptr()), C.c_void_p(Z.data_ptr()),
                                      C.c_void_p(gp.data_ptr()), 64), "train_step")
    g = g.cpu().numpy().astype(np.float64)
    m = 0.1 * g
    v = 0.001 * g * g
    want = w0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    got = net.w.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-6 * max(1.0, np.abs(w0).max())
    # the BF16 operand copy follows the new weights
    wq = net.wq.float().cpu().numpy()
    assert np.array_equal(wq[1024:1024 + 4096], bf16(got[W2:B2]))


def test_training_reduces_loss(ctx):
    from sod2d_amd.surrogate import SurrogateMLP
    net = SurrogateMLP(ctx, seed=1)
    r = np.random.default_rng(2)
    X = np.zeros((20000, NFP), dtype=np.float32)
    X[:, :NF] = r.uniform(0, 1, (20000, NF))
    t = (1.0 / (1.0 + np.exp(-(3 * X[:, 0] - 2 * X[:, 1] + X[:, 5] * X[:, 7] - 0.5)))).astype(np.float32)
    hist = net.train(torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda(), steps=1500,
                     batch=2048, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0], hist
    y = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    assert np.sqrt(np.mean((y - t) ** 2)) < 0.03


def _tgv_state(n=6, steps=3):
    from sod2d_amd.solver import TGVParams, TGVRun
    run = TGVRun(3, n, params=TGVParams(M=0.3), dt_mode="fixed")
    for _ in range(steps):
        run.step()
    return run


def test_features_match_definition(ctx):
    """The features of the sensor's own inputs, restated in NumPy."""
    from oracle import oracle as O
    from sod2d_amd.surrogate import features
    run = _tgv_state()
    n, ne = run.npoin, run.nelem
    X, beta = features(run.ctx, 3, ne, n, run.npoin_w, run.connec, run.lpoin_w, run.auxReta,
                       run.t.gamma_gas, run.rho[1], run.u[1], run.Tem[1], run.eta[n:],
                       run.helem_l, run.helem, run.tabs[0], run.vp)
    X, beta = X.cpu().numpy(), beta.cpu().numpy()
    connec = run.geo.mesh.connec - 1
    lw = run.geo.mesh.lpoin_w - 1
    Reta, rho = run.auxReta.cpu().numpy(), run.rho[1].cpu().numpy()
    u = run.u[1].cpu().numpy().reshape(3, n).T
    T, eta = run.Tem[1].cpu().numpy(), run.eta[n:].cpu().numpy()
    hl = run.helem_l.cpu().numpy().reshape(64, ne).T
    hk = run.helem.cpu().numpy()
    vp = run.vp
    mean = eta[lw].sum() / lw.size
    norm = np.abs(eta[lw] - mean).max()
    Tb = O.Tables(3)
    # the 16 nodes read: vertices (sod2d indices 1..2 in every direction) and
    # interior nodes (3..4)
    corner = (Tb.gI <= 2) & (Tb.gJ <= 2) & (Tb.gK <= 2)
    inner = (Tb.gI >= 3) & (Tb.gJ >= 3) & (Tb.gK >= 3)
    sub = corner | inner
    assert corner.sum() == 8 and inner.sum() == 8
    for e in range(ne):
        nd = connec[e][sub]
        um = np.sqrt((u[nd] ** 2).sum(1))
        c = np.sqrt(vp.Rgas * run.t.gamma_gas * T[nd])
        be = (rho[nd] * hk[e] * (vp.cmax / 3) * (um + c)).max()
        r = vp.ce * rho[nd] * np.abs(Reta[nd]) / norm * hl[e][sub] ** 2 / be
        m = np.minimum(r, 1.0)
        mc, mi = m[corner[sub]], m[inner[sub]]
        g = lambda x: 1 + np.log10(x + 1e-10) / 10  # noqa: E731
        f = [g(m.max()), g(m.mean()), g(m.min()), g(np.sqrt(max(np.mean(m * m) - m.mean() ** 2, 0))),
             g(mc.mean()), g(mi.mean()), np.mean(r >= 1), np.log10(1 + r.max()) / 4,
             (um / c).max(), np.mean(um / (um + c)),
             10 * (rho[nd].max() - rho[nd].min()) / rho[nd].mean(),
             10 * (T[nd].max() - T[nd].min()) / T[nd].mean()]
        assert np.allclose(X[e, :NF], f, rtol=2e-6, atol=1e-7), e
        assert np.all(X[e, NF:] == 0)
        assert abs(beta[e] - be) <= 1e-13 * be


def test_surrogate_sensor_output_and_fallback(ctx):
    """Inside the feature range the element's mu_e is cglob beta_e y_e at every
    node; with an empty range every element falls back to the plain sensor,
    bit for bit."""
    from sod2d_amd.surrogate import SurrogateMLP
    run = _tgv_state()
    n, ne = run.npoin, run.nelem
    net = SurrogateMLP(run.ctx, seed=3)
    with torch.no_grad():
        net.w[B3] = 0.5  # a bias that keeps y inside [0, 1] for small inputs
        net.w[W3:B3] *= 0.01
    net.load(net.state_dict())
    args = (run.auxReta, run.rho[1], run.u[1], run.Tem[1], run.eta[n:])
    run.surrogate = net
    run._sx = torch.empty(ne * 16, device="cuda")
    run._sbeta = torch.empty(ne, dtype=torch.float64, device="cuda")
    run._sy = torch.empty(ne, device="cuda")
    run._sflag = torch.zeros(ne, dtype=torch.int32, device="cuda")
    run._snfall = torch.zeros(1, dtype=torch.int32, device="cuda")
    mu = torch.full_like(run.mu_e, float("nan"))
    run._surrogate_visc(*args, mu)
    y = run._sy.cpu().numpy()
    assert int(run._snfall.item()) == 0 and np.all((y >= 0) & (y <= 1))
    # the fused kernel's features and output are those of the standalone
    # feature and forward entry points, bit for bit
    from sod2d_amd.surrogate import features
    X, beta0 = features(run.ctx, 3, ne, n, run.npoin_w, run.connec, run.lpoin_w, *args[:1],
                        run.t.gamma_gas, *args[1:], run.helem_l, run.helem, run.tabs[0], run.vp)
    assert torch.equal(run._sx.view(ne, 16), X) and torch.equal(run._sbeta, beta0)
    assert np.array_equal(net.forward(X).cpu().numpy(), y)
    assert np.all(run._sflag.cpu().numpy() == 0)
    beta = run._sbeta.cpu().numpy()
    ratio = np.maximum(10.0 ** (10.0 * (y.astype(np.float64) - 1.0)) - 1e-10, 0.0)
    want = np.repeat((run.vp.cglob * beta * ratio)[None, :], 64, 0).ravel()
    assert np.allclose(mu.cpu().numpy(), want, rtol=1e-13, atol=0)
    # every element out of range: the plain sensor's output
    net.lo = np.full(12, 1e30, dtype=np.float32)
    net.hi = np.full(12, -1e30, dtype=np.float32)
    mu2 = torch.full_like(run.mu_e, float("nan"))
    run._surrogate_visc(*args, mu2)
    assert int(run._snfall.item()) == ne
    plain = torch.empty_like(run.mu_e)
    import ctypes as C
    from sod2d_amd._lib import check, lib
    from sod2d_amd.device import _ptr
    check(lib().sodhip_smart_visc_spectral_f64(
        run.ctx.h, 3, ne, n, run.npoin_w, _ptr(run.connec), _ptr(run.lpoin_w), _ptr(args[0]),
        run.t.gamma_gas, _ptr(args[1]), _ptr(args[2]), _ptr(args[3]), _ptr(args[4]),
        _ptr(run.helem_l), _ptr(run.helem), _ptr(run.tabs[0]), C.byref(run.vp), _ptr(plain),
        _ptr(torch.empty_like(plain))), "smart_visc_spectral")
    assert np.array_equal(mu2.cpu().numpy(), plain.cpu().numpy())


def test_surrogate_tgv_run(ctx):
    """A surrogate trained on plain-sensor samples (random operator fields and
    12^3 TGV runs, with scaled entropy residuals) deployed in a 16^3 TGV at Ma 0.1, Re 1600, judged by the plain
    sensor every 10 steps: over 60 steps the kinetic energy stays within 1e-4
    and its dissipation rate within 0.2 % of the plain run; on the judge steps
    after the first the surrogate's mu_e lies within 50 % (2-norm) of the plain
    sensor's and under a fifth of the elements fall back on average."""
    from sod2d_amd.solver import TGVParams, TGVRun
    from sod2d_amd.surrogate import trained_surrogate
    net, hist, _ = trained_surrogate(ctx, steps=1500)
    assert hist[-1] < hist[0]
    plain = TGVRun(3, 16, params=TGVParams(), dt_mode="fixed")
    sur = TGVRun(3, 16, params=TGVParams(), dt_mode="fixed", sensor_path="surrogate",
                 surrogate=net)
    ek_p, ek_s = [], []
    for _ in range(60):
        plain.step()
        sur.step()
        ek_p.append(plain.diagnostics()["EK"])
        ek_s.append(sur.diagnostics()["EK"])
    ek_p, ek_s = np.array(ek_p), np.array(ek_s)
    assert np.abs(ek_s - ek_p).max() <= 1e-4 * np.abs(ek_p).max()
    dt = plain.dt
    diss_p, diss_s = -np.diff(ek_p) / dt, -np.diff(ek_s) / dt
    assert abs(diss_s[-10:].mean() - diss_p[-10:].mean()) <= 2e-3 * abs(diss_p[-10:].mean())
    later = [j for j in sur.judge_log if j[0] >= 10]
    assert len(later) >= 5
    assert max(j[1] for j in later) <= 0.5
    assert np.mean([j[2] for j in later]) <= 0.2 * sur.nelem
