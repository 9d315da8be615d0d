"""Pins of the oracle's time-step sequencing and end-of-step entropy residual.

The oracle's TGV step (oracle.TGVOracle.step) runs its stages through
oracle.ls_rk_step, the restatement of rk_4_ls_main's stage sequence
(time_integ_ls.f90:270-361), and forms the entropy residual with
oracle.entropy_residual (time_integ_ls.f90:436-444).  Both are pinned here by
mathematics the reference's sequence must satisfy, not by restating it:

* the first step (fresh residual) is the third-order SSP(5,3) scheme;
* every later step reuses, in stage 1, the residual of the previous step's
  last updateF, i.e. the derivative at the last stage's abscissa c_S instead
  of at t_n.  First-order error analysis gives the global error of y' = f(y)
  in closed form: e(T) = -K dt int_0^T Phi(T,t) y''(t) dt with
  K = gamma_21 * prod_{s>=2} lambda_{s+1,s} * (1 - c_S), where the abscissae
  follow from the tables alone (c_2 = gamma_21,
  c_{s+1} = lambda_{s+1,s} c_s + gamma_{s+1,s});
* the AB2 entropy residual is second order in dt at a fixed evaluation time,
  and converges to the semi-discrete residual of the entropy equation.

Each check is also run on planted mistakes of the oracle (on the CPU) to show
that it fails them.
"""
import math

import numpy as np
import pytest

from helpers import make_case


def _tables(oracle, stages=5, variant=1):
    return oracle.rk_ls_coeffs(stages, variant)


def _abscissae(L, G, S):
    c = np.zeros(S + 2)
    c[2] = G[1, 0]
    for s in range(2, S + 1):
        c[s + 1] = L[s, s - 1] * c[s] + G[s, s - 1]
    return c


def _ode_run(step_fn, L, G, S, f, y0, T, N, parity):
    """N steps of the sequence on y' = f(y) (complex scalars allowed)."""
    dt = T / N
    y1 = np.array([y0], dtype=complex)
    y2 = y1.copy()
    R = np.zeros(1, dtype=complex)

    def update_f():
        R[:] = f(y2)

    def combine(l1, ls, gdt):
        y2[:] = l1 * y1 + ls * y2 + gdt * R

    first = True
    for _ in range(N):
        y1[:] = y2  # CFDSolverBase.f90:1952-1959
        step_fn(update_f, combine, L, G, S, dt, first, parity)
        first = False
    return y2[0]


def _stale_error_ratio(step_fn, L, G, S, lam, N, T=1.0):
    c = _abscissae(L, G, S)
    D0 = np.prod([L[s, s - 1] for s in range(2, S + 1)])
    K = G[1, 0] * D0 * (1.0 - c[S])
    yT = np.exp(lam * T)
    e = _ode_run(step_fn, L, G, S, lambda y: lam * y, 1.0, T, N, True) - yT
    return e / (-K * lam ** 2 * T * (T / N) * yT)


def test_tables_consistent_abscissae(oracle):
    """The final stage of every scheme lands at t_n + dt (c_{S+1} = 1)."""
    for S, v in ((3, 1), (4, 1), (5, 1), (5, 2), (5, 3), (5, 4)):
        L, G = _tables(oracle, S, v)
        assert abs(_abscissae(L, G, S)[S + 1] - 1.0) < 1e-6, (S, v)
    L, G = _tables(oracle)
    assert abs(_abscissae(L, G, 5)[6] - 1.0) < 1e-15


def test_first_step_is_the_ssp_scheme(oracle):
    """One step from a fresh start: amplification P(z) = exp(z) + O(z^4)."""
    L, G = _tables(oracle)
    errs = []
    for z in (0.2, 0.1, 0.05):
        y = _ode_run(oracle.ls_rk_step, L, G, 5, lambda y, z=z: z * y, 1.0, 1.0, 1, True)
        t3 = 1 + z + z * z / 2 + z ** 3 / 6
        errs.append(abs(y - t3))
        assert abs(y - math.exp(z)) < 1e-2 * z ** 3
    assert 14.0 < errs[0] / errs[1] < 18.0 and 14.0 < errs[1] / errs[2] < 18.0, errs


@pytest.mark.parametrize("lam", [-1.0, 1j, -0.5 + 2j])
def test_reference_sequencing_error_constant(oracle, lam):
    """Global error of the reference sequencing on y' = lam y equals the
    closed-form first-order term -K lam^2 T dt y(T) (ratio -> 1 as dt -> 0)."""
    L, G = _tables(oracle)
    r4, r8 = (_stale_error_ratio(oracle.ls_rk_step, L, G, 5, lam, N) for N in (400, 800))
    assert abs(r4 - 1) < 5e-3 and abs(r8 - 1) < 2.5e-3, (r4, r8)
    # the remainder is O(dt): it halves with dt
    assert 1.7 < abs(r4 - 1) / abs(r8 - 1) < 2.3


def test_reference_sequencing_nonlinear(oracle):
    """y' = -y^2, y(0) = 1: e(T) = -K dt 2 ln(1+T)/(1+T)^2 to first order."""
    L, G = _tables(oracle)
    c = _abscissae(L, G, 5)
    K = G[1, 0] * np.prod([L[s, s - 1] for s in range(2, 6)]) * (1.0 - c[5])
    T = 1.0
    for N in (400, 800):
        e = _ode_run(oracle.ls_rk_step, L, G, 5, lambda y: -y * y, 1.0, T, N, True) - 1 / (1 + T)
        pred = -K * (T / N) * 2.0 * math.log(1 + T) / (1 + T) ** 2
        assert abs(e.real / pred - 1) < 6.0 / N, (N, e, pred)


def test_clean_sequencing_is_third_order(oracle):
    L, G = _tables(oracle)
    e = [abs(_ode_run(oracle.ls_rk_step, L, G, 5, lambda y: -y, 1.0, 1.0, N, False) - math.exp(-1))
         for N in (10, 20, 40)]
    assert 7.0 < e[0] / e[1] < 9.0 and 7.0 < e[1] / e[2] < 9.0, e


def _planted_sequencers(oracle):
    """Plausible mistakes in the sequence: the kept residual taken at the
    wrong stage, the stage-1 residual evaluated fresh, the first-step updateF
    dropped, the last combine's two levels swapped."""
    def wrong_stage(update_f, combine, L, G, S, dt, first, parity, on_first=None):
        if first:
            update_f()
        combine(0.0, 1.0, G[1, 0] * dt)
        for s in range(2, S + 1):
            if s < S:
                update_f()
            combine(L[s, 0], L[s, s - 1], G[s, s - 1] * dt)

    def always_fresh(update_f, combine, L, G, S, dt, first, parity, on_first=None):
        return oracle.ls_rk_step(update_f, combine, L, G, S, dt, first, False, on_first)

    def no_first_update(update_f, combine, L, G, S, dt, first, parity, on_first=None):
        return oracle.ls_rk_step(update_f, combine, L, G, S, dt, False, parity, on_first)

    def last_stage_level1(update_f, combine, L, G, S, dt, first, parity, on_first=None):
        if first:
            update_f()
        combine(0.0, 1.0, G[1, 0] * dt)
        for s in range(2, S + 1):
            update_f()
            if s == S:
                combine(L[s, s - 1], L[s, 0], G[s, s - 1] * dt)
            else:
                combine(L[s, 0], L[s, s - 1], G[s, s - 1] * dt)

    return {"wrong_stage": wrong_stage, "always_fresh": always_fresh,
            "no_first_update": no_first_update, "last_stage_level1": last_stage_level1}


def test_sequencing_checks_fail_planted_mistakes(oracle):
    L, G = _tables(oracle)
    for name, fn in _planted_sequencers(oracle).items():
        r = _stale_error_ratio(fn, L, G, 5, -1.0, 400)
        assert abs(r - 1) > 0.05, (name, r)


def test_reference_sequencing_has_no_imaginary_axis_stability(oracle):
    """The kept stage-1 residual turns SSP(5,3) into a two-level scheme on
    (y, kept residual).  Its one-step map, taken from the oracle's sequencer,
    has spectral radius above 1 for purely imaginary z: undamped oscillatory
    modes grow, and only the operator's viscous, entropy-viscosity (or LPS)
    damping bounds them.  With fresh residuals the scheme is stable on the
    imaginary axis up to |z| = 2.398."""
    L, G = _tables(oracle)

    def step_map(z):
        cols = []
        for y0, r0 in ((1.0, 0.0), (0.0, 1.0)):
            y1 = np.array([y0], complex)
            y2 = y1.copy()
            R = np.array([r0], complex)

            def update_f():
                R[:] = z * y2

            def combine(l1, ls, gdt):
                y2[:] = l1 * y1 + ls * y2 + gdt * R

            oracle.ls_rk_step(update_f, combine, L, G, 5, 1.0, False, True)
            cols.append([y2[0], R[0]])
        return np.array(cols).T

    def fresh(z):
        return _ode_run(oracle.ls_rk_step, L, G, 5, lambda y: z * y, 1.0, 1.0, 1, False)

    rho = lambda z: max(abs(np.linalg.eigvals(step_map(z))))  # noqa: E731
    for w in (0.05, 0.5, 1.0, 2.0):
        assert rho(1j * w) > 1.0 + 1e-5 * w, w
    assert abs(rho(0.5j) - 1.021) < 1e-3
    assert rho(-0.5 + 0j) < 1.0 and rho(-0.3 + 0.2j) < 1.0
    assert all(abs(fresh(1j * w)) <= 1.0 for w in np.linspace(0.01, 2.39, 60))
    assert abs(fresh(2.41j)) > 1.0


def test_tgv_oracle_steps_through_the_pinned_sequencer(oracle, monkeypatch):
    """TGVOracle.step calls ls_rk_step (so the pins above apply to it) and
    entropy_residual, once per step each, with the run's flags."""
    calls = []
    real_seq, real_res = oracle.ls_rk_step, oracle.entropy_residual

    def spy_seq(*a, **k):
        calls.append(("seq", a[5], a[6], a[7]))
        return real_seq(*a, **k)

    def spy_res(*a, **k):
        calls.append(("res",))
        return real_res(*a, **k)

    monkeypatch.setattr(oracle, "ls_rk_step", spy_seq)
    monkeypatch.setattr(oracle, "entropy_residual", spy_res)
    g = make_case(2, 2, periodic=True, lengths=(2 * np.pi,) * 3)
    to = oracle.tgv_params()
    orc = oracle.TGVOracle(2, g, oracle.tgv_ic(g.mesh.coord, to), to, dt_mode="fixed")
    orc.step()
    orc.step()
    dt = orc.dt
    assert calls == [("seq", dt, True, True), ("res",), ("seq", dt, False, True), ("res",)]


# ============================================================== AB2 entropy residual
def _aux_at(oracle, g, ic, to, tau, N, residual=None, monkeypatch=None):
    """auxReta after N clean steps whose last residual is centred at tau."""
    if residual is not None:
        monkeypatch.setattr(oracle, "entropy_residual", residual)
    orc = oracle.TGVOracle(3, g, ic, to, parity=False, dt_mode="fixed", sensor=False)
    orc.dt = tau / (N - 0.5)
    for _ in range(N):
        orc.step()
    return orc.auxReta.copy(), orc.Reta[:, 1].copy()


def _ab2_ratios(oracle, residual=None, monkeypatch=None):
    g = make_case(3, 3, periodic=True, lengths=(2 * np.pi,) * 3)
    to = oracle.tgv_params(Re=1e12)  # inviscid: the entropy equation holds for smooth flow
    ic = oracle.tgv_ic(g.mesh.coord, to)
    a = {N: _aux_at(oracle, g, ic, to, 0.04, N, residual, monkeypatch) for N in (8, 16, 32)}
    d1 = np.linalg.norm(a[8][0] - a[16][0])
    d2 = np.linalg.norm(a[16][0] - a[32][0])
    return d1 / d2, np.linalg.norm(a[32][0]), np.linalg.norm(a[32][1])


def test_entropy_residual_second_order(oracle):
    """1.5 Reta_n - 0.5 Reta_{n-1} + (eta_{n+1} - eta_n)/dt is centred at
    t_{n+1/2}: at a fixed centre time successive halvings of dt change it by
    factors of 4, and it stays a fraction of |Reta| (the semi-discrete
    residual of the entropy equation, a spatial error on 3^3 elements)."""
    ratio, aux, reta = _ab2_ratios(oracle)
    assert 3.6 < ratio < 4.6, ratio
    assert aux < 0.7 * reta


@pytest.mark.parametrize("mistake", ["euler_weights", "swapped_levels", "sign"])
def test_entropy_residual_check_fails_planted_mistakes(oracle, monkeypatch, mistake):
    real = oracle.entropy_residual

    def planted(lpoin_w, Reta, eta, dt, factor_comp=1.0):
        lw = lpoin_w - 1
        aux = np.zeros(Reta.shape[0])
        if mistake == "euler_weights":
            aux[lw] = Reta[lw, 1] + (eta[lw, 1] - eta[lw, 0]) / dt
        elif mistake == "swapped_levels":
            aux[lw] = (1.5 * Reta[lw, 1] - 0.5 * Reta[lw, 0]) + (eta[lw, 0] - eta[lw, 1]) / dt
        else:
            aux[lw] = -(1.5 * Reta[lw, 1] - 0.5 * Reta[lw, 0]) + (eta[lw, 1] - eta[lw, 0]) / dt
        Reta[lw, 0] = Reta[lw, 1]
        return aux

    assert real is not planted
    ratio, aux, reta = _ab2_ratios(oracle, planted, monkeypatch)
    assert not (3.6 < ratio < 4.6 and aux < 0.7 * reta), (mistake, ratio, aux, reta)


# ============================================================== KEP split form
def _divergence_form(T, connec, He, gpvol, u, q, pr):
    """gpvol * (div q, div(q_i u) + grad_i p) per element node, assembled: the
    plain divergence form of the same equations on the same derivative table
    (the form the split-form check must tell apart)."""
    nn = T.nnode
    p = round(nn ** (1 / 3)) - 1
    dl = T.dlxigp_ip  # (ngaus, 3, p+1)
    inv = T.invAtoIJK - 1  # (p+1, p+1, p+1), Fortran order (i, j, k)
    gI, gJ, gK = T.gmshAtoI - 1, T.gmshAtoJ - 1, T.gmshAtoK - 1
    npoin = u.shape[0]
    Rm, Rq = np.zeros(npoin), np.zeros((npoin, 3))
    for e in range(connec.shape[0]):
        nodes = connec[e] - 1
        def dxi(f):  # (nn,) nodal -> (nn, 3) reference-space derivatives
            out = np.zeros((nn, 3))
            for a in range(nn):
                I, J, K = gI[a], gJ[a], gK[a]
                for ii in range(p + 1):
                    out[a, 0] += dl[a, 0, ii] * f[inv[ii, J, K]]
                    out[a, 1] += dl[a, 1, ii] * f[inv[I, ii, K]]
                    out[a, 2] += dl[a, 2, ii] * f[inv[I, J, ii]]
            return out
        H = He[:, :, :, e]  # (3, 3, nn)
        grad = lambda f: np.einsum("ijg,gj->gi", H, dxi(f))  # noqa: E731
        ue, qe, pe = u[nodes], q[nodes], pr[nodes]
        vol = gpvol[0, :, e]
        divq = sum(grad(qe[:, j])[:, j] for j in range(3))
        Rm[nodes] += vol * divq
        for i in range(3):
            r = grad(pe)[:, i] + sum(grad(qe[:, i] * ue[:, j])[:, j] for j in range(3))
            Rq[nodes, i] += vol * r
    return Rm, Rq


def _ke_balance(lw, u, Rm, Rq):
    w = lw - 1
    bal = np.sum(u[w] * Rq[w]) - 0.5 * np.sum((u[w] ** 2).sum(1) * Rm[w])
    return bal / np.sum(np.abs(u[w] * Rq[w]))


@pytest.mark.parametrize("p,n", [(3, 3), (2, 4), (4, 2)])
def test_convection_is_kinetic_energy_preserving(oracle, p, n):
    """With a uniform pressure on a periodic affine mesh the discrete kinetic
    energy balance sum_n (u . Rmom - |u|^2/2 Rmass) vanishes for the reference's
    quarter-weight split form (elem_convec.f90:425-443): GLL differentiation is
    summation-by-parts with the lumped mass, and the split form is the
    skew-symmetric (kinetic-energy preserving) one.  The plain divergence form
    on the same derivative table leaves an O(1) imbalance, as would a dropped
    or mis-weighted split term."""
    from helpers import random_state
    g = make_case(p, n, periodic=True)
    m, T = g.mesh, g.tables
    s = random_state(m.npoin, m.nelem, T.nnode, seed=3)
    OT = oracle.Tables(p)
    pr = np.full(m.npoin, 0.7)
    Rm, Rq, _ = oracle.full_convec_ijk(OT, m.connec, g.He, g.gpvol, s["u"], s["q"], s["rho"], pr,
                                       s["h"], fact=1.0)
    assert abs(_ke_balance(m.lpoin_w, s["u"], Rm, Rq)) < 1e-14
    # the divergence form, checked first to be consistent (a uniform state
    # gives zero), fails the balance
    if p == 3:
        one = np.ones(m.npoin)
        z0 = _divergence_form(T, m.connec, g.He, g.gpvol, np.asfortranarray(np.ones((m.npoin, 3))),
                              np.asfortranarray(np.ones((m.npoin, 3))), one)
        assert np.abs(z0[0]).max() < 1e-12 and np.abs(z0[1]).max() < 1e-12
        Dm, Dq = _divergence_form(T, m.connec, g.He, g.gpvol, s["u"], s["q"], pr)
        assert abs(_ke_balance(m.lpoin_w, s["u"], Dm, Dq)) > 1e-3


# ---------------------------------------------------------------- classical RK4
def _rk4_amplification(oracle, A, dt, a=None, b=None, seq=None):
    """The matrix M with y_{n+1} = M y_n of one oracle RK4 step on y' = A y,
    column by column (R = -A y: rk_4_main's y advances by -dt sum b_i R)."""
    a0, b0, _ = oracle.rk4_coeffs()
    a = a0 if a is None else a
    b = b0 if b is None else b
    n = A.shape[0]
    M = np.zeros((n, n))
    lw = np.arange(1, n + 1, dtype=np.int32)
    for j in range(n):
        y = (np.eye(n)[:, j].copy(), np.zeros((n, 3), order="F"), np.zeros(n))
        aux = tuple(np.zeros_like(v) for v in y)
        R = tuple(np.zeros_like(v) for v in y)
        Rs = tuple(np.zeros_like(v) for v in y)

        def update_f(i):
            st = y if i == 0 else aux
            R[0][:] = -(A @ st[0])

        def node_update(an, bi, first, last):
            oracle.rk4_update(lw, dt, an, bi, first, last, y, aux, R, Rs)

        (seq or oracle.rk4_step)(update_f, node_update, a, b)
        M[:, j] = y[0]
    return M


def _rk4_poly(A, dt):
    Z = A * dt
    out, term = np.eye(A.shape[0]), np.eye(A.shape[0])
    for k in range(1, 5):
        term = term @ Z / k
        out = out + term
    return out


@pytest.mark.parametrize("A", [np.array([[-1.3]]), np.array([[0.7]]),
                               np.array([[0.0, -2.0], [2.0, 0.0]]),
                               np.array([[-0.5, -1.5], [1.5, -0.5]])])
def test_rk4_step_is_the_classical_fourth_order_polynomial(oracle, A):
    """One step on y' = A y multiplies y by 1 + Z + Z^2/2 + Z^3/6 + Z^4/24
    (Z = A dt), for real, imaginary and complex spectra, exactly to rounding;
    the tableau is rk_4_main's (time_integ.f90:84-87)."""
    for dt in (0.1, 0.37):
        M = _rk4_amplification(oracle, A, dt)
        assert np.abs(M - _rk4_poly(A, dt)).max() < 4e-16 * max(1.0, np.abs(M).max())


def test_rk4_planted_mistakes_fail(oracle):
    """A wrong weight, a wrong abscissa, or a stage built from the previous
    stage state instead of y each move the amplification off the RK4
    polynomial by far more than rounding."""
    A = np.array([[-0.5, -1.5], [1.5, -0.5]])
    dt = 0.3
    ref = _rk4_poly(A, dt)
    a, b, _ = oracle.rk4_coeffs()
    bad_b = np.array([0.25, 0.25, 0.25, 0.25])
    bad_a = np.array([0.0, 0.5, 1.0, 1.0])
    assert np.abs(_rk4_amplification(oracle, A, dt, b=bad_b) - ref).max() > 1e-3
    assert np.abs(_rk4_amplification(oracle, A, dt, a=bad_a) - ref).max() > 1e-3

    def chained(update_f, node_update, a_, b_, order=4):
        # stage state from the previous stage state: aux = aux - dt a R
        for i in range(order):
            update_f(i)
            last = i == order - 1
            node_update(0.0 if last else a_[i + 1] - (a_[i] if i else 0.0), b_[i], i == 0, last)

    # a chain that reaches the same abscissae through increments is still not RK4
    assert np.abs(_rk4_amplification(oracle, A, dt, seq=chained) - ref).max() > 1e-4


def test_rk4_tgv_step_goes_through_the_sequencer(oracle, monkeypatch):
    """TGVOracle with the classical integrator runs its stages through
    oracle.rk4_step, once per step, with rk_4_main's tableau."""
    g = make_case(3, 2, periodic=True, lengths=(2 * np.pi,) * 3)
    t = oracle.tgv_params()
    ic = oracle.tgv_ic(g.mesh.coord, t)
    orc = oracle.TGVOracle(3, g, ic, t, integrator="rk4", dt_mode="fixed")
    calls = []
    real = oracle.rk4_step

    def spy(update_f, node_update, a, b, order=4):
        calls.append((tuple(a), tuple(b)))
        return real(update_f, node_update, a, b, order)

    monkeypatch.setattr(oracle, "rk4_step", spy)
    orc.step()
    orc.step()
    assert calls == [((0.0, 0.5, 0.5, 1.0), (1 / 6, 1 / 3, 1 / 3, 1 / 6))] * 2
    assert np.isfinite(orc.rho[1]).all()
