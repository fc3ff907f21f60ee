"""Closed-form pins of the oracle's linear terms, error norms, step control
and multistep coefficients.

Each ``pin_*`` function takes the oracle module as its argument and raises
AssertionError when the module disagrees with a result that the mathematics
fixes independently of the oracle's own formulas:

  * pin_shear_identity      - a perturbation whose streamfunction is A*U(z)
                              in one Fourier mode has RHS = -i kx A U (beta -
                              k^2 U), whatever the stretching matrix
                              (QGN_Module.f90:270-273 with qyBar = -L uBar,
                              :148);
  * pin_phillips            - two-layer eigenfrequencies against the roots
                              of the textbook Phillips-model dispersion
                              relation;
  * pin_ark_error_control   - ARK43 / ARK43_QGLeith on one barotropic Fourier
                              mode: the embedded error norm against the
                              stability functions of the Butcher tableau
                              (ARK43_QGLeith.f90:121-124, ARK43.f90:107-110),
                              one rejection then one acceptance with each
                              variant's reject factor and safety factor
                              (ARK43_QGLeith.f90:125-139, ARK43.f90:111-123),
                              and the accepted state R(z) q;
  * pin_bsrk3_error_control - BS3(2): error norm (z^3 + z^4)/48 |q|,
                              rejection/acceptance and the new state
                              (BSRK3.f90:28-49);
  * pin_ab3_local_error     - AB3 fed the exact history of q' = lambda q
                              leaves the local error 3/8 z^4 q (AB3.f90:15-17).

``tests/test_oracle_pins.py`` runs them on the oracle, and
``tests/test_oracle_mutations.py`` runs them on copies of the oracle with
one planted mistake each, which must fail.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

from qgn_fortran_amd import background

LEITH_C32 = float(np.float32(1.3))   # ARK43_QGLeith.f90:72, a default-real literal


def _exp_model(orc, nx=48, nz=4, **kw):
    H, S, u = background.exponential(nz, 4000.0, 1e-4, 5e-3, 800.0, 0.1)
    Hs, L, Modes, EVals, qy = orc.vertical_modes(H, S, u, 4000.0)
    args = dict(nx=nx, ny=nx, nz=nz, Lx=1e6, Ly=1e6, Htot=4000.0, beta=1.6e-11)
    args.update(kw)
    p = orc.Params(**args)
    return orc.Model(p, Hs, Modes, EVals, u, qy), L, np.asarray(u, float)


# ---------------------------------------------------------------- RHS -------
def pin_shear_identity(orc):
    """psi_hat = A U (layer-varying U of the exponential profile, so
    Qy = -L U != 0) in one Fourier mode: q_hat = (-k^2 I + L) psi_hat and
    RHS = -i kx [U q_hat + (beta + Qy) psi_hat] = -i kx A U (beta - k^2 U),
    in which L cancels.  A single mode has no Jacobian."""
    nx = 48
    m, L, U = _exp_model(orc, nx=nx, r0=0.0, C_d=0.0)
    assert np.ptp(U) > 0.5 * np.max(np.abs(U))          # the shear is real
    for (j, i) in ((3, 2), (45, 5), (0, 7)):
        kx, ky = m.kx[j, i], m.ky[j, i]
        k2 = kx * kx + ky * ky
        A = 1e4 * (1.0 - 0.3j)
        psi = A * U
        qh = np.zeros((4, nx, nx // 2 + 1), complex)
        qh[:, j, i] = (-k2 * np.eye(4) + L) @ psi
        rhs = m.get_rhs(qh)
        exp = -1j * kx * A * U * (1.6e-11 - k2 * U)
        scale = np.abs(exp).max()
        assert np.abs(rhs[:, j, i] - exp).max() < 1e-10 * scale, (j, i, rhs[:, j, i], exp)
        rest = rhs.copy()
        rest[:, j, i] = 0
        assert np.abs(rest).max() < 1e-10 * scale


def pin_phillips(orc):
    """Two layers with U = (U1, U2): the linear operator's eigenvalues at a
    Fourier mode are -i kx c with c the roots of the Phillips-model relation
    det[[(U1-c)(-K^2-F1) + Q1, (U1-c) F1], [(U2-c) F2, (U2-c)(-K^2-F2) + Q2]]
    = 0, F1 = f0^2/(g' H1), F2 = f0^2/(g' H2), Q1 = beta + F1 (U1-U2),
    Q2 = beta - F2 (U1-U2)."""
    H1, H2, Rd, U1, U2, beta = 1000.0, 3000.0, 25e3, 0.05, -0.01, 1.6e-11
    H, S, u = background.two_layer(H1, H2, Rd, U1, U2)
    Hs, L, Modes, EVals, qy = orc.vertical_modes(H, S, u, H1 + H2)
    # f0^2/g' from Rd^2 = g' H1 H2 / (f0^2 (H1 + H2))
    f2g = H1 * H2 / ((H1 + H2) * Rd * Rd)
    F1, F2 = f2g / H1, f2g / H2
    nx = 32
    p = orc.Params(nx=nx, ny=nx, nz=2, Lx=1e6, Ly=1e6, Htot=H1 + H2, beta=beta, r0=0.0, C_d=0.0)
    m = orc.Model(p, Hs, Modes, EVals, u, qy)
    for (j, i) in ((1, 1), (3, 2), (30, 4)):
        kx, ky = m.kx[j, i], m.ky[j, i]
        K2 = kx * kx + ky * ky
        M = np.zeros((2, 2), complex)
        for col in range(2):
            qh = np.zeros((2, nx, nx // 2 + 1), complex)
            qh[col, j, i] = 1e3
            M[:, col] = m.get_rhs(qh)[:, j, i] / 1e3
        Q1, Q2 = beta + F1 * (U1 - U2), beta - F2 * (U1 - U2)

        def det(c):
            return ((U1 - c) * (-K2 - F1) + Q1) * ((U2 - c) * (-K2 - F2) + Q2) - (U1 - c) * F1 * (U2 - c) * F2
        # det is quadratic in c: fit it exactly through three points
        c0, c1, cm = det(0.0), det(1.0), det(-1.0)
        a2, a1 = 0.5 * (c1 + cm) - c0, 0.5 * (c1 - cm)
        roots = np.roots([a2, a1, c0])
        want = np.sort_complex(-1j * kx * roots)
        got = np.sort_complex(np.linalg.eigvals(M))
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max()), (j, i, got, want)


# ----------------------------------------------------------- steppers -------
def _ark_tab(orc):
    """The Butcher tables as exact rationals (the ones the coefficient
    identities of the oracle's tests pin)."""
    return orc.ARK_AE, orc.ARK_AI, orc.ARK_B, orc.ARK_BE


def ark_stability(orc, zE, zI):
    """(R, R - R_hat) of ARK4(3)6L[2]SA for y' = (lE + lI) y, complex zE, zI:
    Y_i = (1 + sum_j (ae_ij zE + ai_ij zI) Y_j) / (1 - zI/4),
    R = 1 + sum b_j (zE + zI) Y_j, R - R_hat = sum be_j (zE + zI) Y_j."""
    ae, ai, b, be = _ark_tab(orc)
    f = lambda x: float(x)
    Y = [1.0 + 0j]
    for i in range(2, 7):
        acc = 1.0 + 0j
        for j in range(1, i):
            acc += (f(ae.get((i, j), Fraction(0))) * zE + f(ai.get((i, j), Fraction(0))) * zI) * Y[j - 1]
        Y.append(acc / (1.0 - 0.25 * zI))
    R = 1.0 + sum(f(b[j]) * (zE + zI) * Y[j] for j in range(6))
    D = sum(f(be[j]) * (zE + zI) * Y[j] for j in range(6))
    return R, D


def _grid_max(nx, ny, j, i, coef):
    """max over the grid of |2 Re(coef exp(+2 pi i (i x/nx + j y/ny)))|, the
    c2r of one half-spectrum bin 0 < i < nx/2, computed point by point."""
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    jj = j if j <= ny // 2 else j - ny
    ph = np.exp(2j * np.pi * (i * x / nx + jj * y / ny))
    return np.max(np.abs(2.0 * np.real(coef * ph)))


def _barotropic_mode_state(m, j, i, q_phys_amp):
    """q_hat in the barotropic vertical mode (constant over layers), one
    Fourier bin, physical amplitude q_phys_amp (q = c2r(q_hat)/(nx ny))."""
    p = m.p
    nn = p.nx * p.ny
    bt = int(np.argmax(m.EVals))
    assert abs(m.EVals[bt]) < 1e-15 and np.ptp(m.Modes[:, bt]) < 1e-12
    Q = q_phys_amp * nn / 2.0 * np.exp(0.4j)
    qh = np.zeros((p.nz, p.ny, p.nx // 2 + 1), complex)
    qh[:, j, i] = Q * m.Modes[:, bt]
    omega = -p.beta * m.kx[j, i] / (m.k2[j, i] - m.EVals[bt])   # Rossby frequency of the mode
    return qh, Q * m.Modes[0, bt], omega


def pin_ark_error_control(orc, leith):
    """One barotropic Fourier mode (no Jacobian, U = 0): explicit part
    -i omega q, implicit part -(A2 k^2 + A8 k^8) q with A8 fixed (ARK43) or
    the QG-Leith value of the state (ARK43_QGLeith).  The first attempt's
    error norm is max|c2r((R - R_hat) q)|/(nx ny); TOL is placed so that the
    first dt is rejected and the reference's reject factor (0.5 with
    QG-Leith, 0.75 without) gives an accepted step, after which
    dt_new = min(dt (s TOL/err)^0.075 (TOL/err)^0.1, dt_max) with the
    variant's safety s (0.2 with QG-Leith, 0.75 without) and the state is
    R(z) q."""
    nx = 48
    j, i = 2, 3
    A2, A8 = 20.0, 0.0 if leith else 3e20
    m, _L, _U = _exp_model(orc, nx=nx, nz=3, r0=0.0, C_d=0.0, A2=A2, A8=A8, dt_max=1e12)
    m.uBar = np.zeros(3)
    m.qyBar = np.zeros(3)
    amp = 2e-6
    qh, Q, omega = _barotropic_mode_state(m, j, i, amp)
    k2 = m.k2[j, i]
    if leith:
        # mean(q del^8 q) = amp^2 k^8 / 2 for q = amp cos(theta)
        a8 = (LEITH_C32 * 1e6 / (nx * math.pi)) ** 12 * math.sqrt(amp * amp * k2 ** 4 / 2.0)
    else:
        a8 = A8
    lam = A2 * k2 + a8 * k2 ** 4
    nn = nx * nx

    def err(dt):
        R, D = ark_stability(orc, -1j * omega * dt, -lam * dt)
        return _grid_max(nx, nx, j, i, D * Q) / nn, R

    dt0 = 0.9 / abs(omega)
    f_rej = 0.5 if leith else 0.75
    s = 0.2 if leith else 0.75
    e_full, _ = err(dt0)
    e_acc, R_acc = err(f_rej * dt0)
    # TOL between the error at the reduced dt and the next larger candidate
    e_up, _ = err((0.75 if leith else 0.9) * dt0)
    assert e_acc < e_up < e_full
    TOL = math.sqrt(e_acc * e_up)
    m.p.TOL = TOL
    st = orc.ARK43(m, leith=leith)
    q1, t1, dt1 = st.step(qh, 0.0, dt0)
    assert len(st.log) == 2 and not st.log[0][2] and st.log[1][2], st.log
    assert abs(st.log[0][1] - e_full) < 1e-9 * e_full, (st.log[0][1], e_full)
    assert st.log[1][0] == f_rej * dt0
    assert abs(st.log[1][1] - e_acc) < 1e-9 * e_acc
    assert t1 == f_rej * dt0
    e1 = st.log[1][1]    # pinned to the closed form just above
    want_dt = f_rej * dt0 * (s * TOL / e1) ** 0.075 * (TOL / e1) ** 0.1
    assert abs(dt1 - want_dt) < 1e-12 * want_dt, (dt1, want_dt)
    want_q = R_acc * qh
    assert np.abs(q1 - want_q).max() < 1e-11 * np.abs(qh).max()


def pin_bsrk3_error_control(orc):
    """BS3(2) on one barotropic Fourier mode with RHS = -i omega q (no
    viscosity, is_explicit false): with z = -i omega dt the error estimate is
    (z^3 + z^4)/48 q (the tableau of BSRK3.f90:21-31 applied to y' = lambda
    y), the new state (1 + z + z^2/2 + z^3/6) q; one rejection with
    dt *= 0.75, then dt_new = min(dt (0.75 TOL/r1)^0.1 (TOL/r1)^(0.4/3),
    dt_max)."""
    nx = 48
    j, i = 5, 2
    m, _L, _U = _exp_model(orc, nx=nx, nz=3, r0=0.0, C_d=0.0, dt_max=1e12)
    m.uBar = np.zeros(3)
    m.qyBar = np.zeros(3)
    qh, Q, omega = _barotropic_mode_state(m, j, i, 3e-6)
    nn = nx * nx

    def err(dt):
        z = -1j * omega * dt
        return _grid_max(nx, nx, j, i, (z ** 3 + z ** 4) / 48.0 * Q) / nn, 1 + z + z * z / 2 + z ** 3 / 6

    dt0 = 0.8 / abs(omega)
    e_full, _ = err(dt0)
    e_acc, R_acc = err(0.75 * dt0)
    e_up, _ = err(0.9 * dt0)
    assert e_acc < e_up < e_full
    TOL = math.sqrt(e_acc * e_up)
    m.p.TOL = TOL
    st = orc.BSRK3(m)
    q1, t1, dt1 = st.step(qh, 0.0, dt0)
    assert len(st.log) == 2 and not st.log[0][2] and st.log[1][2], st.log
    assert abs(st.log[0][1] - e_full) < 1e-9 * e_full, (st.log[0][1], e_full)
    assert abs(st.log[1][1] - e_acc) < 1e-9 * e_acc
    assert t1 == 0.75 * dt0
    e1 = st.log[1][1]    # pinned to the closed form just above
    want_dt = 0.75 * dt0 * (0.75 * TOL / e1) ** 0.1 * (TOL / e1) ** (0.4 / 3.0)
    assert abs(dt1 - want_dt) < 1e-12 * want_dt, (dt1, want_dt)
    assert np.abs(q1 - R_acc * qh).max() < 1e-11 * np.abs(qh).max()


def pin_ab3_local_error(orc):
    """AB3 (AB3.f90:15-17) given the exact history of q' = -i omega q (one
    barotropic Fourier mode) takes a step whose local error is 3/8 z^4 q +
    O(z^5), z = -i omega dt, the error constant of the three-step
    Adams-Bashforth method; Richardson extrapolation over dt and dt/2
    removes the O(z^5) term."""
    nx = 32
    m, _L, _U = _exp_model(orc, nx=nx, nz=3, r0=0.0, C_d=0.0)
    m.uBar = np.zeros(3)
    m.qyBar = np.zeros(3)
    j, i = 1, 2
    qh, Q, omega = _barotropic_mode_state(m, j, i, 1e-6)

    def const(dt):
        z = -1j * omega * dt
        st = orc.AB3(m)
        st.first = st.second = False
        st.r0 = m.get_rhs(np.exp(-2 * z) * qh)    # RHS at t - 2 dt
        st.r1 = m.get_rhs(np.exp(-z) * qh)        # RHS at t - dt
        q1, _t, _dt = st.step(qh, 0.0, dt)
        e = (np.exp(z) * Q - q1[0, j, i]) / Q
        return e / z ** 4

    dt = 0.04 / abs(omega)
    c1, c2 = const(dt), const(dt / 2)
    c0 = 2.0 * c2 - c1
    assert abs(c0 - 3.0 / 8.0) < 2e-3, (c1, c2, c0)


PINS = {
    "shear_identity": pin_shear_identity,
    "phillips": pin_phillips,
    "ark43_error_control": lambda orc: pin_ark_error_control(orc, leith=False),
    "ark43_qgleith_error_control": lambda orc: pin_ark_error_control(orc, leith=True),
    "bsrk3_error_control": pin_bsrk3_error_control,
    "ab3_local_error": pin_ab3_local_error,
}
