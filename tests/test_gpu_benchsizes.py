"""Parity at the sizes the throughput benchmark times: the RHS kernels are
compiled per transform length, so the 1024^2 x 16 and 2048^2 x 32
instantiations get their own checks against the FP64 oracle
(QGN_Module.f90:258-296, ARK43_QGLeith.f90).

  expstrat_1024x16  FP64 RHS per layer <= 1e-12; mixed RHS per layer <= 1e-5;
                    3 FP64 ARK43_QGLeith steps (state <= 1e-10, t and dt
                    <= 1e-8); 3 mixed steps, the oracle taking the dt
                    sequence the mixed controller chose (state <= 1e-5,
                    the FP32 error norm within 3e-2 of the oracle's).
  expstrat_2048x32  mixed RHS per layer <= 1e-5 and FP64 RHS <= 1e-12.

The state is the workload's seeded initial PV with the spectrum broadened
to every kept mode (peak at k = 16 plus a k^-1 tail out to the cutoff), so
every bin of the pruned transforms carries data.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(name, tail=True):
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import configs, initial
    from qgn_fortran_amd.model import vertical_modes
    w = configs.WORKLOADS[name]()
    p = w.params
    Hs, _L, Modes, EVals, qy = vertical_modes(w.H, w.S, w.uBar, p.Htot)
    po = orc.Params(nx=p.nx, ny=p.ny, nz=p.nz, Lx=p.Lx, Ly=p.Ly, Htot=p.Htot, beta=p.beta, A2=p.A2,
                    A8=p.A8, r0=p.r0, C_d=p.C_d, TOL=p.TOL, is_explicit=p.is_explicit)
    om = orc.Model(po, Hs, Modes, EVals, w.uBar, qy, workers=16)
    q0 = initial.random_phase_pv(p.nx, p.ny, p.nz, w.k_peak, w.rms, w.seed, p.Lx, p.Ly)
    qh = om.initialize(q0)
    if tail:
        # a k^-1 tail of random phases on every kept mode, 10% of the RMS
        rng = np.random.default_rng(w.seed + 1)
        k = np.sqrt(om.k2) / (2 * np.pi / p.Lx)
        amp = np.where(k > 0, 1.0 / np.maximum(k, 1.0), 0.0)
        extra = orc.dealias(amp * np.exp(2j * np.pi * rng.random(qh.shape)), p.nx, p.ny)
        qh = qh + 0.1 * np.linalg.norm(qh) / np.linalg.norm(extra) * extra
    return w, om, qh, dict(H=Hs, Modes=Modes, EVals=EVals, uBar=w.uBar, qyBar=qy)


def _gpu(w, vs, **over):
    import dataclasses
    from qgn_fortran_amd import QGN
    return QGN(dataclasses.replace(w.params, **over), vs["H"], vs["Modes"], vs["EVals"], vs["uBar"], vs["qyBar"])


def test_expstrat_1024x16_rhs_and_steps():
    from oracle import qgn_oracle as orc
    print("1024x16: oracle RHS", flush=True)
    w, om, qh, vs = _setup("expstrat_1024x16")
    ref = om.get_rhs(qh)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        print(f"1024x16: {prec} RHS", flush=True)
        g = _gpu(w, vs, precision=prec)
        err = rel_l2(g.get_rhs(qh), ref, axis=(1, 2))
        print(f"  per-layer rel L2 max {err.max():.2e}", flush=True)
        assert np.all(err < tol), err
        g.close()
    del ref

    # three FP64 ARK43_QGLeith steps against the oracle
    print("1024x16: 3 FP64 ARK43_QGLeith steps", flush=True)
    g = _gpu(w, vs, precision="fp64")
    g.set_state(qh)
    st = orc.ARK43(om, leith=True)
    q, t, dt = qh, 0.0, w.dt0
    tg, dtg = 0.0, w.dt0
    for n in range(3):
        q, t, dt = st.step(q, t, dt)
        tg, dtg = g.time_step(tg, dtg)
        err = rel_l2(g.get_state(), q)
        print(f"  step {n + 1}: t {tg:.6g} / {t:.6g}, dt {dtg:.6g} / {dt:.6g}, state rel L2 {err:.2e}", flush=True)
        assert err < 1e-10
        assert abs(tg - t) <= 1e-8 * t and abs(dtg - dt) <= 1e-8 * dt
    g.close()

    # three mixed steps; the oracle takes the same dt sequence (it accepts
    # every step at an unreachable tolerance)
    print("1024x16: 3 mixed ARK43_QGLeith steps", flush=True)
    g = _gpu(w, vs, precision="mixed")
    g.set_state(qh)
    om.p.TOL = 1e30
    st = orc.ARK43(om, leith=True)
    q, t = qh, 0.0
    tg, dtg = 0.0, w.dt0
    for n in range(3):
        t0 = tg
        tg, dtg = g.time_step(tg, dtg)
        q, t, _ = st.step(q, t, tg - t0)
        err = rel_l2(g.get_state(), q)
        # the mixed error norm (FP32 column and row c2r of the error field)
        # against the oracle's FP64 one for the same step
        e_gpu, e_orc = g.stats()["last_err"], st.log[-1][1]
        e_rel = abs(e_gpu - e_orc) / e_orc
        print(f"  step {n + 1}: dt {tg - t0:.6g}, state rel L2 vs oracle {err:.2e}, err1 {e_gpu:.6e} / {e_orc:.6e} "
              f"(rel {e_rel:.1e})", flush=True)
        assert err < 1e-5
        assert e_rel < 3e-2   # measured 7.7e-3, 5.8e-5, 3.2e-4
    assert abs(t - tg) <= 1e-12 * tg
    g.close()


def test_expstrat_2048x32_rhs():
    print("2048x32: oracle RHS", flush=True)
    w, om, qh, vs = _setup("expstrat_2048x32")
    ref = om.get_rhs(qh)
    for prec, tol in (("mixed", 1e-5), ("fp64", 1e-12)):
        print(f"2048x32: {prec} RHS", flush=True)
        g = _gpu(w, vs, precision=prec)
        err = rel_l2(g.get_rhs(qh), ref, axis=(1, 2))
        print(f"  per-layer rel L2 max {err.max():.2e}", flush=True)
        assert np.all(err < tol), err
        g.close()
