"""Batches and ensembles: members stacked along z with their own vertical
structure, Ekman drag and dt, stepped by device-controlled ARK43 attempts in
a HIP graph.  Each member is checked against the same model run alone
through qgn_time_step (whose parity with the oracle the other GPU tests
establish), and one member directly against the oracle."""
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _single(p, m):
    import dataclasses
    from qgn_fortran_amd import QGN
    return QGN(dataclasses.replace(p, r0=m["r0"]), m["H"], m["Modes"], m["EVals"], m["uBar"], m["qyBar"])


def _check_members(batch, group, q0_hat, which, tol_state, tol_t, same_tiles=False, exact=False):
    """Run members `which` alone for as many accepted steps as the batch took
    (same_tiles: with the batch's FFT tile choices, which on the mixed path
    decide the FP32 rounding; exact: the states and times must agree to the
    bit)."""
    info = batch.times()
    state = batch.get_state()
    nz = group.params.nz
    for m in which:
        steps = int(info["steps"][m])
        assert info["status"][m] == 0
        single = _single(group.params, group.members[m])
        if same_tiles:
            tb = batch.tuning()
            single.set_tuning(tb["col_inv"], tb["col_rhs"], tb["row"])
        single.set_state(q0_hat[m * nz:(m + 1) * nz])
        t, dt = 0.0, group.dt0
        for _ in range(steps):
            t, dt = single.time_step(t, dt)
        err = rel_l2(state[m * nz:(m + 1) * nz], single.get_state())
        print(f"member {m}: {steps} steps, {int(info['rejects'][m])} rejects, t {info['t'][m]:.6e}, "
              f"state rel L2 {err:.2e}", flush=True)
        assert steps >= 1
        assert abs(info["t"][m] - t) <= tol_t * t
        assert err < tol_state
        if exact:
            assert np.array_equal(state[m * nz:(m + 1) * nz], single.get_state()) and info["t"][m] == t
        single.close()
    return info


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_members_match_single_runs(precision):
    """4 members at 96^2 x 4 with different deformation radii, Ekman drag
    and initial states, 10 attempts from dt = 60000 s (rejections occur).
    The batch's graph attempts (candidate and QG-Leith partial sums from the
    update kernel, state alternating between two buffers) and the single
    runs' qgn_time_step (a Leith pass over q per step) agree to the bit;
    the mixed single runs take the batch's FFT tiles."""
    from qgn_fortran_amd import Batch, configs
    print("batch", precision, flush=True)
    group = configs.ensemble96(sizes=(96,), per_size=4, nz=4, precision=precision, seed=11)[0]
    group.params.TOL = 1e-11
    group.dt0 = 60000.0
    group.q0 *= 10.0
    b = Batch(group.params, group.members)
    q0_hat = b.initialize(group.q0)
    b.set_time(0.0, group.dt0)
    b.step(10)
    tol = 1e-10 if precision == "fp64" else 1e-6
    info = _check_members(b, group, q0_hat, range(4), tol, 1e-12 if precision == "fp64" else 1e-6,
                          same_tiles=precision == "mixed", exact=True)
    assert info["rejects"].sum() > 0       # the reject branch ran
    if precision == "fp64":
        # member 2 against the oracle directly
        from oracle import qgn_oracle as orc
        p, m = group.params, group.members[2]
        po = orc.Params(nx=p.nx, ny=p.ny, nz=p.nz, Lx=p.Lx, Ly=p.Ly, Htot=p.Htot, beta=p.beta, r0=m["r0"],
                        C_d=p.C_d, TOL=p.TOL)
        om = orc.Model(po, m["H"], m["Modes"], m["EVals"], m["uBar"], m["qyBar"])
        st = orc.ARK43(om, leith=True)
        q, t, dt = q0_hat[8:12].copy(), 0.0, group.dt0
        for _ in range(int(info["steps"][2])):
            q, t, dt = st.step(q, t, dt)
        err = rel_l2(b.get_state()[8:12], q)
        print("member 2 vs oracle:", err, flush=True)
        assert err < 1e-10 and abs(info["t"][2] - t) <= 1e-12 * t
    b.close()


def test_batch_set_time_restarts_controller():
    """A second set_time resets counters and the controller memory: the same
    attempts from the same state give the same result."""
    from qgn_fortran_amd import Batch, configs
    group = configs.ensemble96(sizes=(64,), per_size=2, nz=3, seed=5)[0]
    b = Batch(group.params, group.members)
    q0_hat = b.initialize(group.q0)
    out = []
    for _ in range(2):
        b.set_state(q0_hat)
        b.set_time([0.0, 100.0], [750.0, 500.0])
        b.step(4)
        info = b.times()
        out.append((b.get_state(), info))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1]["t"], out[1][1]["t"]) and np.array_equal(out[0][1]["steps"], out[1][1]["steps"])
    assert out[0][1]["t"][1] > 100.0
    b.close()


def test_time_step_then_batch_step_on_one_member():
    """qgn_time_step and qgn_batch_step on the same one-member batch: the
    graph reads the state wherever qgn_time_step's buffer swaps left it, so
    the attempts that follow match a fresh batch started from that state."""
    from qgn_fortran_amd import Batch, configs
    group = configs.ensemble96(sizes=(64,), per_size=1, nz=3, seed=9)[0]
    b = Batch(group.params, group.members)
    b.initialize(group.q0)
    b.set_time(0.0, 750.0)
    b.step(1)                                   # captures the graph
    info = b.times()
    t, dt = b.time_step(float(info["t"][0]), float(info["dt"][0]))   # swaps the state buffers
    s2 = b.get_state()
    b.set_time(t, dt)
    b.step(0)
    assert np.array_equal(b.get_state(), s2)
    b.step(3)
    fresh = Batch(group.params, group.members)
    fresh.set_state(s2)
    fresh.set_time(t, dt)
    fresh.step(3)
    assert np.array_equal(b.get_state(), fresh.get_state())
    assert np.array_equal(b.times()["t"], fresh.times()["t"])
    b.close()
    fresh.close()


def _oracle_replay(group, m, q0_hat, dts):
    """Member m of a group through the oracle's ARK43_QGLeith
    (ARK43_QGLeith.f90) with the accepted step sizes `dts` (an unreachable
    tolerance, so every step is accepted)."""
    from oracle import qgn_oracle as orc
    p, mm = group.params, group.members[m]
    po = orc.Params(nx=p.nx, ny=p.ny, nz=p.nz, Lx=p.Lx, Ly=p.Ly, Htot=p.Htot, beta=p.beta, r0=mm["r0"],
                    C_d=p.C_d, TOL=1e30)
    om = orc.Model(po, mm["H"], mm["Modes"], mm["EVals"], mm["uBar"], mm["qyBar"])
    st = orc.ARK43(om, leith=True)
    nz = p.nz
    q, t = q0_hat[m * nz:(m + 1) * nz].copy(), 0.0
    for dt in dts:
        q, t, _ = st.step(q, t, dt)
    return q, t


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble96(precision):
    """The 96-member ensemble (32 members of 8 layers on each of 128^2,
    256^2, 384^2; mixed batches pick their FFT tiles with the autotuner at
    creation).  The first 3 attempts run one at a time: member 7 of every
    size is then replayed through the oracle with the step sizes it took
    (state <= 1e-10 FP64, <= 1e-5 mixed).  Then 20 attempts run as chained
    launches (8 attempts per graph launch, the rest single) and member 7
    of each size is checked against its own single run (mixed: with the
    batch's FFT tiles, so with the same FP32 rounding)."""
    from qgn_fortran_amd import Batch, Ensemble, configs
    print("ensemble96", precision, flush=True)
    groups = configs.ensemble96(precision=precision)
    batches, q0s = [], []
    for g in groups:
        b = Batch(g.params, g.members)
        q0s.append(b.initialize(g.q0))
        b.set_time(0.0, g.dt0)
        batches.append(b)
    ens = Ensemble(batches)
    m = 7
    dts = [[] for _ in groups]
    t_prev = [0.0 for _ in groups]
    for _ in range(3):
        ens.step(1)
        for k, b in enumerate(batches):
            t = float(b.times()["t"][m])
            if t != t_prev[k]:
                dts[k].append(t - t_prev[k])
                t_prev[k] = t
    tol = 1e-10 if precision == "fp64" else 1e-5
    for k, (g, b) in enumerate(zip(groups, batches)):
        q, t = _oracle_replay(g, m, q0s[k], dts[k])
        err = rel_l2(b.get_state()[m * g.params.nz:(m + 1) * g.params.nz], q)
        print(f"  {g.params.nx}^2 member {m}: {len(dts[k])} steps vs oracle, state rel L2 {err:.2e}", flush=True)
        assert len(dts[k]) >= 1 and err < tol and abs(t - t_prev[k]) <= 1e-12 * t
    t0 = time.perf_counter()
    attempts = 20
    ens.step(attempts)
    infos = [b.times() for b in batches]
    wall = time.perf_counter() - t0
    steps = sum(int(i["steps"].sum()) for i in infos)
    days = sum(float(i["t"].sum()) for i in infos) / 86400.0
    print(f"ensemble96 {precision}: {attempts + 3} attempts, {steps} accepted member-steps, "
          f"{attempts} attempts in {wall * 1e3:.1f} ms ({attempts / wall:.1f} attempts/s)", flush=True)
    for g, b, q0, i in zip(groups, batches, q0s, infos):
        assert np.all(i["status"] == 0) and np.all(i["steps"] >= 1)
        if precision == "fp64":
            _check_members(b, g, q0, [m], 1e-10, 1e-12)
        else:
            _check_members(b, g, q0, [m], 1e-12, 1e-12, same_tiles=True)
    assert days > 0
    ens.close()
    for b in batches:
        b.close()


def test_controller_error_paths():
    """The device step control's failure reports: a NaN in the state gives
    QGN_ENONFINITE (the reference would accept the step) and a dt that keeps
    being rejected gives QGN_EREJECT after max_rejects attempts, for
    qgn_time_step and, per member, for a batch (the other member goes on)."""
    import dataclasses
    from qgn_fortran_amd import Batch, configs
    from qgn_fortran_amd._lib import QGN_ENONFINITE, QGN_EREJECT, QgnError
    group = configs.ensemble96(sizes=(64,), per_size=2, nz=3, seed=4)[0]
    p = dataclasses.replace(group.params, max_rejects=2, TOL=1e-14)
    b = Batch(p, group.members)
    q0 = b.initialize(group.q0)
    bad = q0.copy()
    bad[1, 3, 2] = np.nan                       # member 0, layer 1
    one = Batch(p, group.members[:1])
    one.set_state(bad[:3])
    with pytest.raises(QgnError) as e:
        one.time_step(0.0, 750.0)
    assert e.value.code == QGN_ENONFINITE
    one.set_state(q0[:3])
    with pytest.raises(QgnError) as e:
        one.time_step(0.0, 1.0e7)               # TOL 1e-14 and dt 1e7 s: every attempt is rejected
    assert e.value.code == QGN_EREJECT
    b.set_state(bad)
    b.set_time(0.0, 750.0)
    b.step(3)
    with pytest.raises(QgnError) as e:
        b.times()
    assert e.value.code == QGN_ENONFINITE
    st = np.empty(2, dtype=np.int32)
    import ctypes as C
    b.lib.qgn_batch_get_time(b._ctx, None, None, None, None, st.ctypes.data_as(C.POINTER(C.c_int)))
    assert st[0] == 1 and st[1] == 0
    one.close()
    b.close()


def test_closure_in_batch_graph_matches_single_runs():
    """The closure network inside the graph-captured batch attempt: each member
    of a 2-member batch steps exactly as the same model alone with the same
    network (qgn_time_step); an ensemble whose batches share one network is
    refused."""
    from oracle import cnn_oracle as co
    from qgn_fortran_amd import Batch, Ensemble, configs
    from qgn_fortran_amd._lib import QGN_EINVAL, QgnError
    from qgn_fortran_amd.cnn import CNN
    group = configs.ensemble96(sizes=(64,), per_size=2, nz=8, seed=7, scheme="ARK43")[0]
    group.q0 *= 10.0
    net = CNN(16, 8, co.init_params(16, 8, seed=2, last_scale=0.05))
    s_in = np.full(16, 1.0)
    s_in[:8] = 1.0 / np.std(group.q0)
    s_in[8:] = 1.0 / 3.0e3
    s_out = np.full(8, 1e-12)
    b = Batch(group.params, group.members)
    q0 = b.initialize(group.q0)
    b.set_closure(net, s_in, s_out)
    b.set_time(0.0, 2000.0)
    b.step(4)
    info = b.times()
    state = b.get_state()
    for m in range(2):
        single = Batch(group.params, group.members[m:m + 1])
        single.set_closure(net, s_in, s_out)
        single.set_state(q0[8 * m:8 * (m + 1)])
        t, dt = 0.0, 2000.0
        for _ in range(int(info["steps"][m])):
            t, dt = single.time_step(t, dt)
        err = rel_l2(state[8 * m:8 * (m + 1)], single.get_state())
        print(f"member {m}: {int(info['steps'][m])} steps, state rel L2 {err:.2e}", flush=True)
        assert err == 0.0 and t == info["t"][m]
        single.set_closure(None)
        single.close()
    other = Batch(group.params, group.members)
    other.set_closure(net, s_in, s_out)
    other.set_time(0.0, 2000.0)
    with pytest.raises(QgnError) as e:
        Ensemble([b, other])
    assert e.value.code == QGN_EINVAL
    for x in (b, other):
        x.set_closure(None)
        x.close()
    net.close()
