"""The C++/OpenMP restatement (oracle/qgn_cpu.cpp, the CPU baseline, FFTs by
MKL DFTI as in the reference) against the numpy oracle: GetRHS per layer
<= 1e-12 (QGN_Module.f90:258-296), ARK43 and ARK43_QGLeith trajectories
with rejections <= 1e-10, t and dt <= 1e-8.  Skipped where the library is
not built or MKL does not load."""
import numpy as np
import pytest

from conftest import make_case, rel_l2


def _models(case, **kw):
    from oracle import qgn_cpu
    from oracle import qgn_oracle as orc
    ok, why = qgn_cpu.available()
    if not ok:
        pytest.skip(why)
    om = orc.Model(case["po"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])
    try:
        cm = qgn_cpu.Model(case["po"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"], **kw)
    except RuntimeError as e:
        pytest.skip(str(e))
    return om, cm


@pytest.mark.parametrize("kw", [dict(nx=96, ny=96, nz=4, C_d=1e-5),
                                dict(nx=128, ny=128, nz=2, profile="two", A8=1e23, is_explicit=True, C_d=0.0),
                                dict(nx=192, ny=128, nz=3, C_d=1e-5, Ly=0.7e6)],
                         ids=lambda k: f"{k['nx']}x{k['ny']}x{k['nz']}")
def test_rhs_matches_oracle(kw):
    case = make_case(**kw)
    om, cm = _models(case, threads=4)
    qh = om.initialize(case["q0"])
    err = rel_l2(cm.get_rhs(qh), om.get_rhs(qh), axis=(1, 2))
    assert np.all(err < 1e-12), err
    cm.close()


@pytest.mark.parametrize("leith", [True, False], ids=["ARK43_QGLeith", "ARK43"])
def test_ark43_trajectory_matches_oracle(leith):
    from oracle import qgn_oracle as orc
    case = make_case(96, 96, 4, C_d=1e-5, A8=0.0 if leith else 2e20, TOL=1e-11, rms=1e-5)
    om, cm = _models(case, threads=4)
    q = om.initialize(case["q0"])
    st = orc.ARK43(om, leith=leith)
    qc, t, dt = q.copy(), 0.0, 60000.0
    tc, dtc = 0.0, 60000.0
    for _ in range(4):
        q, t, dt = st.step(q, t, dt)
        qc, tc, dtc = cm.ark43_step(qc, tc, dtc, leith=leith)
        assert rel_l2(qc, q) < 1e-10
        assert abs(tc - t) <= 1e-8 * t and abs(dtc - dt) <= 1e-8 * dt
    assert st.rejects > 0 and cm.rejects == st.rejects
    cm.close()
