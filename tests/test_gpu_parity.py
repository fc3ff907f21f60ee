"""GPU parity: the HIP library through its C-ABI against the FP64 oracle on the
same seeded inputs.  Tolerances: FP64 path per-layer relative L2 <= 1e-12 for
one RHS and for transforms; trajectories <= 1e-10 after a few steps; mixed
precision RHS relative L2 <= 1e-5 against FP64."""
import numpy as np
import pytest

from conftest import make_case, rel_l2

pytestmark = pytest.mark.gpu


def _gpu_model(case):
    from qgn_fortran_amd import QGN
    return QGN(case["pp"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])


def _oracle_model(case):
    from oracle import qgn_oracle as orc
    return orc.Model(case["po"], case["H"], case["Modes"], case["EVals"], case["uBar"], case["qyBar"])


@pytest.mark.parametrize("nx,ny,howmany", [(24, 18, 1), (96, 96, 3), (128, 128, 2), (384, 384, 2),
                                            (1024, 1024, 2), (160, 120, 2), (2048, 64, 1)])
def test_dft_forward_backward(nx, ny, howmany):
    """DFTI-equivalent r2c / c2r (MKL_DFTI_Setup.f90) vs pocketfft; backward
    on non-Hermitian input follows the real-part rule for kx = 0, nx/2."""
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import Dfti
    print(f"dft {nx}x{ny}x{howmany}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny)
    g = rng.standard_normal((howmany, ny, nx))
    d = Dfti(nx, ny, howmany)
    s = d.forward(g)
    ref = orc.rfft2(g)
    assert rel_l2(s, ref) < 1e-13
    spec = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    back = d.backward(spec)
    refb = orc.irfft2(spec, nx)
    assert rel_l2(back, refb) < 1e-13
    d.close()


RHS_CASES = [
    dict(nx=128, ny=128, nz=2, profile="two", A8=1e23, is_explicit=True, C_d=0.0, k_peak=8),
    dict(nx=96, ny=96, nz=4, C_d=1e-5),
    dict(nx=192, ny=128, nz=3, C_d=1e-5, Ly=0.7e6),
    dict(nx=384, ny=384, nz=8, C_d=1e-5, k_peak=16),
    dict(nx=256, ny=256, nz=16, C_d=0.0, k_peak=16),
    # sizes outside the compiled transform lengths: the generic path
    dict(nx=576, ny=576, nz=4, C_d=1e-5, k_peak=12),
    dict(nx=320, ny=320, nz=3, C_d=1e-5),
    dict(nx=160, ny=90, nz=2, profile="two", A8=1e21, is_explicit=True, C_d=1e-5, Ly=0.6e6),
]


@pytest.mark.parametrize("kw", RHS_CASES, ids=lambda k: f"{k['nx']}x{k['ny']}x{k['nz']}")
def test_rhs_fp64(kw):
    """GetRHS (QGN_Module.f90:258-296): per-layer relative L2 <= 1e-12."""
    print("rhs", kw, flush=True)
    case = make_case(**kw)
    om = _oracle_model(case)
    gm = _gpu_model(case)
    q_hat = om.initialize(case["q0"])
    ref = om.get_rhs(q_hat)
    out = gm.get_rhs(q_hat)
    err = rel_l2(out, ref, axis=(1, 2))
    print("per-layer rel L2", err)
    assert np.all(err < 1e-12)
    psi = gm.get_psi(q_hat)
    assert np.all(rel_l2(psi, om.get_psi(q_hat), axis=(1, 2)) < 1e-13)
    gm.close()


def test_initialize_and_fields():
    """Initialize (QGN_Module.f90:177-183) and WriteQP's fields (:353-361)."""
    case = make_case(192, 192, 4, k_peak=10)
    om = _oracle_model(case)
    gm = _gpu_model(case)
    qh = gm.initialize(case["q0"])
    ref = om.initialize(case["q0"])
    assert rel_l2(qh, ref) < 1e-13
    q, psi = gm.fields()
    qr, pr = om.get_psi_phys(ref)
    assert rel_l2(q, qr) < 1e-13
    assert rel_l2(psi, pr) < 1e-13
    st = gm.get_state()
    assert np.array_equal(st, qh)
    gm.close()


def test_leith_coefficients():
    """A8L (ARK43_QGLeith.f90:62-73): Parseval on the device vs physical-space
    sums in the oracle, including the non-Hermitian kx=0 column at n=128."""
    case = make_case(128, 128, 4, k_peak=12)
    om = _oracle_model(case)
    gm = _gpu_model(case)
    qh = om.initialize(case["q0"])
    gm.set_state(qh)
    a = gm.leith()
    r = om.leith_a8(qh)
    assert np.max(np.abs(a / r - 1)) < 1e-12
    gm.close()


@pytest.mark.parametrize("nx,ny,nz", [(128, 128, 2), (192, 128, 4)])
def test_modal_fields(nx, ny, nz):
    """WriteQPmodal's fields (QGN_Module.f90:381-416): modal q, psi and the
    full (not dealiased) flux-form Jacobian, relative L2 <= 1e-12."""
    print("modal", nx, ny, nz, flush=True)
    kw = dict(profile="two") if nz == 2 else dict(Ly=0.7e6)
    case = make_case(nx, ny, nz, k_peak=10, **kw)
    om = _oracle_model(case)
    gm = _gpu_model(case)
    qh = om.initialize(case["q0"])
    gm.set_state(qh)
    got = gm.modal_fields()
    ref = om.modal_fields(qh)
    for name, a, b in zip(("q_mode", "psi_mode", "J_mode"), got, ref):
        err = rel_l2(a, b, axis=(1, 2))
        print(name, err)
        assert np.all(err < 1e-12), name
    # a second call reuses the work arrays
    again = gm.modal_fields()
    assert all(np.array_equal(a, b) for a, b in zip(got, again))
    gm.close()


@pytest.mark.parametrize("nx,ny,nz", [(128, 128, 3), (192, 128, 4)])
def test_spectra(nx, ny, nz):
    """KE and enstrophy shell spectra (the diagnostics hook of
    QGN_Driver.f90:52-55) against the oracle's full-plane FFT of the
    physical fields, relative L2 per layer <= 1e-12; n = 128 keeps a
    non-Hermitian kx = 0 column."""
    from qgn_fortran_amd.model import spectra_bins
    print("spectra", nx, ny, nz, flush=True)
    case = make_case(nx, ny, nz, k_peak=12, Ly=(1e6 if nx == ny else 0.7e6))
    om = _oracle_model(case)
    gm = _gpu_model(case)
    qh = om.initialize(case["q0"])
    gm.set_state(qh)
    nb = spectra_bins(case["pp"])
    ke, ens = gm.spectra()
    rke, rens = om.spectra(qh, nb)
    assert ke.shape == (nz, nb)
    assert np.all(rel_l2(ke, rke, axis=1) < 1e-12)
    assert np.all(rel_l2(ens, rens, axis=1) < 1e-12)
    ke2, _ = gm.spectra(nbins=7)        # truncated: the first bins only
    assert np.allclose(ke2, rke[:, :7], rtol=1e-12, atol=0)
    gm.close()


STEP_CASES = [
    # (scheme, case, dt0, steps); adaptive cases start above the stable dt so
    # rejections occur and err1 sits near TOL (not at round-off level)
    ("RK4", dict(nx=64, ny=64, nz=2, profile="two", A8=1e23, is_explicit=True, C_d=0.0), 1800.0, 6),
    ("AB3", dict(nx=64, ny=64, nz=2, profile="two", A8=1e23, is_explicit=True, C_d=0.0), 1800.0, 6),
    ("BSRK32", dict(nx=96, ny=96, nz=3, C_d=1e-5, TOL=1e-11, rms=1e-5), 3000.0, 6),
    ("ARK43", dict(nx=96, ny=96, nz=3, C_d=1e-5, A8=1e22, TOL=1e-11, rms=1e-5), 8000.0, 6),
    ("ARK43_QGLeith", dict(nx=96, ny=96, nz=4, C_d=1e-5, TOL=1e-11, rms=1e-5), 8000.0, 6),
    ("ARK43_QGLeith", dict(nx=128, ny=128, nz=4, C_d=1e-5, TOL=1e-12, rms=1e-5), 20000.0, 6),
    # generic-path sizes (576 = 3 x 192, as parameters.f90:11 recommends; 320 = 2^6 x 5)
    ("ARK43_QGLeith", dict(nx=576, ny=576, nz=4, C_d=1e-5, TOL=1e-11, rms=1e-5), 8000.0, 4),
    ("BSRK32", dict(nx=320, ny=320, nz=3, C_d=1e-5, TOL=1e-11, rms=1e-5), 3000.0, 4),
]


@pytest.mark.parametrize("scheme,kw,dt0,nsteps", STEP_CASES, ids=lambda x: str(x)[:20])
def test_stepper_trajectory(scheme, kw, dt0, nsteps):
    """TimeStep with each include fragment: state, t and dt after each step
    agree with the oracle (accept/reject sequence included)."""
    from oracle import qgn_oracle as orc
    print("step", scheme, kw, flush=True)
    case = make_case(scheme=scheme, **kw)
    om = _oracle_model(case)
    gm = _gpu_model(case)
    stepper = orc.SCHEMES[scheme](om)
    q = om.initialize(case["q0"])
    gm.set_state(q)
    t_o = t_g = 0.0
    dt_o = dt_g = dt0
    for n in range(nsteps):
        q, t_o, dt_o = stepper.step(q, t_o, dt_o)
        t_g, dt_g = gm.time_step(t_g, dt_g)
        qg = gm.get_state()
        e = rel_l2(qg, q)
        assert e < 1e-10, (n, e)
        # the embedded error is a cancellation of nearly equal stage sums, so
        # its max-norm carries round-off ~1e-10 relative into the controller
        assert abs(t_g - t_o) <= 1e-8 * max(1.0, abs(t_o)), (n, t_g, t_o)
        assert abs(dt_g - dt_o) <= 1e-8 * dt_o, (n, dt_g, dt_o)
    rej = getattr(stepper, "rejects", 0)
    assert gm.stats()["rejects"] == rej
    gm.close()


@pytest.mark.parametrize("n,nz", [(256, 8), (128, 24), (128, 32)])
def test_mixed_precision_rhs(n, nz):
    """FP32 FFT/Jacobian path with the split-BF16 MFMA projection (16-layer
    tiles; nz = 24 leaves a partial tile): RHS relative L2 vs the FP64 RHS
    <= 1e-5."""
    case = make_case(n, n, nz, k_peak=16, C_d=1e-5)
    om = _oracle_model(case)
    q_hat = om.initialize(case["q0"])
    ref = om.get_rhs(q_hat)
    case["pp"].precision = "mixed"
    gm = _gpu_model(case)
    out = gm.get_rhs(q_hat)
    err = rel_l2(out, ref, axis=(1, 2))
    print("mixed per-layer rel L2", err)
    assert np.all(err < 1e-5)
    gm.close()


@pytest.mark.parametrize("nx,ny,nz", [(576, 576, 4), (320, 320, 3), (1152, 1152, 4)])
def test_generic_size_mixed_rhs(nx, ny, nz):
    """The generic path in FP32 (run-time-plan transforms, MFMA projection):
    RHS relative L2 per layer vs the FP64 oracle <= 1e-5."""
    case = make_case(nx, ny, nz, k_peak=16, C_d=1e-5)
    om = _oracle_model(case)
    q_hat = om.initialize(case["q0"])
    ref = om.get_rhs(q_hat)
    case["pp"].precision = "mixed"
    gm = _gpu_model(case)
    err = rel_l2(gm.get_rhs(q_hat), ref, axis=(1, 2))
    print(f"generic {nx}x{ny}x{nz} mixed per-layer rel L2 max {err.max():.2e}", flush=True)
    assert np.all(err < 1e-5)
    gm.close()


def test_autotuned_launches_and_every_variant():
    """The autotuner's choices are valid tiles/groups with measured times, and
    every compiled variant of the mixed path's column tiles and row groups
    gives the same RHS (relative L2 <= 1e-5 against FP64, as the fixed rule)."""
    case = make_case(256, 256, 8, k_peak=16, C_d=1e-5)
    om = _oracle_model(case)
    q_hat = om.initialize(case["q0"])
    ref = om.get_rhs(q_hat)
    case["pp"].precision = "mixed"
    gm = _gpu_model(case)
    t = gm.tuning()
    print("tuning", t, flush=True)
    assert t["col_inv"] in (2, 4, 8, 16) and t["col_rhs"] in (2, 4, 8, 16) and t["row"] in (2, 4, 8)
    assert np.all(t["times_us"][0][t["times_us"][0] > 0] > 0)
    base = gm.get_rhs(q_hat)
    assert np.all(rel_l2(base, ref, axis=(1, 2)) < 1e-5)
    for ci, cr, rw in [(2, 16, 2), (4, 8, 8), (16, 2, 4), (8, 4, 2)]:
        gm.set_tuning(ci, cr, rw)
        out = gm.get_rhs(q_hat)
        err = rel_l2(out, ref, axis=(1, 2))
        print(ci, cr, rw, "max rel L2", err.max(), flush=True)
        assert np.all(err < 1e-5)
        assert rel_l2(out, base) < 1e-6
    gm.close()
