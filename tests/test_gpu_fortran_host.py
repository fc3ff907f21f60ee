"""The Fortran host (qgn_fortran_amd/host) driving the HIP library end to end:
profile files in, restart file in, q/p checkpoint files out, compared with
the oracle stepping the same ARK43_QGLeith sequence (QGN_Driver.f90
semantics: restart from q.N0.dat, steps N0..N0+Nt, WriteQP and WriteQPmodal
every step, WriteSpectra at the diagnostics cadence)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qgn_fortran_amd", "host",
                    "qgn_driver")


def test_fortran_driver_restart_and_checkpoints(tmp_path):
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import background, initial
    from qgn_fortran_amd.model import vertical_modes
    if not os.path.exists(HOST):
        pytest.skip("Fortran host not built (flang missing at build time)")
    nx = ny = 96
    nz = 4
    H, S, u = background.exponential(nz, 4000.0, 1e-4, 5e-3, 800.0, 0.1)
    background.write_profiles(tmp_path, H, S, u)
    q0 = initial.random_phase_pv(nx, ny, nz, 6, 1e-5, 3, 1e6, 1e6)
    initial.write_restart(tmp_path, 2, q0)
    (tmp_path / "qgn.nml").write_text(
        "&qgn\n Lx=1.0e6, Ly=1.0e6, Htot=4000.0, nx=96, ny=96, nz=4, beta=1.6e-11, A2=0.0, A8=0.0,\n"
        " r0=1.0e-7, C_d=1.0e-5, TOL=1.0e-11, is_explicit=.false., precision=0, sync_host=.false. /\n"
        "&driver\n dt=30000.0, Nt=2, chkpt_freq=1, out_freq=1, diag_freq=2 /\n")
    print("running Fortran driver", flush=True)
    r = subprocess.run([HOST, "2"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    # oracle with the same vertical structure (H2: same EVals/Modes)
    Hs, _L, Modes, EVals, qy = vertical_modes(H, S, u, 4000.0)
    po = orc.Params(nx=nx, ny=ny, nz=nz, Lx=1e6, Ly=1e6, Htot=4000.0, beta=1.6e-11, r0=1e-7, C_d=1e-5, TOL=1e-11)
    om = orc.Model(po, Hs, Modes, EVals, u, qy)
    st = orc.ARK43(om, leith=True)
    q = om.initialize(q0)
    qp, pp = om.get_psi_phys(q)
    got = initial.read_field(tmp_path / "q.000000000.dat", nx, ny, nz)
    assert rel_l2(got, qp) < 1e-13
    t, dt = 0.0, 30000.0
    for n in (2, 3, 4):
        q, t, dt = st.step(q, t, dt)
        qp, pp = om.get_psi_phys(q)
        gq = initial.read_field(tmp_path / initial.restart_name("q", n), nx, ny, nz)
        gp = initial.read_field(tmp_path / initial.restart_name("p", n), nx, ny, nz)
        assert rel_l2(gq, qp) < 1e-10, n
        assert rel_l2(gp, pp) < 1e-10, n
        # WriteQPmodal's files (QGN_Driver.f90:59)
        for pre, ref in zip(("q_mode", "p_mode", "J_mode"), om.modal_fields(q)):
            gm = initial.read_field(tmp_path / f"{pre}.{n:09d}.dat", nx, ny, nz)
            assert rel_l2(gm, ref) < 1e-10, (pre, n)
    # WriteSpectra at the diagnostics hook (QGN_Driver.f90:52-55), step 4
    from qgn_fortran_amd.model import spectra_bins
    nb = spectra_bins(po)
    raw = np.fromfile(tmp_path / "spec.000000004.dat")
    assert raw.size == 2 * nb * nz
    ke, ens = raw[:nb * nz].reshape(nz, nb), raw[nb * nz:].reshape(nz, nb)
    rke, rens = om.spectra(q, nb)
    assert rel_l2(ke, rke) < 1e-10 and rel_l2(ens, rens) < 1e-10
    # ARK43_QGLeith prints each rejected attempt (ARK43_QGLeith.f90:126)
    assert st.rejects > 0 and r.stdout.count("Rejected step") == st.rejects
    for name in ("L.dat", "modes.dat", "Qy.dat"):
        assert (tmp_path / name).exists()
    assert np.allclose(np.fromfile(tmp_path / "Qy.dat"), qy, rtol=1e-14, atol=0)


DFTI_CHECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build", "dfti_check")


def test_fortran_dfti_module(tmp_path):
    """tests/fortran/dfti_check.f90 configures two descriptors through the
    qgn_dfti module in the sequence of MKL_DFTI_Setup.f90:11-50, transforms
    a batch forward and a non-Hermitian spectrum backward (relative L2
    <= 1e-13 against the oracle's DFTI-convention transforms), and shows
    that an in-place descriptor and a forward compute on a backward
    descriptor are refused with the library's messages."""
    from oracle import qgn_oracle as orc
    if not os.path.exists(DFTI_CHECK):
        pytest.skip("dfti_check not built (flang missing at build time)")
    nx, ny, nz = 192, 96, 3
    rng = np.random.default_rng(5)
    g = rng.standard_normal((nz, ny, nx))
    sp = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    g.tofile(tmp_path / "grid.bin")
    sp.tofile(tmp_path / "spec.bin")
    print("running dfti_check", flush=True)
    r = subprocess.run([DFTI_CHECK, str(nx), str(ny), str(nz), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    fwd = np.fromfile(tmp_path / "fwd.bin", dtype=np.complex128).reshape(nz, ny, nx // 2 + 1)
    bwd = np.fromfile(tmp_path / "bwd.bin").reshape(nz, ny, nx)
    assert rel_l2(fwd, orc.rfft2(g)) < 1e-13
    assert rel_l2(bwd, orc.irfft2(sp, nx)) < 1e-13
    assert "refused in place: DFTI: transforms are DFTI_NOT_INPLACE" in r.stdout
    assert "refused direction: DFTI: the descriptor's strides describe a backward transform" in r.stdout
