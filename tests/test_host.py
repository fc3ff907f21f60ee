"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/qgn.h declares, the library's eigensolver (the
DGEEV replacement) reproduces the notebook's stored radii and the oracle's
modes, file formats round-trip, and the Fortran host compiles."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    from qgn_fortran_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "qgn.h")).read()
    names = set(re.findall(r"\b(qgn_[a-z0-9_]+)\s*\(", hdr))
    assert names == set(_lib.EXPORTS)
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    assert b"gfx950" in lib.qgn_version()


def test_vertical_modes_library_vs_notebook_and_oracle():
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import background, vertical_modes
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "notebook_radii.json")))
    H, S, u = background.notebook_profile(gold["nz"], gold["alpha"])
    Hs, L, Modes, EVals, qy = vertical_modes(H, S, u, gold["H_tot"])
    r = np.sort(background.deformation_radii(EVals[EVals < -1e-12])) / 1e3
    g = np.array([v for v in gold["radii_sorted"] if v is not None])
    assert np.max(np.abs(r / g - 1)) < 2e-8
    Ho, Lo, Mo, Eo, qyo = orc.vertical_modes(H, S, u, gold["H_tot"])
    assert np.allclose(Hs, Ho, rtol=1e-15) and np.allclose(L, Lo, rtol=1e-15, atol=0)
    # qyBar = -L uBar is a cancellation of O(|L| |u|) terms
    assert np.abs(qy - qyo).max() < 1e-13 * np.abs(L).max() * np.abs(u).max()
    # same eigenpairs up to ordering: match by eigenvalue
    for k in range(len(EVals)):
        if EVals[k] > -1e-12:
            continue
        m = np.argmin(np.abs(Eo - EVals[k]))
        assert abs(Eo[m] / EVals[k] - 1) < 1e-10
        assert np.allclose(Modes[:, k], Mo[:, m], rtol=0, atol=1e-8 * np.abs(Mo[:, m]).max())
    G = Modes.T @ np.diag(Hs) @ Modes / gold["H_tot"]
    assert np.allclose(G, np.eye(len(Hs)), atol=1e-12)
    assert np.all(Modes[0] > 0)


def test_profile_and_restart_formats(tmp_path):
    from qgn_fortran_amd import background, initial
    H, S, u = background.exponential(8, 4000.0, 1e-4, 5e-3, 800.0, 0.1)
    background.write_profiles(tmp_path, H, S, u)
    H2, S2, u2 = background.read_profiles(tmp_path, 8)
    assert np.array_equal(H, H2) and np.array_equal(S, S2) and np.array_equal(u, u2)
    assert os.path.getsize(tmp_path / "S.dat") == 9 * 8
    q = initial.random_phase_pv(48, 32, 3, 5, 2e-6, 9)
    assert np.allclose(q.mean(axis=(1, 2)), 0, atol=1e-20)
    assert np.allclose(np.sqrt((q ** 2).mean(axis=(1, 2))), 2e-6)
    p = initial.write_restart(tmp_path, 12, q)
    assert os.path.basename(p) == "q.000000012.dat"
    assert np.array_equal(initial.read_field(p, 48, 32, 3), q)
    # Fortran order: element (i, j, k) (1-based) at offset (i-1) + nx*((j-1) + ny*(k-1))
    raw = np.fromfile(p)
    assert raw[5 + 48 * (7 + 32 * 2)] == q[2, 7, 5]


def test_random_phase_spectrum_peak():
    from qgn_fortran_amd import initial
    q = initial.random_phase_pv(128, 128, 1, 10, 1.0, 4)
    s = np.abs(np.fft.rfft2(q[0])) ** 2
    i = np.arange(65)[None, :]
    j = np.fft.fftfreq(128, 1 / 128)[:, None]
    k = np.rint(np.sqrt(i ** 2 + j ** 2)).astype(int)
    shell = np.bincount(k.ravel(), s.ravel(), minlength=100)[:60] / np.maximum(
        np.bincount(k.ravel(), minlength=100)[:60], 1)
    assert 8 <= np.argmax(shell) <= 12


def test_params_defaults():
    from qgn_fortran_amd import Params
    p = Params(nx=128, ny=128, nz=2, Lx=1e6, Ly=1e6, Htot=4000.0, beta=1.6e-11)
    assert abs(p.dt_max - np.pi / (1.6e-11 * 1e6)) < 1e-9
    c = p.to_c()
    assert c.scheme == 4 and c.precision == 0


@pytest.mark.skipif(not os.path.exists("/opt/rocm/lib/llvm/bin/flang"), reason="flang not installed")
def test_fortran_host_compiles(tmp_path):
    host = os.path.join(ROOT, "qgn_fortran_amd", "host")
    srcs = [os.path.join(host, f) for f in ("qgn_c.f90", "qgn_host.f90", "qgn_driver.f90")]
    r = subprocess.run(["/opt/rocm/lib/llvm/bin/flang", "-c", "-J", str(tmp_path), "-I", host] + srcs,
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
