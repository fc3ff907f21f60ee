import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libqgn.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def make_case(nx, ny, nz, *, beta=1.6e-11, Lx=1.0e6, Ly=None, r0=1e-7, C_d=1e-5, A2=0.0, A8=0.0,
              is_explicit=False, TOL=1e-9, seed=0, rms=1e-6, k_peak=6, scheme="ARK43_QGLeith",
              precision="fp64", profile="exp"):
    """Build matching (package Params, oracle Params, vertical structure, IC)."""
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd import background, initial
    from qgn_fortran_amd.model import Params

    Ly = Lx if Ly is None else Ly
    if profile == "exp" and nz > 1:
        H, S, u = background.exponential(nz, 4000.0, 1e-4, 5e-3, 800.0, 0.1)
        Htot = 4000.0
    else:
        H, S, u = background.two_layer(1000.0, 3000.0, 25e3, 0.05, 0.0)
        Htot = 4000.0
    Hs, L, Modes, EVals, qy = orc.vertical_modes(H, S, u, Htot)
    pp = Params(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Htot=Htot, beta=beta, A2=A2, A8=A8, r0=r0, C_d=C_d,
                TOL=TOL, is_explicit=is_explicit, scheme=scheme, precision=precision)
    po = orc.Params(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Htot=Htot, beta=beta, A2=A2, A8=A8, r0=r0, C_d=C_d,
                    TOL=TOL, is_explicit=is_explicit)
    q0 = initial.random_phase_pv(nx, ny, nz, k_peak, rms, seed, Lx, Ly)
    return dict(pp=pp, po=po, H=Hs, Modes=Modes, EVals=EVals, uBar=u, qyBar=qy, q0=q0)


def rel_l2(a, b, axis=None):
    num = np.sqrt(np.sum(np.abs(a - b) ** 2, axis=axis))
    den = np.sqrt(np.sum(np.abs(b) ** 2, axis=axis))
    return num / np.where(den == 0, 1.0, den)
