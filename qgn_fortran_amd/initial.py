"""Seeded initial conditions and the restart-file format.

QGN reads a physical PV restart ``q.%09d.dat`` (raw float64, Fortran order
(nx,ny,nz)) when started with N0 > 1 (QGN_Driver.f90:21-31); WriteQP writes
``q.%09d.dat`` / ``p.%09d.dat`` in the same format (QGN_Module.f90:364-377).
"""
from __future__ import annotations

import math
import os

import numpy as np


def random_phase_pv(nx, ny, nz, k_peak, rms, seed, Lx=1.0, Ly=1.0):
    """Random-phase PV with an isotropic amplitude spectrum
    |q_hat| ~ k^2 exp(-(k/k_peak)^2) (peak at wavenumber index k_peak), each
    layer with zero mean and RMS ``rms``.  numpy PCG64(seed); returns a
    physical field of shape (nz, ny, nx)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nxh = nx // 2 + 1
    i = np.arange(nxh)[None, :].astype(np.float64)
    j = np.arange(ny)[:, None]
    jj = np.where(j <= ny // 2, j, j - ny).astype(np.float64) * (Lx / Ly)
    kk = np.sqrt(i ** 2 + jj ** 2)
    amp = kk ** 2 * np.exp(-(kk / float(k_peak)) ** 2)
    q = np.empty((nz, ny, nx))
    for z in range(nz):
        ph = rng.uniform(0.0, 2.0 * math.pi, size=(ny, nxh))
        f = np.fft.irfft2(amp * np.exp(1j * ph), s=(ny, nx))
        f -= f.mean()
        f *= rms / math.sqrt(np.mean(f ** 2))
        q[z] = f
    return q


def restart_name(prefix, n):
    """'q.%09d.dat' naming of QGN_Driver.f90:23 / QGN_Module.f90:364."""
    return f"{prefix}.{n:09d}.dat"


def write_field(path, field):
    """Raw float64 stream in Fortran (nx,ny,nz) order == numpy (nz,ny,nx) C order."""
    np.ascontiguousarray(field, dtype=np.float64).tofile(path)


def read_field(path, nx, ny, nz):
    a = np.fromfile(path, dtype=np.float64, count=nx * ny * nz)
    if a.size != nx * ny * nz:
        raise ValueError(f"{path}: expected {nx * ny * nz} float64 values")
    return a.reshape(nz, ny, nx)


def write_restart(directory, n, q_phys):
    os.makedirs(directory, exist_ok=True)
    p = os.path.join(directory, restart_name("q", n))
    write_field(p, q_phys)
    return p
