"""Background profiles: layer thicknesses H, interface S = f0^2/N^2 and mean
zonal flow uBar, in the binary formats QGN reads (QGN_Module.f90:84-91,
145-147; written by set_background_profiles.ipynb:326-342).

File layout (all raw native-endian float64, top layer first):
  H.dat     nz      layer thicknesses (QGN rescales them to Htot)
  S.dat     nz+1    f0^2/N^2 at interfaces 0 (surface) .. nz (bottom)
  UBar.dat  nz      mean zonal velocity per layer
"""
from __future__ import annotations

import os

import numpy as np


def write_profiles(directory, H, S, uBar):
    """Write H.dat, S.dat, UBar.dat as QGN's Initialize reads them."""
    os.makedirs(directory, exist_ok=True)
    np.asarray(H, dtype=np.float64).tofile(os.path.join(directory, "H.dat"))
    np.asarray(S, dtype=np.float64).tofile(os.path.join(directory, "S.dat"))
    np.asarray(uBar, dtype=np.float64).tofile(os.path.join(directory, "UBar.dat"))


def read_profiles(directory, nz):
    H = np.fromfile(os.path.join(directory, "H.dat"), dtype=np.float64, count=nz)
    S = np.fromfile(os.path.join(directory, "S.dat"), dtype=np.float64, count=nz + 1)
    u = np.fromfile(os.path.join(directory, "UBar.dat"), dtype=np.float64, count=nz)
    if H.size != nz or S.size != nz + 1 or u.size != nz:
        raise ValueError("profile files are shorter than nz")
    return H, S, u


def two_layer(H1, H2, Rd, U1, U2):
    """Two layers with first deformation radius Rd for the stretching matrix
    of QGN_Module.f90:102-113, whose non-zero eigenvalue is -2 S1/(H1 H2)."""
    # -lambda = 2 S1/(H1+H2) * (1/H1 + 1/H2) = 2 S1 / (H1 H2)  ->  S1 = H1 H2 / (2 Rd^2)
    S1 = H1 * H2 / (2.0 * Rd ** 2)
    return np.array([H1, H2]), np.array([0.0, S1, 0.0]), np.array([U1, U2])


def exponential(nz, depth, f0, N0, scale, U0, uscale=None):
    """Equal layers over `depth`; N^2(z) = N0^2 exp(z/scale) (z <= 0 up);
    S = f0^2/N^2 at interfaces; uBar = U0 exp(z_c/uscale) at layer centres."""
    uscale = scale if uscale is None else uscale
    H = np.full(nz, depth / nz)
    zi = -np.arange(nz + 1) * (depth / nz)           # interfaces, 0 at the surface
    S = f0 ** 2 / (N0 ** 2 * np.exp(zi / scale))
    zc = 0.5 * (zi[:-1] + zi[1:])
    uBar = U0 * np.exp(zc / uscale)
    return H, S, uBar


def notebook_profile(nz=32, alpha=1.0):
    """The demonstration profile of set_background_profiles.ipynb (cells 2-15):
    N^2(t) = (1 + 9 t + 50 exp(-75 (t - 0.8)^2))^2 on unit depth (t = 0 at
    the bottom), layer interfaces at equal increments of int sqrt(N^2)^alpha,
    S = 1/N^2 at interfaces, uBar = 0.22 * (first baroclinic mode) - 0.08 *
    (ocean-Charney profile).  Returns (H, S, uBar) in QGN's top-first order."""
    from scipy import integrate, optimize

    def N2(t):
        return (1.0 + 9.0 * t + 50.0 * np.exp(-75.0 * (t - 0.8) ** 2)) ** 2

    def SS(t):
        return np.sqrt(N2(t)) ** alpha

    norm = integrate.quad(SS, 0, 1)[0]
    z = np.zeros(nz + 1)
    z[nz] = 1.0
    for i in range(1, nz):
        z[i] = optimize.fsolve(lambda t: integrate.quad(SS, 0, t[0])[0] - norm * i / nz, i / nz)[0]
    h = z[1:] - z[:-1]
    # level-based stretching operator (bottom-first), as the notebook builds it
    DS = np.diag(1.0 / N2(z[1:-1]))
    D0 = np.zeros((nz - 1, nz))
    D0[:, :-1] = 2.0 * np.diag(1.0 / (h[:-1] + h[1:]))
    D0[:, 1:] = D0[:, 1:] - 2.0 * np.diag(1.0 / (h[:-1] + h[1:]))
    D = DS @ D0
    L = (1.0 / h) * (np.vstack((np.zeros((1, nz)), D)) - np.vstack((D, np.zeros((1, nz)))))
    L = L.T
    vals, vecs = np.linalg.eig(L)
    ind = np.argsort(-vals.real)
    bc1 = vecs[:, ind[1]].real.copy()
    bc1 /= bc1[-1]
    L_oc = L[:-1, 1:]
    oc = np.linalg.solve(L_oc, np.ones(nz - 1))
    oc = np.concatenate((np.zeros(1), oc))
    oc /= oc[-1]
    u_bar = 0.22 * bc1 - 0.08 * oc
    S = np.zeros(nz + 1)
    for i in range(nz + 1):
        S[nz - i] = 1.0 / N2(z[i])
    return np.flip(h).copy(), S, np.flip(u_bar).copy()


def deformation_radii(EVals):
    """1/sqrt(-EVals) as printed by Initialize (QGN_Module.f90:141); NaN for
    the barotropic eigenvalue when it rounds to >= 0."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return 1.0 / np.sqrt(-np.asarray(EVals, dtype=np.float64))


__all__ = ["write_profiles", "read_profiles", "two_layer", "exponential", "notebook_profile",
           "deformation_radii"]
