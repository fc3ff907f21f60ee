"""FP64 CPU restatement of QGN's spectral right-hand side and time steppers.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it: as the reference the GPU results are compared against,
or as the timed CPU baseline.

Arrays use numpy C order ``(nz, ny, nx//2+1)`` for spectra and ``(nz, ny, nx)``
for physical fields.  That is byte-for-byte the reference's Fortran
column-major ``complex(nx/2+1, ny, nz)`` / ``real(nx, ny, nz)`` layout
(``QGN_Module.f90:30-41``).  Index 0 here is Fortran index 1.

FFTs are scipy's pocketfft with the MKL DFTI conventions of
``MKL_DFTI_Setup.f90:1-52``: forward r2c and backward c2r are both
unnormalised; the half spectrum lies along x; the backward transform is a
complex inverse along y followed by an x-c2r that uses only the real part of
the kx=0 and kx=nx/2 bins (the convention MKL follows for the non-Hermitian
kx=0 column that the asymmetric 2/3 mask leaves when n is not a multiple
of 3).

Pinning (each function is checked in ``tests/test_oracle.py`` against
something other than its own formula):
  * rfft2/irfft2      - brute-force O(n^4) DFT with explicit symmetrisation,
                        closed-form single modes, Parseval.
  * dealias           - closed-form kept index sets for n = 96..2048.
  * get_psi           - per-wavenumber dense solve (-k^2 I + L) psi = q,
                        H-orthonormality of the modes.
  * jacobian          - closed form for two single Fourier modes,
                        antisymmetry, energy/enstrophy conservation at n%3==0.
  * get_rhs linear    - Rossby-wave dispersion relation, Ekman closed form.
  * quadratic drag    - closed form for a single barotropic Fourier mode
                        (|u| then has a closed-form product).
  * vertical_modes    - the deformation radii stored in
                        set_background_profiles.ipynb:180-185, 2-layer closed
                        form, L Modes = Modes diag(EVals).
  * steppers          - Butcher-coefficient identities, linear-test-equation
                        amplification factors, observed convergence order.
  * leith / pi34      - Leith sum against a single-mode closed form and the
                        binary32 literal 1.3; controller on hand-computed scalars.
"""
from __future__ import annotations

import dataclasses
import math
from fractions import Fraction

import numpy as np
import scipy.fft as sfft

PI = 4.0 * math.atan(1.0)          # parameters.f90:5
# parameters.f90 default-real literal 1.3 promoted to double
# (ARK43_QGLeith.f90:72): 1.3 rounded to binary32 first.
LEITH_C = float(np.float32(1.3))


# ---------------------------------------------------------------------------
# Parameters and set-up  (parameters.f90:1-31, QGN_Module.f90:53-188)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Params:
    nx: int
    ny: int
    nz: int
    Lx: float
    Ly: float
    Htot: float
    beta: float
    A2: float = 0.0
    A8: float = 0.0
    r0: float = 0.0
    C_d: float = 0.0
    TOL: float = 1e-10
    is_explicit: bool = False
    dt_max: float | None = None

    def __post_init__(self):
        if self.dt_max is None:
            # parameters.f90:31  CFL for the fastest Rossby wave
            self.dt_max = PI / (self.beta * self.Lx) if self.beta != 0 else float("inf")


def stretching_matrix(H, S):
    """Tridiagonal L of QGN_Module.f90:102-113 ([Del + L] psi = q)."""
    nz = len(H)
    L = np.zeros((nz, nz))
    L[0, 1] = 2.0 * S[1] / (H[0] * (H[0] + H[1]))
    L[0, 0] = -L[0, 1]
    for k in range(1, nz - 1):
        L[k, k - 1] = 2.0 * S[k] / (H[k] * (H[k] + H[k - 1]))
        L[k, k + 1] = 2.0 * S[k + 1] / (H[k] * (H[k] + H[k + 1]))
        L[k, k] = -(L[k, k - 1] + L[k, k + 1])
    L[nz - 1, nz - 2] = 2.0 * S[nz - 1] / (H[nz - 1] * (H[nz - 1] + H[nz - 2]))
    L[nz - 1, nz - 1] = -L[nz - 1, nz - 2]
    return L


def vertical_modes(H_raw, S, uBar, Htot):
    """QGN_Module.f90:84-148: rescale H, build L, eigen-decompose, normalise.

    ``S`` has nz+1 entries (interfaces 0..nz, only 1..nz-1 used).  The
    reference calls LAPACK DGEEV (QGN_Module.f90:121); numpy.linalg.eig is
    LAPACK geev as well.  Eigenvalues are real for this matrix (it is similar
    to a symmetric one), so the imaginary parts are dropped.  Column order is
    LAPACK's and affects only the order of modal output.
    Returns (H, L, Modes, EVals, qyBar).
    """
    H = np.asarray(H_raw, dtype=np.float64)
    H = Htot * H / np.sum(H)                          # :87
    L = stretching_matrix(H, np.asarray(S, dtype=np.float64))
    w, V = np.linalg.eig(L)                           # :121
    EVals = np.real(w).copy()
    Modes = np.real(V).copy()
    for k in range(len(H)):                           # :125-127
        Modes[:, k] = math.copysign(1.0, Modes[0, k]) * Modes[:, k] / math.sqrt(
            np.sum(H * Modes[:, k] ** 2) / Htot)
    qyBar = L @ (-np.asarray(uBar, dtype=np.float64))  # :148
    return H, L, Modes, EVals, qyBar


def wavenumbers(p: Params):
    """QGN_Module.f90:161-172. Arrays of shape (ny, nx//2+1)."""
    nxh = p.nx // 2 + 1
    i = np.arange(nxh)
    j = np.arange(p.ny)
    kx1 = (2.0 * PI / p.Lx) * i.astype(np.float64)
    jj = np.where(j <= p.ny // 2, j, j - p.ny)       # rows ny/2+2..ny (1-based) negative
    ky1 = (2.0 * PI / p.Ly) * jj.astype(np.float64)
    kx = np.broadcast_to(kx1[None, :], (p.ny, nxh)).copy()
    ky = np.broadcast_to(ky1[:, None], (p.ny, nxh)).copy()
    k2 = kx ** 2 + ky ** 2
    return kx, ky, k2


# ---------------------------------------------------------------------------
# FFTs with DFTI conventions (MKL_DFTI_Setup.f90:1-52)
# ---------------------------------------------------------------------------
def rfft2(grid, workers=1):
    """DftiComputeForward(g2s_q): unnormalised r2c, (nz,ny,nx)->(nz,ny,nx/2+1)."""
    return sfft.rfft2(grid, axes=(-2, -1), workers=workers)


def irfft2(spec, nx, workers=1):
    """DftiComputeBackward(s2g_q): unnormalised c2r, (nz,ny,nx/2+1)->(nz,ny,nx)."""
    ny = spec.shape[-2]
    return sfft.irfft2(spec, s=(ny, nx), axes=(-2, -1), norm="forward", workers=workers)


# ---------------------------------------------------------------------------
# Spectral core  (QGN_Module.f90:194-341)
# ---------------------------------------------------------------------------
def dealias_mask(nx, ny):
    """Boolean (ny, nx//2+1) mask of the modes DeAlias keeps (QGN_Module.f90:194-204)."""
    keep = np.ones((ny, nx // 2 + 1), dtype=bool)
    keep[:, nx // 3:] = False                 # i = nx/3+1 .. nx/2+1 (1-based)
    keep[ny // 3: 2 * (ny // 3) + 1, :] = False  # j = ny/3+1 .. 2*(ny/3)+1
    keep[0, 0] = False
    return keep


def dealias(field, nx, ny):
    out = np.array(field, dtype=np.complex128, copy=True)
    out[..., :, nx // 3:] = 0.0
    out[..., ny // 3: 2 * (ny // 3) + 1, :] = 0.0
    out[..., 0, 0] = 0.0
    return out


class Model:
    """Module state of QGN_Module.f90:11-41 plus the procedures that use it."""

    def __init__(self, p: Params, H, Modes, EVals, uBar, qyBar, workers=1):
        self.p = p
        self.H = np.asarray(H, dtype=np.float64)
        self.Modes = np.asarray(Modes, dtype=np.float64)
        self.EVals = np.asarray(EVals, dtype=np.float64)
        self.uBar = np.asarray(uBar, dtype=np.float64)
        self.qyBar = np.asarray(qyBar, dtype=np.float64)
        self.kx, self.ky, self.k2 = wavenumbers(p)
        self.k8 = self.k2 ** 4                      # ARK43_QGLeith.f90:16
        self.workers = workers
        # outputs that the reference keeps as module side effects
        self.psi_hat = None
        self.v_hat = None
        self.u_phys = None
        self.v_phys = None

    # ---- FFT helpers -----------------------------------------------------
    def c2r(self, spec):
        return irfft2(spec, self.p.nx, self.workers)

    def r2c(self, grid):
        return rfft2(grid, self.workers)

    # ---- QGN_Module.f90:315-341 ------------------------------------------
    def get_psi(self, q_hat):
        p = self.p
        Hm = self.H[:, None] * self.Modes                  # H(kk)*Modes(kk,k)
        q_mode = np.einsum("ak,ayx->kyx", Hm, q_hat) / p.Htot
        # the (0,0) mode divides by the barotropic eigenvalue, a round-off
        # number that may be 0; that mode is zeroed just below (:328)
        with np.errstate(divide="ignore", invalid="ignore"):
            psi_mode = q_mode / (-self.k2[None, :, :] + self.EVals[:, None, None])
        psi_mode[:, 0, 0] = 0.0
        return np.einsum("ka,ayx->kyx", self.Modes, psi_mode)

    # ---- QGN_Module.f90:215-249 ------------------------------------------
    def jacobian(self, q_hat):
        nn = float(self.p.nx * self.p.ny)
        psi_hat = self.get_psi(q_hat)
        u_hat = -1j * self.ky * psi_hat
        v_hat = 1j * self.kx * psi_hat
        q_phys = self.c2r(q_hat) / nn
        u_phys = self.c2r(u_hat) / nn
        v_phys = self.c2r(v_hat) / nn
        uq_hat = self.r2c(u_phys * q_phys)
        vq_hat = self.r2c(v_phys * q_phys)
        jaco_hat = 1j * self.kx * uq_hat + 1j * self.ky * vq_hat
        self.psi_hat, self.v_hat = psi_hat, v_hat
        self.u_phys, self.v_phys, self.q_phys = u_phys, v_phys, q_phys
        return jaco_hat

    # ---- QGN_Module.f90:258-296 ------------------------------------------
    def get_rhs(self, q_hat):
        p = self.p
        nz = p.nz
        jaco_hat = self.jacobian(q_hat)
        rhs = -jaco_hat
        rhs = rhs - self.uBar[:, None, None] * (1j * self.kx) * q_hat \
            - (p.beta + self.qyBar)[:, None, None] * self.v_hat          # :270-273
        if p.is_explicit:                                                   # :275-280
            rhs = rhs - (p.A2 * self.k2 + p.A8 * self.k8) * q_hat
        rhs[nz - 1] += (p.r0 * p.Htot / self.H[nz - 1]) * self.k2 * self.psi_hat[nz - 1]  # :282
        if p.C_d != 0.0:                                                    # :284-292
            # Only layer nz of the two transforms is used (H7): transforming
            # that layer alone gives the same numbers.
            ub, vb = self.u_phys[nz - 1:], self.v_phys[nz - 1:]
            urms = np.sqrt(ub ** 2 + vb ** 2)
            drag_hat = 1j * self.ky * self.r2c(urms * ub)[0]
            drag_hat = drag_hat - 1j * self.kx * self.r2c(urms * vb)[0]
            rhs[nz - 1] += (p.C_d * p.Htot / self.H[nz - 1]) * drag_hat
        return dealias(rhs, p.nx, p.ny)                                     # :294

    # ---- QG-Leith coefficient, ARK43_QGLeith.f90:62-73 -------------------
    def leith_a8(self, q_hat):
        p = self.p
        nn = float(p.nx * p.ny)
        d_phys = self.c2r(self.k8 * q_hat) / nn
        q_phys = self.c2r(q_hat) / nn
        s = np.sum(q_phys * d_phys, axis=(1, 2))
        return (LEITH_C * p.Lx / (p.nx * PI)) ** 12 * np.sqrt(s / nn)

    def initialize(self, q0):
        """QGN_Module.f90:177-183: forward FFT of the physical IC, then DeAlias."""
        return dealias(self.r2c(np.asarray(q0, dtype=np.float64)), self.p.nx, self.p.ny)

    def modal_fields(self, q_hat):
        """WriteQPmodal's fields (QGN_Module.f90:381-416): modal q, psi and
        the full flux-form Jacobian, each c2r(.)/(nx ny)."""
        p = self.p
        nn = float(p.nx * p.ny)
        jaco_hat = self.jacobian(q_hat)
        Hm = self.H[:, None] * self.Modes
        q_mode = np.einsum("ak,ayx->kyx", Hm, q_hat) / p.Htot
        j_mode = np.einsum("ak,ayx->kyx", Hm, jaco_hat) / p.Htot
        # the (0,0) mode divides by the barotropic eigenvalue, a round-off
        # number that may be 0; that mode is zeroed just below (:328)
        with np.errstate(divide="ignore", invalid="ignore"):
            psi_mode = q_mode / (-self.k2[None, :, :] + self.EVals[:, None, None])
        psi_mode[:, 0, 0] = 0.0
        return self.c2r(q_mode) / nn, self.c2r(psi_mode) / nn, self.c2r(j_mode) / nn

    def spectra(self, q_hat, nbins):
        """Isotropic KE and enstrophy spectra per layer from the physical
        fields of q_hat (the diagnostics hook QGN_Driver.f90:52-55 leaves
        open): full-plane fft2 of psi and q, bin floor(|k|/dk + 1/2),
        dk = min(2 pi/Lx, 2 pi/Ly), KE = 1/2 k^2 |psi_hat|^2 and
        Z = 1/2 |q_hat|^2, both over (nx ny)^2."""
        p = self.p
        nn = float(p.nx * p.ny)
        q_phys, psi_phys = self.get_psi_phys(q_hat)
        Q = np.fft.fft2(q_phys, axes=(-2, -1))
        Ps = np.fft.fft2(psi_phys, axes=(-2, -1))
        kx1 = (2.0 * PI / p.Lx) * np.fft.fftfreq(p.nx, 1.0 / p.nx)
        ky1 = (2.0 * PI / p.Ly) * np.fft.fftfreq(p.ny, 1.0 / p.ny)
        k2 = kx1[None, :] ** 2 + ky1[:, None] ** 2
        dk = min(2.0 * PI / p.Lx, 2.0 * PI / p.Ly)
        b = np.floor(np.sqrt(k2) / dk + 0.5).astype(np.int64).ravel()
        keep = b < nbins
        ke = np.zeros((p.nz, nbins))
        ens = np.zeros((p.nz, nbins))
        for z in range(p.nz):
            e = (0.5 * k2 * np.abs(Ps[z]) ** 2 / nn ** 2).ravel()
            w = (0.5 * np.abs(Q[z]) ** 2 / nn ** 2).ravel()
            ke[z] = np.bincount(b[keep], weights=e[keep], minlength=nbins)[:nbins]
            ens[z] = np.bincount(b[keep], weights=w[keep], minlength=nbins)[:nbins]
        return ke, ens

    def get_psi_phys(self, q_hat):
        """WriteQP's psi field (QGN_Module.f90:353-361)."""
        nn = float(self.p.nx * self.p.ny)
        return self.c2r(q_hat) / nn, self.c2r(self.get_psi(q_hat)) / nn


# ---------------------------------------------------------------------------
# Stepper coefficients  (ARK43.f90:17-57, identical in ARK43_QGLeith.f90:18-58)
# ---------------------------------------------------------------------------
ARK_AE = {
    (2, 1): Fraction(1, 2),
    (3, 1): Fraction(13861, 62500), (3, 2): Fraction(6889, 62500),
    (4, 1): Fraction(-116923316275, 2393684061468),
    (4, 2): Fraction(-2731218467317, 15368042101831),
    (4, 3): Fraction(9408046702089, 11113171139209),
    (5, 1): Fraction(-451086348788, 2902428689909),
    (5, 2): Fraction(-2682348792572, 7519795681897),
    (5, 3): Fraction(12662868775082, 11960479115383),
    (5, 4): Fraction(3355817975965, 11060851509271),
    (6, 1): Fraction(647845179188, 3216320057751),
    (6, 2): Fraction(73281519250, 8382639484533),
    (6, 3): Fraction(552539513391, 3454668386233),
    (6, 4): Fraction(3354512671639, 8306763924573),
    (6, 5): Fraction(4040, 17871),
}
ARK_B = [Fraction(82889, 524892), Fraction(0), Fraction(15625, 83664),
         Fraction(69875, 102672), Fraction(-2260, 8211), Fraction(1, 4)]
ARK_AI = {
    (2, 1): Fraction(1, 4),
    (3, 1): Fraction(8611, 62500), (3, 2): Fraction(-1743, 31250),
    (4, 1): Fraction(5012029, 34652500), (4, 2): Fraction(-654441, 2922500),
    (4, 3): Fraction(174375, 388108),
    (5, 1): Fraction(15267082809, 155376265600), (5, 2): Fraction(-71443401, 120774400),
    (5, 3): Fraction(730878875, 902184768), (5, 4): Fraction(2285395, 8070912),
}
for _j in range(1, 6):
    ARK_AI[(6, _j)] = ARK_B[_j - 1]     # ai(6,:) = b  (ARK43.f90:51)
ARK_BE = [Fraction(31666707, 9881966720), Fraction(0), Fraction(-256875, 105007616),
          Fraction(-2768025, 128864768), Fraction(169839, 3864644), Fraction(-5247, 225920)]


def _fd(x):
    """Fortran evaluates a._dp/b._dp in double: round each operand, then divide."""
    return float(x.numerator) / float(x.denominator)


AE = {k: _fd(v) for k, v in ARK_AE.items()}
AI = {k: _fd(v) for k, v in ARK_AI.items()}
B = [_fd(v) for v in ARK_B]
BE = [_fd(v) for v in ARK_BE]


def pi34(dt, err0, err1, target, a, b, dt_max):
    """PI step-size controller: dt (target/err1)^a (err0/err1)^b, capped at
    dt_max.  ARK43*.f90 use a = 0.075, b = 0.1 (ARK43_QGLeith.f90:137,
    ARK43.f90:122); BSRK3.f90:87 uses a = 0.1, b = 0.4/3."""
    return min(dt * ((target / err1) ** a) * ((err0 / err1) ** b), dt_max)


class ARK43:
    """ARK4(3)6L[2]SA with fixed A8 (ARK43.f90) or QG-Leith A8L (ARK43_QGLeith.f90).

    One instance holds the stepper's ``save`` state (err0, FirstCall).
    ``step`` returns (q_hat, t, dt) after one accepted step, as TimeStep does,
    including the final DeAlias of QGN_Module.f90:307.
    """

    def __init__(self, model: Model, leith: bool, max_rejects: int = 10_000):
        self.m = model
        self.leith = leith
        self.err0 = model.p.TOL            # :17
        self.rejects = 0
        self.last_err = None
        self.max_rejects = max_rejects
        self.log = []                      # (dt tried, err1, accepted)

    def step(self, q_hat, t, dt):
        m, p = self.m, self.m.p
        nz = p.nz
        if self.leith:
            a8 = m.leith_a8(q_hat)         # ARK43_QGLeith.f90:63-73
        else:
            a8 = np.full(nz, p.A8)
        lam = p.A2 * m.k2[None] + a8[:, None, None] * m.k8[None]   # A2 k2 + A8 k8
        N = [None] * 7
        L = [None] * 7
        N[1] = m.get_rhs(q_hat)
        L[1] = -lam * q_hat
        reject_factor = 0.5 if self.leith else 0.75
        safety = 0.2 if self.leith else 0.75
        tries = 0
        while True:
            Mq = 1.0 / (1.0 + 0.25 * dt * lam)
            for i in range(2, 7):
                acc = 0.0
                for j in range(1, i):
                    acc = acc + (AE.get((i, j), 0.0) * N[j] + AI.get((i, j), 0.0) * L[j])
                q_tmp = Mq * (q_hat + dt * acc)
                N[i] = m.get_rhs(q_tmp)
                L[i] = -lam * q_tmp
            spec = (BE[0] * (N[1] + L[1]) + BE[2] * (N[3] + L[3]) + BE[3] * (N[4] + L[4])
                    + BE[4] * (N[5] + L[5]) + BE[5] * (N[6] + L[6]))
            grid = m.c2r(spec)
            err1 = dt * np.max(np.abs(grid)) / float(p.nx * p.ny)
            self.last_err = err1
            if err1 > p.TOL:
                self.log.append((dt, err1, False))
                dt = reject_factor * dt
                self.rejects += 1
                tries += 1
                if tries > self.max_rejects:
                    raise RuntimeError("ARK43: too many rejected steps")
                continue
            self.log.append((dt, err1, True))
            q_new = q_hat + dt * (B[0] * (N[1] + L[1]) + B[2] * (N[3] + L[3])
                                  + B[3] * (N[4] + L[4]) + B[4] * (N[5] + L[5])
                                  + B[5] * (N[6] + L[6]))
            q_new[:, 0, 0] = 0.0
            t = t + dt
            dt = pi34(dt, self.err0, err1, safety * p.TOL, 0.075, 0.1, p.dt_max)
            self.err0 = err1
            return dealias(q_new, p.nx, p.ny), t, dt


class RK4:
    """RK4.f90:1-18, fixed dt."""

    def __init__(self, model: Model):
        self.m = model

    def step(self, q_hat, t, dt):
        g = self.m.get_rhs
        r0 = g(q_hat)
        r1 = g(q_hat + (0.5 * dt) * r0)
        r2 = g(q_hat + (0.5 * dt) * r1)
        r3 = g(q_hat + dt * r2)
        q = q_hat + (dt / 6.0) * (r0 + 2.0 * r1 + 2.0 * r2 + r3)
        return dealias(q, self.m.p.nx, self.m.p.ny), t + dt, dt


class AB3:
    """AB3.f90:1-23: Euler, AB2, then AB3 with saved history."""

    def __init__(self, model: Model):
        self.m = model
        self.first = True
        self.second = False
        self.r0 = self.r1 = None

    def step(self, q_hat, t, dt):
        g = self.m.get_rhs
        if self.first:
            self.r0 = g(q_hat)
            q = q_hat + dt * self.r0
            self.first, self.second = False, True
        elif self.second:
            self.r1 = g(q_hat)
            q = q_hat + dt * (1.5 * self.r1 - 0.5 * self.r0)
            self.second = False
        else:
            r2 = g(q_hat)
            q = q_hat + (dt / 12.0) * (23.0 * r2 - 16.0 * self.r1 + 5.0 * self.r0)
            self.r0, self.r1 = self.r1, r2
        return dealias(q, self.m.p.nx, self.m.p.ny), t + dt, dt


class BSRK3:
    """BSRK3.f90:1-51: Bogacki-Shampine 3(2), FSAL, adaptive."""

    def __init__(self, model: Model, max_rejects: int = 10_000):
        self.m = model
        self.first = True
        self.rhs0 = None
        self.r0 = None
        self.rejects = 0
        self.max_rejects = max_rejects
        self.log = []

    def step(self, q_hat, t, dt):
        m, p = self.m, self.m.p
        g = m.get_rhs
        if self.first:
            self.rhs0 = g(q_hat)
            self.r0 = p.TOL
            self.first = False
        tries = 0
        while True:
            rhs1 = g(q_hat + 0.5 * dt * self.rhs0)
            rhs2 = g(q_hat + 0.75 * dt * rhs1)
            q_tmp = q_hat + (2.0 * dt / 9.0) * self.rhs0 + (1.0 * dt / 3.0) * rhs1 + (4.0 * dt / 9.0) * rhs2
            rhs3 = g(q_tmp)
            err_spec = ((5.0 / 72.0) * self.rhs0 - (1.0 / 12.0) * rhs1
                        - (1.0 / 9.0) * rhs2 + (1.0 / 8.0) * rhs3)
            err_phys = m.c2r(err_spec) / float(p.nx * p.ny)
            r1 = dt * np.max(np.abs(err_phys))
            if r1 > p.TOL:
                self.log.append((dt, r1, False))
                dt = 0.75 * dt
                self.rejects += 1
                tries += 1
                if tries > self.max_rejects:
                    raise RuntimeError("BSRK3: too many rejected steps")
                continue
            self.log.append((dt, r1, True))
            t = t + dt
            self.rhs0 = rhs3
            dt = pi34(dt, self.r0, r1, 0.75 * p.TOL, 0.1, 0.4 / 3.0, p.dt_max)
            self.r0 = r1
            return dealias(q_tmp, p.nx, p.ny), t, dt


SCHEMES = {"AB3": AB3, "RK4": RK4, "BSRK32": BSRK3,
           "ARK43": lambda m: ARK43(m, leith=False),
           "ARK43_QGLeith": lambda m: ARK43(m, leith=True)}
