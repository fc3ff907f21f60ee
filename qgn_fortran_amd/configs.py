"""Named workloads.  Values the workload descriptions leave open are builder
choices and are marked so.

  two_layer_128     2-layer QG, 128x128, 1000 km box, H = 1000/3000 m, Rd = 25 km,
                    beta = 1.6e-11, U = (0.05, 0), Ekman r0 = 1e-7, RK4 dt = 1800 s
                    x 1000 steps, explicit A8 hyperviscosity, PV peaked at k = 8,
                    RMS 1e-6 per layer.
  expstrat_512x16   16 layers on 512^2, N^2 ~ exp(z/800 m) over 4000 m, surface-
                    intensified mean flow with the same scale, ARK43_QGLeith.
  expstrat_1024x16  the same physics at 1024^2 x 16 (throughput workload).
  expstrat_2048x32  2048^2 x 32 (HBM roofline workload).
  ensemble96        96 members of 8 layers, 32 on each of 128^2, 256^2 and
                    384^2, per-member seeds, first deformation radius drawn
                    from U[15, 40] km and Ekman drag r0 from U[5e-8, 2e-7] /s,
                    ARK43_QGLeith.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import background
from .model import Params


@dataclasses.dataclass
class Workload:
    name: str
    params: Params
    H: object
    S: object
    uBar: object
    k_peak: float
    rms: float
    seed: int
    dt0: float
    steps: int
    notes: str = ""


def two_layer_128(scheme="RK4", precision="fp64"):
    H, S, u = background.two_layer(1000.0, 3000.0, 25.0e3, 0.05, 0.0)
    p = Params(nx=128, ny=128, nz=2, Lx=1.0e6, Ly=1.0e6, Htot=4000.0, beta=1.6e-11,
               A2=0.0, A8=1.0e23,          # builder choice: A8 (reference default value)
               r0=1.0e-7, C_d=0.0, TOL=1.0e-10,
               is_explicit=True,           # needed for RK4 to see any hyperviscosity
               scheme=scheme, precision=precision)
    return Workload("two_layer_128", p, H, S, u, k_peak=8, rms=1.0e-6, seed=1, dt0=1800.0, steps=1000,
                    notes="fixed dt = 1800 s")


def _expstrat(n, nz, scheme, precision, seed=2024):
    # builder choices: f0 = 1e-4 /s, N0 = 5e-3 /s (first Rd ~ 23 km), U0 = 0.1 m/s,
    # box 1000 km, Ekman r0 = 1e-7, quadratic drag C_d = 1e-5 (reference default),
    # TOL = 1e-9 in PV units (the IC has RMS 1e-6 /s and grows to ~1e-5 /s).
    H, S, u = background.exponential(nz, 4000.0, 1.0e-4, 5.0e-3, 800.0, 0.1)
    p = Params(nx=n, ny=n, nz=nz, Lx=1.0e6, Ly=1.0e6, Htot=4000.0, beta=1.6e-11,
               A2=0.0, A8=0.0, r0=1.0e-7, C_d=1.0e-5, TOL=1.0e-9, is_explicit=False,
               scheme=scheme, precision=precision)
    return Workload(f"expstrat_{n}x{nz}", p, H, S, u, k_peak=16, rms=1.0e-6, seed=seed, dt0=750.0,
                    steps=2000, notes="dt0 = 750 s as QGN_Driver.f90:6")


def expstrat_512x16(scheme="ARK43_QGLeith", precision="fp64"):
    return _expstrat(512, 16, scheme, precision)


def expstrat_1024x16(scheme="ARK43_QGLeith", precision="fp64"):
    return _expstrat(1024, 16, scheme, precision)


def expstrat_2048x32(scheme="ARK43_QGLeith", precision="fp64"):
    return _expstrat(2048, 32, scheme, precision)


@dataclasses.dataclass
class EnsembleGroup:
    params: Params
    members: list      # dicts: H, Modes, EVals, uBar, qyBar, r0, seed, Rd
    q0: object         # (members * nz, n, n) physical PV, member after member
    dt0: float


def _radius_scaled_profile(nz, Rd, f0=1.0e-4, N0_ref=5.0e-3):
    """Exponential stratification with N0 scaled so that the first
    deformation radius is Rd (radii scale linearly with N0)."""
    from .model import vertical_modes
    H, S, u = background.exponential(nz, 4000.0, f0, N0_ref, 800.0, 0.1)
    Hs, _L, Modes, EVals, qy = vertical_modes(H, S, u, 4000.0)
    rd_ref = 1.0 / math.sqrt(-np.sort(EVals)[-2])
    N0 = N0_ref * Rd / rd_ref
    H, S, u = background.exponential(nz, 4000.0, f0, N0, 800.0, 0.1)
    Hs, _L, Modes, EVals, qy = vertical_modes(H, S, u, 4000.0)
    return dict(H=Hs, Modes=Modes, EVals=EVals, uBar=u, qyBar=qy)


def ensemble96(sizes=(128, 256, 384), per_size=32, nz=8, precision="fp64", seed=96, scheme="ARK43_QGLeith"):
    """The 96-member ensemble as one group per grid size.  Builder choices:
    the exponential stratification of _expstrat with N0 set per member from
    its Rd, C_d = 1e-5, TOL = 1e-9, PV peaked at k = 8 with RMS 1e-6,
    dt0 = 750 s."""
    from . import initial
    rng = np.random.default_rng(seed)
    groups = []
    for n in sizes:
        p = Params(nx=n, ny=n, nz=nz, Lx=1.0e6, Ly=1.0e6, Htot=4000.0, beta=1.6e-11, A2=0.0, A8=0.0,
                   r0=1.0e-7, C_d=1.0e-5, TOL=1.0e-9, is_explicit=False, scheme=scheme, precision=precision)
        members, q0 = [], []
        for _ in range(per_size):
            Rd = rng.uniform(15.0e3, 40.0e3)
            r0 = rng.uniform(5.0e-8, 2.0e-7)
            s = int(rng.integers(1 << 31))
            m = _radius_scaled_profile(nz, Rd)
            m.update(r0=r0, seed=s, Rd=Rd)
            members.append(m)
            q0.append(initial.random_phase_pv(n, n, nz, 8, 1.0e-6, s, p.Lx, p.Ly))
        groups.append(EnsembleGroup(p, members, np.concatenate(q0, axis=0), 750.0))
    return groups


WORKLOADS = {
    "two_layer_128": two_layer_128,
    "expstrat_512x16": expstrat_512x16,
    "expstrat_1024x16": expstrat_1024x16,
    "expstrat_2048x32": expstrat_2048x32,
}
