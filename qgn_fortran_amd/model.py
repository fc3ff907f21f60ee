"""Host-side mirror of QGN_Module's public interface over the HIP library.

``QGN`` follows QGN_Module.f90:7 (Initialize, TimeStep, WriteQP, Cleanup) plus
the internal GetRHS/GetPsi used by the parity tests; ``Dfti`` follows the two
MKL DFTI descriptors of MKL_DFTI_Setup.f90.  Arrays are numpy, in the
reference's memory layout: spectra complex128 of shape (nz, ny, nx//2+1)
(= Fortran complex(nx/2+1,ny,nz)), physical fields float64 (nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib
from ._lib import check, dptr

PI = 4.0 * math.atan(1.0)


@dataclasses.dataclass
class Params:
    """parameters.f90:1-31, sizes at run time."""
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
    scheme: str = "ARK43_QGLeith"
    precision: str = "fp64"
    max_rejects: int = 1000

    def __post_init__(self):
        if self.dt_max is None:
            self.dt_max = PI / (self.beta * self.Lx) if self.beta != 0 else float("inf")   # parameters.f90:31

    def to_c(self) -> _lib.QgnParams:
        return _lib.QgnParams(
            self.nx, self.ny, self.nz, self.Lx, self.Ly, self.Htot, self.beta, self.A2, self.A8,
            self.r0, self.C_d, self.TOL, self.dt_max, int(bool(self.is_explicit)),
            _lib.SCHEMES[self.scheme], _lib.PRECISIONS[self.precision], int(self.max_rejects))


def vertical_modes(H_raw, S, uBar, Htot):
    """Vertical structure of Initialize (QGN_Module.f90:84-148) via the
    library's symmetric eigensolver.  Returns (H, L, Modes, EVals, qyBar);
    Modes[:, k] is mode k (numpy view of Fortran Modes(:,k))."""
    lib = _lib.load()
    H_raw = np.ascontiguousarray(H_raw, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    uBar = np.ascontiguousarray(uBar, dtype=np.float64)
    nz = H_raw.size
    if S.size != nz + 1 or uBar.size != nz:
        raise ValueError("S needs nz+1 entries and uBar nz entries")
    H = np.empty(nz)
    Lf = np.empty(nz * nz)
    Mf = np.empty(nz * nz)
    ev = np.empty(nz)
    qy = np.empty(nz)
    check(lib.qgn_vertical_modes(nz, float(Htot), dptr(H_raw), dptr(S), dptr(uBar), dptr(H), dptr(Lf),
                                 dptr(Mf), dptr(ev), dptr(qy)))
    # column-major (Fortran) -> numpy [row, col]
    L = Lf.reshape(nz, nz).T.copy()
    Modes = Mf.reshape(nz, nz).T.copy()
    return H, L, Modes, ev, qy


def spectra_bins(p: Params):
    """Number of shells of width dk = min(2 pi/Lx, 2 pi/Ly) that covers every
    wavenumber of the (nx, ny) grid."""
    dkx, dky = 2.0 * np.pi / p.Lx, 2.0 * np.pi / p.Ly
    kmax = np.hypot(0.5 * p.nx * dkx, 0.5 * p.ny * dky)
    return int(np.floor(kmax / min(dkx, dky) + 0.5)) + 1


class QGN:
    """One QGN model instance on the GPU (module state of QGN_Module.f90:11-41)."""

    members = 1

    def __init__(self, p: Params, H, Modes, EVals, uBar, qyBar):
        self.lib = _lib.load()
        self.p = p
        self.nzt = p.nz
        self.nxh = p.nx // 2 + 1
        self.H = np.ascontiguousarray(H, dtype=np.float64)
        modes_f = np.ascontiguousarray(np.asarray(Modes, dtype=np.float64).T)  # Fortran order
        self._ctx = C.c_void_p()
        cp = p.to_c()
        check(self.lib.qgn_create(C.byref(self._ctx), C.byref(cp), dptr(self.H), dptr(modes_f),
                                  dptr(np.ascontiguousarray(EVals, dtype=np.float64)),
                                  dptr(np.ascontiguousarray(uBar, dtype=np.float64)),
                                  dptr(np.ascontiguousarray(qyBar, dtype=np.float64))))

    @classmethod
    def from_profiles(cls, p: Params, H_raw, S, uBar):
        H, _L, Modes, EVals, qyBar = vertical_modes(H_raw, S, uBar, p.Htot)
        return cls(p, H, Modes, EVals, uBar, qyBar)

    # ---- shapes ----------------------------------------------------------
    def spec_shape(self):
        return (self.nzt, self.p.ny, self.nxh)

    def phys_shape(self):
        return (self.nzt, self.p.ny, self.p.nx)

    def _spec_in(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != self.spec_shape():
            raise ValueError(f"spectrum shape {a.shape} != {self.spec_shape()}")
        return a

    # ---- QGN_Module procedures ------------------------------------------
    def initialize(self, q0):
        """Initialize's IC transform (QGN_Module.f90:177-183); returns q_hat."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        if q0.shape != self.phys_shape():
            raise ValueError("q0 shape")
        out = np.empty(self.spec_shape(), dtype=np.complex128)
        check(self.lib.qgn_initialize(self._ctx, dptr(q0), dptr(out)))
        return out

    def set_state(self, q_hat):
        check(self.lib.qgn_set_state(self._ctx, dptr(self._spec_in(q_hat))))

    def get_state(self):
        out = np.empty(self.spec_shape(), dtype=np.complex128)
        check(self.lib.qgn_get_state(self._ctx, dptr(out)))
        return out

    def time_step(self, t, dt):
        """TimeStep (QGN_Module.f90:302-309); returns the new (t, dt)."""
        tt, dd = C.c_double(t), C.c_double(dt)
        check(self.lib.qgn_time_step(self._ctx, C.byref(tt), C.byref(dd)))
        return tt.value, dd.value

    def get_rhs(self, q_hat):
        """GetRHS (QGN_Module.f90:258-296)."""
        out = np.empty(self.spec_shape(), dtype=np.complex128)
        check(self.lib.qgn_rhs(self._ctx, dptr(self._spec_in(q_hat)), dptr(out)))
        return out

    def get_psi(self, q_hat):
        """GetPsi (QGN_Module.f90:315-341)."""
        out = np.empty(self.spec_shape(), dtype=np.complex128)
        check(self.lib.qgn_get_psi(self._ctx, dptr(self._spec_in(q_hat)), dptr(out)))
        return out

    def fields(self):
        """WriteQP's q_phys and psi_phys of the current state (QGN_Module.f90:353-361)."""
        q = np.empty(self.phys_shape())
        psi = np.empty(self.phys_shape())
        check(self.lib.qgn_fields(self._ctx, dptr(q), dptr(psi)))
        return q, psi

    def modal_fields(self):
        """WriteQPmodal's q_mode, psi_mode, J_mode of the current state
        (QGN_Module.f90:381-416)."""
        out = [np.empty(self.phys_shape()) for _ in range(3)]
        check(self.lib.qgn_modal_fields(self._ctx, dptr(out[0]), dptr(out[1]), dptr(out[2])))
        return tuple(out)

    def spectra(self, nbins=None):
        """Isotropic KE and enstrophy spectra per layer, each (nz, nbins);
        bin b holds |k| in [(b - 1/2) dk, (b + 1/2) dk), dk = min(2 pi/Lx, 2 pi/Ly)."""
        p = self.p
        if nbins is None:
            nbins = spectra_bins(p)
        ke = np.empty((self.nzt, nbins))
        ens = np.empty((self.nzt, nbins))
        check(self.lib.qgn_spectra(self._ctx, int(nbins), dptr(ke), dptr(ens)))
        return ke, ens

    def tuning(self):
        """Launch choices of the mixed path (column inverse tile, column
        forward tile, row group; 0 = fixed rule) and the measured times (us)."""
        ci, cr, rw = C.c_int(), C.c_int(), C.c_int()
        t = np.zeros((3, 4), dtype=np.float32)
        check(self.lib.qgn_get_tuning(self._ctx, C.byref(ci), C.byref(cr), C.byref(rw),
                                      t.ctypes.data_as(C.POINTER(C.c_float))))
        return dict(col_inv=ci.value, col_rhs=cr.value, row=rw.value, times_us=t)

    def set_tuning(self, col_inv=0, col_rhs=0, row=0):
        check(self.lib.qgn_set_tuning(self._ctx, int(col_inv), int(col_rhs), int(row)))

    def set_closure(self, net, in_scale=None, out_scale=None):
        """Attach a closure network (qgn_fortran_amd.cnn.CNN) as subgrid PV
        forcing of every RHS evaluation; None detaches it."""
        if net is None:
            check(self.lib.qgn_set_closure(self._ctx, None, None, None))
            self._net = None
            return
        ins = np.ascontiguousarray(in_scale, dtype=np.float64)
        outs = np.ascontiguousarray(out_scale, dtype=np.float64)
        check(self.lib.qgn_set_closure(self._ctx, net._h, dptr(ins), dptr(outs)))
        self._net = net      # keep the network alive while attached

    def leith(self):
        """A8L(nz) of the current state (ARK43_QGLeith.f90:62-73)."""
        out = np.empty(self.nzt)
        check(self.lib.qgn_leith(self._ctx, dptr(out)))
        return out

    def stats(self):
        s = _lib.QgnStats()
        check(self.lib.qgn_get_stats(self._ctx, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def stream(self):
        p = C.c_void_p()
        check(self.lib.qgn_get_stream(self._ctx, C.byref(p)))
        return p.value

    def prof_enable(self, kernel):
        check(self.lib.qgn_prof_enable(self._ctx, _lib.KERNELS[kernel] if isinstance(kernel, str) else kernel))

    def prof_read(self):
        n = C.c_int()
        ms = C.c_double()
        check(self.lib.qgn_prof_read(self._ctx, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def prof_read_all(self):
        """{kernel name: (launches, total ms)} since prof_enable("all")."""
        k = len(_lib.KERNELS) - 1
        n = (C.c_int * k)()
        ms = np.zeros(k)
        check(self.lib.qgn_prof_read_all(self._ctx, n, dptr(ms)))
        return {name: (n[i], float(ms[i])) for name, i in _lib.KERNELS.items() if i >= 0}

    def close(self):
        """Cleanup (QGN_Module.f90:448-458); raises (and keeps the context)
        while an ensemble holds it."""
        if self._ctx:
            check(self.lib.qgn_destroy(self._ctx))
            self._ctx = C.c_void_p()
            self._net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(QGN):
    """Members of one grid size stacked along z, each with its own vertical
    structure and Ekman coefficient, stepped by device-controlled ARK43
    attempts captured in a HIP graph (no counterpart in the reference).

    members: sequence of dicts with H, Modes, EVals, uBar, qyBar and
    optionally r0.  State arrays hold members * nz layers, member after
    member."""

    def __init__(self, p: Params, members):
        self.lib = _lib.load()
        self.p = p
        self.members = len(members)
        self.nzt = p.nz * self.members
        self.nxh = p.nx // 2 + 1
        cat = lambda key: np.ascontiguousarray(np.concatenate([np.ravel(m[key]) for m in members]), dtype=np.float64)
        modes_f = np.ascontiguousarray(np.concatenate(
            [np.asarray(m["Modes"], dtype=np.float64).T.ravel() for m in members]))   # Fortran order each
        r0 = np.array([m.get("r0", p.r0) for m in members], dtype=np.float64)
        self.H = cat("H")
        self._ctx = C.c_void_p()
        cp = p.to_c()
        check(self.lib.qgn_create_batch(C.byref(self._ctx), C.byref(cp), self.members, dptr(self.H), dptr(modes_f),
                                        dptr(cat("EVals")), dptr(cat("uBar")), dptr(cat("qyBar")), dptr(r0)))

    def set_time(self, t, dt):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.members,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.members,)))
        check(self.lib.qgn_batch_set_time(self._ctx, dptr(t), dptr(dt)))

    def step(self, attempts):
        """Enqueue `attempts` ARK43 attempts of every member (asynchronous)."""
        check(self.lib.qgn_batch_step(self._ctx, int(attempts)))

    def times(self):
        """Per-member t, dt, accepted steps, rejected attempts, status (waits)."""
        n = self.members
        t, dt = np.empty(n), np.empty(n)
        steps, rej = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int64)
        st = np.empty(n, dtype=np.int32)
        check(self.lib.qgn_batch_get_time(self._ctx, dptr(t), dptr(dt), steps.ctypes.data_as(C.POINTER(C.c_longlong)),
                                          rej.ctypes.data_as(C.POINTER(C.c_longlong)),
                                          st.ctypes.data_as(C.POINTER(C.c_int))))
        return dict(t=t, dt=dt, steps=steps, rejects=rej, status=st)


class Ensemble:
    """Several batches advanced by one graph launch per attempt."""

    def __init__(self, batches):
        self.lib = _lib.load()
        self.batches = list(batches)
        arr = (C.c_void_p * len(self.batches))(*[b._ctx.value for b in self.batches])
        self._h = C.c_void_p()
        check(self.lib.qgn_ensemble_create(C.byref(self._h), arr, len(self.batches)))

    def step(self, attempts):
        check(self.lib.qgn_ensemble_step(self._h, int(attempts)))

    def close(self):
        if self._h:
            self.lib.qgn_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dfti:
    """Batched 2-D real FFT with the descriptor of MKL_DFTI_Setup.f90:11-50.

    forward(grid) = DftiComputeForward(g2s_q, ...), backward(spec) =
    DftiComputeBackward(s2g_q, ...); both unnormalised."""

    def __init__(self, nx, ny, howmany, precision="fp64"):
        self.lib = _lib.load()
        self.nx, self.ny, self.howmany = nx, ny, howmany
        self._h = C.c_void_p()
        check(self.lib.qgn_dft_create(C.byref(self._h), _lib.PRECISIONS[precision], nx, ny, howmany))

    def forward(self, grid):
        grid = np.ascontiguousarray(grid, dtype=np.float64).reshape(self.howmany, self.ny, self.nx)
        out = np.empty((self.howmany, self.ny, self.nx // 2 + 1), dtype=np.complex128)
        check(self.lib.qgn_dft_forward(self._h, dptr(grid), dptr(out)))
        return out

    def backward(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128).reshape(self.howmany, self.ny, self.nx // 2 + 1)
        out = np.empty((self.howmany, self.ny, self.nx))
        check(self.lib.qgn_dft_backward(self._h, dptr(spec), dptr(out)))
        return out

    def forward_device(self, grid_ptr, spec_ptr, stream=None):
        check(self.lib.qgn_dft_forward_device(self._h, grid_ptr, spec_ptr, stream))

    def backward_device(self, spec_ptr, grid_ptr, stream=None):
        check(self.lib.qgn_dft_backward_device(self._h, spec_ptr, grid_ptr, stream))

    def close(self):
        if self._h:
            self.lib.qgn_dft_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
