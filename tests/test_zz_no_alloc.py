"""qgn_dft_compute_noalloc, the transform entry point that allocates nothing,
against the handle variants it shares its kernels with
(qgn_dft_forward_device / qgn_dft_backward_device): FP64 within relative L2
1e-13 (the DFT parity bar; the twiddle tables are formed on the device
instead of the host), FP32 within 2e-6.  Two calls in a row, the second on a
smaller batch and grid, so that the second runs on scratch the first left
behind; a batch too large for the fixed scratch is refused before any GPU
work."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _case(lib, precision, nx, ny, howmany, seed):
    import torch
    from qgn_fortran_amd._lib import check
    dt = torch.float64 if precision == 0 else torch.float32
    rng = np.random.default_rng(seed)
    grid = torch.from_numpy(rng.standard_normal((howmany, ny, nx))).to("cuda", dt)
    spec = torch.from_numpy(rng.standard_normal((howmany, ny, nx // 2 + 1, 2))).to("cuda", dt)
    outs = {k: torch.full(spec.shape, float("nan"), device="cuda", dtype=dt) for k in ("fh", "fn")}
    outs.update({k: torch.full(grid.shape, float("nan"), device="cuda", dtype=dt) for k in ("bh", "bn")})
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    h = C.c_void_p()
    check(lib.qgn_dft_create(C.byref(h), precision, nx, ny, howmany))
    print(f"noalloc precision {precision} {nx}x{ny}x{howmany}", flush=True)
    check(lib.qgn_dft_forward_device(h, C.c_void_p(grid.data_ptr()), C.c_void_p(outs["fh"].data_ptr()), s))
    check(lib.qgn_dft_backward_device(h, C.c_void_p(spec.data_ptr()), C.c_void_p(outs["bh"].data_ptr()), s))
    check(lib.qgn_dft_compute_noalloc(precision, 1, nx, ny, howmany, C.c_void_p(grid.data_ptr()),
                                      C.c_void_p(outs["fn"].data_ptr()), s))
    check(lib.qgn_dft_compute_noalloc(precision, 0, nx, ny, howmany, C.c_void_p(spec.data_ptr()),
                                      C.c_void_p(outs["bn"].data_ptr()), s))
    torch.cuda.synchronize()
    check(lib.qgn_dft_free(h))
    return {k: v.cpu().numpy().astype(np.float64) for k, v in outs.items()}


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_noalloc_matches_handle_variants(precision):
    from qgn_fortran_amd import _lib
    lib = _lib.load()
    tol = 1e-13 if precision == 0 else 2e-6
    for (nx, ny, howmany) in ((256, 192, 6), (96, 64, 2)):
        o = _case(lib, precision, nx, ny, howmany, seed=nx + precision)
        for a, b in (("fn", "fh"), ("bn", "bh")):
            assert np.all(np.isfinite(o[a]))
            err = rel_l2(o[a], o[b])
            print(f"  {a} vs {b}: rel L2 {err:.2e}", flush=True)
            assert err < tol


def test_noalloc_refuses_oversized_batch():
    import torch
    from qgn_fortran_amd import _lib
    lib = _lib.load()
    buf = torch.empty(16, device="cuda", dtype=torch.float64)
    p = C.c_void_p(buf.data_ptr())
    # 4 FP64 transforms of 2048^2: 134 MB of intermediate spectrum > 64 MiB
    rc = lib.qgn_dft_compute_noalloc(0, 1, 2048, 2048, 4, p, p, None)
    assert rc == _lib.QGN_EINVAL
    assert b"scratch" in lib.qgn_last_error()
