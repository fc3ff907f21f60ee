// GetRHS for grid sizes outside the compiled transform lengths of the
// column and row passes (kernels_ct.hpp): the same arithmetic, staged
// through full-layout spectra and physical fields so that the batched
// transforms with run-time radix plans (k_col_fft, k_row_c2r, k_row_r2c) do
// the FFTs.  Every nx/2 and ny that factor into 2, 3 and 5 (nx <= 2048,
// ny <= 4096) run here.
//   k_gen_expand    compact -> full (nz, ny, nx/2+1) spectrum with the
//                   derivative of the field applied
//   k_gen_products  J = u q_x + v q_y in physical space (advective form, as
//                   the row pass; the reference's flux form is the same
//                   Galerkin product on the kept modes) and, on each
//                   member's bottom layer, |u| u and |u| v for the
//                   quadratic drag (QGN_Module.f90:284-292)
//   k_gen_combine   the spectral terms of GetRHS (QGN_Module.f90:266-294)
//                   on the kept modes, as the forward column pass forms them
//   k_gen_maxabs    per-member maximum of |x| over physical fields (the
//                   adaptive steppers' error norm), as ordered bit patterns
#pragma once
#include "kernels.hpp"

namespace qgn {

// mult: 0 the field, 1 -i ky (u from psi), 2 +i kx, 3 +i ky
template <typename R>
__global__ void k_gen_expand(const void* __restrict__ src, int src_is_double, int mult, cplx<R>* __restrict__ full,
                             Geom g) {
  using Cx = cplx<R>;
  const long n = (long)g.nz * g.ny * g.nxh;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nxh);
    const long r = e / g.nxh;
    const int j = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    const int jc = full_row_to_compact(j, g);
    Cx x = cmk<Cx>((R)0, (R)0);
    if (jc >= 0 && i < g.Kx) {
      const long c = ((long)z * g.Ky + jc) * g.Kxp + i;
      if (src_is_double) x = to_c<R>(static_cast<const double2*>(src)[c]);
      else x = static_cast<const Cx*>(src)[c];
      if (mult == 1) {
        const R ky = ky_r<R>(j, g);
        x = cmk<Cx>(ky * x.y, -ky * x.x);
      } else if (mult == 2) {
        const R kx = kx_r<R>(i, g);
        x = cmk<Cx>(-kx * x.y, kx * x.x);
      } else if (mult == 3) {
        const R ky = ky_r<R>(j, g);
        x = cmk<Cx>(-ky * x.y, ky * x.x);
      }
    }
    full[e] = x;
  }
}

// J into qx's array; with drag, |u| u and |u| v of each member's bottom
// layer into du, dv (members, ny, nx)
template <typename R>
__global__ void k_gen_products(const R* __restrict__ u, const R* __restrict__ v, R* __restrict__ qx,
                               const R* __restrict__ qy, R* __restrict__ du, R* __restrict__ dv, int drag, Geom g) {
  const long nxy = (long)g.nx * g.ny;
  const long n = (long)g.nz * nxy;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const R uu = u[e], vv = v[e];
    qx[e] = uu * qx[e] + vv * qy[e];
    const int z = (int)(e / nxy);
    if (drag && z % g.nzm == g.nzm - 1) {
      const R urms = sqrt(uu * uu + vv * vv);
      const long d = (long)(z / g.nzm) * nxy + e % nxy;
      du[d] = urms * uu;
      dv[d] = urms * vv;
    }
  }
}

// RHS = -J_hat - Ubar_k i kx q - (beta + Qy_k) i kx psi [- (A2 k2 + A8 k8) q]
//       [+ r0 Htot/H_nz k2 psi + C_d Htot/H_nz (i ky du - i kx dv) on layer nz]
// on the kept modes, (0,0) and the padding zero.  jh, duh, dvh: full
// spectra; q (or q32), psi compact.
template <typename R>
__global__ void k_gen_combine(ColRhsArgs a, const cplx<R>* __restrict__ jh, const cplx<R>* __restrict__ duh,
                              const cplx<R>* __restrict__ dvh, Geom g) {
  using Cx = cplx<R>;
  const long n = (long)g.nz * g.Ky * g.Kxp;
  Cx* out = static_cast<Cx*>(a.out);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.Kxp);
    const long r = e / g.Kxp;
    const int jc = (int)(r % g.Ky);
    const int z = (int)(r / g.Ky);
    const int j = compact_row_to_full(jc, g);
    if ((i == 0 && j == 0) || i >= g.Kx) {
      out[e] = cmk<Cx>((R)0, (R)0);
      continue;
    }
    const long f = ((long)z * g.ny + j) * g.nxh + i;
    const Cx v = jh[f];
    const Cx q = a.q32 ? static_cast<const Cx*>(a.q32)[e] : to_c<R>(a.q[e]);
    const Cx ps = static_cast<const Cx*>(a.psi)[e];
    const R kx = kx_r<R>(i, g), ky = ky_r<R>(j, g);
    const R ub = (R)a.ubar[z];
    const R bq = (R)(a.beta + a.qybar[z]);
    Cx o = cmk<Cx>(-v.x, -v.y);
    o.x += ub * kx * q.y + bq * kx * ps.y;
    o.y -= ub * kx * q.x + bq * kx * ps.x;
    const R k2 = kx * kx + ky * ky;
    if (a.explicit_visc) {
      const R k4 = k2 * k2;
      const R lam = (R)a.A2 * k2 + (R)a.A8 * (k4 * k4);
      o.x -= lam * q.x;
      o.y -= lam * q.y;
    }
    if (z % g.nzm == g.nzm - 1) {
      const int mb = z / g.nzm;
      const R ek = (R)a.ekman[mb];
      o.x += ek * k2 * ps.x;
      o.y += ek * k2 * ps.y;
      if (a.drag) {
        const long d = ((long)mb * g.ny + j) * g.nxh + i;
        const R cd = (R)a.cdrag[mb];
        const Cx x = duh[d], y = dvh[d];
        o.x += -cd * ky * x.y + cd * kx * y.y;
        o.y += cd * ky * x.x - cd * kx * y.x;
      }
    }
    out[e] = o;
  }
}

// Per-member maximum of |x| over n = nzm ny nx values per member: block b
// of member blockIdx.y writes partial[member * gridDim.x + b] (the bit
// pattern of a non-negative double orders as the value; NaN orders above
// every finite value, so the step control sees it)
__global__ void __launch_bounds__(256) k_gen_maxabs(const double* __restrict__ x, long n,
                                                    unsigned long long* __restrict__ partial) {
  __shared__ unsigned long long red[4];
  const double* xm = x + (long)blockIdx.y * n;
  unsigned long long m = 0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const unsigned long long b = __double_as_longlong(fabs(xm[e]));
    m = b > m ? b : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long mm = 0;
    for (int w = 0; w < 4; ++w) mm = red[w] > mm ? red[w] : mm;
    partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = mm;
  }
}

}  // namespace qgn
