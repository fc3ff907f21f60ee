// Device kernels of the QGN spectral right-hand side and time steppers.
//
// Storage (all complex arrays interleaved re,im):
//   full spectrum   (nz, ny, nx/2+1)  - the reference layout (QGN_Module.f90:33)
//   compact         (nz, Ky, Kx)      - only the modes DeAlias keeps
//                                        (QGN_Module.f90:194-204): kx index
//                                        i < nx/3, ky rows j < ny/3 or
//                                        j > 2*(ny/3).  Every prognostic and
//                                        stage array is dealiased, so the
//                                        solver keeps them in this form.
//   mixed           (nz, ny, K)       - y transformed to physical space, kx
//                                        still spectral, K = Kx (or nx/2+1 for
//                                        the general FFT entry points).
//   physical        (nz, ny, nx)      - reals.
#pragma once
#include "fft_lds.hpp"

namespace qgn {

struct Geom {
  int nx, ny, nz;   // nz: layers of all members together
  int nzm;          // layers per member (a batch stacks members along z)
  int nxh;   // nx/2 + 1
  int N2;    // nx/2
  int Kx;    // kept kx indices: 0 .. nx/3-1
  int Kxp;   // row stride of compact and mixed arrays: Kx padded to 16
  int Ky;    // kept ky rows
  int ya;    // rows [0, ya) kept
  int yb;    // rows [yb, ny) kept
  double dkx, dky;  // 2*pi/Lx, 2*pi/Ly
};

__device__ __forceinline__ int compact_row_to_full(int jc, const Geom& g) {
  return jc < g.ya ? jc : jc - g.ya + g.yb;
}
__device__ __forceinline__ int full_row_to_compact(int j, const Geom& g) {
  return j < g.ya ? j : (j >= g.yb ? j - g.yb + g.ya : -1);
}
// QGN_Module.f90:161-172
__device__ __forceinline__ double kx_of(int i, const Geom& g) { return g.dkx * (double)i; }
__device__ __forceinline__ double ky_of(int j, const Geom& g) {
  return g.dky * (double)(j <= g.ny / 2 ? j : j - g.ny);
}

// The same in the pass's arithmetic type (a single rounding of dkx, dky to
// FP32 on the mixed path instead of FP64 products converted per element).
template <typename R> __device__ __forceinline__ R kx_r(int i, const Geom& g) { return (R)g.dkx * (R)i; }
template <typename R> __device__ __forceinline__ R ky_r(int j, const Geom& g) {
  return (R)g.dky * (R)(j <= g.ny / 2 ? j : j - g.ny);
}

// Mixed arrays (y physical, kx spectral; nz planes of ny rows of Kxp bins).
// FP32 planes store their rows in interleaved pairs: bin k of row y sits at
// ((y/2) Kxp + k) 2 + y%2, so a column tile of 8 FP32 columns moves whole
// 128-byte lines (two rows of 64 bytes) and a row pass over two rows reads
// one contiguous block.  Against the plain row-major layout this takes the
// column passes from 64-byte (8 FP32 columns) or 32-byte (4) row segments,
// which the HBM serves at about 4.8 TB/s on MI355X in a copy of P1's
// shape, to whole lines (6.7 TB/s).  FP64 planes stay row-major.  Offsets
// are relative to the start of an even row (every plane, row group and tile
// starts on one).
template <typename R> constexpr bool mix_pairs() { return sizeof(R) == 4; }
template <typename R> __device__ __forceinline__ int mix_at(int y, int k, int Kxp) {
  if constexpr (mix_pairs<R>()) return ((y >> 1) * Kxp + k) * 2 + (y & 1);
  else return y * Kxp + k;
}
// distance between bins k and k+1 of one row
template <typename R> constexpr int mix_kstep() { return mix_pairs<R>() ? 2 : 1; }

template <typename T> __device__ __forceinline__ cplx<T> to_c(double2 a) { return cmk<cplx<T>>((T)a.x, (T)a.y); }
template <typename T> __device__ __forceinline__ cplx<T> to_c(float2 a) { return cmk<cplx<T>>((T)a.x, (T)a.y); }

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// ---------------------------------------------------------------------------
// Layout conversion: full <-> compact (DeAlias, QGN_Module.f90:194-204)
// ---------------------------------------------------------------------------
__global__ void k_pack(const double2* __restrict__ full, double2* __restrict__ cmp, Geom g) {
  const long n = (long)g.nz * g.Ky * g.Kxp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.Kxp);
    const long r = e / g.Kxp;
    const int jc = (int)(r % g.Ky);
    const int z = (int)(r / g.Ky);
    const int j = compact_row_to_full(jc, g);
    double2 v = make_double2(0.0, 0.0);
    if (i < g.Kx && !(i == 0 && j == 0)) v = full[((long)z * g.ny + j) * g.nxh + i];
    cmp[e] = v;
  }
}

__global__ void k_unpack(const double2* __restrict__ cmp, double2* __restrict__ full, Geom g) {
  const long n = (long)g.nz * g.ny * g.nxh;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nxh);
    const long r = e / g.nxh;
    const int j = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    const int jc = full_row_to_compact(j, g);
    double2 v = make_double2(0.0, 0.0);
    if (jc >= 0 && i < g.Kx && !(i == 0 && j == 0)) v = cmp[((long)z * g.Ky + jc) * g.Kxp + i];
    full[e] = v;
  }
}

// ---------------------------------------------------------------------------
// Layer -> vertical mode -> layer projection and modal Helmholtz inversion
// (GetPsi, QGN_Module.f90:315-341).  One thread per (point, layer); the
// points of a workgroup exchange their modal amplitudes through LDS.
//   HM[k*nz + kk]    = H(kk) * Modes(kk, k)
//   MD[k*nz + kk]    = Modes(k, kk)
// with nz the layers of one member; a batch holds one such pair of
// matrices (and nz eigenvalues) per member, and grid.y runs over members.
// ---------------------------------------------------------------------------
template <typename Tout>
__global__ void k_project(const double2* __restrict__ q, cplx<Tout>* __restrict__ psi,
                          const double* __restrict__ HM, const double* __restrict__ MD,
                          const double* __restrict__ evals, double inv_htot, int pp_per_block, Geom g) {
  extern __shared__ __align__(16) double lds_d[];
  const int nz = g.nzm;
  {  // member blockIdx.y: its layers, matrices and eigenvalues
    const int m = blockIdx.y;
    const long off = (long)m * nz * g.Ky * g.Kxp;
    q += off;
    psi += off;
    HM += m * nz * nz;
    MD += m * nz * nz;
    evals += m * nz;
  }
  double* sHM = lds_d;
  double* sMD = sHM + nz * nz;
  double2* sq = reinterpret_cast<double2*>(sMD + nz * nz);  // [nz][pp]
  double2* sm = sq + nz * pp_per_block;                      // [nz][pp]
  const int tid = threadIdx.x;
  for (int e = tid; e < nz * nz; e += blockDim.x) { sHM[e] = HM[e]; sMD[e] = MD[e]; }
  const long plane = (long)g.Ky * g.Kxp;
  const int pp = tid % pp_per_block;
  const int k = tid / pp_per_block;
  const long pt = (long)blockIdx.x * pp_per_block + pp;
  const bool active = (k < nz) && (pt < plane);
  if (active) sq[k * pp_per_block + pp] = q[k * plane + pt];
  __syncthreads();
  if (active) {
    double2 acc = make_double2(0.0, 0.0);
    for (int kk = 0; kk < nz; ++kk) {
      const double w = sHM[k * nz + kk];
      const double2 v = sq[kk * pp_per_block + pp];
      acc.x += w * v.x;
      acc.y += w * v.y;
    }
    const int i = (int)pt % g.Kxp;
    const int j = compact_row_to_full((int)pt / g.Kxp, g);
    const double kx = kx_of(i, g), ky = ky_of(j, g);
    const double den = -(kx * kx + ky * ky) + evals[k];
    double2 pm = make_double2(acc.x * inv_htot / den, acc.y * inv_htot / den);
    if ((i == 0 && j == 0) || i >= g.Kx) pm = make_double2(0.0, 0.0);
    sm[k * pp_per_block + pp] = pm;
  }
  __syncthreads();
  if (active) {
    double2 acc = make_double2(0.0, 0.0);
    for (int kk = 0; kk < nz; ++kk) {
      const double w = sMD[k * nz + kk];
      const double2 v = sm[kk * pp_per_block + pp];
      acc.x += v.x * w;
      acc.y += v.y * w;
    }
    psi[k * plane + pt] = cmk<cplx<Tout>>((Tout)acc.x, (Tout)acc.y);
  }
}

// ---------------------------------------------------------------------------
// Argument blocks of the RHS passes (kernels_ct.hpp).
//
// Column pass P1, spectral -> mixed: y-inverse FFT of fields built from
// compact inputs.  For the Jacobian these are u = -i ky psi, v = i kx psi,
// q_x = i kx q, q_y = i ky q (QGN_Module.f90:222-225 builds u, v the same
// way).  mult: 0 -> x1, 1 -> -i ky, 2 -> +i kx, 3 -> +i ky.
// ---------------------------------------------------------------------------
struct ColInvArgs {
  const void* src[6];   // compact inputs (double2 when src_is_double, else cplx<R>)
  int src_is_double[6];
  int mult[6];
  void* dst[6];         // mixed outputs (cplx<R>)
  int nf;
};

// Row pass P2: x-c2r of u, v, q_x, q_y; J = u q_x + v q_y (and |u| u,
// |u| v on the bottom layer); x-r2c keeping the Kx dealiased bins.
struct RowJacArgs {
  const void* U; const void* V; const void* QX; const void* QY;  // mixed (nz, ny, Kxp)
  void* J;                                                       // mixed (nz, ny, Kxp)
  void* DU; void* DV;                                            // (members, ny, Kxp), bottom layers
  int drag;
};

// Column pass P3: y-forward FFT of J and the rest of GetRHS.
// With drag set, the bottom layer of each member also gets the quadratic
// drag cdrag (i ky du_hat - i kx dv_hat) of the row pass's |u| u, |u| v.
struct ColRhsArgs {
  const void* J;         // mixed cplx<R>
  const void* M;         // mixed cplx<R>: the closure network's forcing, or null (RHS += FFT_y(M))
  const double2* q;      // compact
  const void* q32;       // compact cplx<R>, q rounded to R (null: read q)
  const void* psi;       // compact cplx<R>
  void* out;             // compact cplx<S>
  const double* ubar; const double* qybar;   // per layer
  const double* ekman;   // per member: r0 Htot / H_nz, applied on its bottom layer
  const void* DU; const void* DV;   // (members, ny, Kxp) cplx<R>
  const double* cdrag;   // per member: C_d Htot / H_nz
  double beta, A2, A8;
  int explicit_visc;
  int drag;
};

// ---------------------------------------------------------------------------
// x-c2r pre-processing: Z[k] = (X[k] + conj X[N2-k]) + i e^{+2 pi i k/nx}
// (X[k] - conj X[N2-k]), k in [0, N2); the imaginary parts of the kx = 0
// and kx = nx/2 bins are ignored, as DFTI does.  z = IFFT_N2(Z) then holds
// x[2m] + i x[2m+1].
// ---------------------------------------------------------------------------
// From the already loaded bins a = X[k], b = X[N2-k] and table value
// w = twr[k] = exp(-2 pi i k / nx).
template <typename R>
__device__ __forceinline__ cplx<R> c2r_pre_w(cplx<R> a, cplx<R> b, int k, cplx<R> w, R scale = (R)1) {
  using Cx = cplx<R>;
  if (k == 0) { a.y = (R)0; b.y = (R)0; }
  const Cx s = cmk<Cx>(a.x + b.x, a.y - b.y);   // X[k] + conj X[N2-k]
  const Cx d = cmk<Cx>(a.x - b.x, a.y + b.y);   // X[k] - conj X[N2-k]
  w.y = -w.y;                                   // e^{+2 pi i k / nx}
  const Cx wd = cmul(w, d);
  return cmk<Cx>((s.x - wd.y) * scale, (s.y + wd.x) * scale);  // s + i*wd
}
// c2r_pre_w with scale 1 and the twiddle product folded into FMAs:
// s + i conj(w) d as four FMAs instead of a complex product and two adds
template <typename R>
__device__ __forceinline__ cplx<R> c2r_pre_fma(cplx<R> a, cplx<R> b, int k, cplx<R> w) {
  using Cx = cplx<R>;
  if (k == 0) { a.y = (R)0; b.y = (R)0; }
  const Cx s = cmk<Cx>(a.x + b.x, a.y - b.y);   // X[k] + conj X[N2-k]
  const Cx d = cmk<Cx>(a.x - b.x, a.y + b.y);   // X[k] - conj X[N2-k]
  // conj(w) d = (w.x d.x + w.y d.y, w.x d.y - w.y d.x); s + i conj(w) d
  return cmk<Cx>(fma_r(-w.x, d.y, fma_r(w.y, d.x, s.x)), fma_r(w.x, d.x, fma_r(w.y, d.y, s.y)));
}
// r2c_post's bin from Z[k] = aa and Z[N2-k] = bb with the twiddle w:
// (S + w O) / 2 with S = aa + conj bb, O = -i (aa - conj bb), the product
// folded into FMAs and the halving applied once
template <typename R>
__device__ __forceinline__ cplx<R> r2c_post_fma(cplx<R> aa, cplx<R> bb, cplx<R> w) {
  using Cx = cplx<R>;
  const R sx = aa.x + bb.x, sy = aa.y - bb.y;
  const R ox = aa.y + bb.y, oy = bb.x - aa.x;   // -i (aa - conj bb)
  return cmk<Cx>((R)0.5 * fma_r(w.x, ox, fma_r(-w.y, oy, sx)), (R)0.5 * fma_r(w.x, oy, fma_r(w.y, ox, sy)));
}
template <typename R>
__device__ __forceinline__ cplx<R> c2r_pre_ab(cplx<R> a, cplx<R> b, int k, const cplx<R>* __restrict__ twr,
                                              R scale) {
  return c2r_pre_w<R>(a, b, k, twr[k], scale);
}

template <typename R>
__device__ __forceinline__ cplx<R> c2r_pre(const cplx<R>* __restrict__ row, int Kin, int k, int N2,
                                           const cplx<R>* __restrict__ twr, R scale) {
  using Cx = cplx<R>;
  Cx a = ld_or_zero(row, k, k < Kin);
  const int kk = N2 - k;
  Cx b = ld_or_zero(row, kk, kk < Kin);
  return c2r_pre_ab<R>(a, b, k, twr, scale);
}

// The same, mixed row r (a lane-dependent row of a workgroup's block of
// rows, mix_at layout) of the buffer rs: out-of-range bins load as zero.
template <typename R>
__device__ __forceinline__ cplx<R> c2r_pre_rs(__amdgpu_buffer_rsrc_t rs, int r, int Kxp, int Kin, int k, int N2,
                                              const cplx<R>* __restrict__ twr, R scale) {
  using Cx = cplx<R>;
  constexpr int SZ = (int)sizeof(Cx);
  const int kk = N2 - k;
  const Cx a = rsrc_ld<Cx>(rs, k < Kin ? mix_at<R>(r, k, Kxp) * SZ : kRowOob);
  const Cx b = rsrc_ld<Cx>(rs, kk < Kin ? mix_at<R>(r, kk, Kxp) * SZ : kRowOob);
  return c2r_pre_ab<R>(a, b, k, twr, scale);
}

// x-r2c post-processing: X[k] = (Z[k] + conj Z[N2-k])/2
//                              - i/2 e^{-2 pi i k/nx} (Z[k] - conj Z[N2-k])
template <typename R>
__device__ __forceinline__ cplx<R> r2c_post(const cplx<R>* __restrict__ zs, int k, int N2,
                                            const cplx<R>* __restrict__ twr) {
  using Cx = cplx<R>;
  const Cx a = zs[k == N2 ? 0 : k];
  const Cx b = zs[k == 0 ? 0 : N2 - k];
  const Cx e = cmk<Cx>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
  const Cx dd = cmk<Cx>((R)0.5 * (a.x - b.x), (R)0.5 * (a.y + b.y));  // (Z - conj Zr)/2
  const Cx o = cmk<Cx>(dd.y, -dd.x);                                   // -i * dd
  return cadd(e, cmul(twr[k], o));
}

// ---------------------------------------------------------------------------
// IMEX stage combination (ARK43*.f90 stage blocks):
//   q_tmp = Mq (q + dt * (sum_j ae_j N_j + sum_j ai_j L_j)),
//   L_j = -(A2 k2 + A8L_k k8) Y_j, Mq = 1 / (1 + dt/4 (A2 k2 + A8L_k k8)),
// with Y_1 = q.  Also serves the explicit steppers (lam = 0, Mq = 1).
// ---------------------------------------------------------------------------
struct StageArgs {
  const double2* q;
  const void* N[6];       // cplx<S>
  const double2* Y[6];    // Y[0] unused (= q)
  double ae[6], ai[6];
  int ns;                 // number of previous stages used
  double dt;
  const double* dtv;      // per-member dt on the device (nullptr -> dt)
  const double* a8l;      // per-layer A8 (nullptr -> A8 below)
  double A8, A2;
  int implicit_visc;      // use Mq and L_j terms
  double2* out;           // may be null where nothing reads the stage value in FP64
  float2* out32;          // MFMA projection kernels: FP32 copy of the stage value (may be null)
  // recompute != 0 (implicit_visc only): the earlier stage values Y_2..Y_ns
  // are recomputed from q and N instead of read from Y, with the tables
  // ye[j][k] = ae, yi[j][k] = ai of stage j+1 on stage k+1 (k < j)
  int recompute;
  double ye[6][6], yi[6][6];
};

// One element's stage combination over the first n earlier stages:
//   mq (q + dt (sum_k ae_k N_k + sum_k ai_k (-lam Y_k))),  Y_0 = q,
// explicit sum first, as the stage kernel that stored a stage value formed
// it, so a recomputed Y_k repeats its arithmetic.
__device__ __forceinline__ double2 stage_combine(int n, const double* ae, const double* ai, const double2* nd,
                                                 const double2* y, double2 q, bool impl, double lam, double mq,
                                                 double dt) {
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (k >= n) break;
    sx += ae[k] * nd[k].x;
    sy += ae[k] * nd[k].y;
  }
  if (impl) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k >= n) break;
      const double2 yk = k == 0 ? q : y[k];
      sx += ai[k] * (-lam * yk.x);
      sy += ai[k] * (-lam * yk.y);
    }
  }
  return make_double2(mq * (q.x + dt * sx), mq * (q.y + dt * sy));
}

// lam = A2 k2 + A8 k2^4 of compact point pt, layer kl of the batch
__device__ __forceinline__ double visc_lambda_ij(const Geom& g, int i, int j, int kl, const double* a8l, double A8,
                                                 double A2) {
  const double kx = kx_of(i, g), ky = ky_of(j, g);
  const double k2 = kx * kx + ky * ky;
  const double k4 = k2 * k2;
  const double a8 = a8l ? a8l[kl] : A8;
  return A2 * k2 + a8 * (k4 * k4);
}
__device__ __forceinline__ double visc_lambda(const Geom& g, long pt, int kl, const double* a8l, double A8,
                                              double A2) {
  return visc_lambda_ij(g, (int)pt % g.Kxp, compact_row_to_full((int)pt / g.Kxp, g), kl, a8l, A8, A2);
}

// ARK stage value of element e (layer kl of the batch, point pt of its
// plane), the stage blocks of ARK43*.f90:
//   q_tmp = Mq (q + dt sum_j (ae_j N_j + ai_j L_j)),  L_j = -lam Y_j (Y_0 = q),
//   Mq = 1 / (1 + dt lam / 4),  lam = A2 k2 + A8 k2^4,
// with Mq = 1 and no L_j without implicit viscosity.  NS, the number of
// previous stages, is a compile-time count, so all of the element's loads
// are issued before the first sum.
// The same with the element's kx index i and full ky row j given.
template <typename S, int NS>
__device__ __forceinline__ double2 stage_value_ij(const StageArgs& a, const Geom& g, long e, int i, int j, int kl,
                                                  double dt);
template <typename S, int NS>
__device__ __forceinline__ double2 stage_value(const StageArgs& a, const Geom& g, long e, long pt, int kl,
                                               double dt) {
  return stage_value_ij<S, NS>(a, g, e, (int)pt % g.Kxp, compact_row_to_full((int)pt / g.Kxp, g), kl, dt);
}
template <typename S, int NS>
__device__ __forceinline__ double2 stage_value_ij(const StageArgs& a, const Geom& g, long e, int i, int j, int kl,
                                                  double dt) {
  constexpr int M = NS > 0 ? NS : 1;
  const double2 q = a.q[e];
  cplx<S> nv[M];
#pragma unroll
  for (int jj = 0; jj < NS; ++jj) nv[jj] = reinterpret_cast<const cplx<S>*>(a.N[jj])[e];
  double2 yv[M];
  const bool rec = a.recompute != 0;
  if (a.implicit_visc && !rec) {
#pragma unroll
    for (int jj = 1; jj < NS; ++jj) yv[jj] = a.Y[jj][e];
  }
  double2 nd[M];
#pragma unroll
  for (int jj = 0; jj < NS; ++jj) nd[jj] = make_double2((double)nv[jj].x, (double)nv[jj].y);
  double lam = 0.0, mq = 1.0;
  if (a.implicit_visc) {
    lam = visc_lambda_ij(g, i, j, kl, a.a8l, a.A8, a.A2);
    mq = 1.0 / (1.0 + 0.25 * dt * lam);
    if (rec) {
      yv[0] = q;
#pragma unroll
      for (int jj = 1; jj < NS; ++jj) yv[jj] = stage_combine(jj, a.ye[jj], a.yi[jj], nd, yv, q, true, lam, mq, dt);
    }
  }
  return stage_combine(NS, a.ae, a.ai, nd, yv, q, a.implicit_visc != 0, lam, mq, dt);
}

template <typename S, int NS>
__global__ void k_stage(StageArgs a, Geom g) {
  const long plane = (long)g.Ky * g.Kxp;
  const long n = plane * g.nz;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int z = (int)(e / plane);
    const double dt = a.dtv ? a.dtv[z / g.nzm] : a.dt;
    a.out[e] = stage_value<S, NS>(a, g, e, e % plane, z, dt);
  }
}

// Stage combination fused with GetPsi: a workgroup holds PP spectral points
// x nz layers, one thread per (point, layer).  Each thread forms q_tmp of its
// layer (as k_stage) and writes it; the workgroup then applies the
// layer -> mode -> layer projection of k_project through LDS and writes psi.
template <typename S, typename Tp, int NS>
__global__ void __launch_bounds__(256) k_stage_proj(StageArgs a, Geom g, const double* __restrict__ HM,
                                                    const double* __restrict__ MD, const double* __restrict__ evals,
                                                    double inv_htot, int pp_per_block, cplx<Tp>* __restrict__ psi) {
  extern __shared__ __align__(16) double lds_d[];
  const int nz = g.nzm;
  const int mb = blockIdx.y;           // member
  HM += mb * nz * nz;
  MD += mb * nz * nz;
  evals += mb * nz;
  const double dt = a.dtv ? a.dtv[mb] : a.dt;
  double* sHM = lds_d;
  double* sMD = sHM + nz * nz;
  double2* sq = reinterpret_cast<double2*>(sMD + nz * nz);  // [nz][pp]
  double2* sm = sq + nz * pp_per_block;                      // [nz][pp]
  const int tid = threadIdx.x;
  for (int e = tid; e < nz * nz; e += blockDim.x) { sHM[e] = HM[e]; sMD[e] = MD[e]; }
  const long plane = (long)g.Ky * g.Kxp;
  const int pp = tid % pp_per_block;
  const int k = tid / pp_per_block;
  const int kl = mb * nz + k;          // layer within the batch
  const long pt = (long)blockIdx.x * pp_per_block + pp;
  const bool active = (k < nz) && (pt < plane);
  int i = 0, j = 0;
  double k2 = 0.0;
  if (active) {
    i = (int)pt % g.Kxp;
    j = compact_row_to_full((int)pt / g.Kxp, g);
    const double kx = kx_of(i, g), ky = ky_of(j, g);
    k2 = kx * kx + ky * ky;
    const long e = kl * plane + pt;
    const double2 qt = stage_value<S, NS>(a, g, e, pt, kl, dt);
    a.out[e] = qt;
    sq[k * pp_per_block + pp] = qt;
  }
  __syncthreads();
  if (active) {
    double2 acc = make_double2(0.0, 0.0);
    for (int kk = 0; kk < nz; ++kk) {
      const double w = sHM[k * nz + kk];
      const double2 v = sq[kk * pp_per_block + pp];
      acc.x += w * v.x;
      acc.y += w * v.y;
    }
    const double den = -k2 + evals[k];
    double2 pm = make_double2(acc.x * inv_htot / den, acc.y * inv_htot / den);
    if ((i == 0 && j == 0) || i >= g.Kx) pm = make_double2(0.0, 0.0);
    sm[k * pp_per_block + pp] = pm;
  }
  __syncthreads();
  if (active) {
    double2 acc = make_double2(0.0, 0.0);
    for (int kk = 0; kk < nz; ++kk) {
      const double w = sMD[k * nz + kk];
      const double2 v = sm[kk * pp_per_block + pp];
      acc.x += v.x * w;
      acc.y += v.y * w;
    }
    psi[kl * plane + pt] = cmk<cplx<Tp>>((Tp)acc.x, (Tp)acc.y);
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 MFMA projection (mixed precision, nz <= 32).
// psi = MD diag(1/(-k2 + EVals)) (HM / Htot) q as two 16x16x16 contractions
// on v_mfma_f32_16x16x16bf16_1k.  Each FP32 operand x is split into
// hi = bf16(x) and lo = bf16(x - hi); a product is accumulated in FP32 as
// hi*hi + hi*lo + lo*hi (3 MFMAs), about 2^-16 relative per product.
// Operand layouts (checked on gfx950): A[i][k] in lane l at i = l%16,
// k = 4(l/16)+r; B[k][j] at j = l%16, k = 4(l/16)+r; D[i][j] at
// i = 4(l/16)+r, j = l%16.  D of the first product is the B operand of the
// second as it stands.
// ---------------------------------------------------------------------------
typedef short qgn_short4 __attribute__((ext_vector_type(4)));
typedef float qgn_float4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }
__device__ __forceinline__ short bf16_lo(float x, short hi) { return bf16_rne(x - bf16_to_f(hi)); }
// The split of four floats into an MFMA operand pair (hi = bf16(x), lo =
// bf16(x - hi)) with gfx950's packed conversion (v_cvt_pk_bf16_f32, round
// to nearest even like bf16_rne): two conversions per pair of values.
__device__ __forceinline__ void bf16_split4(const float (&x)[4], qgn_short4& hi, qgn_short4& lo) {
  typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
  typedef float f2v __attribute__((ext_vector_type(2)));
  unsigned h[2], l[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const bf2v hv = __builtin_convertvector((f2v){x[2 * p], x[2 * p + 1]}, bf2v);
    __builtin_memcpy(&h[p], &hv, 4);
    const float h0 = __uint_as_float(h[p] << 16), h1 = __uint_as_float(h[p] & 0xffff0000u);
    const bf2v lv = __builtin_convertvector((f2v){x[2 * p] - h0, x[2 * p + 1] - h1}, bf2v);
    __builtin_memcpy(&l[p], &lv, 4);
  }
  __builtin_memcpy(&hi, h, 8);
  __builtin_memcpy(&lo, l, 8);
}
__device__ __forceinline__ qgn_float4 mfma3(const qgn_short4& ah, const qgn_short4& al, const qgn_short4& bh,
                                            const qgn_short4& bl, qgn_float4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
  return c;
}

// The mixed path's stage values alone, as a streaming pass (the projection
// follows in k_proj32_mfma16): grid (plane / 512, layers of the batch),
// each thread two adjacent points of its layer.  Writes the stage
// value rounded to FP32 (and the FP64 value where a.out is set); every
// stage kernel forms the sums with stage_value_ij, so the values agree to
// the bit.
template <int NS>
__global__ void __launch_bounds__(256) k_stage_ew(StageArgs a, Geom g) {
  const int z = blockIdx.y;
  const int pt = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);   // Kxp is even: pt, pt + 1 share a row
  if (pt >= g.Ky * g.Kxp) return;
  // compact row and column: a single-precision quotient (pt < 2^24),
  // corrected to the exact one
  int jc = (int)((float)pt * (1.0f / (float)g.Kxp));
  int i = pt - jc * g.Kxp;
  if (i < 0) { --jc; i += g.Kxp; }
  else if (i >= g.Kxp) { ++jc; i -= g.Kxp; }
  const int j = compact_row_to_full(jc, g);
  const long e = (long)z * g.Ky * g.Kxp + pt;
  const double dt = a.dtv ? a.dtv[z / g.nzm] : a.dt;
  const double2 q0 = stage_value_ij<float, NS>(a, g, e, i, j, z, dt);
  const double2 q1 = stage_value_ij<float, NS>(a, g, e + 1, i + 1, j, z, dt);
  if (a.out) {
    a.out[e] = q0;
    a.out[e + 1] = q1;
  }
  *reinterpret_cast<float4*>(a.out32 + e) = make_float4((float)q0.x, (float)q0.y, (float)q1.x, (float)q1.y);
}

// GetPsi of the FP32 stage value q32 (nz <= 16) as a split-BF16 MFMA
// projection (v_mfma_f32_16x16x16bf16_1k): each FP32 operand x is split
// into hi = bf16(x) and lo = bf16(x - hi), and a product accumulates
// hi*hi + hi*lo + lo*hi in FP32.  Per tile of 32 points x 16 layers each
// of the 4 waves projects 8 points (16 columns: re, im): modal amplitudes
// D1[mode][col] = (HM/Htot)[mode][layer] Q[layer][col], the modal Helmholtz
// inversion D1 / (lambda_mode - k^2), then psi[layer][col] = Modes D.
// Groups of T tiles are walked grid-stride, the next group's loads in
// flight while this one is projected; psi goes out through LDS in whole
// 256-byte rows of a layer.
template <int T>
__global__ void __launch_bounds__(256) k_proj32_mfma16(const float2* __restrict__ q32, Geom g,
                                                       const double* __restrict__ HM, const double* __restrict__ MD,
                                                       const double* __restrict__ evals, double inv_htot,
                                                       float2* __restrict__ psi) {
  __shared__ float2 sq[T][16][32];    // [tile][layer][point]: q32, then psi
  const int nz = g.nzm;
  const int mb = blockIdx.y;         // member
  HM += mb * nz * nz;
  MD += mb * nz * nz;
  evals += mb * nz;
  const long moff = (long)mb * nz * g.Ky * g.Kxp;
  const int tid = threadIdx.x;
  const long plane = (long)g.Ky * g.Kxp;
  const long ngroups = (plane + 32 * T - 1) / (32 * T);
  const float kxp_inv = 1.0f / (float)g.Kxp;
  const int l = tid & 63;
  const int wv = tid >> 6;
  const int row = l & 15, kb = 4 * (l >> 4);
  qgn_short4 a1h, a1l, a2h, a2l, bh, bl;
  double ev[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int kk = kb + r;
    const float m1 = (row < nz && kk < nz) ? (float)(HM[row * nz + kk] * inv_htot) : 0.f;
    const float m2 = (row < nz && kk < nz) ? (float)MD[row * nz + kk] : 0.f;
    a1h[r] = bf16_rne(m1);
    a1l[r] = bf16_lo(m1, a1h[r]);
    a2h[r] = bf16_rne(m2);
    a2l[r] = bf16_lo(m2, a2h[r]);
    ev[r] = kk < nz ? evals[kk] : 1.0;
  }
  // tile t, layer k of a group: 32 points = 256 bytes; 16 threads per row,
  // 2 points each (plane is even, so pairs never straddle its end).  The
  // next group's loads are in flight while this group is projected.
  const int lk = tid >> 4, lp2 = 2 * (tid & 15);
  float4 vn[T];
  auto fetch = [&](long grp) {
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const long pt = grp * (32 * T) + 32 * t + lp2;
      vn[t] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lk < nz && pt < plane) vn[t] = *reinterpret_cast<const float4*>(q32 + moff + lk * plane + pt);
    }
  };
  if (blockIdx.x < ngroups) fetch(blockIdx.x);
  for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const long p0 = grp * (32 * T);
    __syncthreads();   // the previous group's psi rows have been read out of sq
#pragma unroll
    for (int t = 0; t < T; ++t) *reinterpret_cast<float4*>(&sq[t][lk][lp2]) = vn[t];
    if (grp + gridDim.x < ngroups) fetch(grp + gridDim.x);
    __syncthreads();
    qgn_float4 pv[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      float xs[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float2 v = sq[t][kb + r][8 * wv + (row >> 1)];
        xs[r] = (row & 1) ? v.y : v.x;
      }
      bf16_split4(xs, bh, bl);
      qgn_float4 d = {0.f, 0.f, 0.f, 0.f};
      d = mfma3(a1h, a1l, bh, bl, d);
      const int pt = (int)(p0 + 32 * t + 8 * wv + (row >> 1));
      bool zero_mode = pt >= plane;
      double k2 = 0.0;
      if (!zero_mode) {
        // compact row and column of pt: a single-precision quotient
        // (pt < 2^24), corrected to the exact one
        int jc = (int)((float)pt * kxp_inv);
        int i = pt - jc * g.Kxp;
        if (i < 0) { --jc; i += g.Kxp; }
        else if (i >= g.Kxp) { ++jc; i -= g.Kxp; }
        const int j = compact_row_to_full(jc, g);
        const double kx = kx_of(i, g), ky = ky_of(j, g);
        k2 = kx * kx + ky * ky;
        zero_mode = (i == 0 && j == 0) || i >= g.Kx;
      }
      float pms[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mode = kb + r;
        // modal inversion by the hardware reciprocal (1 ulp): well inside
        // the split-BF16 products' 2^-16
        pms[r] = (zero_mode || mode >= nz) ? 0.f : d[r] * __builtin_amdgcn_rcpf((float)(-k2 + ev[r]));
      }
      bf16_split4(pms, bh, bl);
      qgn_float4 pz = {0.f, 0.f, 0.f, 0.f};
      pv[t] = mfma3(a2h, a2l, bh, bl, pz);
    }
    __syncthreads();   // every wave has read its B operands out of sq
    float* sp = reinterpret_cast<float*>(&sq[0][0][0]);
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) sp[2 * ((t * 16 + kb + r) * 32 + 8 * wv + (row >> 1)) + (row & 1)] = pv[t][r];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int k = tid >> 4, pp2 = 2 * (tid & 15);
      const long pt = p0 + 32 * t + pp2;
      if (k < nz && pt < plane) {
        const float4 v = *reinterpret_cast<const float4*>(&sq[t][k][pp2]);
        float2* dst = psi + moff + k * plane + pt;
        if (pt + 1 < plane) *reinterpret_cast<float4*>(dst) = v;
        else *dst = make_float2(v.x, v.y);
      }
    }
  }
}


// Stage combination (or, with a.ns = 0 and a.out = nullptr, a plain copy of
// q) followed by the MFMA projection, for nz <= NZT layers (NZT = 16 or 32:
// the nz x nz matrices as (NZT/16)^2 tiles of 16 x 16).  A workgroup walks
// tiles of 16 points x NZT layers (grid-stride); per tile each thread forms
// the FP64 stage sums of NZT/16 (point, layer) elements, then waves 0 and 1
// project 8 points each (16 columns: re, im).  The split-BF16 matrix
// operands are built once per workgroup and stay in registers while the
// workgroup walks its tiles (the 32-layer matrices are 4x those of 16).
// The tile buffer alternates between two halves, so one barrier per tile
// separates its writers from its readers (each half is rewritten only after
// the next tile's barrier, by which time its readers are done).
template <typename S, int NZT, int NS>
__global__ void __launch_bounds__(256) k_stage_proj_mfma(StageArgs a, Geom g, const double* __restrict__ HM,
                                                         const double* __restrict__ MD,
                                                         const double* __restrict__ evals, double inv_htot,
                                                         float2* __restrict__ psi) {
  constexpr int NT = NZT / 16;          // 16-row tiles of modes and of layers
  __shared__ double2 sq[2][NZT][16];    // [half][layer][point]
  const int nz = g.nzm;
  const int mb = blockIdx.y;            // member
  HM += mb * nz * nz;
  MD += mb * nz * nz;
  evals += mb * nz;
  const long moff = (long)mb * nz * g.Ky * g.Kxp;   // first element of the member
  const double dt = a.dtv ? a.dtv[mb] : a.dt;
  const int tid = threadIdx.x;
  const long plane = (long)g.Ky * g.Kxp;
  const long ntiles = (plane + 15) / 16;
  const int l = tid & 63;
  const int wv = tid >> 6;                     // waves 0, 1: points 8*wv .. 8*wv+7 of a tile
  const int row = l & 15, kb = 4 * (l >> 4);   // A: i = row; B, D: j = row
  qgn_short4 a1h[NT][NT], a1l[NT][NT], a2h[NT][NT], a2l[NT][NT];
  double ev[NT][4];
  auto build_ops = [&]() {
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
      for (int tk = 0; tk < NT; ++tk)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * ti + row, k = 16 * tk + kb + r;
          const bool in = i < nz && k < nz;
          const float m1 = in ? (float)(HM[i * nz + k] * inv_htot) : 0.f;   // (HM/Htot)[mode i][layer k]
          const float m2 = in ? (float)MD[i * nz + k] : 0.f;                // Modes[layer i][mode k]
          a1h[ti][tk][r] = bf16_rne(m1);
          a1l[ti][tk][r] = bf16_lo(m1, a1h[ti][tk][r]);
          a2h[ti][tk][r] = bf16_rne(m2);
          a2l[ti][tk][r] = bf16_lo(m2, a2h[ti][tk][r]);
        }
#pragma unroll
    for (int mi = 0; mi < NT; ++mi)
#pragma unroll
      for (int r = 0; r < 4; ++r) ev[mi][r] = 16 * mi + kb + r < nz ? evals[16 * mi + kb + r] : 1.0;
  };
  if (tid < 128) build_ops();
  int half = 0;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x, half ^= 1) {
#pragma unroll
    for (int h = 0; h < NT; ++h) {
      const int pp = tid & 15, k = (tid >> 4) + 16 * h;
      const long pt = tile * 16 + pp;
      double2 qt = make_double2(0.0, 0.0);
      if (k < nz && pt < plane) {
        const long e = moff + k * plane + pt;
        qt = stage_value<S, NS>(a, g, e, pt, mb * nz + k, dt);
        if (a.out) a.out[e] = qt;
        if (a.out32) a.out32[e] = make_float2((float)qt.x, (float)qt.y);
      }
      sq[half][k][pp] = qt;
    }
    __syncthreads();
    if (tid >= 128) continue;
    // first product: modal amplitudes D1[mode][column] = (HM/Htot)[mode][layer] Q[layer][column]
    qgn_float4 d[NT];
#pragma unroll
    for (int mi = 0; mi < NT; ++mi) d[mi] = qgn_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < NT; ++kt) {
      qgn_short4 bh, bl;
      float xs[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double2 v = sq[half][16 * kt + kb + r][8 * wv + (row >> 1)];
        xs[r] = (float)((row & 1) ? v.y : v.x);
      }
      bf16_split4(xs, bh, bl);
#pragma unroll
      for (int mi = 0; mi < NT; ++mi) d[mi] = mfma3(a1h[mi][kt], a1l[mi][kt], bh, bl, d[mi]);
    }
    const long pt = tile * 16 + 8 * wv + (row >> 1);
    bool zero_mode = pt >= plane;
    double k2 = 0.0;
    if (!zero_mode) {
      const int i = (int)pt % g.Kxp;
      const int j = compact_row_to_full((int)pt / g.Kxp, g);
      const double kx = kx_of(i, g), ky = ky_of(j, g);
      k2 = kx * kx + ky * ky;
      zero_mode = (i == 0 && j == 0) || i >= g.Kx;
    }
    // modal Helmholtz inversion; D1 as it stands is the B operand of the second product
    qgn_short4 ph[NT], pl[NT];
#pragma unroll
    for (int mi = 0; mi < NT; ++mi) {
      float pms[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mode = 16 * mi + kb + r;
        pms[r] = (zero_mode || mode >= nz) ? 0.f : d[mi][r] / (float)(-k2 + ev[mi][r]);
      }
      bf16_split4(pms, ph[mi], pl[mi]);
    }
    // second product: psi[layer][column] = Modes[layer][mode] pm[mode][column]
#pragma unroll
    for (int li = 0; li < NT; ++li) {
      qgn_float4 p = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mt = 0; mt < NT; ++mt) p = mfma3(a2h[li][mt], a2l[li][mt], ph[mt], pl[mt], p);
      if (pt < plane) {
        float* out = reinterpret_cast<float*>(psi);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int layer = 16 * li + kb + r;
          if (layer < nz) out[2 * (moff + layer * plane + pt) + (row & 1)] = p[r];
        }
      }
    }
  }
}

// Embedded error field and candidate update of ARK43 (ARK43*.f90 error and
// update blocks), or of any scheme written as q + dt * sum_j b_j (N_j + L_j):
//   E  = sum_j be_j (N_j + L_j),   qn = q + dt sum_j b_j (N_j + L_j), qn(0,0)=0
// QG-Leith partial sums (k_leith_partial, and k_update for the candidate of
// a graph attempt): one term 2 k^8 |v|^2 of a kx > 0 mode, with every
// product and sum rounded on its own (no contraction), so both kernels form
// bit-identical sums over the same decomposition (grid (nblk, nz) of 256
// threads, thread t of block b taking points b*256 + t + k*nblk*256).
__device__ __forceinline__ double leith_add(double s, const double2 v, long pt, const Geom& g) {
  const int i = (int)pt % g.Kxp;
  if (i == 0 || i >= g.Kx) return s;
  const int j = compact_row_to_full((int)pt / g.Kxp, g);
  const double kx = kx_of(i, g), ky = ky_of(j, g);
  const double k2 = __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
  const double k4 = __dmul_rn(k2, k2);
  const double m = __dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y));
  return __dadd_rn(s, __dmul_rn(__dmul_rn(2.0, __dmul_rn(k4, k4)), m));
}
// the workgroup's sum of s (256 threads) into *out, in a fixed order
__device__ __forceinline__ void leith_block_store(double s, double* out) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = ((red[0] + red[1]) + red[2]) + red[3];
}

struct UpdateArgs {
  const double2* q;
  const void* N[6];
  const double2* Y[6];
  double b[6], be[6];
  int ns;
  double dt;
  const double* dtv;  // per-member dt on the device (nullptr -> dt)
  const double* a8l;
  double A8, A2;
  int implicit_visc;
  double2* E;     // may be null
  int e32;        // E as float2 (the mixed path's error norm runs in FP32)
  double2* qn;
  double* lpart;  // QG-Leith partial sums of qn, lnblk per layer (null: none; grid (lnblk, nz) of 256)
  int lnblk;
  int recompute;  // as in StageArgs: Y_2..Y_ns from q and N
  double ye[6][6], yi[6][6];
};

// grid (x, nz): blockIdx.y is the layer, x strides over its plane
template <typename S>
__global__ void k_update(UpdateArgs a, Geom g) {
  const int plane = g.Ky * g.Kxp;
  const int z = blockIdx.y;
  const bool rec = a.recompute != 0 && a.implicit_visc;
  const double dt = a.dtv ? a.dtv[z / g.nzm] : a.dt;
  double ls = 0.0;   // QG-Leith partial sum of the candidate
  for (int pt = blockIdx.x * blockDim.x + threadIdx.x; pt < plane; pt += gridDim.x * blockDim.x) {
    const long e = (long)z * plane + pt;
    const double2 q = a.q[e];
    // all of the element's loads first
    double2 nd[6], yv[6];
#pragma unroll
    for (int jj = 0; jj < 6; ++jj) {
      if (jj >= a.ns) break;
      const cplx<S> v = reinterpret_cast<const cplx<S>*>(a.N[jj])[e];
      nd[jj] = make_double2((double)v.x, (double)v.y);
      if (a.implicit_visc && !rec && jj > 0) yv[jj] = a.Y[jj][e];
    }
    double lam = 0.0;
    if (a.implicit_visc) {
      lam = visc_lambda(g, pt, z, a.a8l, a.A8, a.A2);
      if (rec) {
        const double mq = 1.0 / (1.0 + 0.25 * dt * lam);
#pragma unroll
        for (int jj = 1; jj < 6; ++jj) {
          if (jj >= a.ns) break;
          yv[jj] = stage_combine(jj, a.ye[jj], a.yi[jj], nd, yv, q, true, lam, mq, dt);
        }
      }
    }
    double ex = 0.0, ey = 0.0, ux = 0.0, uy = 0.0;
#pragma unroll
    for (int jj = 0; jj < 6; ++jj) {
      if (jj >= a.ns) break;
      if (a.b[jj] == 0.0 && a.be[jj] == 0.0) continue;
      double tx = nd[jj].x, ty = nd[jj].y;
      if (a.implicit_visc) {
        const double2 y = jj == 0 ? q : yv[jj];
        tx += -lam * y.x;
        ty += -lam * y.y;
      }
      ex += a.be[jj] * tx;
      ey += a.be[jj] * ty;
      ux += a.b[jj] * tx;
      uy += a.b[jj] * ty;
    }
    if (a.E) {
      if (a.e32) reinterpret_cast<float2*>(a.E)[e] = make_float2((float)ex, (float)ey);
      else a.E[e] = make_double2(ex, ey);
    }
    double2 r = make_double2(q.x + dt * ux, q.y + dt * uy);
    if (pt == 0) r = make_double2(0.0, 0.0);
    a.qn[e] = r;
    if (a.lpart) ls = leith_add(ls, r, pt, g);
  }
  if (a.lpart) leith_block_store(ls, a.lpart + (long)z * a.lnblk + blockIdx.x);
}

// Plain axpy-type combination for the explicit steppers:
//   out = q + sum_j c_j X_j   (X_j compact cplx<S>), out(0,0)=0 optional
struct AxpyArgs {
  const double2* q;
  const void* X[4];
  double c[4];
  int nx;
  double2* out;
};
template <typename S>
__global__ void k_axpy(AxpyArgs a, Geom g) {
  const long n = (long)g.Ky * g.Kxp * g.nz;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double2 r = a.q[e];
    for (int jj = 0; jj < a.nx; ++jj) {
      const cplx<S> v = reinterpret_cast<const cplx<S>*>(a.X[jj])[e];
      r.x += a.c[jj] * (double)v.x;
      r.y += a.c[jj] * (double)v.y;
    }
    a.out[e] = r;
  }
}

// Linear combination into a compact cplx<S> array: out = sum_j c_j X_j.
template <typename S>
__global__ void k_lincomb(const void* X0, const void* X1, const void* X2, const void* X3, double c0, double c1,
                          double c2, double c3, int nxs, void* out, Geom g) {
  const long n = (long)g.Ky * g.Kxp * g.nz;
  const cplx<S>* xs[4] = {reinterpret_cast<const cplx<S>*>(X0), reinterpret_cast<const cplx<S>*>(X1),
                          reinterpret_cast<const cplx<S>*>(X2), reinterpret_cast<const cplx<S>*>(X3)};
  const double cs[4] = {c0, c1, c2, c3};
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double rx = 0.0, ry = 0.0;
    for (int jj = 0; jj < nxs; ++jj) {
      const cplx<S> v = xs[jj][e];
      rx += cs[jj] * (double)v.x;
      ry += cs[jj] * (double)v.y;
    }
    reinterpret_cast<double2*>(out)[e] = make_double2(rx, ry);
  }
}

// ---------------------------------------------------------------------------
// QG-Leith coefficient (ARK43_QGLeith.f90:62-73) by Parseval:
//   sum_xy q d / (nx ny) with q, d the physical fields of q_hat and
//   k8 q_hat equals sum_full k8 |Q|^2 / (nx ny)^2 over the Hermitian
//   extension Q of the half spectrum.  kx > 0 modes count twice; the kx = 0
//   column is first symmetrised, s(ky) = (q(ky) + conj q(-ky)) / 2, which is
//   what the c2r of a non-Hermitian kx = 0 column evaluates.
// Stage 1: per-block partial sums over kx > 0; stage 2: per layer, adds the
// kx = 0 column and the partials in a fixed order (deterministic).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_leith_partial(const double2* __restrict__ q, Geom g,
                                                     double* __restrict__ partial, int nblk) {
  const int z = blockIdx.y;
  const long plane = (long)g.Ky * g.Kxp;
  double s = 0.0;
  for (long pt = blockIdx.x * (long)blockDim.x + threadIdx.x; pt < plane; pt += (long)nblk * blockDim.x)
    s = leith_add(s, q[z * plane + pt], pt, g);
  leith_block_store(s, partial + z * nblk + blockIdx.x);
}

// Per layer: the kx = 0 column and the nblk partials (one per thread,
// nblk <= 256) in a fixed order, then A8L.  With cand (a graph attempt):
// the members that accepted take the candidate's partials in cand, which
// then replace theirs in `partial`; q is the state after the attempt.
__global__ void __launch_bounds__(256) k_leith_final(const double2* __restrict__ q, Geom g, double* partial,
                                                   const double* __restrict__ cand, const int* __restrict__ accept,
                                                   int nblk, double c12, double* __restrict__ a8l) {
  __shared__ double red[16];
  const int z = blockIdx.x;
  const long plane = (long)g.Ky * g.Kxp;
  const bool take = cand && accept[z / g.nzm];
  double s = 0.0;
  if ((int)threadIdx.x < nblk) {
    s = (take ? cand : partial)[z * nblk + threadIdx.x];
    if (take) partial[z * nblk + threadIdx.x] = s;
  }
  for (int j = threadIdx.x; j < g.ny; j += blockDim.x) {
    const int jc = full_row_to_compact(j, g);
    const int jm = (g.ny - j) % g.ny;
    const int jmc = full_row_to_compact(jm, g);
    double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
    if (jc >= 0) a = q[z * plane + (long)jc * g.Kxp];
    if (jmc >= 0) b = q[z * plane + (long)jmc * g.Kxp];
    if (j == 0) a = b = make_double2(0.0, 0.0);  // (0,0) is always zero
    const double sx = 0.5 * (a.x + b.x), sy = 0.5 * (a.y - b.y);
    const double ky = ky_of(j, g);
    const double k2 = ky * ky;
    const double k4 = k2 * k2;
    s += (k4 * k4) * (sx * sx + sy * sy);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += red[w];
    const double nn = (double)g.nx * (double)g.ny;
    a8l[z] = c12 * sqrt(t) / nn;
  }
}

// ---------------------------------------------------------------------------
// Full-spectrum helpers for the modal output of WriteQPmodal
// (QGN_Module.f90:381-416); arrays (nz, ny, nx/2+1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void full_k(long pt, const Geom& g, double* kx, double* ky) {
  const int i = (int)(pt % g.nxh);
  const int j = (int)(pt / g.nxh);
  *kx = kx_of(i, g);
  *ky = ky_of(j, g);
}

// out = W2 . diag(div ? 1/(-k2 + evals) : 1) . W1 . in over the layers at
// each point (W1, W2 row-major nz x nz, nullptr = identity); (0,0) of the
// divided quantity is zero (GetPsi, QGN_Module.f90:326-328).
__global__ void k_layer_mix_full(const double2* __restrict__ in, double2* __restrict__ out, Geom g,
                                 const double* __restrict__ W1, double s1, int div,
                                 const double* __restrict__ evals, const double* __restrict__ W2) {
  const long plane = (long)g.ny * g.nxh;
  const int nz = g.nz;
  for (long pt = blockIdx.x * (long)blockDim.x + threadIdx.x; pt < plane; pt += (long)gridDim.x * blockDim.x) {
    double kx, ky;
    full_k(pt, g, &kx, &ky);
    const double k2 = kx * kx + ky * ky;
    double2 m[64];
    for (int k = 0; k < nz; ++k) {
      double2 acc = make_double2(0.0, 0.0);
      if (W1) {
        for (int kk = 0; kk < nz; ++kk) {
          const double w = W1[k * nz + kk];
          const double2 v = in[kk * plane + pt];
          acc.x += w * v.x;
          acc.y += w * v.y;
        }
      } else {
        acc = in[k * plane + pt];
      }
      acc.x *= s1;
      acc.y *= s1;
      if (div) {
        const double den = -k2 + evals[k];
        acc = pt == 0 ? make_double2(0.0, 0.0) : make_double2(acc.x / den, acc.y / den);
      }
      m[k] = acc;
    }
    for (int k = 0; k < nz; ++k) {
      double2 acc = make_double2(0.0, 0.0);
      if (W2) {
        for (int kk = 0; kk < nz; ++kk) {
          const double w = W2[k * nz + kk];
          acc.x += m[kk].x * w;
          acc.y += m[kk].y * w;
        }
      } else {
        acc = m[k];
      }
      out[k * plane + pt] = acc;
    }
  }
}

// u = -i ky psi, v = i kx psi on the full spectrum (QGN_Module.f90:222-225)
__global__ void k_uv_full(const double2* __restrict__ psi, double2* __restrict__ u, double2* __restrict__ v,
                          Geom g) {
  const long plane = (long)g.ny * g.nxh;
  const long n = plane * g.nz;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double kx, ky;
    full_k(e % plane, g, &kx, &ky);
    const double2 p = psi[e];
    u[e] = make_double2(ky * p.y, -ky * p.x);
    v[e] = make_double2(-kx * p.y, kx * p.x);
  }
}

// physical products u q, v q (QGN_Module.f90:237,240)
__global__ void k_prod_phys(const double* __restrict__ u, const double* __restrict__ v,
                            const double* __restrict__ q, double* __restrict__ uq, double* __restrict__ vq,
                            long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    uq[e] = u[e] * q[e];
    vq[e] = v[e] * q[e];
  }
}

// jaco_hat = i kx uq_hat + i ky vq_hat (QGN_Module.f90:244-247)
__global__ void k_jac_full(const double2* __restrict__ uq, const double2* __restrict__ vq,
                           double2* __restrict__ jac, Geom g) {
  const long plane = (long)g.ny * g.nxh;
  const long n = plane * g.nz;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double kx, ky;
    full_k(e % plane, g, &kx, &ky);
    const double2 a = uq[e], b = vq[e];
    jac[e] = make_double2(-kx * a.y - ky * b.y, kx * a.x + ky * b.x);
  }
}

// ---------------------------------------------------------------------------
// Device-side ARK43 step control of a batch (ARK43*.f90 accept/reject and
// PI.3.4 blocks), one thread per member, after the error norm of an attempt:
//   err1 = dt max|c2r(E)| / (nx ny)
//   err1 > TOL: dt *= reject_factor                              (reject)
//   else: t += dt, dt = min(dt (safety TOL/err1)^0.075 (err0/err1)^0.1,
//         dt_max), err0 = err1                                   (accept)
// with the arithmetic of the single-model host controller.  A non-finite
// err1 sets status 1 and more than max_rejects consecutive rejections
// status 2; a member with a status keeps its state and time from then on.
// ---------------------------------------------------------------------------
struct BatchCtl {
  double *t, *dt, *err0, *last;
  long long *steps, *attempts, *rejects;
  int *nrej, *status, *accept;
  unsigned long long* maxbits;         // per member, from k_max_final, or formed here from partial
  const unsigned long long* partial;   // npart per-block maxima per member (null: maxbits is set)
  int npart;
  double* report;   // single model: accept, t, dt, err1, status, err0 of member 0 (or nullptr)
  double TOL, safety, reject_factor, dt_max, nn;
  int max_rejects, nm;
};

__device__ void batch_control_member(const BatchCtl& c, int m) {
  c.accept[m] = 0;
  if (c.status[m]) return;
  const double mx = __longlong_as_double((long long)c.maxbits[m]);
  const double dt = c.dt[m];
  const double err1 = dt * mx / c.nn;
  c.last[m] = err1;
  c.attempts[m] += 1;
  if (!isfinite(err1)) { c.status[m] = 1; return; }
  if (err1 > c.TOL) {
    c.dt[m] = c.reject_factor * dt;
    c.rejects[m] += 1;
    if (++c.nrej[m] > c.max_rejects) c.status[m] = 2;
    return;
  }
  c.nrej[m] = 0;
  c.accept[m] = 1;
  c.t[m] = c.t[m] + dt;
  c.dt[m] = fmin(dt * pow(c.safety * c.TOL / err1, 0.075) * pow(c.err0[m] / err1, 0.1), c.dt_max);
  c.err0[m] = err1;
  c.steps[m] += 1;
}

// one workgroup per member: the max over the error norm's per-block
// partials (when given), then the member's step control on thread 0
__global__ void __launch_bounds__(1024) k_batch_control(BatchCtl c) {
  const int m = blockIdx.x;
  if (c.partial) {
    __shared__ unsigned long long red[16];
    const unsigned long long* pm = c.partial + (long)m * c.npart;
    unsigned long long x = 0;
    for (int i = threadIdx.x; i < c.npart; i += blockDim.x) x = pm[i] > x ? pm[i] : x;
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_down(x, off, 64);
      x = o > x ? o : x;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w) x = red[w] > x ? red[w] : x;
      c.maxbits[m] = x;
    }
  }
  if (threadIdx.x != 0) return;
  batch_control_member(c, m);
  if (c.report && m == 0) {
    c.report[0] = c.accept[0];
    c.report[1] = c.t[0];
    c.report[2] = c.dt[0];
    c.report[3] = c.last[0];
    c.report[4] = c.status[0];
    c.report[5] = c.err0[0];
  }
}

// Single model: the caller's t and dt (and, on the first step, err0 = TOL)
// into the controller state of member 0; clears its reject count and status.
__global__ void k_control_set(double* t, double* dt, double* err0, int* nrej, int* status, double tv, double dtv,
                              double err0v) {
  if (threadIdx.x != 0) return;
  t[0] = tv;
  dt[0] = dtv;
  if (err0v >= 0.0) err0[0] = err0v;
  nrej[0] = 0;
  status[0] = 0;
}

// End of a graph attempt: the members that did not accept copy their state
// q into qn, so that qn holds every member's state (the next attempt's
// input); grid (blocks, members)
__global__ void k_batch_settle(const double2* __restrict__ q, double2* __restrict__ qn,
                               const int* __restrict__ accept, Geom g) {
  const int m = blockIdx.y;
  if (accept[m]) return;
  const long n = (long)g.nzm * g.Ky * g.Kxp;
  const long off = (long)m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    qn[off + e] = q[off + e];
}

// ---------------------------------------------------------------------------
// Isotropic KE and enstrophy spectra of the dealiased state, per layer:
//   KE(b) = 1/2 sum w k^2 |psi_hat|^2 / (nx ny)^2,  Z(b) = 1/2 sum w |q_hat|^2 / (nx ny)^2
// over the half plane with w = 2 for kx > 0; bin b = floor(|k|/dk + 1/2).
// The kx = 0 column enters symmetrised, s(ky) = (x(ky) + conj x(-ky))/2 at
// every ky with weight 1, which is the spectrum of the physical field that
// c2r produces from a non-Hermitian column.  Bins accumulate in LDS, then in
// out (ke then z, each nz x nbins), which the caller zeroes.
// ---------------------------------------------------------------------------
__global__ void k_spectra(const double2* __restrict__ q, const double2* __restrict__ psi, Geom g, double dk,
                          int nbins, double norm, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* hk = reinterpret_cast<double*>(lds_raw);
  double* hz = hk + nbins;
  const int z = blockIdx.y;
  for (int b = threadIdx.x; b < 2 * nbins; b += blockDim.x) hk[b] = 0.0;
  __syncthreads();
  const long plane = (long)g.Ky * g.Kxp;
  const double2* qz = q + z * plane;
  const double2* pz = psi + z * plane;
  for (long pt = blockIdx.x * (long)blockDim.x + threadIdx.x; pt < plane; pt += (long)gridDim.x * blockDim.x) {
    const int i = (int)pt % g.Kxp;
    if (i == 0 || i >= g.Kx) continue;
    const int j = compact_row_to_full((int)pt / g.Kxp, g);
    const double kx = kx_of(i, g), ky = ky_of(j, g);
    const double k2 = kx * kx + ky * ky;
    const int b = (int)floor(sqrt(k2) / dk + 0.5);
    if (b >= nbins) continue;
    const double2 a = pz[pt], c = qz[pt];
    atomicAdd(&hk[b], k2 * (a.x * a.x + a.y * a.y) * norm);
    atomicAdd(&hz[b], (c.x * c.x + c.y * c.y) * norm);
  }
  if (blockIdx.x == 0) {  // kx = 0 column over every full row
    for (int j = threadIdx.x; j < g.ny; j += blockDim.x) {
      if (j == 0) continue;
      const int jc = full_row_to_compact(j, g);
      const int jmc = full_row_to_compact((g.ny - j) % g.ny, g);
      double2 pa = make_double2(0.0, 0.0), pb = pa, qa = pa, qb = pa;
      if (jc >= 0) { pa = pz[(long)jc * g.Kxp]; qa = qz[(long)jc * g.Kxp]; }
      if (jmc >= 0) { pb = pz[(long)jmc * g.Kxp]; qb = qz[(long)jmc * g.Kxp]; }
      const double psx = 0.5 * (pa.x + pb.x), psy = 0.5 * (pa.y - pb.y);
      const double qsx = 0.5 * (qa.x + qb.x), qsy = 0.5 * (qa.y - qb.y);
      const double ky = ky_of(j, g);
      const int b = (int)floor(fabs(ky) / dk + 0.5);
      if (b >= nbins) continue;
      atomicAdd(&hk[b], 0.5 * ky * ky * (psx * psx + psy * psy) * norm);
      atomicAdd(&hz[b], 0.5 * (qsx * qsx + qsy * qsy) * norm);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    if (hk[b] != 0.0) atomicAdd(&out[(long)z * nbins + b], hk[b]);
    if (hz[b] != 0.0) atomicAdd(&out[(long)(g.nz + z) * nbins + b], hz[b]);
  }
}

// ---------------------------------------------------------------------------
// General batched 2-D FFT pieces with the DFTI layout (MKL_DFTI_Setup.f90),
// used by Initialize's forward transform and by the output fields.
// ---------------------------------------------------------------------------
// rows of nx reals -> Kout half-spectrum bins
template <typename R, int EPT>
__global__ void __launch_bounds__(256) k_row_r2c(const R* __restrict__ in, cplx<R>* __restrict__ out, int Kout, Geom g, FftPlan px,
                          const cplx<R>* __restrict__ twx, const cplx<R>* __restrict__ twr) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  const int N2 = g.N2;
  const Cx* row = reinterpret_cast<const Cx*>(in + (long)blockIdx.x * g.nx);
  for (int k = threadIdx.x; k < N2; k += blockDim.x) buf[k] = row[k];
  __syncthreads();
  lds_fft<R, EPT>(buf, LdsBatch{1, N2, 1, false}, px, twx, -1);
  Cx* o = out + (long)blockIdx.x * Kout;
  for (int k = threadIdx.x; k < Kout; k += blockDim.x) o[k] = r2c_post<R>(buf, k, N2, twr);
}

// rows of Kin half-spectrum bins -> nx reals, times scale
template <typename R, int EPT>
__global__ void __launch_bounds__(256) k_row_c2r(const cplx<R>* __restrict__ in, int Kin, R* __restrict__ out, R scale, Geom g,
                          FftPlan px, const cplx<R>* __restrict__ twx, const cplx<R>* __restrict__ twr) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  const int N2 = g.N2;
  const Cx* row = in + (long)blockIdx.x * Kin;
  for (int k = threadIdx.x; k < N2; k += blockDim.x) buf[k] = c2r_pre<R>(row, Kin, k, N2, twr, scale);
  __syncthreads();
  lds_fft<R, EPT>(buf, LdsBatch{1, N2, 1, false}, px, twx, +1);
  Cx* o = reinterpret_cast<Cx*>(out + (long)blockIdx.x * g.nx);
  for (int k = threadIdx.x; k < N2; k += blockDim.x) o[k] = buf[k];
}

// columns: (nseq, ny, K) -> (nseq, ny, K), y transform with sign; grid (tiles, nseq)
template <typename R, int EPT>
__global__ void __launch_bounds__(256) k_col_fft(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int K, int sign, Geom g,
                          FftPlan py, const cplx<R>* __restrict__ twy, int C) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  const int i0 = blockIdx.x * C;
  const long zoff = (long)blockIdx.y * g.ny;
  const int tot = C * g.ny;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int c = e % C, j = e / C, i = i0 + c;
    buf[e] = (i < K) ? in[(zoff + j) * K + i] : cmk<Cx>((R)0, (R)0);
  }
  __syncthreads();
  lds_fft<R, EPT>(buf, LdsBatch{C, 1, C, true}, py, twy, sign);
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int c = e % C, j = e / C, i = i0 + c;
    if (i < K) out[(zoff + j) * K + i] = buf[e];
  }
}

}  // namespace qgn
