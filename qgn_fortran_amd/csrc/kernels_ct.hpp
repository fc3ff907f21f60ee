// The FFT passes of GetRHS with compile-time sizes (array layouts in
// kernels.hpp).  NY = ny for the column passes, N2 = nx/2 for the row
// passes; C columns per workgroup, NTHR threads.  Twiddle tables are copied
// into LDS beside the data.
//
// The Jacobian is formed in advective form, J = u q_x + v q_y, where the
// reference uses the flux form i kx (u q)^ + i ky (v q)^
// (QGN_Module.f90:237-247).  Both are the 2/3-truncated Galerkin product of
// band-limited fields; they differ by the truncation of q (u_x + v_y), and
// u = -i ky psi, v = i kx psi make that divergence zero exactly in spectral
// space, so the two agree to round-off (checked against the oracle's flux
// form in the GPU parity tests).  The advective form leaves one product
// field per row instead of two, so the column pass needs no accumulators.
#pragma once
#include "fft_ct.hpp"
#include "kernels.hpp"

namespace qgn {

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Column tiles of 64-byte row segments share 128-byte lines with their
// neighbour, so map consecutive dispatch slots of one XCD to consecutive
// tiles: XCD x takes the x-th eighth of the tiles in order.  A permutation of
// [0, total); the last total % 8 ids map to themselves.
__device__ __forceinline__ int xcd_tile(int w, int total) {
  const int per = total >> 3;
  if (w >= (per << 3)) return w;
  return (w & 7) * per + (w >> 3);
}

// P1: y-inverse FFT of up to four fields built from compact inputs.
template <typename R, int NY, int C, int NTHR>
__global__ void __launch_bounds__(NTHR) k_col_inv_ct(ColInvArgs a, Geom g, const cplx<R>* __restrict__ twy) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* tw = buf + C * NY;
  using L = CtLayout<C, 1, C, true>;
  constexpr int TOT = C * NY;
  const int ntile = g.Kxp / C;
  // grid = nf * ntile * nz, field fastest: the fields that share a source
  // (u, v from psi; q_x, q_y from q) and neighbouring tiles run together on
  // one XCD and share its L2
  const int w = xcd_tile(blockIdx.x, gridDim.x);
  const int f = w % a.nf;
  const int i0 = ((w / a.nf) % ntile) * C;
  const int z = w / (a.nf * ntile);
  // FP32: the twiddle table's loads in flight with the tile's loads below
  // (FP64 tiles would spill the extra registers at 2048 rows)
  constexpr bool TAB_OVERLAP = sizeof(R) == 4;
  cplx<R> tab[TAB_OVERLAP ? table_per_thread<ct_tw_size(NY), NTHR>() : 1];
  if constexpr (TAB_OVERLAP) table_fetch<R, ct_tw_size(NY), NTHR>(tab, twy);
  else load_table<R>(tw, twy, ct_tw_size(NY));
  {
    const int mult = a.mult[f];
    const bool isd = a.src_is_double[f];
    // the layer's compact plane as a raw buffer: the dropped ky rows (and
    // lanes past the tile) load as zero, and all EPT loads of a thread are
    // in flight together
    constexpr int EPT = (TOT + NTHR - 1) / NTHR;
    // FP32 tiles move two adjacent columns per 16-byte access: element u of
    // a thread is column pair (tid + (u/2) NTHR), member u & 1
    constexpr bool PAIR = sizeof(Cx) == 8 && EPT % 2 == 0 && C % 2 == 0;
    auto elem = [&](int u) {
      return PAIR ? 2 * ((int)threadIdx.x + (u >> 1) * NTHR) + (u & 1) : (int)threadIdx.x + u * NTHR;
    };
    const int lbytes = g.Ky * g.Kxp * (isd ? (int)sizeof(double2) : (int)sizeof(Cx));
    const __amdgpu_buffer_rsrc_t rs =
        row_rsrc(static_cast<const unsigned char*>(a.src[f]) + (long)z * lbytes, lbytes);
    int off[EPT];
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = elem(u);
      const int c = e % C, j = e / C;
      const int jc = e < TOT ? full_row_to_compact(j, g) : -1;
      off[u] = jc >= 0 ? jc * g.Kxp + i0 + c : -1;
    }
    Cx v[EPT];
    if (isd) {
#pragma unroll
      for (int u = 0; u < EPT; ++u)
        v[u] = to_c<R>(rsrc_ld<double2>(rs, off[u] >= 0 ? off[u] * (int)sizeof(double2) : kRowOob));
    } else if constexpr (PAIR) {
#pragma unroll
      for (int u = 0; u < EPT; u += 2) {
        const double2 d = rsrc_ld<double2>(rs, off[u] >= 0 ? off[u] * (int)sizeof(Cx) : kRowOob);
        const float4 p4 = __builtin_bit_cast(float4, d);
        v[u] = cmk<Cx>(p4.x, p4.y);
        v[u + 1] = cmk<Cx>(p4.z, p4.w);
      }
    } else {
#pragma unroll
      for (int u = 0; u < EPT; ++u) v[u] = rsrc_ld<Cx>(rs, off[u] >= 0 ? off[u] * (int)sizeof(Cx) : kRowOob);
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = elem(u);
      if (EPT * NTHR != TOT && e >= TOT) continue;
      const int c = e % C, j = e / C, i = i0 + c;
      Cx x = v[u];
      if (mult == 1) {          // -i ky
        const R ky = ky_r<R>(j, g);
        x = cmk<Cx>(ky * x.y, -ky * x.x);
      } else if (mult == 2) {   // +i kx
        const R kx = kx_r<R>(i, g);
        x = cmk<Cx>(-kx * x.y, kx * x.x);
      } else if (mult == 3) {   // +i ky
        const R ky = ky_r<R>(j, g);
        x = cmk<Cx>(-ky * x.y, ky * x.x);
      }
      buf[e] = x;
    }
    if constexpr (TAB_OVERLAP) table_store<R, ct_tw_size(NY), NTHR>(tab, tw);
    __syncthreads();
    ct_fft<R, NY, L, NTHR>(buf, tw, +1);
    Cx* dst = reinterpret_cast<Cx*>(a.dst[f]) + (long)z * NY * g.Kxp;
    if constexpr (mix_pairs<R>()) {
      // one 16-byte store per column and row pair (rows 2p, 2p+1 of column
      // c are adjacent in the mixed layout): 8 columns fill a 128-byte line
      for (int e2 = threadIdx.x; 2 * e2 < TOT; e2 += NTHR) {
        const int c = e2 % C, p = e2 / C;
        const Cx lo = buf[2 * p * C + c], hi = buf[(2 * p + 1) * C + c];
        *reinterpret_cast<float4*>(dst + mix_at<R>(2 * p, i0 + c, g.Kxp)) = make_float4(lo.x, lo.y, hi.x, hi.y);
      }
    } else {
      for (int e = threadIdx.x; e < TOT; e += NTHR) {
        const int c = e % C, j = e / C;
        dst[mix_at<R>(j, i0 + c, g.Kxp)] = buf[e];
      }
    }
    __syncthreads();
  }
}

// Padding of the contiguous-row LDS layout: one slot per 16 complex values,
// FP32 and FP64 (the first radix-8 pass stores 8 consecutive values per
// lane; a slot per 32 left lanes t and t + 32 on one bank: the fused row
// pass took 76.6 us at 1024^2 x 16 with it, 74.1 with this spacing)
template <typename R> constexpr int row_pad_shift() { return 4; }
template <typename R> constexpr int row_lds_slots(int n) { return n + (n >> row_pad_shift<R>()) + 1; }

// x-r2c post-processing on a padded LDS sequence starting at slot `base`.
template <typename R, int PS>
__device__ __forceinline__ cplx<R> r2c_post_pad(const cplx<R>* buf, int base, int k, int N2,
                                                const cplx<R>* __restrict__ twr) {
  using Cx = cplx<R>;
  const Cx a = buf[padded<PS>(base + (k == N2 ? 0 : k))];
  const Cx b = buf[padded<PS>(base + (k == 0 ? 0 : N2 - k))];
  const Cx e = cmk<Cx>((R)0.5 * (a.x + b.x), (R)0.5 * (a.y - b.y));
  const Cx dd = cmk<Cx>((R)0.5 * (a.x - b.x), (R)0.5 * (a.y + b.y));
  const Cx o = cmk<Cx>(dd.y, -dd.x);
  return cadd(e, cmul(twr[k], o));
}

// P2: a persistent grid; each workgroup takes groups of ROWS consecutive rows
// (z, y) (ny is a multiple of ROWS, so a group never spans two layers) and
// prefetches the next group's bins into registers while it transforms the
// current one.  LDS: 4 padded sequences of N2 per row, the per-pass FFT
// tables and the real-FFT post-processing table.
template <typename R, int N2, int ROWS, int NTHR>
__global__ void __launch_bounds__(NTHR) k_row_jac_ct(RowJacArgs a, Geom g, const cplx<R>* __restrict__ twx,
                                                     const cplx<R>* __restrict__ twr, R scale, int ngroups) {
  using Cx = cplx<R>;
  constexpr int PS = row_pad_shift<R>();
  constexpr int EL = (ROWS * N2 + NTHR - 1) / NTHR;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* ts = buf + row_lds_slots<R>(ROWS * 4 * N2);   // per-pass tables of the N2-point FFT
  Cx* tr = ts + ct_tw_size(N2);                     // exp(-2 pi i k / nx), k <= N2
  load_table<R>(ts, twx, ct_tw_size(N2));
  load_table<R>(tr, twr, N2 + 1);
  const int Kx = g.Kx;
  const long Kxp = g.Kxp;
  const Cx* srcs[4] = {reinterpret_cast<const Cx*>(a.U), reinterpret_cast<const Cx*>(a.V),
                       reinterpret_cast<const Cx*>(a.QX), reinterpret_cast<const Cx*>(a.QY)};
  Cx ra[4][EL], rb[4][EL];
  auto fetch = [&](int grp) {
    const long row0 = (long)grp * ROWS;
#pragma unroll
    for (int b = 0; b < EL; ++b) {
      const int e = threadIdx.x + b * NTHR;
      const int r = e / N2, k = e - r * N2;
      const bool ok = (ROWS * N2) % NTHR == 0 || e < ROWS * N2;
      const int oa = (ok && k < Kx) ? mix_at<R>(r, k, (int)Kxp) * (int)sizeof(Cx) : kRowOob;
      const int ob = (ok && N2 - k < Kx) ? mix_at<R>(r, N2 - k, (int)Kxp) * (int)sizeof(Cx) : kRowOob;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const __amdgpu_buffer_rsrc_t rs = row_rsrc(srcs[s] + row0 * Kxp, ROWS * (int)Kxp * (int)sizeof(Cx));
        ra[s][b] = rsrc_ld<Cx>(rs, oa);
        rb[s][b] = rsrc_ld<Cx>(rs, ob);
      }
    }
  };
  int grp = blockIdx.x;
  if (grp < ngroups) fetch(grp);
  __syncthreads();
  for (; grp < ngroups; grp += gridDim.x) {
    const int row0 = grp * ROWS;
    const int z = row0 / g.ny;
    const int j0 = row0 - z * g.ny;
    const bool bottom = a.drag && (z % g.nzm == g.nzm - 1);
#pragma unroll
    for (int b = 0; b < EL; ++b) {
      const int e = threadIdx.x + b * NTHR;
      if ((ROWS * N2) % NTHR == 0 || e < ROWS * N2) {
        const int r = e / N2, k = e - r * N2;
#pragma unroll
        for (int s = 0; s < 4; ++s)
          buf[padded<PS>((4 * r + s) * N2 + k)] = c2r_pre_ab<R>(ra[s][b], rb[s][b], k, tr, scale);
      }
    }
    if (grp + (int)gridDim.x < ngroups) fetch(grp + gridDim.x);   // in flight during the FFTs below
    __syncthreads();
    ct_fft<R, N2, CtLayout<4 * ROWS, N2, 1, false, PS>, NTHR>(buf, ts, +1);
    // physical values: element m of a sequence holds x[2m] + i x[2m+1]
    for (int e = threadIdx.x; e < ROWS * N2; e += NTHR) {
      const int r = e / N2, m = e - r * N2;
      const int b0 = 4 * r * N2 + m;
      const Cx u = buf[padded<PS>(b0)], v = buf[padded<PS>(b0 + N2)];
      const Cx qx = buf[padded<PS>(b0 + 2 * N2)], qy = buf[padded<PS>(b0 + 3 * N2)];
      buf[padded<PS>(b0)] = cmk<Cx>(u.x * qx.x + v.x * qy.x, u.y * qx.y + v.y * qy.y);
      if (bottom) {   // URMS = sqrt(u^2 + v^2) (QGN_Module.f90:285-289)
        const R r0 = sqrt(u.x * u.x + v.x * v.x), r1 = sqrt(u.y * u.y + v.y * v.y);
        buf[padded<PS>(b0 + N2)] = cmk<Cx>(r0 * u.x, r1 * u.y);
        buf[padded<PS>(b0 + 2 * N2)] = cmk<Cx>(r0 * v.x, r1 * v.y);
      }
    }
    __syncthreads();
    if (bottom) ct_fft<R, N2, CtLayout<4 * ROWS, N2, 1, false, PS>, NTHR>(buf, ts, -1);
    else ct_fft<R, N2, CtLayout<ROWS, 4 * N2, 1, false, PS>, NTHR>(buf, ts, -1);
    Cx* jo = reinterpret_cast<Cx*>(a.J) + (long)row0 * Kxp;
    for (int e = threadIdx.x; e < ROWS * Kx; e += NTHR) {
      const int r = e / Kx, k = e - r * Kx;
      jo[mix_at<R>(r, k, (int)Kxp)] = r2c_post_pad<R, PS>(buf, 4 * r * N2, k, N2, tr);
    }
    if (bottom) {
      Cx* du = reinterpret_cast<Cx*>(a.DU) + (long)((z / g.nzm) * g.ny + j0) * Kxp;
      Cx* dv = reinterpret_cast<Cx*>(a.DV) + (long)((z / g.nzm) * g.ny + j0) * Kxp;
      for (int e = threadIdx.x; e < ROWS * Kx; e += NTHR) {
        const int r = e / Kx, k = e - r * Kx;
        du[mix_at<R>(r, k, (int)Kxp)] = r2c_post_pad<R, PS>(buf, (4 * r + 1) * N2, k, N2, tr);
        dv[mix_at<R>(r, k, (int)Kxp)] = r2c_post_pad<R, PS>(buf, (4 * r + 2) * N2, k, N2, tr);
      }
    }
    __syncthreads();   // buf is refilled by the next group
  }
}

// Radix-RAD Stockham pass over the row sequences of the fused row pass:
// ROWS rows of N2 values, row r's element i at padded<PS>(r N2 + i).  Every
// LDS address is a per-thread base plus a compile-time offset: with
// 2^PS | d, padded(x + d) = padded(x) + d + d / 2^PS, which covers the reads
// (stride N2/RAD) and the writes of passes with 2^PS | P; a write stride of
// P = 8 under PS = 4 adds the x-dependent carry (x/8 & 1 + m) / 2, which
// takes two bases.  Same arithmetic as ct_pass on the same layout.
template <int PS> constexpr int pad_step(int d) { return d + (d >> PS); }

template <typename R, int N2, int RAD, int P, int NTHR, int PS, int OFF>
__device__ __forceinline__ void row_pass(cplx<R>* buf, const cplx<R>* __restrict__ tw, int sign) {
  using C = cplx<R>;
  constexpr int NB = N2 / RAD;
  constexpr int G = NB * (NTHR / (N2 / 8));   // butterflies: NB per row, NTHR / (N2/8) rows
  constexpr int MAXB = G / NTHR;
  static_assert(G % NTHR == 0 && NB % (1 << PS) == 0, "row pass shape");
  const int tid = threadIdx.x;
  C v[MAXB][RAD];
  int wbase[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int gg = tid + b * NTHR;
    const int t = gg % NB, s = gg / NB;
    const int k = t % P;
    const C* rp = buf + padded<PS>(s * N2 + t);
#pragma unroll
    for (int j = 0; j < RAD; ++j) v[b][j] = rp[j * pad_step<PS>(NB)];
    if constexpr (P > 1) {
      const C* tp = tw + OFF + k;
      DftTw<C, RAD>::run(v[b], [&](int j) {
        C w = tp[(j - 1) * P];
        if (sign > 0) w.y = -w.y;
        return w;
      }, sign);
    } else {
      Dft<C, RAD>::run(v[b], sign);
    }
    wbase[b] = s * N2 + (t - k) * RAD + k;
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int x = wbase[b];
    if constexpr (P % (1 << PS) == 0) {
      C* wp = buf + padded<PS>(x);
#pragma unroll
      for (int m = 0; m < RAD; ++m) wp[m * pad_step<PS>(P)] = v[b][m];
    } else if constexpr (P == 8 && PS == 4) {
      // padded(x + 8m) = padded(x) + 8m + (((x >> 3) & 1) + m) / 2
      C* w0 = buf + padded<PS>(x);
      C* w1 = w0 + ((x >> 3) & 1);
#pragma unroll
      for (int m = 0; m < RAD; ++m) {
        if (m & 1) w1[8 * m + (m >> 1)] = v[b][m];
        else w0[8 * m + (m >> 1)] = v[b][m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < RAD; ++m) buf[padded<PS>(x + m * P)] = v[b][m];
    }
  }
  __syncthreads();
}

// Passes [PASS, END) of the length-N2 plan on the row layout; passes whose
// butterflies do not divide evenly over the threads (radix 3, 5) or whose
// read stride is not a multiple of 2^PS take ct_pass.
template <typename R, int N2, int PASS, int END, int P, int NTHR, int PS>
__device__ __forceinline__ void row_passes_range(cplx<R>* buf, const cplx<R>* __restrict__ tw, int sign) {
  constexpr CtPlan PL = make_ct_plan(N2);
  if constexpr (PASS < END) {
    constexpr int RAD = PL.rad[PASS];
    constexpr int NB = N2 / RAD, ROWS = NTHR / (N2 / 8);
    if constexpr ((NB * ROWS) % NTHR == 0 && NB % (1 << PS) == 0)
      row_pass<R, N2, RAD, P, NTHR, PS, ct_tw_offset(N2, PASS)>(buf, tw, sign);
    else
      ct_pass<R, N2, RAD, P, CtLayout<ROWS, N2, 1, false, PS>, NTHR, ct_tw_offset(N2, PASS)>(buf, tw, sign);
    row_passes_range<R, N2, PASS + 1, END, P * RAD, NTHR, PS>(buf, tw, sign);
  }
}

// P2, register-fused variant for nx/2 whose plan starts and ends with a
// radix-8 pass (64 | nx/2), ROWS rows per workgroup, nx/16 threads per row,
// one padded LDS sequence per row.  Each inverse transform's first pass reads
// global memory into registers and its last pass leaves the physical values
// in registers, at the positions where the forward transform's first pass
// takes them, so the products skip an LDS round trip.  u and v are
// transformed first (the bottom layer's |u| u and |u| v go out right away),
// then q_x and q_y are folded into J = u q_x + v q_y one at a time, so at
// most two fields and the accumulator are live.  LDS addresses are bases
// plus compile-time offsets throughout (row_pass).
template <typename R, int N2, int ROWS>
__global__ void __launch_bounds__(ROWS * (N2 / 8)) k_row_jac_fused2(RowJacArgs a, Geom g,
                                                                   const cplx<R>* __restrict__ twx,
                                                                   const cplx<R>* __restrict__ twr, R scale) {
  using Cx = cplx<R>;
  constexpr CtPlan PL = make_ct_plan(N2);
  static_assert(N2 % 64 == 0 && PL.rad[0] == 8 && PL.rad[PL.np - 1] == 8, "needs radix-8 end passes");
  constexpr int NB = N2 / 8;
  constexpr int NTHR = ROWS * NB;
  constexpr int PS = row_pad_shift<R>();
  constexpr int LAST = PL.np - 1;
  constexpr int OFF_LAST = ct_tw_offset(N2, LAST);
  static_assert(N2 % (1 << PS) == 0, "row layout");
  // the last pass's reads (stride NB) and the output loop's (stride NB) as
  // base + offset when 2^PS | NB, else in full
  constexpr bool AFF = NB % (1 << PS) == 0;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* ts = buf + row_lds_slots<R>(ROWS * N2);
  Cx* tr = ts + ct_tw_size(N2);
  load_table<R>(ts, twx, ct_tw_size(N2));
  load_table<R>(tr, twr, N2 + 1);
  const int r = threadIdx.x / NB;
  const int t = threadIdx.x - r * NB;
  // last rows first: the bottom layer's heavier groups (drag) start early
  // (the group count from the kernel arguments keeps row0 in scalar registers)
  const int row0 = (g.nz * g.ny / ROWS - 1 - (int)blockIdx.x) * ROWS;
  const int z = row0 / g.ny;
  const int j0 = row0 - z * g.ny;
  const bool bottom = a.drag && (z % g.nzm == g.nzm - 1);
  const int Kx = g.Kx;
  const int Kxp = g.Kxp;
  // this thread's slots: the first pass's outputs 8t .. 8t+7 of row r, and
  // the last pass's inputs t + j NB
  Cx* const w8p = buf + padded<PS>(r * N2 + 8 * t);   // padded(r N2 + 8t + m) = this + m, m < 8
  Cx* const lp = buf + padded<PS>(r * N2 + t);
  const Cx* const twl = ts + OFF_LAST + t;
  // x-c2r of one mixed row in two halves.  fetch: the kept bins X[k] and
  // X[N2 - k] of this thread's 8 first-pass inputs (out-of-range bins read
  // zero).  head: the c2r pre-step and the first radix-8 pass, into LDS
  // (the fetched bins are dead after it, so the next field's fetch can
  // follow).  tail: the middle passes and the last pass, which leaves
  // o[j] = x[2m], x[2m+1] at m = t + j NB, unnormalised (every product
  // below is of two such fields: s2 = scale^2 normalises it once)
  const R s2 = scale * scale;
  auto fetch = [&](const void* src, Cx(&xa)[8], Cx(&xb)[8]) {
    const __amdgpu_buffer_rsrc_t rs =
        row_rsrc(reinterpret_cast<const Cx*>(src) + (long)row0 * Kxp, ROWS * Kxp * (int)sizeof(Cx));
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = t + j * NB, kk = N2 - k;
      xa[j] = rsrc_ld<Cx>(rs, k < Kx ? mix_at<R>(r, k, Kxp) * (int)sizeof(Cx) : kRowOob);
      xb[j] = rsrc_ld<Cx>(rs, kk < Kx ? mix_at<R>(r, kk, Kxp) * (int)sizeof(Cx) : kRowOob);
    }
  };
  auto head = [&](const Cx(&xa)[8], const Cx(&xb)[8]) {
    Cx v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = c2r_pre_fma<R>(xa[j], xb[j], t + j * NB, tr[t + j * NB]);
    Dft<Cx, 8>::run(v, +1);
    __syncthreads();   // earlier readers of buf are done
#pragma unroll
    for (int m = 0; m < 8; ++m) w8p[m] = v[m];
    __syncthreads();
  };
  auto tail = [&](Cx(&o)[8]) {
    row_passes_range<R, N2, 1, LAST, 8, NTHR, PS>(buf, ts, +1);
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = AFF ? lp[j * pad_step<PS>(NB)] : buf[padded<PS>(r * N2 + t + j * NB)];
    DftTw<Cx, 8>::run(o, [&](int j) {
      Cx w = twl[(j - 1) * NB];
      w.y = -w.y;
      return w;
    }, +1);
  };
  // x-r2c of the products in w8 (same positions) to the Kx kept bins of ROWS rows at out
  auto forward = [&](Cx(&w8)[8], Cx* out) {
    Dft<Cx, 8>::run(w8, -1);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) w8p[m] = w8[m];
    __syncthreads();
    row_passes_range<R, N2, 1, PL.np, 8, NTHR, PS>(buf, ts, -1);
    // rows fastest: with row pairs interleaved, neighbouring lanes store
    // neighbouring 8-byte words.  Thread e = tid + i NTHR takes bin
    // k = e / ROWS of row e % ROWS; its LDS reads step by pad_step(NTHR / ROWS)
    constexpr int KSTEP = NTHR / ROWS;
    constexpr int NIT = (ROWS * (2 * N2 / 3 + 1) + NTHR - 1) / NTHR;   // Kx <= 2 N2 / 3 + 1
    const int rr = threadIdx.x % ROWS, k0 = threadIdx.x / ROWS;
    const Cx* za = buf + padded<PS>(rr * N2 + k0);            // Z[k], k = k0 + i KSTEP
    const Cx* zb = buf + padded<PS>(rr * N2 + N2 - k0);       // Z[N2 - k]
    const Cx* z0 = buf + padded<PS>(rr * N2);                 // Z[0] (for k = 0)
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int k = k0 + i * KSTEP;
      if (k < Kx) {
        const Cx aa = AFF ? za[i * pad_step<PS>(KSTEP)] : buf[padded<PS>(rr * N2 + k)];
        const Cx bb = k == 0 ? *z0 : (AFF ? zb[-i * pad_step<PS>(KSTEP)] : buf[padded<PS>(rr * N2 + N2 - k)]);
        out[mix_at<R>(rr, k, Kxp)] = r2c_post_fma<R>(aa, bb, tr[k]);
      }
    }
  };
  // each field's bins load while the previous field runs its LDS passes;
  // on bottom-layer rows q_x loads after the drag products' transforms
  Cx fu[8], fv[8], acc[8];
  {
    Cx xa[8], xb[8];
    fetch(a.U, xa, xb);   // in flight with the twiddle tables' loads
    __syncthreads();
    head(xa, xb);
    fetch(a.V, xa, xb);
    tail(fu);
    head(xa, xb);
    if (!bottom) fetch(a.QX, xa, xb);
    tail(fv);
    if (bottom) {   // URMS = sqrt(u^2 + v^2) (QGN_Module.f90:285-289)
      const long drow = (long)((z / g.nzm) * g.ny + j0) * Kxp;   // bottom layer of member z / nzm
      Cx w8[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const R r0 = s2 * sqrt(fu[m].x * fu[m].x + fv[m].x * fv[m].x);
        const R r1 = s2 * sqrt(fu[m].y * fu[m].y + fv[m].y * fv[m].y);
        w8[m] = cmk<Cx>(r0 * fu[m].x, r1 * fu[m].y);
      }
      forward(w8, reinterpret_cast<Cx*>(a.DU) + drow);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const R r0 = s2 * sqrt(fu[m].x * fu[m].x + fv[m].x * fv[m].x);
        const R r1 = s2 * sqrt(fu[m].y * fu[m].y + fv[m].y * fv[m].y);
        w8[m] = cmk<Cx>(r0 * fv[m].x, r1 * fv[m].y);
      }
      forward(w8, reinterpret_cast<Cx*>(a.DV) + drow);
      fetch(a.QX, xa, xb);
    }
    head(xa, xb);
    fetch(a.QY, xa, xb);
    tail(acc);
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = cmk<Cx>(fu[m].x * acc[m].x, fu[m].y * acc[m].y);
    head(xa, xb);
  }
  {
    Cx qy[8];
    tail(qy);
#pragma unroll
    for (int m = 0; m < 8; ++m)
      acc[m] = cmk<Cx>(s2 * (acc[m].x + fv[m].x * qy[m].x), s2 * (acc[m].y + fv[m].y * qy[m].y));
  }
  forward(acc, reinterpret_cast<Cx*>(a.J) + (long)row0 * Kxp);
}

// Tile of P3 block b of total = ntile * nz: (layer z, column tile t).  With
// drag, bottom-layer tiles do three column FFTs instead of one; each XCD's
// range of xcd_tile positions then starts with an equal share of them
// (heavy tiles first, then the other layers in order), so no XCD ends with
// a run of heavy tiles.  Without drag, or when the shares do not fit, tiles
// follow xcd_tile in layer order.
__device__ __forceinline__ void rhs_tile(int b, int total, int ntile, int nzm, int drag, int* z, int* t) {
  const int w = xcd_tile(b, total);
  const int per = total >> 3;
  const int heavy = total / nzm;          // one bottom layer per member
  if (!drag || nzm < 2 || (heavy + 7) / 8 > per) {
    *z = w / ntile; *t = w % ntile;
    return;
  }
  int h = -1, l;
  if (w < (per << 3)) {
    const int x = w / per, p = w - x * per;
    const int hs = x * heavy / 8, hx = (x + 1) * heavy / 8 - hs;
    if (p < hx) h = hs + p;
    else l = x * per - hs + (p - hx);
  } else {
    l = w - heavy;
  }
  if (h >= 0) {
    *z = (h / ntile) * nzm + nzm - 1; *t = h % ntile;
  } else {
    const int lz = l / ntile;             // over the members' upper layers
    *z = (lz / (nzm - 1)) * nzm + lz % (nzm - 1); *t = l % ntile;
  }
}

// P3: y-forward FFT of J (of J - M with the closure network's forcing M),
// then
//   RHS = -J_hat - Ubar_k i kx q - (beta + Qy_k) i kx psi
//         [- (A2 k2 + A8 k8) q  if is_explicit]  [+ r0 Htot/H_nz k2 psi on layer nz]
//         [+ C_d Htot/H_nz (i ky du - i kx dv) on layer nz, QGN_Module.f90:284-292]
// on the kept modes (DeAlias, QGN_Module.f90:294), (0,0) and padding zero.
// Each thread keeps its modes' values in registers (at most C Ky / NTHR <=
// 8 Ky / NY < 6 of them); on a bottom-layer tile with drag, |u| u and |u| v
// then go through the same LDS buffer.
// A column tile of a mixed slab into LDS: C columns from i0, NY = TOT / C
// rows of `stride` complex values; each thread's loads are all issued before
// its LDS stores.
template <typename R, int C, int NTHR, int TOT>
__device__ __forceinline__ void load_cols(cplx<R>* buf, const cplx<R>* __restrict__ src, int Kxp, int i0) {
  constexpr int EPT = (TOT + NTHR - 1) / NTHR;
  constexpr bool EXACT = EPT * NTHR == TOT;
  if constexpr (mix_pairs<R>() && EPT % 2 == 0 && EXACT) {
    // FP32: rows 2p and 2p+1 of column c per 16-byte load
    float4 v[EPT / 2];
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int e2 = (int)threadIdx.x + k * NTHR, c = e2 % C, p = e2 / C;
      v[k] = *reinterpret_cast<const float4*>(src + mix_at<R>(2 * p, i0 + c, Kxp));
    }
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int e2 = (int)threadIdx.x + k * NTHR, c = e2 % C, p = e2 / C;
      buf[2 * p * C + c] = cmk<cplx<R>>(v[k].x, v[k].y);
      buf[(2 * p + 1) * C + c] = cmk<cplx<R>>(v[k].z, v[k].w);
    }
    return;
  }
  cplx<R> v[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    const int ee = (EXACT || e < TOT) ? e : TOT - 1;
    v[u] = src[mix_at<R>(ee / C, i0 + ee % C, Kxp)];
  }
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    if (EXACT || e < TOT) buf[e] = v[u];
  }
}

// The same tile of a - b.
template <typename R, int C, int NTHR, int TOT>
__device__ __forceinline__ void load_cols_sub(cplx<R>* buf, const cplx<R>* __restrict__ a,
                                              const cplx<R>* __restrict__ b, int Kxp, int i0) {
  constexpr int EPT = (TOT + NTHR - 1) / NTHR;
  constexpr bool EXACT = EPT * NTHR == TOT;
  if constexpr (mix_pairs<R>() && EPT % 2 == 0 && EXACT) {
    float4 v[EPT / 2], w[EPT / 2];
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int e2 = (int)threadIdx.x + k * NTHR, c = e2 % C, p = e2 / C;
      const int o = mix_at<R>(2 * p, i0 + c, Kxp);
      v[k] = *reinterpret_cast<const float4*>(a + o);
      w[k] = *reinterpret_cast<const float4*>(b + o);
    }
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int e2 = (int)threadIdx.x + k * NTHR, c = e2 % C, p = e2 / C;
      buf[2 * p * C + c] = cmk<cplx<R>>(v[k].x - w[k].x, v[k].y - w[k].y);
      buf[(2 * p + 1) * C + c] = cmk<cplx<R>>(v[k].z - w[k].z, v[k].w - w[k].w);
    }
    return;
  }
  cplx<R> v[EPT], w[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    const int ee = (EXACT || e < TOT) ? e : TOT - 1;
    const int o = mix_at<R>(ee / C, i0 + ee % C, Kxp);
    v[u] = a[o];
    w[u] = b[o];
  }
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    if (EXACT || e < TOT) buf[e] = cmk<cplx<R>>(v[u].x - w[u].x, v[u].y - w[u].y);
  }
}

// FP32 power-of-two lengths with at most 512 threads: at least 6 waves per
// SIMD (at most 85 VGPRs; these fit without spills, the radix-3 lengths do
// not), so three workgroups of 8 waves share a CU within its LDS.
template <typename R, typename S, int NY, int C, int NTHR>
__global__ void __launch_bounds__(NTHR)
    __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && (NY & (NY - 1)) == 0 && NTHR <= 512 ? 6 : 1)))
    k_col_rhs_ct(ColRhsArgs a, Geom g, const cplx<R>* __restrict__ twy) {
  using Cx = cplx<R>;
  constexpr int MAXO = 6;   // kept modes per thread: C Ky <= 6 NTHR since Ky < 3 ny / 4
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* tw = buf + C * NY;
  using L = CtLayout<C, 1, C, true>;
  constexpr int TOT = C * NY;
  const int ntile = g.Kxp / C;
  int z, i0;
  rhs_tile(blockIdx.x, gridDim.x, ntile, g.nzm, a.drag, &z, &i0);   // grid = ntile * nz
  i0 *= C;
  // the tile of J (of J - M with the closure network's forcing M) into LDS
  auto load_tile = [&]() {
    if (a.M)   // the forcing enters as -FFT_y(J - M)
      load_cols_sub<R, C, NTHR, TOT>(buf, reinterpret_cast<const Cx*>(a.J) + (long)z * NY * g.Kxp,
                                     reinterpret_cast<const Cx*>(a.M) + (long)z * NY * g.Kxp, g.Kxp, i0);
    else
      load_cols<R, C, NTHR, TOT>(buf, reinterpret_cast<const Cx*>(a.J) + (long)z * NY * g.Kxp, g.Kxp, i0);
  };
  if constexpr (sizeof(R) == 4 && NTHR <= 512) {
    // FP32 with up to 512 threads: the twiddle table's loads in flight with
    // the tile's (FP64 tiles, and the 1024-thread tiles of 2048 rows, would
    // lose occupancy to the extra registers)
    cplx<R> tab[table_per_thread<ct_tw_size(NY), NTHR>()];
    table_fetch<R, ct_tw_size(NY), NTHR>(tab, twy);
    load_tile();
    table_store<R, ct_tw_size(NY), NTHR>(tab, tw);
  } else {
    load_table<R>(tw, twy, ct_tw_size(NY));
    load_tile();
  }
  __syncthreads();
  ct_fft<R, NY, L, NTHR>(buf, tw, -1);
  cplx<S>* out = reinterpret_cast<cplx<S>*>(a.out);
  const R ub = (R)a.ubar[z];
  const R bq = (R)(a.beta + a.qybar[z]);
  const bool bottom = (z % g.nzm == g.nzm - 1);
  const R ek = bottom ? (R)a.ekman[z / g.nzm] : (R)0;
  const int nown = C * g.Ky;
  // q and psi of the owned modes through raw buffers over the layer's
  // compact plane (all loads in flight together; lanes past the tile read
  // zero)
  const long lay = (long)z * g.Ky * g.Kxp;
  const int plane = g.Ky * g.Kxp;
  // (q rounded to R, when the stage kernel wrote that copy: half the bytes)
  const bool q_r = a.q32 != nullptr;
  const __amdgpu_buffer_rsrc_t rq =
      q_r ? row_rsrc(reinterpret_cast<const Cx*>(a.q32) + lay, plane * (int)sizeof(Cx))
          : row_rsrc(a.q + lay, plane * (int)sizeof(double2));
  const __amdgpu_buffer_rsrc_t rp = row_rsrc(reinterpret_cast<const Cx*>(a.psi) + lay, plane * (int)sizeof(Cx));
  Cx qv[MAXO];
  Cx pv[MAXO];
  if (q_r) {
#pragma unroll
    for (int u = 0; u < MAXO; ++u) {
      const int o = threadIdx.x + u * NTHR;
      const int off = o < nown ? (o / C) * g.Kxp + i0 + o % C : -1;
      qv[u] = rsrc_ld<Cx>(rq, off >= 0 ? off * (int)sizeof(Cx) : kRowOob);
      pv[u] = rsrc_ld<Cx>(rp, off >= 0 ? off * (int)sizeof(Cx) : kRowOob);
    }
  } else {
#pragma unroll
    for (int u = 0; u < MAXO; ++u) {
      const int o = threadIdx.x + u * NTHR;
      const int off = o < nown ? (o / C) * g.Kxp + i0 + o % C : -1;
      qv[u] = to_c<R>(rsrc_ld<double2>(rq, off >= 0 ? off * (int)sizeof(double2) : kRowOob));
      pv[u] = rsrc_ld<Cx>(rp, off >= 0 ? off * (int)sizeof(Cx) : kRowOob);
    }
  }
  Cx keep[MAXO];
#pragma unroll
  for (int u = 0; u < MAXO; ++u) {
    const int o = threadIdx.x + u * NTHR;
    if (o >= nown) continue;
    const int c = o % C, jc = o / C, i = i0 + c;
    const int j = compact_row_to_full(jc, g);
    const Cx v = buf[j * C + c];
    const Cx q = qv[u];
    const Cx ps = pv[u];
    const R kx = kx_r<R>(i, g), ky = ky_r<R>(j, g);
    Cx r = cmk<Cx>(-v.x, -v.y);
    r.x += ub * kx * q.y + bq * kx * ps.y;
    r.y -= ub * kx * q.x + bq * kx * ps.x;
    if (a.explicit_visc) {
      const R k2 = kx * kx + ky * ky;
      const R k4 = k2 * k2;
      const R lam = (R)a.A2 * k2 + (R)a.A8 * (k4 * k4);
      r.x -= lam * q.x;
      r.y -= lam * q.y;
    }
    if (bottom) {
      const R k2 = kx * kx + ky * ky;
      r.x += ek * k2 * ps.x;
      r.y += ek * k2 * ps.y;
    }
    keep[u] = r;
  }
  if (bottom && a.drag) {
    const int mb = z / g.nzm;
    const R cd = (R)a.cdrag[mb];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {   // |u| u, then |u| v
      const Cx* d = reinterpret_cast<const Cx*>(pass ? a.DV : a.DU) + (long)mb * NY * g.Kxp;
      __syncthreads();   // the previous contents of buf are read
      load_cols<R, C, NTHR, TOT>(buf, d, g.Kxp, i0);
      __syncthreads();
      ct_fft<R, NY, L, NTHR>(buf, tw, -1);
#pragma unroll
      for (int u = 0; u < MAXO; ++u) {
        const int o = threadIdx.x + u * NTHR;
        if (o >= nown) continue;
        const int c = o % C, jc = o / C, i = i0 + c;
        const int j = compact_row_to_full(jc, g);
        const Cx x = buf[j * C + c];
        const R kx = kx_r<R>(i, g), ky = ky_r<R>(j, g);
        if (pass == 0) {   // + cd i ky du
          keep[u].x -= cd * ky * x.y;
          keep[u].y += cd * ky * x.x;
        } else {           // - cd i kx dv
          keep[u].x += cd * kx * x.y;
          keep[u].y -= cd * kx * x.x;
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < MAXO; ++u) {
    const int o = threadIdx.x + u * NTHR;
    if (o >= nown) continue;
    const int c = o % C, jc = o / C, i = i0 + c;
    const int j = compact_row_to_full(jc, g);
    Cx r = keep[u];
    if ((i == 0 && j == 0) || i >= g.Kx) r = cmk<Cx>((R)0, (R)0);
    out[lay + (long)jc * g.Kxp + i] = cmk<cplx<S>>((S)r.x, (S)r.y);
  }
}

// x-c2r pre-processing of ROWS mixed rows (stride complex values apart,
// Kin kept bins) read from the raw buffer rs into padded LDS rows; all loads
// of a thread are issued before its LDS stores.
template <typename R, int N2, int ROWS, int NTHR, int PS>
__device__ __forceinline__ void c2r_pre_rows(cplx<R>* buf, __amdgpu_buffer_rsrc_t rs, int stride, int Kin,
                                             const cplx<R>* __restrict__ tr, R scale) {
  using Cx = cplx<R>;
  constexpr int TOT = ROWS * N2, EPT = (TOT + NTHR - 1) / NTHR;
  constexpr bool EXACT = EPT * NTHR == TOT;
  constexpr int SZ = (int)sizeof(Cx);
  Cx va[EPT], vb[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    const bool ok = EXACT || e < TOT;
    const int r = e / N2, k = e - r * N2, kk = N2 - k;
    va[u] = rsrc_ld<Cx>(rs, ok && k < Kin ? (r * stride + k) * SZ : kRowOob);
    vb[u] = rsrc_ld<Cx>(rs, ok && kk < Kin ? (r * stride + kk) * SZ : kRowOob);
  }
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    if (EXACT || e < TOT) buf[padded<PS>(e)] = c2r_pre_ab<R>(va[u], vb[u], e % N2, tr, scale);
  }
}

// Max-norm of the c2r of mixed rows: max |x| over the grid into one 64-bit
// slot (a non-negative double's bit pattern orders like its value; a NaN
// anywhere stores a NaN, which orders above +Inf).  ARK43_QGLeith.f90:121-124
// take the max over all layers.
template <typename R, int N2, int ROWS, int NTHR>
__global__ void __launch_bounds__(NTHR) k_row_maxabs_ct(const cplx<R>* __restrict__ mixed, int Kin, int stride,
                                                        Geom g, const cplx<R>* __restrict__ twx,
                                                        const cplx<R>* __restrict__ twr,
                                                        unsigned long long* __restrict__ partial) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int PS = row_pad_shift<R>();
  constexpr int TOT = ROWS * N2, EPT = (TOT + NTHR - 1) / NTHR;
  constexpr bool EXACT = EPT * NTHR == TOT;
  constexpr int SZ = (int)sizeof(Cx);
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* ts = buf + row_lds_slots<R>(ROWS * N2);
  __shared__ double red[16];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  // the rows' bins and each element's pre-processing twiddle exp(+2 pi i k/nx)
  // are loaded straight into registers, all in flight while the FFT's
  // per-pass tables are copied into LDS
  const long row0 = (long)blockIdx.x * ROWS;
  const __amdgpu_buffer_rsrc_t rs = row_rsrc(mixed + row0 * stride, ROWS * stride * SZ);
  const __amdgpu_buffer_rsrc_t rt = row_rsrc(twr, (N2 + 1) * SZ);
  Cx va[EPT], vb[EPT], w[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    const bool ok = EXACT || e < TOT;
    const int r = e / N2, k = e - r * N2, kk = N2 - k;
    va[u] = rsrc_ld<Cx>(rs, ok && k < Kin ? mix_at<R>(r, k, stride) * SZ : kRowOob);
    vb[u] = rsrc_ld<Cx>(rs, ok && kk < Kin ? mix_at<R>(r, kk, stride) * SZ : kRowOob);
    w[u] = rsrc_ld<Cx>(rt, ok ? k * SZ : kRowOob);
  }
  load_table<R>(ts, twx, ct_tw_size(N2));
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    if (EXACT || e < TOT) buf[padded<PS>(e)] = c2r_pre_w<R>(va[u], vb[u], e % N2, w[u]);
  }
  __syncthreads();
  ct_fft<R, N2, CtLayout<ROWS, N2, 1, false, PS>, NTHR>(buf, ts, +1);
  double m = 0.0;
  bool nan = false;
  for (int e = threadIdx.x; e < ROWS * N2; e += NTHR) {
    const Cx v = buf[padded<PS>(e)];
    nan |= (v.x != v.x) || (v.y != v.y);
    m = fmax(m, fmax(fabs((double)v.x), fabs((double)v.y)));
  }
  if (nan) bad = 1;
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < (NTHR + 63) / 64; ++w) mm = fmax(mm, red[w]);
    unsigned long long bits = (unsigned long long)__double_as_longlong(mm);
    if (bad) bits = 0x7FF8000000000000ull;   // NaN orders above every finite value and +Inf
    partial[blockIdx.x] = bits;
  }
}

// Closure network input.  One workgroup per (member, row y, group of G
// channels); channel ch < nzm is layer ch of q, else layer ch - nzm of psi,
// so a group's rows all come from one field, one layer apart.  The rows'
// x-c2r times nn (the 1/(nx ny) of the unnormalised transforms), times the
// channel's input scale, rounded to bf16 and stored channels-last (G = 8:
// one 16-byte store per pixel):
//   x[((m ny + y) nx + px) C + ch],  C = 2 nzm.
// The rows' bins and their pre-processing twiddles are loaded into
// registers, all in flight while the FFT tables are copied into LDS.
template <typename R, int N2, int G, int NTHR>
__global__ void __launch_bounds__(NTHR) k_closure_rows_in(const cplx<R>* __restrict__ mq,
                                                          const cplx<R>* __restrict__ mp,
                                                          const double* __restrict__ scale, short* __restrict__ x,
                                                          R nn, Geom g, const cplx<R>* __restrict__ twx,
                                                          const cplx<R>* __restrict__ twr) {
  static_assert(G == 8, "one 16-byte store of 8 bf16 channels per pixel");
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int PS = row_pad_shift<R>();
  constexpr int TOT = G * N2, EPT = (TOT + NTHR - 1) / NTHR;
  constexpr bool EXACT = EPT * NTHR == TOT;
  constexpr int SZ = (int)sizeof(Cx);
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* ts = buf + row_lds_slots<R>(TOT);
  const int C = 2 * g.nzm, ngrp = C / G;
  const int cg = blockIdx.x % ngrp;
  const int my = blockIdx.x / ngrp;            // member * ny + y
  const int m = my / g.ny, y = my - m * g.ny;
  const int c0 = cg * G;
  const bool isq = c0 < g.nzm;
  const int l0 = m * g.nzm + (isq ? c0 : c0 - g.nzm);   // the group's first layer in the batch
  const int stride = g.ny * g.Kxp;                       // complex values from one layer's row to the next's
  constexpr int KS = mix_kstep<R>();
  const __amdgpu_buffer_rsrc_t rs =
      row_rsrc((isq ? mq : mp) + (long)l0 * stride + mix_at<R>(y, 0, g.Kxp), G * stride * SZ);
  const __amdgpu_buffer_rsrc_t rt = row_rsrc(twr, (N2 + 1) * SZ);
  const int Kin = g.Kx;
  Cx va[EPT], vb[EPT], w[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    const bool ok = EXACT || e < TOT;
    const int r = e / N2, k = e - r * N2, kk = N2 - k;
    va[u] = rsrc_ld<Cx>(rs, ok && k < Kin ? (r * stride + k * KS) * SZ : kRowOob);
    vb[u] = rsrc_ld<Cx>(rs, ok && kk < Kin ? (r * stride + kk * KS) * SZ : kRowOob);
    w[u] = rsrc_ld<Cx>(rt, ok ? k * SZ : kRowOob);
  }
  load_table<R>(ts, twx, ct_tw_size(N2));
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * NTHR;
    if (EXACT || e < TOT) buf[padded<PS>(e)] = c2r_pre_w<R>(va[u], vb[u], e % N2, w[u], nn);
  }
  __syncthreads();
  ct_fft<R, N2, CtLayout<G, N2, 1, false, PS>, NTHR>(buf, ts, +1);
  // pixel px is component px & 1 of complex entry px / 2 of each row
  short* xo = x + (long)my * g.nx * C + c0;
  for (int px = threadIdx.x; px < 2 * N2; px += NTHR) {
    unsigned wd[G / 2];
#pragma unroll
    for (int k = 0; k < G / 2; ++k) {
      short h[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int ch = 2 * k + t;
        const Cx z = buf[padded<PS>(ch * N2 + (px >> 1))];
        const R v = (px & 1) ? z.y : z.x;
        h[t] = bf16_rne((float)((double)v * scale[c0 + ch]));
      }
      wd[k] = (unsigned)(unsigned short)h[0] | ((unsigned)(unsigned short)h[1] << 16);
    }
    *reinterpret_cast<uint4*>(xo + (long)px * C) = make_uint4(wd[0], wd[1], wd[2], wd[3]);
  }
}

// Closure network output.  One workgroup per (member, row y, group of G
// channels): the network's channels-last output y[((m ny + y) nx + px) nzm
// + ch] times the channel's output scale, x-r2c, into the Kx bins of the
// mixed rows of layers m nzm + ch.
template <typename R, int N2, int G, int NTHR>
__global__ void __launch_bounds__(NTHR) k_closure_rows_out(const float* __restrict__ yn,
                                                           const double* __restrict__ scale,
                                                           cplx<R>* __restrict__ mixed, Geom g,
                                                           const cplx<R>* __restrict__ twx,
                                                           const cplx<R>* __restrict__ twr) {
  using Cx = cplx<R>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int PS = row_pad_shift<R>();
  Cx* buf = reinterpret_cast<Cx*>(lds_raw);
  Cx* ts = buf + row_lds_slots<R>(G * N2);
  Cx* tr = ts + ct_tw_size(N2);
  load_table<R>(ts, twx, ct_tw_size(N2));
  load_table<R>(tr, twr, N2 + 1);
  const int ngrp = g.nzm / G;
  const int cg = blockIdx.x % ngrp;
  const int my = blockIdx.x / ngrp;            // member * ny + y
  const int m = my / g.ny, y = my - m * g.ny;
  const int c0 = cg * G;
  // pixel px goes to component px & 1 of complex entry px / 2 of each row
  R* bre = reinterpret_cast<R*>(buf);
  const float* yi = yn + (long)my * g.nx * g.nzm + c0;
  for (int px = threadIdx.x; px < 2 * N2; px += NTHR) {
    float v[G];
#pragma unroll
    for (int k = 0; k < G / 4; ++k) {
      const float4 f = *reinterpret_cast<const float4*>(yi + (long)px * g.nzm + 4 * k);
      v[4 * k] = f.x; v[4 * k + 1] = f.y; v[4 * k + 2] = f.z; v[4 * k + 3] = f.w;
    }
#pragma unroll
    for (int ch = 0; ch < G; ++ch)
      bre[2 * padded<PS>(ch * N2 + (px >> 1)) + (px & 1)] = (R)((double)v[ch] * scale[c0 + ch]);
  }
  __syncthreads();
  ct_fft<R, N2, CtLayout<G, N2, 1, false, PS>, NTHR>(buf, ts, -1);
  const int Kx = g.Kx;
  for (int e = threadIdx.x; e < G * Kx; e += NTHR) {
    const int r = e / Kx, k = e - r * Kx;
    mixed[(long)(m * g.nzm + c0 + r) * g.ny * g.Kxp + mix_at<R>(y, k, g.Kxp)] =
        r2c_post_pad<R, PS>(buf, r * N2, k, N2, tr);
  }
}


// Final max over the per-block partials: workgroup b reduces partials
// [b n, (b+1) n) into slot[b] (one per member of a batch).
__global__ void __launch_bounds__(1024) k_max_final(const unsigned long long* __restrict__ partial, int n,
                                                    unsigned long long* __restrict__ slot) {
  __shared__ unsigned long long red[16];
  partial += (long)blockIdx.x * n;
  unsigned long long m = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = partial[i] > m ? partial[i] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long mm = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) mm = red[w] > mm ? red[w] : mm;
    slot[blockIdx.x] = mm;
  }
}

}  // namespace qgn
