// Compile-time-sized batched FFTs in LDS.
//
// Same transform and Stockham self-sorting scheme as fft_lds.hpp, but the
// length, radix plan, batch layout and thread count are template constants,
// so every pass unrolls, index arithmetic reduces to shifts and multiplies,
// and the twiddles come from an LDS copy of the table.
#pragma once
#include "fft_lds.hpp"

namespace qgn {

struct CtPlan {
  int n = 0, np = 0;
  int rad[kMaxPasses] = {};
};

// Radix plan for length n.  When 64 divides n the first and the last pass
// are radix 8 (the row passes keep those two in registers); the rest is
// 3s, 5s, then 8s, 4s and 2s.
constexpr int ct_factor_middle(int m, int* rad, int k) {
  int e2 = 0;
  while (m % 2 == 0) { m /= 2; ++e2; }
  int n3 = 0, n5 = 0;
  while (m % 3 == 0) { m /= 3; ++n3; }
  while (m % 5 == 0) { m /= 5; ++n5; }
  for (int i = 0; i < n3; ++i) rad[k++] = 3;
  for (int i = 0; i < n5; ++i) rad[k++] = 5;
  int n8 = e2 / 3, rem = e2 % 3;
  if (rem == 1 && n8 > 0) { n8 -= 1; rad[k++] = 4; rad[k++] = 4; }
  else if (rem == 1) rad[k++] = 2;
  else if (rem == 2) rad[k++] = 4;
  for (int i = 0; i < n8; ++i) rad[k++] = 8;
  return m == 1 ? k : -1;
}

constexpr CtPlan make_ct_plan(int n) {
  CtPlan P{};
  P.n = n;
  if (n % 64 == 0) {
    P.rad[0] = 8;
    int k = ct_factor_middle(n / 64, P.rad, 1);
    if (k < 0) { P.np = -1; return P; }
    P.rad[k++] = 8;
    P.np = k;
    return P;
  }
  P.np = ct_factor_middle(n, P.rad, 0);
  return P;
}

constexpr bool ct_supported(int n) { return make_ct_plan(n).np > 0; }

// Per-pass twiddle tables: pass q of radix R with P = product of the earlier
// radices needs exp(-2 pi i j k / (P R)) for j in [1, R), k in [0, P); they
// are stored at offset(q) + (j-1)*P + k, so lanes with consecutive k read
// consecutive LDS words.  Passes with P = 1 need none.
constexpr int ct_tw_offset(int n, int pass) {
  const CtPlan PL = make_ct_plan(n);
  int off = 0, P = 1;
  for (int q = 0; q < pass && q < PL.np; ++q) {
    if (P > 1) off += (PL.rad[q] - 1) * P;
    P *= PL.rad[q];
  }
  return off;
}
constexpr int ct_tw_size(int n) { return ct_tw_offset(n, make_ct_plan(n).np); }

// LDS index padding: one empty slot after every 2^PADSHIFT slots breaks the
// power-of-two strides of the radix passes across the LDS banks
// (PADSHIFT = 0: no padding).
template <int PADSHIFT>
__device__ __forceinline__ constexpr int padded(int i) {
  if constexpr (PADSHIFT == 0) return i;
  else return i + (i >> PADSHIFT);
}

// Batch layout: element t of sequence s at buf[padded(s*SS + t*ES)];
// SEQ_FAST spreads consecutive threads over sequences first.
template <int NSEQ, int SS, int ES, bool SEQ_FAST, int PADSHIFT = 0> struct CtLayout {
  static constexpr int nseq = NSEQ, ss = SS, es = ES;
  static constexpr bool seq_fast = SEQ_FAST;
  static __device__ __forceinline__ int at(int s, int t) { return padded<PADSHIFT>(s * SS + t * ES); }
};

template <typename Rt, int N, int RAD, int P, class L, int NTHR, int OFF>
__device__ __forceinline__ void ct_pass(cplx<Rt>* buf, const cplx<Rt>* __restrict__ tw, int sign) {
  using C = cplx<Rt>;
  constexpr int NB = N / RAD;
  constexpr int G = NB * L::nseq;
  constexpr int MAXB = (G + NTHR - 1) / NTHR;
  const int tid = threadIdx.x;
  C v[MAXB][RAD];
  int obase[MAXB], otb[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int g = tid + b * NTHR;
    obase[b] = -1;
    if (G % NTHR == 0 || g < G) {
      int s, t;
      if constexpr (L::seq_fast) { s = g % L::nseq; t = g / L::nseq; }
      else { t = g % NB; s = g / NB; }
      const int k = t % P;
#pragma unroll
      for (int j = 0; j < RAD; ++j) v[b][j] = buf[L::at(s, t + j * NB)];
      if constexpr (P > 1) {
#pragma unroll
        for (int j = 1; j < RAD; ++j) {
          C w = tw[OFF + (j - 1) * P + k];
          if (sign > 0) w.y = -w.y;
          v[b][j] = cmul(v[b][j], w);
        }
      }
      Dft<C, RAD>::run(v[b], sign);
      obase[b] = s;
      otb[b] = (t - k) * RAD + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (G % NTHR == 0 || obase[b] >= 0) {
#pragma unroll
      for (int m = 0; m < RAD; ++m) buf[L::at(obase[b], otb[b] + m * P)] = v[b][m];
    }
  }
  __syncthreads();
}

template <typename Rt, int N, int PASS, int P, class L, int NTHR>
__device__ __forceinline__ void ct_passes(cplx<Rt>* buf, const cplx<Rt>* __restrict__ tw, int sign) {
  constexpr CtPlan PL = make_ct_plan(N);
  if constexpr (PASS < PL.np) {
    constexpr int RAD = PL.rad[PASS];
    ct_pass<Rt, N, RAD, P, L, NTHR, ct_tw_offset(N, PASS)>(buf, tw, sign);
    ct_passes<Rt, N, PASS + 1, P * RAD, L, NTHR>(buf, tw, sign);
  }
}

// Passes [PASS, END) of the length-N plan; P = product of the radices of the
// passes before PASS.
template <typename Rt, int N, int PASS, int END, int P, class L, int NTHR>
__device__ __forceinline__ void ct_passes_range(cplx<Rt>* buf, const cplx<Rt>* __restrict__ tw, int sign) {
  constexpr CtPlan PL = make_ct_plan(N);
  if constexpr (PASS < END) {
    constexpr int RAD = PL.rad[PASS];
    ct_pass<Rt, N, RAD, P, L, NTHR, ct_tw_offset(N, PASS)>(buf, tw, sign);
    ct_passes_range<Rt, N, PASS + 1, END, P * RAD, L, NTHR>(buf, tw, sign);
  }
}

// Full unnormalised transform of every sequence of the batch; the caller has
// synchronised after filling buf, and the workgroup is synchronised on return.
// tw: the per-pass twiddle tables of length N (ct_tw_size(N) entries).
template <typename Rt, int N, class L, int NTHR>
__device__ __forceinline__ void ct_fft(cplx<Rt>* buf, const cplx<Rt>* __restrict__ tw, int sign) {
  static_assert(ct_supported(N), "FFT length must factor into 2, 3, 5");
  ct_passes<Rt, N, 0, 1, L, NTHR>(buf, tw, sign);
}

// copy a table of `count` complex values into LDS (caller synchronises)
template <typename Rt>
__device__ __forceinline__ void load_table(cplx<Rt>* dst, const cplx<Rt>* __restrict__ src, int count) {
  // four loads in flight per thread before their LDS stores
  const int step = blockDim.x;
  int i = threadIdx.x;
  for (; i + 3 * step < count; i += 4 * step) {
    const cplx<Rt> a = src[i], b = src[i + step], c = src[i + 2 * step], d = src[i + 3 * step];
    dst[i] = a;
    dst[i + step] = b;
    dst[i + 2 * step] = c;
    dst[i + 3 * step] = d;
  }
  for (; i < count; i += step) dst[i] = src[i];
}

// The same copy in two halves for a compile-time count and thread count:
// table_fetch() issues the thread's loads into registers, table_store()
// writes them to LDS.  A kernel fetches the table, then issues its own data
// loads, so both sets are in flight together before either is stored.
template <int COUNT, int NTHR> constexpr int table_per_thread() { return (COUNT + NTHR - 1) / NTHR; }
template <typename Rt, int COUNT, int NTHR, int PER>
__device__ __forceinline__ void table_fetch(cplx<Rt> (&v)[PER], const cplx<Rt>* __restrict__ src) {
  static_assert(PER == table_per_thread<COUNT, NTHR>(), "table share");
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int i = (int)threadIdx.x + u * NTHR;
    v[u] = src[PER * NTHR == COUNT || i < COUNT ? i : 0];
  }
}
template <typename Rt, int COUNT, int NTHR, int PER>
__device__ __forceinline__ void table_store(const cplx<Rt> (&v)[PER], cplx<Rt>* dst) {
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int i = (int)threadIdx.x + u * NTHR;
    if (PER * NTHR == COUNT || i < COUNT) dst[i] = v[u];
  }
}

}  // namespace qgn
