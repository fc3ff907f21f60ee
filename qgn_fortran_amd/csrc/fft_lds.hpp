// Batched complex FFTs staged in LDS for gfx950 (64-wide wavefronts).
//
// Several sequences of one length n live in one LDS buffer; a workgroup
// transforms all of them together with mixed-radix Stockham passes
// (radices 2, 3, 4, 5, 8).  Each pass is register-staged: every thread
// loads the R inputs of up to MAXB butterflies, the workgroup barriers,
// each thread writes its R outputs to their self-sorting positions, and the
// workgroup barriers again, so a single LDS buffer suffices.
//
// Transform definition: X[m] = sum_j x[j] exp(sign * 2*pi*i * j*m / n), no
// scaling, which is what MKL DFTI computes for DftiComputeForward
// (sign = -1) and DftiComputeBackward (sign = +1) along each axis
// (MKL_DFTI_Setup.f90:1-52).
#pragma once
#include <hip/hip_runtime.h>

namespace qgn {

template <typename R> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r; r.x = a; r.y = b; return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
// Raw buffer resource over the nbytes at base (base wave-uniform; dword 3
// is the gfx9-family default format word).  A load whose byte offset is at
// or past nbytes returns zero, so masked loads need no branch: pass
// kRowOob as the offset of a lane that has nothing to load.
constexpr int kRowOob = 0x7ffffff0;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, int nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nbytes, 0x00020000);
}
template <typename C> __device__ __forceinline__ C rsrc_ld(__amdgpu_buffer_rsrc_t rs, int off);
template <> __device__ __forceinline__ float2 rsrc_ld<float2>(__amdgpu_buffer_rsrc_t rs, int off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
  float2 c;
  __builtin_memcpy(&c, &v, sizeof(c));
  return c;
}
template <> __device__ __forceinline__ double2 rsrc_ld<double2>(__amdgpu_buffer_rsrc_t rs, int off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
  double2 c;
  __builtin_memcpy(&c, &v, sizeof(c));
  return c;
}

// p[k] if ok, else zero.  The load is unconditional, from index 0 when !ok
// (p must point at one valid element), so it stays one vector load with a
// select after it instead of a branch around each component.
template <typename C> __device__ __forceinline__ C ld_or_zero(const C* p, int k, bool ok) {
  const C v = p[ok ? k : 0];
  return ok ? v : cmk<C>(0, 0);
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }
// multiply by sign*i  (sign = +1 or -1)
template <typename C> __device__ __forceinline__ C cmul_si(C a, int sign) {
  return sign > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
}

// ---- small DFT kernels on registers (in place), X[m] = sum_j a[j] w^(j m),
//      w = exp(sign*2*pi*i/R) ------------------------------------------------
template <typename C, int RAD> struct Dft;

template <typename C> struct Dft<C, 2> {
  static __device__ __forceinline__ void run(C* a, int) {
    C t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <typename C> struct Dft<C, 4> {
  static __device__ __forceinline__ void run(C* a, int s) {
    C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    C t2 = cadd(a[1], a[3]), t3 = cmul_si(csub(a[1], a[3]), s);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <typename C> struct Dft<C, 8> {
  static __device__ __forceinline__ void run(C* a, int s) {
    using R = decltype(C::x);
    const R h = (R)0.70710678118654752440084436210484903928;
    C e[4] = {a[0], a[2], a[4], a[6]};
    C o[4] = {a[1], a[3], a[5], a[7]};
    Dft<C, 4>::run(e, s);
    Dft<C, 4>::run(o, s);
    // o[m] *= w8^m, w8 = exp(s*i*pi/4) = (1 + s i)/sqrt(2): the factor h
    // of m = 1, 3 goes into the FMAs of the last butterflies
    const C d1 = cmk<C>(o[1].x - s * o[1].y, o[1].y + s * o[1].x);
    const C o2 = cmul_si(o[2], s);
    const C d3 = cmk<C>(-o[3].x - s * o[3].y, -o[3].y + s * o[3].x);
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cmk<C>(fma_r(h, d1.x, e[1].x), fma_r(h, d1.y, e[1].y));
    a[5] = cmk<C>(fma_r(-h, d1.x, e[1].x), fma_r(-h, d1.y, e[1].y));
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cmk<C>(fma_r(h, d3.x, e[3].x), fma_r(h, d3.y, e[3].y));
    a[7] = cmk<C>(fma_r(-h, d3.x, e[3].x), fma_r(-h, d3.y, e[3].y));
  }
};

// ---- the same DFTs of twiddled inputs, X[m] = sum_j a[j] W^j w^(j m) --------
// The twiddle products are folded into the butterflies: a radix-2 step
// a' = a + t b, b' = a - t b takes four FMAs for a' and one each for
// b' = 2a - a', six instructions against eight for a complex product and
// two complex additions.  Radix 8 splits into two radix-4 steps over the
// even and odd inputs (base W^2) and a last radix-2 step with W w^m; it
// needs the table values W, W^2 and W^4 only.
template <typename C> __device__ __forceinline__ void bfly_tw(C& a, C& b, const C& t) {
  const auto px = fma_r(t.x, b.x, fma_r(-t.y, b.y, a.x));
  const auto py = fma_r(t.x, b.y, fma_r(t.y, b.x, a.y));
  b = cmk<C>(fma_r(decltype(px)(2), a.x, -px), fma_r(decltype(px)(2), a.y, -py));
  a = cmk<C>(px, py);
}
// radix 4: y[r] W^r with W = w1, W^2 = w2; results in y[0..3]
template <typename C> __device__ __forceinline__ void dft4_tw(C& y0, C& y1, C& y2, C& y3, const C& w1, const C& w2,
                                                              int s) {
  bfly_tw(y0, y2, w2);   // y0 +- W^2 y2
  bfly_tw(y1, y3, w2);   // y1 +- W^2 y3
  bfly_tw(y0, y1, w1);   // X0, X2
  bfly_tw(y2, y3, cmul_si(w1, s));   // X1, X3
  const C x1 = y2;
  y2 = y1;
  y1 = x1;
}
template <typename C, int RAD> struct DftTw {
  // generic radix: the complex products, then the plain DFT
  template <typename TW> static __device__ __forceinline__ void run(C* a, TW twj, int s) {
#pragma unroll
    for (int j = 1; j < RAD; ++j) a[j] = cmul(a[j], twj(j));
    Dft<C, RAD>::run(a, s);
  }
};
template <typename C> struct DftTw<C, 2> {
  template <typename TW> static __device__ __forceinline__ void run(C* a, TW twj, int) { bfly_tw(a[0], a[1], twj(1)); }
};
template <typename C> struct DftTw<C, 4> {
  template <typename TW> static __device__ __forceinline__ void run(C* a, TW twj, int s) {
    dft4_tw(a[0], a[1], a[2], a[3], twj(1), twj(2), s);
  }
};
template <typename C> struct DftTw<C, 8> {
  template <typename TW> static __device__ __forceinline__ void run(C* a, TW twj, int s) {
    using R = decltype(C::x);
    const R h = (R)0.70710678118654752440084436210484903928;
    const C w1 = twj(1), w2 = twj(2), w4 = twj(4);
    C e[4] = {a[0], a[2], a[4], a[6]};
    C o[4] = {a[1], a[3], a[5], a[7]};
    dft4_tw(e[0], e[1], e[2], e[3], w2, w4, s);
    dft4_tw(o[0], o[1], o[2], o[3], w2, w4, s);
    // X[m] = E[m] + W w8^m O[m], X[m+4] = E[m] - W w8^m O[m]
    const C t1 = cmk<C>(h * (w1.x - s * w1.y), h * (w1.y + s * w1.x));
    bfly_tw(e[0], o[0], w1);
    bfly_tw(e[1], o[1], t1);
    bfly_tw(e[2], o[2], cmul_si(w1, s));
    bfly_tw(e[3], o[3], cmul_si(t1, s));
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      a[m] = e[m];
      a[m + 4] = o[m];
    }
  }
};

template <typename C> struct Dft<C, 3> {
  static __device__ __forceinline__ void run(C* a, int s) {
    using R = decltype(C::x);
    const R c3 = (R)0.86602540378443864676372317075293618347;  // sqrt(3)/2
    C t = cadd(a[1], a[2]);
    C d = csub(a[1], a[2]);
    C m = cmk<C>(a[0].x - (R)0.5 * t.x, a[0].y - (R)0.5 * t.y);
    C r = cmul_si(cscale(d, c3), s);
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <typename C> struct Dft<C, 5> {
  static __device__ __forceinline__ void run(C* a, int s) {
    using R = decltype(C::x);
    const R c1 = (R)0.30901699437494742410229341718281905886;   // cos(2pi/5)
    const R c2 = (R)-0.80901699437494742410229341718281905886;  // cos(4pi/5)
    const R s1 = (R)0.95105651629515357211643933337938214340;   // sin(2pi/5)
    const R s2 = (R)0.58778525229247312916870595463907276860;   // sin(4pi/5)
    C t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    C d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    C m1 = cmk<C>(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
    C m2 = cmk<C>(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
    C r1 = cmul_si(cmk<C>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), s);
    C r2 = cmul_si(cmk<C>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), s);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, r1);
    a[4] = csub(m1, r1);
    a[2] = cadd(m2, r2);
    a[3] = csub(m2, r2);
  }
};

// ---- FFT plan: list of radix passes for one length --------------------------
constexpr int kMaxPasses = 16;
struct FftPlan {
  int n;
  int npass;
  int radix[kMaxPasses];
  int p[kMaxPasses];        // product of the radices of the earlier passes
  int twstride[kMaxPasses]; // n / (p * radix): table stride of the pass twiddle
};

// Layout of a batch of sequences in LDS: element t of sequence s sits at
// buf[s * sstride + t * estride].  seq_fast selects how consecutive threads
// are spread: over sequences first (column tiles, where sequences are adjacent
// in memory) or over elements first (rows).
struct LdsBatch {
  int nseq;
  int sstride;
  int estride;
  bool seq_fast;
};

// One radix-RAD Stockham pass over every sequence of the batch.
// tw: table of exp(-2*pi*i*m/n), m in [0, n).  EPT: elements per thread bound.
template <typename Rt, int RAD, int EPT>
__device__ __forceinline__ void stockham_pass(cplx<Rt>* buf, const LdsBatch& B, int n, int p, int twstride,
                                              const cplx<Rt>* __restrict__ tw, int sign) {
  using C = cplx<Rt>;
  constexpr int MAXB = (EPT + RAD - 1) / RAD;
  const int nb = n / RAD;
  const int G = nb * B.nseq;
  const int nthr = blockDim.x;
  const int tid = threadIdx.x;
  C v[MAXB][RAD];
  int obase[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int g = tid + b * nthr;
    obase[b] = -1;
    if (g < G) {
      int s, t;
      if (B.seq_fast) { s = g % B.nseq; t = g / B.nseq; }
      else { t = g % nb; s = g / nb; }
      const int k = t % p;
      C* base = buf + s * B.sstride;
#pragma unroll
      for (int j = 0; j < RAD; ++j) v[b][j] = base[(t + j * nb) * B.estride];
      if (p > 1) {
#pragma unroll
        for (int j = 1; j < RAD; ++j) {
          C w = tw[j * k * twstride];
          if (sign > 0) w.y = -w.y;
          v[b][j] = cmul(v[b][j], w);
        }
      }
      Dft<C, RAD>::run(v[b], sign);
      obase[b] = s * B.sstride + ((t - k) * RAD + k) * B.estride;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (obase[b] >= 0) {
#pragma unroll
      for (int m = 0; m < RAD; ++m) buf[obase[b] + m * p * B.estride] = v[b][m];
    }
  }
  __syncthreads();
}

// Full transform of every sequence in the batch.  Caller must have
// synchronised after filling buf; on return buf holds the result and the
// workgroup is synchronised.
template <typename Rt, int EPT>
__device__ __forceinline__ void lds_fft(cplx<Rt>* buf, const LdsBatch& B, const FftPlan& P,
                                        const cplx<Rt>* __restrict__ tw, int sign) {
  for (int ps = 0; ps < P.npass; ++ps) {
    const int r = P.radix[ps], p = P.p[ps], ts = P.twstride[ps];
    switch (r) {
      case 8: stockham_pass<Rt, 8, EPT>(buf, B, P.n, p, ts, tw, sign); break;
      case 4: stockham_pass<Rt, 4, EPT>(buf, B, P.n, p, ts, tw, sign); break;
      case 2: stockham_pass<Rt, 2, EPT>(buf, B, P.n, p, ts, tw, sign); break;
      case 3: stockham_pass<Rt, 3, EPT>(buf, B, P.n, p, ts, tw, sign); break;
      default: stockham_pass<Rt, 5, EPT>(buf, B, P.n, p, ts, tw, sign); break;
    }
  }
}

}  // namespace qgn
