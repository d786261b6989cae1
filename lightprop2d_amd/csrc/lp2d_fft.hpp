// Device-side FFT core for gfx950: in-register DFT butterflies and a Stockham
// (autosort) line FFT whose inter-pass exchanges go through LDS.
//
// Layout convention for one line of N points processed by TL = N/E threads, E
// elements per thread: on entry thread t holds x[t + TL*m] (m = 0..E-1), on exit it
// holds X[t + TL*m].  Both ends are therefore stride-TL patterns, which is what makes
// the global loads and stores of every pass coalesced and the output naturally
// ordered.  Forward = exp(-2*pi*i*jk/N), unnormalised (numpy.fft convention).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace lp2d {

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) {
  cx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
// a * b
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return mk(a.re * s, a.im * s); }

// compile-time loop
template <int I, int N, typename F>
__device__ __forceinline__ void sfor_impl(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor_impl<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) { sfor_impl<0, N>(f); }

// cos(2*pi*k/64), k = 0..16 (first quadrant); the rest follows by symmetry.
__host__ __device__ constexpr double kQuarterCos64(int k) {
  constexpr double q[17] = {1.0,
                            0.9951847266721969,
                            0.9807852804032304,
                            0.9569403357322088,
                            0.9238795325112867,
                            0.881921264348355,
                            0.8314696123025452,
                            0.773010453362737,
                            0.7071067811865476,
                            0.6343932841636455,
                            0.5555702330196023,
                            0.4713967368259978,
                            0.38268343236508984,
                            0.29028467725446233,
                            0.19509032201612833,
                            0.09801714032956077,
                            0.0};
  return q[k];
}
__host__ __device__ constexpr double kCos64(int k) {
  k &= 63;
  return k <= 16 ? kQuarterCos64(k) : k <= 32 ? -kQuarterCos64(32 - k)
                                   : k <= 48 ? -kQuarterCos64(k - 32)
                                             : kQuarterCos64(64 - k);
}
__host__ __device__ constexpr double kSin64(int k) { return kCos64(k - 16); }

// a * exp(-+2*pi*i*E/R) with E a compile-time exponent (forward: minus sign).
template <typename T, int R, int E, bool INV>
__device__ __forceinline__ cx<T> twc(cx<T> a) {
  static_assert(64 % R == 0, "radix must divide 64");
  constexpr int k = ((E % R) + R) % R * (64 / R);
  if constexpr (k == 0) {
    return a;
  } else if constexpr (k == 16) {  // fwd: -i, inv: +i
    return INV ? mk(-a.im, a.re) : mk(a.im, -a.re);
  } else if constexpr (k == 32) {
    return mk(-a.re, -a.im);
  } else if constexpr (k == 48) {  // fwd: +i, inv: -i
    return INV ? mk(a.im, -a.re) : mk(-a.im, a.re);
  } else {
    constexpr double c = kCos64(k);
    constexpr double s = INV ? kSin64(k) : -kSin64(k);
    return cmul(a, mk((T)c, (T)s));
  }
}

// In-register DFT of length R (power of two <= 64), natural order in and out.
// Cooley-Tukey decimation in time, R = R1*R2 with R1 = 4 (or 2).
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cx<T> (&x)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    cx<T> a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = a - b;
  } else if constexpr (R == 4) {
    cx<T> t0 = x[0] + x[2], t1 = x[0] - x[2], t2 = x[1] + x[3];
    cx<T> t3 = twc<T, 4, 1, INV>(x[1] - x[3]);
    x[0] = t0 + t2;
    x[2] = t0 - t2;
    x[1] = t1 + t3;
    x[3] = t1 - t3;
  } else {
    constexpr int R1 = (R == 8) ? 2 : 4;
    constexpr int R2 = R / R1;
    // step 1: length-R2 DFTs over the stride-R1 subsequences, then twiddle
    sfor<R1>([&](auto n1c) {
      constexpr int n1 = decltype(n1c)::value;
      cx<T> s[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) s[n2] = x[n2 * R1 + n1];
      dft<T, R2, INV>(s);
      sfor<R2>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        x[k2 * R1 + n1] = twc<T, R, n1 * k2, INV>(s[k2]);
      });
    });
    // step 2: length-R1 DFTs across, output X[k1*R2 + k2]
    cx<T> out[R];
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      cx<T> s[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) s[n1] = x[k2 * R1 + n1];
      dft<T, R1, INV>(s);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) out[k1 * R2 + k2] = s[k1];
    }
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = out[i];
  }
}

// ---------------------------------------------------------------------------
// LDS accessors.  `pad(i)` inserts one element every 2^PS elements so that the
// stride-R writes of the early Stockham passes spread over the banks.
template <int PS>
__device__ __forceinline__ int pad_idx(int i) { return i + (i >> PS); }

template <int N, int PS>
struct PaddedLen {
  static constexpr int value = N + (N >> PS);
};

// One line stored contiguously (rows).
template <typename T, int PS>
struct LineAcc {
  cx<T>* base;
  __device__ __forceinline__ void st(int i, cx<T> v) const { base[pad_idx<PS>(i)] = v; }
  __device__ __forceinline__ cx<T> ld(int i) const { return base[pad_idx<PS>(i)]; }
};

// W lines interleaved element by element (column strips): element i of line c sits
// at pad(i)*W + c, so lanes that differ in c touch consecutive addresses.
template <typename T, int PS, int W>
struct InterAcc {
  cx<T>* base;  // already offset by the line index c
  __device__ __forceinline__ void st(int i, cx<T> v) const { base[pad_idx<PS>(i) * W] = v; }
  __device__ __forceinline__ cx<T> ld(int i) const { return base[pad_idx<PS>(i) * W]; }
};

// Strided access inside a 2-D LDS tile (column c of a row-major tile with pitch P).
template <typename T, int P>
struct StrideAcc {
  cx<T>* base;  // offset to column c (already padded)
  __device__ __forceinline__ void st(int i, cx<T> v) const { base[i * P] = v; }
  __device__ __forceinline__ cx<T> ld(int i) const { return base[i * P]; }
};

// Synchronisation of an exchange among lanes of one wave: LDS accesses of a wave
// execute in order, so ordering the compiler's accesses is enough (no workgroup
// barrier, other waves run on).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WS>
__device__ __forceinline__ void exchange_sync() {
  if constexpr (WS)
    wave_sync();
  else
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Stockham autosort passes.  Pass with stride NS (product of the previous radices)
// and radix R: butterfly j in [0, N/R) reads in[j + r*N/R], multiplies element r by
// w_{NS*R}^{(j mod NS)*r}, runs DFT_R and writes out[(j/NS)*NS*R + (j mod NS) + r*NS].
// Thread t owns the butterflies j = t + TL*q (q < E/R), whose inputs are exactly the
// registers v[q + r*E/R].  The last pass writes back into registers in the stride-TL
// layout, so only the passes before it exchange through LDS.
//
// tw points at the twiddle table of length N for this direction:
// tw[i] = exp(-+2*pi*i*i/N) (minus for the forward transform).
template <typename T, int N, int E, bool INV, int NS, bool WS = false, typename Acc>
__device__ __forceinline__ void stockham(cx<T> (&v)[E], int t, const Acc& acc,
                                         const cx<T>* __restrict__ tw) {
  constexpr int REM = N / NS;
  constexpr int R = REM >= E ? E : REM;
  constexpr int B = E / R;
  constexpr int TL = N / E;
  constexpr bool LAST = (NS * R == N);
  static_assert(E % R == 0, "radix must divide the per-thread element count");
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const int j = t + TL * q;
    const int k = j & (NS - 1);
    cx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + r * B];
    if constexpr (NS > 1) {
      constexpr int TS = N / (NS * R);
      // byte step between this butterfly's twiddles (twiddle r at r * step).  The
      // empty asm makes it opaque, so a kernel running several transforms recomputes
      // the R-1 offsets per pass instead of keeping them live across the kernel.
      unsigned step = (unsigned)(k * TS) * (unsigned)sizeof(cx<T>);
      asm volatile("" : "+v"(step));
      const char* twb = reinterpret_cast<const char*>(tw);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        a[r] = cmul(a[r], *reinterpret_cast<const cx<T>*>(twb + r * step));
      }
    }
    dft<T, R, INV>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * B] = a[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) acc.st(base + r * NS, a[r]);
    }
  }
  if constexpr (!LAST) {
    exchange_sync<WS>();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = acc.ld(t + TL * m);
    exchange_sync<WS>();
    stockham<T, N, E, INV, NS * R, WS>(v, t, acc, tw);
  }
}

// WS: every line's TL = N/E lanes lie in one wave, so the exchanges synchronise that
// wave only.
template <typename T, int N, int E, bool INV, bool WS = false, typename Acc>
__device__ __forceinline__ void line_fft(cx<T> (&v)[E], int t, const Acc& acc,
                                         const cx<T>* __restrict__ tw) {
  static_assert((N & (N - 1)) == 0 && (E & (E - 1)) == 0 && E <= N, "powers of two");
  static_assert(!WS || 64 % (N / E) == 0, "a line's lanes within one wave");
  stockham<T, N, E, INV, 1, WS>(v, t, acc, tw);
}

// ---------------------------------------------------------------------------
// Twiddles staged in LDS as one 2-D table per pass: the pass with stride NS and radix
// R keeps w_{NS*R}^{k*r} at row k, column r, rows padded to R+2 entries (the rows of
// the lanes of one wave then start in different banks, and stay 16-byte aligned).  A
// butterfly reads its R-1 twiddles at one per-lane base plus compile-time offsets, and
// the inverse transform uses the same table conjugated.
//
// CL (compact last pass): when the last pass has radix R <= 4, its table holds only
// w_N^k (one entry per row k) and the butterfly forms w^{2k} = w^k w^k and
// w^{3k} = w^{2k} w^k in registers (a rounding or two more than the table).  For
// n = 4096 with 32 elements per thread this shrinks the tables from 7232 to 2112
// entries, which lets a 4-column strip and its twiddles share the LDS.
// CL = 2 (split compact last pass): w^k = w^{S*(k/S)} * w^{k%S} from two short tables
// (S = 2^floor(log2(NS)/2) entries and NS/S entries, 64 instead of 1024 for n = 4096
// with 32 elements per thread; one product more per butterfly), and last-pass radices
// up to 16, whose w^{rk} come from the product tree w^{rk} = w^{(r-r/2)k} w^{(r/2)k}
// (at most 1 + log2(r) roundings); so that 2-row tiles of 4096 points and their
// tables fit twice in one CU's LDS.
constexpr int twl_radix(int N, int E, int ns) { return N / ns >= E ? E : N / ns; }
constexpr bool twl_compact(int N, int E, int ns, int CL) {
  return CL != 0 && ns > 1 && ns * twl_radix(N, E, ns) == N &&
         twl_radix(N, E, ns) <= (CL == 2 ? 16 : 4);
}
constexpr int twl_split(int ns) {
  int l = 0;
  while ((2 << l) <= ns) ++l;
  return 1 << (l / 2);
}
constexpr int twl_compact_len(int ns, int CL) {
  return CL == 2 ? twl_split(ns) + ns / twl_split(ns) : ns;
}
constexpr int twl_off(int N, int E, int NS, int CL = 0) {
  int off = 0;
  for (int ns = 1; ns < NS; ns *= twl_radix(N, E, ns))
    if (ns > 1)
      off += twl_compact(N, E, ns, CL) ? twl_compact_len(ns, CL)
                                       : ns * (twl_radix(N, E, ns) + 2);
  return off;
}
template <int N, int E, int CL = 0>
struct TwlLen {
  static constexpr int value = twl_off(N, E, N, CL) > 0 ? twl_off(N, E, N, CL) : 1;
};

// Fill the per-pass tables from the concatenated global table (size L at offset L-2).
template <typename T, int N, int E, int THREADS, int NS = 1, int CL = 0>
__device__ __forceinline__ void twl_fill(cx<T>* twl, const cx<T>* __restrict__ tw, int tid) {
  if constexpr (NS < N) {
    constexpr int R = twl_radix(N, E, NS);
    if constexpr (twl_compact(N, E, NS, CL) && CL == 2) {
      // w^i for i < S, then w^{S*h} for h < NS/S
      constexpr int L = NS * R, OFF = twl_off(N, E, NS, CL), S = twl_split(NS);
      for (int i = tid; i < S + NS / S; i += THREADS)
        twl[OFF + i] = tw[(L - 2) + (i < S ? i : (i - S) * S)];
    } else if constexpr (twl_compact(N, E, NS, CL)) {
      constexpr int L = NS * R, OFF = twl_off(N, E, NS, CL);
      for (int i = tid; i < NS; i += THREADS) twl[OFF + i] = tw[(L - 2) + i];
    } else if constexpr (NS > 1) {
      constexpr int L = NS * R, P = R + 2, OFF = twl_off(N, E, NS, CL);
      for (int i = tid; i < L; i += THREADS) {
        const int k = i / R, r = i % R;
        twl[OFF + k * P + r] = tw[(L - 2) + k * r];
      }
    }
    twl_fill<T, N, E, THREADS, NS * R, CL>(twl, tw, tid);
  }
}

// The Stockham passes above with the twiddles read from the LDS tables.  WS: every
// line's TL = N/E lanes lie in one wave, so the exchanges synchronise that wave only.
template <typename T, int N, int E, bool INV, int NS, int CL = 0, bool WS = false,
          typename Acc>
__device__ __forceinline__ void stockham_l(cx<T> (&v)[E], int t, const Acc& acc,
                                           const cx<T>* twl) {
  constexpr int REM = N / NS;
  constexpr int R = REM >= E ? E : REM;
  constexpr int B = E / R;
  constexpr int TL = N / E;
  constexpr bool LAST = (NS * R == N);
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const int j = t + TL * q;
    const int k = j & (NS - 1);
    cx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + r * B];
    if constexpr (twl_compact(N, E, NS, CL)) {
      // w^k from the table; w^{2k}, w^{3k} by products (R <= 4)
      cx<T> w[R > 1 ? R : 2];
      if constexpr (CL == 2) {
        constexpr int S = twl_split(NS), OFF = twl_off(N, E, NS, CL);
        w[1] = cmul(twl[OFF + S + k / S], twl[OFF + (k & (S - 1))]);
      } else {
        w[1] = twl[twl_off(N, E, NS, CL) + k];
      }
#pragma unroll
      for (int r = 2; r < R; ++r) w[r] = cmul(w[r - r / 2], w[r / 2]);
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = INV ? cmulc(a[r], w[r]) : cmul(a[r], w[r]);
    } else if constexpr (NS > 1) {
      const cx<T>* tp = twl + (twl_off(N, E, NS, CL) + k * (R + 2));
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = INV ? cmulc(a[r], tp[r]) : cmul(a[r], tp[r]);
    }
    dft<T, R, INV>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * B] = a[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) acc.st(base + r * NS, a[r]);
    }
  }
  if constexpr (!LAST) {
    exchange_sync<WS>();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = acc.ld(t + TL * m);
    exchange_sync<WS>();
    stockham_l<T, N, E, INV, NS * R, CL, WS>(v, t, acc, twl);
  }
}

template <typename T, int N, int E, bool INV, int CL = 0, bool WS = false, typename Acc>
__device__ __forceinline__ void line_fft_l(cx<T> (&v)[E], int t, const Acc& acc,
                                           const cx<T>* twl) {
  static_assert(!WS || 64 % (N / E) == 0, "a line's lanes within one wave");
  stockham_l<T, N, E, INV, 1, CL, WS>(v, t, acc, twl);
}

}  // namespace lp2d
