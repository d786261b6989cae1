// Kernel bodies of the angular-spectrum path (gfx950).
//
// Three tile shapes cover every operation:
//   rows_body   LINES rows of one field: [spatial prologue] -> row FFT -> [epilogue]
//   cols_body   W columns of one field:  [col FFT] -> [spectral middle] -> [col IFFT]
//   small_body  a whole n x n field resident in LDS (n <= 128 for c64, <= 64 for c128):
//               rows -> columns (+middle) -> rows, one HBM read and one HBM write.
// A field-to-field propagation of a large field is rows(fwd) -> cols(fwd,H,inv) ->
// rows(inv): three sweeps over the field; the intermediate planes live in the output
// buffer, so no workspace is needed.
//
// The transfer function is generated on chip: fx, fy from the fftfreq index, Kz^2 in
// FP64, Kz*z in FP64 reduced to its fractional cycle, then sin/cos in the field's
// precision.  That reduction is the only FP64 step of the c64 path.
#pragma once
#include "lp2d_fft.hpp"

namespace lp2d {

// ----------------------------------------------------------------- flags
enum : int {
  F_MASK0 = 1 << 0,      // spatial prologue: multiply by mask m0
  F_MASK1 = 1 << 1,      //                   multiply by mask m1
  F_LENS = 1 << 2,       //                   reference lens phase
  F_PHASE = 1 << 3,      //                   exp(i * phase[y, x])
  F_FIELD_OUT = 1 << 4,  //                   write the multiplied field
  F_H = 1 << 5,          // spectral middle: transfer function H(z)
  F_APERTURE = 1 << 6,   //                  circular pupil
  F_SMASK0 = 1 << 7,     //                  multiply by mask s0
  F_SMASK1 = 1 << 8,     //                  multiply by mask s1
  F_SPEC_OUT = 1 << 9,   //                  write the spectrum
  F_SCALE = 1 << 10,     // epilogue: multiply by scale
  F_INTENSITY = 1 << 11, //           write |v|^2
  F_GRAD = 1 << 12,      //           adjoint phase-mask gradient
};
constexpr int F_SPATIAL = F_MASK0 | F_MASK1 | F_LENS | F_PHASE | F_FIELD_OUT;
constexpr int F_MID = F_H | F_APERTURE | F_SMASK0 | F_SMASK1 | F_SPEC_OUT;

enum : int { MODE_FWD = 1, MODE_INV = 2 };
enum : int { EV_PASS = 0, EV_ZERO = 1, EV_DECAY = 2 };
enum : int { MK_U8 = 1, MK_F32 = 2, MK_F64 = 3, MK_C64 = 4, MK_C128 = 5 };

struct MaskD {
  const void* ptr;
  int dtype;
  int pad_;
  long long sb, sy, sx;
};

template <typename T>
struct PassArgs {
  const cx<T>* tw;      // concatenated twiddle tables, exp(-2*pi*i*k/N)
  const cx<T>* tw_inv;  // the same, conjugated (inverse passes)
  const cx<T>* src;
  cx<T>* dst;
  long long fstride;  // elements between consecutive fields
  int batch;
  int flags;
  // spatial prologue
  MaskD m0, m1;
  double dL;        // grid step (cm)
  double lens_rev;  // lens phase in revolutions = (X^2 + Y^2) * lens_rev
  const float* phase;
  long long phase_sb;
  cx<T>* field_out;
  // spectral middle
  double k0sq;     // (1/wl)^2
  double inv_ndL;  // 1/(n*dL): fftfreq step
  double z;
  double ap_r2;
  int evan;
  int pad2_;
  MaskD s0, s1;
  cx<T>* spec_out;
  // epilogue
  T scale;
  T* intensity;
  const cx<T>* u_saved;
  long long u_sb;
  float* grad_phase;
};

// ----------------------------------------------------------------- twiddles
// Concatenated per-size tables (PassArgs::tw): the table of size N (a power of two)
// starts at N - 2 and holds exp(-2*pi*i*k/N), k < N; PassArgs::tw_inv holds the
// conjugates.  Separate tables keep the forward and inverse twiddle loads distinct,
// so a kernel running both transforms never holds the forward ones live across the
// spectral stage.
constexpr int kMaxN = 8192;
constexpr int kTwiddleLen = 2 * kMaxN;

// ----------------------------------------------------------------- elementwise math
// exp(2*pi*i*r) for a reduced argument |r| <= 1 (revolutions).
__device__ __forceinline__ cx<float> expi_rev(float r) {
#ifdef LP2D_ACCURATE_SINCOS
  float s, c;
  sincospif(2.0f * r, &s, &c);
  return mk(c, s);
#else
  // v_sin_f32 / v_cos_f32 take their argument in revolutions.
  return mk(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
#endif
}
__device__ __forceinline__ cx<double> expi_rev(double r) {
  double s, c;
  sincospi(2.0 * r, &s, &c);
  return mk(c, s);
}

template <typename T>
__device__ __forceinline__ cx<T> exp_i_revolutions(double r) {
  return expi_rev((T)r);
}

template <typename T>
__device__ __forceinline__ cx<T> mask_mul(const MaskD& m, cx<T> v, int b, int y, int x) {
  const long long off = (long long)b * m.sb + (long long)y * m.sy + (long long)x * m.sx;
  switch (m.dtype) {
    case MK_U8: {
      const T s = (T)(((const uint8_t*)m.ptr)[off] != 0);
      return mk(v.re * s, v.im * s);
    }
    case MK_F32: {
      const T s = (T)((const float*)m.ptr)[off];
      return mk(v.re * s, v.im * s);
    }
    case MK_F64: {
      const T s = (T)((const double*)m.ptr)[off];
      return mk(v.re * s, v.im * s);
    }
    case MK_C64: {
      const cx<float> w = ((const cx<float>*)m.ptr)[off];
      return cmul(v, mk((T)w.re, (T)w.im));
    }
    default: {
      const cx<double> w = ((const cx<double>*)m.ptr)[off];
      return cmul(v, mk((T)w.re, (T)w.im));
    }
  }
}

constexpr double kInv2Pi = 0.15915494309189533577;

// Spatial prologue at (y, x) of field b (n x n).  Order follows the reference:
// f_gen mask, then f_init mask (beam2d.py:338, 341), lens phase (beam2d.py:508-510),
// phase screen.
template <typename T>
__device__ __forceinline__ cx<T> spatial_pro(const PassArgs<T>& a, cx<T> v, int b, int y, int x,
                                             int n) {
  const int f = a.flags;
  if (f & F_MASK0) v = mask_mul(a.m0, v, b, y, x);
  if (f & F_MASK1) v = mask_mul(a.m1, v, b, y, x);
  if (f & F_LENS) {
#pragma clang fp contract(off)
    const double X = (double)(x - n / 2) * a.dL;
    const double Y = (double)(y - n / 2) * a.dL;
    double r = (X * X + Y * Y) * a.lens_rev;
    r -= rint(r);
    v = cmul(v, exp_i_revolutions<T>(r));
  }
  if (f & F_PHASE) {
    double r = (double)a.phase[(long long)b * a.phase_sb + (long long)y * n + x] * kInv2Pi;
    r -= rint(r);
    v = cmul(v, exp_i_revolutions<T>(r));
  }
  return v;
}

// Spectral middle at frequency indices (ky, kx) of field b.
// H = exp(-2*pi*i*frac(Kz*z)) exactly as beam2d.py:485-489 (the integer cycles of
// Kz*z are dropped in FP64); Kz^2 <= 0 is evanescent: Kz := 0 (beam2d.py:307-308) so
// H = 1 unless the caller opted into EV_ZERO / EV_DECAY.
template <typename T>
__device__ __forceinline__ cx<T> spectral_mid(const PassArgs<T>& a, cx<T> v, int b, int ky,
                                              int kx, int n) {
  const int f = a.flags;
  if (f & (F_H | F_APERTURE)) {
#pragma clang fp contract(off)
    const int sx = kx < (n + 1) / 2 ? kx : kx - n;
    const int sy = ky < (n + 1) / 2 ? ky : ky - n;
    const double fx = (double)sx * a.inv_ndL;
    const double fy = (double)sy * a.inv_ndL;
    if (f & F_H) {
      const double kz2 = (a.k0sq - fx * fx) - fy * fy;
      if (kz2 > 0.0) {
        const double d = sqrt(kz2) * a.z;
        const double fr = d - trunc(d);
        v = cmul(v, exp_i_revolutions<T>(-fr));
      } else if (kz2 < 0.0 && a.evan == EV_ZERO) {
        v = mk((T)0, (T)0);
      } else if (kz2 < 0.0 && a.evan == EV_DECAY) {
        const T s = (T)exp(-2.0 * 3.14159265358979323846 * sqrt(-kz2) * fabs(a.z));
        v = mk(v.re * s, v.im * s);
      }
    }
    if (f & F_APERTURE) {
      if (fx * fx + fy * fy > a.ap_r2) v = mk((T)0, (T)0);
    }
  }
  if (f & F_SMASK0) v = mask_mul(a.s0, v, b, ky, kx);
  if (f & F_SMASK1) v = mask_mul(a.s1, v, b, ky, kx);
  return v;
}

// Epilogue at spatial (y, x) of field b; element offset `off` inside the field.
// Returns the value to store in dst.
template <typename T>
__device__ __forceinline__ cx<T> spatial_epi(const PassArgs<T>& a, cx<T> v, int b, int y, int x,
                                             int n, long long goff) {
  const int f = a.flags;
  if (f & F_SCALE) v = mk(v.re * a.scale, v.im * a.scale);
  if (f & F_GRAD) {
    // v = g1, the adjoint of the propagation applied to the output gradient.  Undo
    // the spatial prologue of the forward call: grad_in = conj(phase factor) *
    // conj(lens factor) * g1; grad_phase = Im(conj(u) * grad_in) with u the forward
    // input, which equals Im(conj(u * exp(i*phase)) * g1).
    if (a.phase) {
      double r = (double)a.phase[(long long)b * a.phase_sb + (long long)y * n + x] * kInv2Pi;
      r -= rint(r);
      v = cmulc(v, exp_i_revolutions<T>(r));
    }
    if (a.lens_rev != 0.0) {
#pragma clang fp contract(off)
      const double X = (double)(x - n / 2) * a.dL;
      const double Y = (double)(y - n / 2) * a.dL;
      double r = (X * X + Y * Y) * a.lens_rev;
      r -= rint(r);
      v = cmulc(v, exp_i_revolutions<T>(r));
    }
    if (a.grad_phase) {
      const cx<T> u = a.u_saved[(long long)b * a.u_sb + (long long)y * n + x];
      a.grad_phase[goff] = (float)(u.re * v.im - u.im * v.re);
    }
  }
  if (f & F_INTENSITY) a.intensity[goff] = v.re * v.re + v.im * v.im;
  return v;
}

// ----------------------------------------------------------------- row tiles
// LINES consecutive rows of field b starting at row y0.  Threads: (N/E) * LINES, the
// thread index inside a row is the fast one, so each load/store instruction of a wave
// covers contiguous memory.  lds: LINES * PaddedLen<N,PS> elements.
template <typename T, int N, int E, int LINES, bool INV, int PS>
__device__ __forceinline__ void rows_body(const PassArgs<T>& a, int b, int y0, int nrows,
                                          cx<T>* lds, int tid) {
  constexpr int TL = N / E;
  constexpr int L = PaddedLen<N, PS>::value;
  const int l = tid / TL, t = tid - (tid / TL) * TL;
  const int y = y0 + l;
  const bool valid = l < nrows;
  const long long frow = (long long)b * a.fstride + (long long)y * N;
  cx<T> v[E];
  const cx<T>* __restrict__ src = a.src + frow;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int x = t + TL * m;
    v[m] = valid ? src[x] : mk((T)0, (T)0);
  }
  if (a.flags & F_SPATIAL) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = t + TL * m;
      if (valid) {
        v[m] = spatial_pro(a, v[m], b, y, x, N);
        if (a.flags & F_FIELD_OUT) a.field_out[frow + x] = v[m];
      }
    }
  }
  LineAcc<T, PS> acc{lds + l * L};
  line_fft<T, N, E, INV>(v, t, acc, (INV ? a.tw_inv : a.tw) + (N - 2));
  cx<T>* __restrict__ dst = a.dst + frow;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int x = t + TL * m;
    if (valid) {
      const cx<T> o = (a.flags & (F_SCALE | F_GRAD | F_INTENSITY))
                          ? spatial_epi(a, v[m], b, y, x, N, frow + x)
                          : v[m];
      dst[x] = o;
    }
  }
}

// ----------------------------------------------------------------- column tiles
// W adjacent columns [x0, x0+W) of field b.  Threads: (N/E) * W with the column index
// fastest, so every load/store instruction reads W*sizeof(cx) contiguous bytes from
// each of 64/W rows.  lds: W * PaddedLen<N,PS> elements, interleaved by column.
template <typename T, int N, int E, int W, int MODE, int PS>
__device__ __forceinline__ void cols_body(const PassArgs<T>& a, int b, int x0, cx<T>* lds,
                                          int tid) {
  constexpr int TL = N / E;
  const int c = tid % W, t = tid / W;
  const int x = x0 + c;
  const long long fbase = (long long)b * a.fstride + x;
  cx<T> v[E];
  const cx<T>* __restrict__ src = a.src + fbase;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = src[(long long)(t + TL * m) * N];
  InterAcc<T, PS, W> acc{lds + c};
  if constexpr ((MODE & MODE_FWD) != 0) line_fft<T, N, E, false>(v, t, acc, a.tw + (N - 2));
  if (a.flags & F_MID) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int ky = t + TL * m;
      v[m] = spectral_mid(a, v[m], b, ky, x, N);
      if (a.flags & F_SPEC_OUT) a.spec_out[fbase + (long long)ky * N] = v[m];
    }
  }
  if constexpr ((MODE & MODE_INV) != 0) line_fft<T, N, E, true>(v, t, acc, a.tw_inv + (N - 2));
  cx<T>* __restrict__ dst = a.dst + fbase;
  const bool sc = (a.flags & F_SCALE) != 0;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    cx<T> o = v[m];
    if (sc) o = mk(o.re * a.scale, o.im * a.scale);
    dst[(long long)(t + TL * m) * N] = o;
  }
}

// ----------------------------------------------------------------- whole field in LDS
// Field b fully resident in LDS as an N x P tile (row pitch P, element (y, x) at
// y*P + pad(x)).  Threads: N * (N/E).  mode: MODE_FWD (2-D forward FFT) and/or
// MODE_INV (2-D inverse), with the spectral middle between them.
template <int N, int PS>
struct SmallPitch {
  static constexpr int value = N + (N >> PS);
};

template <typename T, int N, int E, int PS>
__device__ __forceinline__ void small_body(const PassArgs<T>& a, int b, int mode, cx<T>* lds,
                                           int tid) {
  constexpr int TL = N / E;
  constexpr int P = SmallPitch<N, PS>::value;
  const cx<T>* tw = a.tw + (N - 2);
  const cx<T>* twi = a.tw_inv + (N - 2);
  const long long fb = (long long)b * a.fstride;
  cx<T> v[E];
  // phase 1: rows (load, spatial prologue, forward row FFT)
  {
    const int y = tid / TL, t = tid % TL;
    const cx<T>* __restrict__ src = a.src + fb + (long long)y * N;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = src[t + TL * m];
    if (a.flags & F_SPATIAL) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int x = t + TL * m;
        v[m] = spatial_pro(a, v[m], b, y, x, N);
        if (a.flags & F_FIELD_OUT) a.field_out[fb + (long long)y * N + x] = v[m];
      }
    }
    LineAcc<T, PS> acc{lds + y * P};
    if (mode & MODE_FWD) line_fft<T, N, E, false>(v, t, acc, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) acc.st(t + TL * m, v[m]);
  }
  __syncthreads();
  // phase 2: columns (forward FFT, spectral middle, inverse FFT)
  {
    const int x = tid % N, t = tid / N;
    StrideAcc<T, P> acc{lds + pad_idx<PS>(x)};
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = acc.ld(t + TL * m);
    __syncthreads();  // the column FFT below overwrites rows other threads still read
    if (mode & MODE_FWD) line_fft<T, N, E, false>(v, t, acc, tw);
    if (a.flags & F_MID) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ky = t + TL * m;
        v[m] = spectral_mid(a, v[m], b, ky, x, N);
        if (a.flags & F_SPEC_OUT) a.spec_out[fb + (long long)ky * N + x] = v[m];
      }
    }
    if (mode & MODE_INV) line_fft<T, N, E, true>(v, t, acc, twi);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) acc.st(t + TL * m, v[m]);
  }
  __syncthreads();
  // phase 3: rows (inverse row FFT, epilogue, store)
  {
    const int y = tid / TL, t = tid % TL;
    LineAcc<T, PS> acc{lds + y * P};
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = acc.ld(t + TL * m);
    if (mode & MODE_INV) {
      __syncthreads();
      line_fft<T, N, E, true>(v, t, acc, twi);
    }
    cx<T>* __restrict__ dst = a.dst + fb + (long long)y * N;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = t + TL * m;
      const cx<T> o = (a.flags & (F_SCALE | F_GRAD | F_INTENSITY))
                          ? spatial_epi(a, v[m], b, y, x, N, fb + (long long)y * N + x)
                          : v[m];
      dst[x] = o;
    }
  }
}

}  // namespace lp2d
