// Kernel bodies of the angular-spectrum path (gfx950).
//
// Three tile shapes cover every operation:
//   rows_body   LINES rows of one field: [spatial prologue] -> row FFT -> [epilogue]
//   cols_body   W columns of one field:  [col FFT] -> [spectral middle] -> [col IFFT]
//   small_body  a whole n x n field resident in LDS (n <= 128 for c64, <= 64 for c128):
//               rows -> columns (+middle) -> rows, one HBM read and one HBM write.
// The *_pipe_body variants run the same work over a sequence of fields or row tiles,
// loading the next one while the current one is transformed.
// A field-to-field propagation of a large field is rows(fwd) -> cols(fwd,H,inv) ->
// rows(inv): three sweeps over the field; the intermediate planes live in the output
// buffer, so no workspace is needed (the adjoint can take one for its column pass).
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
  F_NO_STORE = 1 << 13,  //           skip the dst store (gradient-only epilogue)
  F_TILED = 1 << 14,     // the intermediate plane between the sweeps is block-tiled
};

// Block-tiled intermediate plane (F_TILED): rows in blocks of R = tile_rows(n), each
// block stored as C = tile_cols(n)-column strips of R x C elements, i.e. element (y, x)
// of an n x n plane at
//   (y/R)*R*n + (x/C)*R*C + (y%R)*C + x%C.
// A block of R rows occupies the same addresses as in row-major order, so a row pass
// over whole blocks converts in place; a column strip of width C then reads R*C
// contiguous elements per block instead of C-element pieces of R different rows.  The
// geometry follows each length's schedules (row tiles of R rows, strips of C columns):
//   8 x 8 up to 2048 points: 512-byte chunks, every 128-byte line written whole by one
//         store instruction (at 2048 the 4-column strips leave the plane row-major);
//   2 x 4 at 4096: each 2-row block stays contiguous, so the 2-row passes keep their
//         access pattern, and the 4-column strips read 64-byte chunks instead of 32.
// (4 x 4 tiles at 2048 sped the column pass up by 25 us of 240, slowed each row pass
// by about 10 us and doubled the adjoint's one-row gradient pass: not used.)
constexpr int tile_rows(int n) { return n >= 4096 ? 2 : 8; }
constexpr int tile_cols(int n) { return n >= 4096 ? 4 : 8; }

// Offset of (y, x) in the block-tiled layout of an N-point plane.
template <int N>
__device__ __forceinline__ unsigned tiled_off(int y, int x) {
  constexpr unsigned R = tile_rows(N), C = tile_cols(N);
  const unsigned uy = (unsigned)y, ux = (unsigned)x;
  return (uy / R) * R * N + (ux / C) * (R * C) + (uy % R) * C + ux % C;
}

// Thread -> (row l, lane t) for reading a block-tiled plane with row tiles of TR-row
// blocks: each wave takes TR rows x 64/TR lanes, ordered so that the wave's 64 loads
// of one element index cover 64 contiguous elements (whole tiles).  Needs TR*TL a
// multiple of 64 and 64/TR of TC.
template <int TL, int TR, int TC>
struct TiledWaveMap {
  static constexpr bool ok = (TR * TL) % 64 == 0 && (64 / TR) % TC == 0;
  static __device__ __forceinline__ void map(int tid, int& l, int& t) {
    const int g = tid / (TR * TL), r = tid - g * (TR * TL);
    const int w = r >> 6, lam = r & 63;
    l = g * TR + (lam / TC) % TR;
    t = w * (64 / TR) + (lam / (TR * TC)) * TC + lam % TC;
  }
};
// Thread -> (row l, lane t) for a row tile of LINES rows that reads (or writes) the
// block-tiled plane and writes (or reads) the natural one.  With 8 x 8 tiles and TL a
// multiple of 16, each wave takes 4 rows x 16 lanes of an 8-row block: every access
// of the tiled plane covers two 256-byte runs and every access of the natural plane
// whole 128-byte lines (with 8 x 8 lanes per wave the two halves of a natural line
// would come from two instructions, and the line would be written back twice).
// Otherwise TiledWaveMap (whole tiles per instruction).
template <int N, int TL, int LINES>
struct TiledRowMap {
  static constexpr int TR = tile_rows(N), TC = tile_cols(N);
  using WM = TiledWaveMap<TL, TR, TC>;
  static constexpr bool LINE16 = TR == 8 && TC == 8 && TL % 16 == 0;
  static constexpr bool ok = LINES % TR == 0 && TL % TC == 0 && (LINE16 || WM::ok);
  static __device__ __forceinline__ void map(int tid, int& l, int& t) {
    if constexpr (LINE16) {
      const int g = tid / (TR * TL), r = tid - g * (TR * TL);
      const int wv = r >> 6, ln = r & 63;
      l = g * TR + (ln >> 4) + 4 * (wv & 1);
      t = (ln & 15) + 16 * (wv >> 1);
    } else {
      WM::map(tid, l, t);
    }
  }
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

// Per-field optics of an uneven batch (device table, one entry per field).
struct FieldOpt {
  double k0sq, inv_ndL, z, ap_r2;
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
  // uneven batches: per-field base pointers and optics (null = strided batch)
  const cx<T>* const* src_tab;
  cx<T>* const* dst_tab;
  const FieldOpt* fopt;
};

// Global-address-space pointers: loads and stores through them are global_* with a
// uniform (SGPR) base and a 32-bit lane offset.  Through a generic pointer they would
// be flat_* with a 64-bit address per lane, which also count against the LDS wait
// counter and so stall every LDS wait behind outstanding memory traffic.
// (The host pass of this single-source file sees the same declarations without the
// qualifier: address spaces exist only in the device compilation.)
#if defined(__HIP_DEVICE_COMPILE__)
#define LP2D_GLB __attribute__((address_space(1)))
#else
#define LP2D_GLB
#endif
template <typename T>
__device__ __forceinline__ const LP2D_GLB T* glb(const T* p) { return (const LP2D_GLB T*)p; }
template <typename T>
__device__ __forceinline__ LP2D_GLB T* glb(T* p) { return (LP2D_GLB T*)p; }
// A pointer that is the same on every lane, moved to scalar registers.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

// Streaming accesses to field planes.  Every element of a plane is read once and
// written once per sweep, and a sweep's planes are far larger than the caches, so the
// loads and stores may carry the nontemporal hint (global_load/store ... nt).  A plain
// 256 MiB copy runs at 6.0 TB/s without the hint and 6.4 TB/s with it on MI355X, but
// the hint changes how the next sweep finds the plane, so it is chosen per access
// site by measurement.  LP2D_NT is a mask of the sites that carry it.  Measured at
// 32 x 1024^2 complex64: the row-pass loads gain 5-7 us per row pass and 8 us in the
// following column pass, the gradient pass 28 us; the inverse row pass's stores (the
// final field) leave the forward step unchanged and speed up the adjoint's column
// pass (130 -> 124 us), the speckle chain (5 %) and the uneven batch (1.5 %).  Hinted
// forward-row and column stores, and hinted column loads, were slower; the
// whole-field kernels are indifferent.
enum : int {
  NT_ROW_LD = 1,                   // row passes: loads
  NT_ROW_ST = 2,                   //             stores of the forward pass
  NT_COL_LD = 4, NT_COL_ST = 8,    // column passes
  NT_GRAD = 16,                    // adjoint gradient row pass
  NT_SMALL = 32,                   // whole-field kernels
  NT_ROWI_ST = 64,                 // row passes: stores of the inverse pass
  NT_TILE_ST = 128,                // whole-tile stores into the block-tiled plane
                                   // (4096 points: LDS-twiddle rows, chained passes)
};
#ifndef LP2D_NT
#define LP2D_NT (NT_ROW_LD | NT_GRAD | NT_ROWI_ST | NT_TILE_ST)
#endif
template <int SITE, typename T>
__device__ __forceinline__ cx<T> gld(const LP2D_GLB cx<T>* p) {
  if constexpr ((LP2D_NT & SITE) != 0) {
    typedef T V __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load((const LP2D_GLB V*)p);
    return mk(v.x, v.y);
  } else {
    return *p;
  }
}
template <int SITE, typename T>
__device__ __forceinline__ void gst(LP2D_GLB cx<T>* p, cx<T> x) {
  if constexpr ((LP2D_NT & SITE) != 0) {
    typedef T V __attribute__((ext_vector_type(2)));
    V v;
    v.x = x.re;
    v.y = x.im;
    __builtin_nontemporal_store(v, (LP2D_GLB V*)p);
  } else {
    *p = x;
  }
}

// Nontemporal loads of a natural-order row only when the TL lanes of one row cover
// whole 128-byte lines: the hint keeps a line out of the caches, so a line read in two
// halves by two instructions is fetched twice (measured 1.23 x at 256 points, 8 lanes).
template <typename T, int TL>
constexpr int nt_row_site() { return TL * (int)sizeof(cx<T>) >= 128 ? NT_ROW_LD : 0; }

// The same rule for natural-order stores of a row pass: `run` lanes of one wave store
// consecutive elements of a row (TL with the plain thread map; 16 or 64/tile_rows with
// the tiled-plane maps).  A line written in two halves with the hint is written back
// twice (measured 1.088x at 256 points, 8 lanes), without it the halves merge in L2.
template <typename T>
constexpr int nt_store_site(int site, int run) {
  return run * (int)sizeof(cx<T>) >= 128 ? site : 0;
}
template <int N, int TL, int LINES>
constexpr int tiled_map_run() {
  return TiledRowMap<N, TL, LINES>::LINE16 ? 16 : 64 / tile_rows(N);
}

// Base of field b's source / destination plane (b uniform across the workgroup).
template <typename T>
__device__ __forceinline__ const LP2D_GLB cx<T>* field_src(const PassArgs<T>& a, int b) {
  return glb(uniform_ptr(a.src_tab ? a.src_tab[b] : a.src + (long long)b * a.fstride));
}
template <typename T>
__device__ __forceinline__ LP2D_GLB cx<T>* field_dst(const PassArgs<T>& a, int b) {
  return glb(uniform_ptr(a.dst_tab ? a.dst_tab[b] : a.dst + (long long)b * a.fstride));
}
// The call's optics, or field b's when the batch carries a per-field table.
template <typename T>
__device__ __forceinline__ PassArgs<T> field_optics(const PassArgs<T>& a, int b) {
  PassArgs<T> r = a;
  if (a.fopt) {
    const FieldOpt o = a.fopt[b];
    r.k0sq = o.k0sq;
    r.inv_ndL = o.inv_ndL;
    r.z = o.z;
    r.ap_r2 = o.ap_r2;
  }
  return r;
}

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
template <typename T, bool MASKS = true>
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
      // branch-free per element (divergent branches in the unrolled loop cost
      // registers); only the uniform evanescent-mode test remains a branch
      const double kz2 = (a.k0sq - fx * fx) - fy * fy;
      const double kz = sqrt(kz2 > 0.0 ? kz2 : 0.0);  // nan_to_num(sqrt(<0)) = 0 -> H = 1
      const double d = kz * a.z;
      const double fr = d - trunc(d);
      cx<T> h = exp_i_revolutions<T>(-fr);
      if (a.evan != EV_PASS) {
        const T s = (a.evan == EV_ZERO)
                        ? (T)0
                        : (T)exp(-2.0 * 3.14159265358979323846 * sqrt(kz2 < 0.0 ? -kz2 : 0.0) *
                                 fabs(a.z));
        h = kz2 < 0.0 ? mk(s, (T)0) : h;
      }
      v = cmul(v, h);
    }
    if (f & F_APERTURE) {
      const bool out = fx * fx + fy * fy > a.ap_r2;
      v = mk(out ? (T)0 : v.re, out ? (T)0 : v.im);
    }
  }
  if constexpr (MASKS) {
    if (f & F_SMASK0) v = mask_mul(a.s0, v, b, ky, kx);
    if (f & F_SMASK1) v = mask_mul(a.s1, v, b, ky, kx);
  }
  return v;
}

// v = g1, the adjoint of the propagation applied to the output gradient.  Undo the
// spatial prologue of the forward call: returns grad_in = conj(phase factor) *
// conj(lens factor) * g1.  The phase gradient is then Im(conj(u) * grad_in) with u
// the forward input, which equals Im(conj(u * exp(i*phase) * lens) * g1).
template <typename T>
__device__ __forceinline__ cx<T> grad_fold(const PassArgs<T>& a, cx<T> v, int b, int y, int x,
                                           int n) {
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
  return v;
}

// ----------------------------------------------------------------- array forms
// The prologue and epilogue applied to the E elements x = t + TL*m of one line of a
// thread.  Each feature is one uniform branch around a branch-free loop over the E
// elements, so the loads of all elements can be issued together.
template <typename T, int E>
__device__ __forceinline__ void mask_mul_all(const MaskD& mk_, cx<T> (&v)[E], int b, int y,
                                             int t, int TL) {
  const long long base = (long long)b * mk_.sb + (long long)y * mk_.sy;
  switch (mk_.dtype) {
    case MK_U8:
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T s = (T)(((const uint8_t*)mk_.ptr)[base + (long long)(t + TL * m) * mk_.sx] != 0);
        v[m] = mk(v[m].re * s, v[m].im * s);
      }
      break;
    case MK_F32:
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T s = (T)((const float*)mk_.ptr)[base + (long long)(t + TL * m) * mk_.sx];
        v[m] = mk(v[m].re * s, v[m].im * s);
      }
      break;
    case MK_F64:
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T s = (T)((const double*)mk_.ptr)[base + (long long)(t + TL * m) * mk_.sx];
        v[m] = mk(v[m].re * s, v[m].im * s);
      }
      break;
    case MK_C64:
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const cx<float> w = ((const cx<float>*)mk_.ptr)[base + (long long)(t + TL * m) * mk_.sx];
        v[m] = cmul(v[m], mk((T)w.re, (T)w.im));
      }
      break;
    default:
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const cx<double> w = ((const cx<double>*)mk_.ptr)[base + (long long)(t + TL * m) * mk_.sx];
        v[m] = cmul(v[m], mk((T)w.re, (T)w.im));
      }
      break;
  }
}

// exp(i * (X^2 + Y^2) * k0/(2f)) for the E pixels of the line (reference lens phase,
// beam2d.py:508), conjugated when CONJ.
template <typename T, int E, bool CONJ>
__device__ __forceinline__ void lens_mul_all(const PassArgs<T>& a, cx<T> (&v)[E], int y, int t,
                                             int TL, int n) {
#pragma clang fp contract(off)
  const double Y = (double)(y - n / 2) * a.dL;
  const double Y2 = Y * Y;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double X = (double)(t + TL * m - n / 2) * a.dL;
    double r = (X * X + Y2) * a.lens_rev;
    r -= rint(r);
    const cx<T> e = exp_i_revolutions<T>(r);
    v[m] = CONJ ? cmulc(v[m], e) : cmul(v[m], e);
  }
}

// exp(i * phase[y, x]) (conjugated when CONJ) for the E pixels of the line.
template <typename T, int E, bool CONJ>
__device__ __forceinline__ void phase_mul_all(const PassArgs<T>& a, cx<T> (&v)[E], int b, int y,
                                              int t, int TL, int n) {
  const float* ph = a.phase + (long long)b * a.phase_sb + (long long)y * n + t;
  float p[E];
#pragma unroll
  for (int m = 0; m < E; ++m) p[m] = ph[TL * m];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    double r = (double)p[m] * kInv2Pi;
    r -= rint(r);
    const cx<T> e = exp_i_revolutions<T>(r);
    v[m] = CONJ ? cmulc(v[m], e) : cmul(v[m], e);
  }
}

template <typename T, int E>
__device__ __forceinline__ void spatial_pro_all(const PassArgs<T>& a, cx<T> (&v)[E], int b, int y,
                                                int t, int TL, int n) {
  const int f = a.flags;
  if (f & F_MASK0) mask_mul_all<T, E>(a.m0, v, b, y, t, TL);
  if (f & F_MASK1) mask_mul_all<T, E>(a.m1, v, b, y, t, TL);
  if (f & F_LENS) lens_mul_all<T, E, false>(a, v, y, t, TL, n);
  if (f & F_PHASE) phase_mul_all<T, E, false>(a, v, b, y, t, TL, n);
}

// Epilogue over the E pixels of one line (scale, gradient fold and phase gradient,
// intensity); `off` = element offset of pixel x = t in field b.
template <typename T, int E>
__device__ __forceinline__ void spatial_epi_all(const PassArgs<T>& a, cx<T> (&v)[E], int b, int y,
                                                int t, int TL, int n, long long off, bool valid) {
  const int f = a.flags;
  if (f & F_SCALE) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = mk(v[m].re * a.scale, v[m].im * a.scale);
  }
  if (f & F_GRAD) {
    if (a.phase) phase_mul_all<T, E, true>(a, v, b, y, t, TL, n);
    if (a.lens_rev != 0.0) lens_mul_all<T, E, true>(a, v, y, t, TL, n);
    if (a.grad_phase) {
      const cx<T>* u = a.u_saved + (long long)b * a.u_sb + (long long)y * n + t;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const cx<T> w = u[TL * m];
        if (valid) a.grad_phase[off + TL * m] = (float)(w.re * v[m].im - w.im * v[m].re);
      }
    }
  }
  if (f & F_INTENSITY) {
#pragma unroll
    for (int m = 0; m < E; ++m)
      if (valid) a.intensity[off + TL * m] = v[m].re * v[m].re + v[m].im * v[m].im;
  }
}

// ----------------------------------------------------------------- row tiles
// Clamp the row of an idle thread (past the end of a partial tile) onto a real row so
// that its loads stay in bounds; its results are never stored.
__device__ __forceinline__ int valid_row(int y, int ylast) { return y <= ylast ? y : ylast; }

// LINES consecutive rows of field b starting at row y0.  Threads: (N/E) * LINES, the
// thread index inside a row is the fast one, so each load/store instruction of a wave
// covers contiguous memory.  lds: LINES * PaddedLen<N,PS> elements.
// FEAT = false compiles the plain pass (load, FFT, optional scale, store): no
// prologue/epilogue code, so its register budget is the FFT's alone.
// LDS row pitch of a row tile: for tiles of 8 rows or more the padded line plus 4
// elements, so that the 8 rows of one wave (the block-tiled read below) start in
// different banks.
template <int N, int PS, int LINES>
struct RowPitch {
  static constexpr int value = PaddedLen<N, PS>::value + (LINES >= 8 ? 4 : 0);
};

// TW < 0: twiddles from the global tables; otherwise from the LDS tables twl of
// twl_fill with compact mode TW, and a forward pass into the block-tiled plane takes
// the whole-tile thread map with nontemporal stores (every store instruction writes
// whole lines; the column pass then reads the plane faster: 786 -> 739 us at 8 x
// 4096^2 for 22 us more here).  PHO (with FEAT off): the phase screen alone, applied
// before the transform as the featured prologue applies it.
template <typename T, int N, int E, int LINES, bool INV, int PS, bool FEAT = true, int TW = -1,
          bool PHO = false>
__device__ __forceinline__ void rows_body(const PassArgs<T>& a, int b, int y0, int nrows,
                                          cx<T>* lds, int tid, const cx<T>* twl = nullptr) {
  constexpr int TL = N / E;
  constexpr int L = RowPitch<N, PS, LINES>::value;
  constexpr int TR = tile_rows(N), TC = tile_cols(N);
  using WM = TiledWaveMap<TL, TR, TC>;
  // Thread -> (row l, lane t).  Row-major planes: each row's TL threads consecutive.
  // An inverse pass reading a block-tiled plane: each wave takes TR rows x 64/TR lanes,
  // so every load instruction of the wave reads whole tiles (8 x 8: one 512-byte tile).
  // (The same map for the forward pass of 2 x 4 tiles measured no different.)
  int l = tid / TL, t = tid - (tid / TL) * TL;
  if constexpr ((INV || TW >= 0) && TiledRowMap<N, TL, LINES>::ok) {
    if (a.flags & F_TILED) TiledRowMap<N, TL, LINES>::map(tid, l, t);
  }
  const int y = valid_row(y0 + l, y0 + nrows - 1);
  const bool valid = l < nrows;
  const long long frow = (long long)b * a.fstride + (long long)y * N;
  // uniform base (SGPRs) + one 32-bit per-lane offset
  const long long ubase = (long long)y0 * N;
  const unsigned lo = (unsigned)((y - y0) * N + t);
  // block-tiled offset of (y, x = t + TL*m): lo_t + TR*TL*m (y0 a multiple of TR, TL of
  // TC)
  const bool tiled = (a.flags & F_TILED) != 0;
  const unsigned lo_t = tiled_off<N>(l, t);
  cx<T> v[E];
  const LP2D_GLB cx<T>* __restrict__ src = field_src(a, b) + ubase;
  if (INV && tiled) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gld<NT_ROW_LD>(src + (lo_t + (unsigned)(tile_rows(N) * TL * m)));
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gld<nt_row_site<T, TL>()>(src + (lo + (unsigned)(TL * m)));
  }
  if constexpr (!FEAT && PHO) phase_mul_all<T, E, false>(a, v, b, y, t, TL, N);
  if constexpr (FEAT) {
    if (a.flags & F_SPATIAL) {
      spatial_pro_all<T, E>(a, v, b, y, t, TL, N);
      if (a.flags & F_FIELD_OUT) {
#pragma unroll
        for (int m = 0; m < E; ++m)
          if (valid) a.field_out[frow + t + TL * m] = v[m];
      }
    }
  }
  LineAcc<T, PS> acc{lds + l * L};
  if constexpr (TW >= 0)
    line_fft_l<T, N, E, INV, TW>(v, t, acc, twl);
  else
    line_fft<T, N, E, INV>(v, t, acc, (INV ? a.tw_inv : a.tw) + (N - 2));
  LP2D_GLB cx<T>* __restrict__ dst = field_dst(a, b) + ubase;
  // whole lines per store instruction: natural map / tiled-plane map
  constexpr bool WN = nt_store_site<T>(1, TL) != 0;
  constexpr bool WT =
      nt_store_site<T>(1, (INV && TiledRowMap<N, TL, LINES>::ok) ? tiled_map_run<N, TL, LINES>()
                                                                  : TL) != 0;
  if constexpr (FEAT) {
    spatial_epi_all<T, E>(a, v, b, y, t, TL, N, frow + t, valid);
    if ((a.flags & F_NO_STORE) == 0) {
      if (!INV && tiled) {
#pragma unroll
        for (int m = 0; m < E; ++m) gst<INV ? NT_ROWI_ST : NT_ROW_ST>(dst + (lo_t + (unsigned)(tile_rows(N) * TL * m)), v[m]);
      } else if (tiled ? WT : WN) {
#pragma unroll
        for (int m = 0; m < E; ++m)
          if (valid) gst<INV ? NT_ROWI_ST : NT_ROW_ST>(dst + (lo + (unsigned)(TL * m)), v[m]);
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m)
          if (valid) gst<0>(dst + (lo + (unsigned)(TL * m)), v[m]);
      }
    }
  } else {
    const T sc = (a.flags & F_SCALE) ? a.scale : (T)1;
    if (!INV && tiled) {
#pragma unroll
      for (int m = 0; m < E; ++m)
        gst<TW >= 0 ? NT_TILE_ST : NT_ROW_ST>(dst + (lo_t + (unsigned)(tile_rows(N) * TL * m)), mk(v[m].re * sc, v[m].im * sc));
    } else if (tiled ? WT : WN) {
#pragma unroll
      for (int m = 0; m < E; ++m)
        if (valid) gst<INV ? NT_ROWI_ST : NT_ROW_ST>(dst + (lo + (unsigned)(TL * m)), mk(v[m].re * sc, v[m].im * sc));
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m)
        if (valid) gst<0>(dst + (lo + (unsigned)(TL * m)), mk(v[m].re * sc, v[m].im * sc));
    }
  }
}

// Chained planes (a multi-plane stack u_{k+1} = P(dz)[u_k exp(i phi_k)]): the inverse
// row pass of plane k and the forward row pass of plane k+1 in one sweep over the
// block-tiled intermediate (F_TILED): load the tile, inverse row FFT, scale (1/n^2),
// store the plane u_{k+1} in natural order when a.field_out is set (F_FIELD_OUT, a
// checkpoint), multiply by exp(i*phase[y, x]) (the next screen), forward row FFT,
// store the tile in place.  The thread map (TiledRowMap) makes every tiled access
// cover whole tiles or 256-byte runs and every natural-order store whole 128-byte
// lines.  Three sweeps per plane become two.
// twl: null for the twiddles of the global table; otherwise the per-pass LDS tables of
// twl_fill (compact last pass when CL).
template <typename T, int N, int E, int LINES, int PS, int CL = 0>
__device__ __forceinline__ void rows_chain_body(const PassArgs<T>& a, int b, int y0,
                                                cx<T>* lds, const cx<T>* twl, int tid) {
  constexpr int TL = N / E;
  constexpr int L = RowPitch<N, PS, LINES>::value;
  constexpr int TR = tile_rows(N), TC = tile_cols(N);
  using RM = TiledRowMap<N, TL, LINES>;
  static_assert(RM::ok && N % LINES == 0, "whole tiles of whole row blocks");
  int l, t;
  RM::map(tid, l, t);
  const int y = y0 + l;
  const long long ubase = (long long)y0 * N;
  const unsigned lo_t = tiled_off<N>(l, t);
  constexpr unsigned TS = TR * TL;  // tiled offset step of element m
  cx<T> v[E];
  const LP2D_GLB cx<T>* __restrict__ src = field_src(a, b) + ubase;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = gld<NT_ROW_LD>(src + (lo_t + TS * (unsigned)m));
  LineAcc<T, PS> acc{lds + l * L};
  if (twl)
    line_fft_l<T, N, E, true, CL>(v, t, acc, twl);
  else
    line_fft<T, N, E, true>(v, t, acc, a.tw_inv + (N - 2));
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
  if (a.flags & F_FIELD_OUT) {
    LP2D_GLB cx<T>* __restrict__ ck =
        glb(uniform_ptr(a.field_out + (long long)b * a.fstride)) + (long long)y * N;
#pragma unroll
    for (int m = 0; m < E; ++m)
      gst<nt_store_site<T>(NT_ROWI_ST, tiled_map_run<N, TL, LINES>())>(ck + (t + TL * m), v[m]);
  }
  phase_mul_all<T, E, false>(a, v, b, y, t, TL, N);
  if (twl)
    line_fft_l<T, N, E, false, CL>(v, t, acc, twl);
  else
    line_fft<T, N, E, false>(v, t, acc, a.tw + (N - 2));
  LP2D_GLB cx<T>* __restrict__ dst = field_dst(a, b) + ubase;
  // (at 4096 points with the hint: this pass 529 -> 540 us, the column pass after it
  // 765-778 -> 736 us)
#pragma unroll
  for (int m = 0; m < E; ++m) gst<N >= 4096 ? NT_TILE_ST : NT_ROW_ST>(dst + (lo_t + TS * (unsigned)m), v[m]);
}

// Adjoint of a multi-plane stack, one sweep per plane.  For the fields [b0, b1) of a
// group and one row tile (whole row blocks of the block-tiled plane, in place): load
// the tile, inverse row FFT, 1/n^2, fold g = conj(exp(i*phase[y, x])) * g1 (the screen
// shared by the batch), add Im(conj(u) * g) to the group's phase-gradient sums (u:
// the plane's saved input, natural order), then either (LAST) store g in natural
// order -- the gradient of the stack's input, unless F_NO_STORE -- or run the forward
// row FFT of g and store the tile in place: the first sweep of the previous plane's
// adjoint.  The group's sums go to `partial` (n x n) once; with partial null no
// gradient is formed (u is not read).  Twiddles from the LDS tables twl, or from the
// global tables when twl is null.
template <typename T, int N, int E, int LINES, int PS, int CL, bool LAST>
__device__ __forceinline__ void rows_adj_chain_body(const PassArgs<T>& a, int b0, int b1, int y0,
                                                    float* __restrict__ partial, cx<T>* lds,
                                                    const cx<T>* twl, int tid) {
  constexpr int TL = N / E;
  constexpr int L = RowPitch<N, PS, LINES>::value;
  constexpr int TR = tile_rows(N);
  using RM = TiledRowMap<N, TL, LINES>;
  static_assert(RM::ok && N % LINES == 0, "whole tiles of whole row blocks");
  int l, t;
  RM::map(tid, l, t);
  const int y = y0 + l;
  const long long ubase = (long long)y0 * N;
  const unsigned lo_t = tiled_off<N>(l, t);
  constexpr unsigned TS = TR * TL;
  const bool store = (a.flags & F_NO_STORE) == 0;
  LineAcc<T, PS> acc{lds + l * L};
  float gsum[E];
#pragma unroll
  for (int m = 0; m < E; ++m) gsum[m] = 0.f;
  for (int b = b0; b < b1; ++b) {
    // opaque per field: otherwise the twiddle reads (and the screen factors), the same
    // for every field, are hoisted out of the loop and held across it
    int tt = t;
    asm volatile("" : "+v"(tt));
    cx<T> v[E];
    const LP2D_GLB cx<T>* __restrict__ src = field_src(a, b) + ubase;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gld<NT_ROW_LD>(src + (lo_t + TS * (unsigned)m));
    if (twl)
      line_fft_l<T, N, E, true, CL>(v, tt, acc, twl);
    else
      line_fft<T, N, E, true>(v, tt, acc, a.tw_inv + (N - 2));
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
    phase_mul_all<T, E, true>(a, v, b, y, tt, TL, N);
    if (partial) {
      const LP2D_GLB cx<T>* __restrict__ us =
          glb(uniform_ptr(a.u_saved + (long long)b * a.u_sb)) + (long long)y * N;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const cx<T> u = gld<NT_GRAD>(us + (t + TL * m));
        gsum[m] += (float)(u.re * v[m].im - u.im * v[m].re);
      }
    }
    if constexpr (LAST) {
      if (store) {
        LP2D_GLB cx<T>* __restrict__ dst = field_dst(a, b) + (long long)y * N;
#pragma unroll
        for (int m = 0; m < E; ++m)
          gst<nt_store_site<T>(NT_ROWI_ST, tiled_map_run<N, TL, LINES>())>(dst + (t + TL * m),
                                                                          v[m]);
      }
    } else {
      int t2 = tt;
      asm volatile("" : "+v"(t2));
      if (twl)
        line_fft_l<T, N, E, false, CL>(v, t2, acc, twl);
      else
        line_fft<T, N, E, false>(v, t2, acc, a.tw + (N - 2));
      LP2D_GLB cx<T>* __restrict__ dst = field_dst(a, b) + ubase;
#pragma unroll
      for (int m = 0; m < E; ++m)
        gst<N >= 4096 ? NT_TILE_ST : NT_ROW_ST>(dst + (lo_t + TS * (unsigned)m), v[m]);
    }
  }
  if (partial) {
#pragma unroll
    for (int m = 0; m < E; ++m) partial[(long long)y * N + t + TL * m] = gsum[m];
  }
}

// Row tiles of LINES rows with software pipelining: the tile of step k+1 is loaded
// into registers while tile k is transformed, as in the pipelined column pass.  A
// workgroup takes tiles u0, u0 + ustep, u0 + 2*ustep, ... of the whole batch, so the
// resident workgroups sweep the planes together.  Twiddles come from LDS (twl), so
// nothing in the loop waits for the prefetch before it is needed.  Strided batches
// of complex64 without spatial features, except (PH, forward pass) a phase screen:
// its values are prefetched with the tile and applied before the next prefetch is
// issued.  TILED: the forward pass stores, the inverse pass loads the block-tiled
// layout.
template <typename T, int N, int E, int LINES, bool INV, int PS, bool TILED, bool PH = false,
          bool WSK = true>
__device__ __forceinline__ void rows_pipe_body(const PassArgs<T>& a, int u0, int ustep,
                                               int nunits, cx<T>* lds, const cx<T>* twl,
                                               int tid) {
  static_assert(sizeof(T) == 4, "complex64 path");
  constexpr int TL = N / E;
  constexpr int L = RowPitch<N, PS, LINES>::value;
  constexpr int TPF = N / LINES;  // row tiles per field
  static_assert(N % LINES == 0, "whole tiles");
  constexpr int TR = tile_rows(N), TC = tile_cols(N);
  using WM = TiledWaveMap<TL, TR, TC>;
  static_assert(!TILED || (LINES % TR == 0 && TL % TC == 0), "tiled shape");
  int l = tid / TL, t = tid - (tid / TL) * TL;
  if constexpr (INV && TILED) {
    if constexpr (TR == 8 && TC == 8 && TL % 16 == 0) {
      // each wave takes 4 rows x 16 lanes of an 8-row block: every load instruction
      // reads two 256-byte runs of the tiled plane, every store writes whole 128-byte
      // lines of the natural one (with 8 x 8 lanes per wave the two halves of a line
      // come from two waves, and the line is written back twice)
      const int g = tid / (TR * TL), r = tid - g * (TR * TL);
      const int wv = r >> 6, ln = r & 63;
      l = g * TR + (ln >> 4) + 4 * (wv & 1);
      t = (ln & 15) + 16 * (wv >> 1);
    } else if constexpr (WM::ok) {
      // TR rows x 64/TR lanes: every load instruction reads whole tiles
      WM::map(tid, l, t);
    }
  }
  // lane offsets inside a tile of LINES rows: natural (y, x) at y*N + x, tiled see
  // F_TILED; element m of the thread is x = t + TL*m
  const unsigned lo_n = (unsigned)(l * N + t);
  const unsigned lo_t = tiled_off<N>(l, t);
  constexpr bool LD_T = INV && TILED, ST_T = !INV && TILED;
  const unsigned ld_o = LD_T ? lo_t : lo_n, st_o = ST_T ? lo_t : lo_n;
  constexpr unsigned LD_S = LD_T ? TR * TL : TL, ST_S = ST_T ? TR * TL : TL;
  // lanes of one wave storing consecutive elements of a natural-order row
  constexpr int ST_RUN = (INV && TILED) ? ((TR == 8 && TC == 8 && TL % 16 == 0) ? 16
                                           : WM::ok ? 64 / TR : TL)
                                        : TL;
  // each row's TL lanes in one wave (then the transform's exchanges synchronise one
  // wave): the natural map up to 64 lanes, the tiled-plane maps when a wave holds
  // whole rows.  Not for a layout conversion in place (WSK false): there the
  // workgroup barriers inside the transform are what keep one wave's stores off
  // data another wave of the workgroup has yet to load.
  constexpr bool WS = WSK && TL <= 64 && 64 % TL == 0 &&
                      (!(INV && TILED) ||
                       ((TR == 8 && TC == 8 && TL % 16 == 0) ? TL == 16
                        : WM::ok ? TL == 64 / TR : true));
  LineAcc<T, PS> acc{lds + l * L};
  const bool scale = (a.flags & F_SCALE) != 0;
  auto tile_off = [&](int u) {
    const int b = u / TPF;
    return (long long)b * a.fstride + (long long)(u - b * TPF) * LINES * N;
  };
  // Step u moves tile u (prefetched by the previous step) from nxt into cur, requests
  // tile u + ustep into nxt, then transforms cur and stores it; the first step is
  // peeled so that the loop head's wait only ever covers the prefetch.
  static_assert(!PH || !INV, "the screen is a forward-pass prologue");
  cx<T> cur[E], nxt[E];
  float phn[PH ? E : 1];
  auto fetch = [&](int u) {
    const LP2D_GLB cx<T>* __restrict__ src = glb(uniform_ptr(a.src + tile_off(u)));
#pragma unroll
    for (int m = 0; m < E; ++m)
      nxt[m] = gld<LD_T ? NT_ROW_LD : nt_row_site<T, TL>()>(src + (ld_o + LD_S * (unsigned)m));
    if constexpr (PH) {
      const int b = u / TPF;
      const LP2D_GLB float* __restrict__ ph =
          glb(uniform_ptr(a.phase + (long long)b * a.phase_sb +
                          (long long)(u - b * TPF) * LINES * N)) + lo_n;
#pragma unroll
      for (int m = 0; m < E; ++m) phn[m] = ph[TL * m];
    }
  };
  auto step = [&](int u) {
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = nxt[m];
    if constexpr (PH) {  // x exp(i*phase): FP64 reduction to a fraction of a turn
#pragma unroll
      for (int m = 0; m < E; ++m) {
        double r = (double)phn[m] * kInv2Pi;
        r -= rint(r);
        cur[m] = cmul(cur[m], exp_i_revolutions<T>(r));
      }
    }
    if (u + ustep < nunits) fetch(u + ustep);  // uniform branch
    line_fft_l<T, N, E, INV, 0, WS>(cur, t, acc, twl);
    if (scale) {
#pragma unroll
      for (int m = 0; m < E; ++m) cur[m] = cscale(cur[m], a.scale);
    }
    LP2D_GLB cx<T>* __restrict__ dst = glb(uniform_ptr(a.dst + tile_off(u)));
#pragma unroll
    for (int m = 0; m < E; ++m)
      gst<ST_T ? NT_ROW_ST : nt_store_site<T>(INV ? NT_ROWI_ST : NT_ROW_ST, ST_RUN)>(
          dst + (st_o + ST_S * (unsigned)m), cur[m]);
  };
  if (u0 >= nunits) return;
  fetch(u0);
  step(u0);
  for (int u = u0 + ustep; u < nunits; u += ustep) step(u);
}

// Thread -> (row l, lane t) of a gradient row tile of LINES rows.  Reading a
// block-tiled plane, each wave takes all LINES rows x 64/LINES lanes, so every load
// instruction reads whole 128-byte lines at 8 x 8 tiles (rows y, y+1 of a tile are
// adjacent); with one row per wave each line would be read in two halves by two waves.
template <int TL, int LINES>
__device__ __forceinline__ void grad_thread_map(int tid, bool tiled, int& l, int& t) {
  l = tid / TL;
  t = tid - (tid / TL) * TL;
  if constexpr (64 % LINES == 0 && TL % (64 / LINES) == 0 && LINES * TL >= 64) {
    if (tiled) {
      constexpr int LPW = 64 / LINES;  // lanes per row in one wave
      const int lane = tid & 63, w = tid >> 6;
      l = lane / LPW;
      t = w * LPW + (lane - (lane / LPW) * LPW);
    }
  }
}

// Last row pass of the adjoint for fields [b0, b1) that share one phase screen: the
// inverse row FFT and gradient fold per field, with the phase gradient of every pixel
// summed over the fields in registers and written once to `partial` (n x n).  The
// screen factor conj(exp(i*phase) * lens) is the same for every field, so it is
// computed once per pixel before the field loop.
// HOIST: the screen factor of the thread's E elements is computed once and kept in
// registers across the group's fields (small E); otherwise it is recomputed per field
// (large E, where holding it would spill).
template <typename T, int N, int E, int LINES, int PS, bool HOIST = true>
__device__ __forceinline__ void rows_grad_body(const PassArgs<T>& a, int b0, int b1, int y0,
                                               int nrows, float* __restrict__ partial,
                                               cx<T>* lds, int tid) {
  constexpr int TL = N / E;
  constexpr int L = PaddedLen<N, PS>::value;
  const bool tiled = (a.flags & F_TILED) != 0;
  int l, t;
  grad_thread_map<TL, LINES>(tid, tiled, l, t);
  const int y = valid_row(y0 + l, y0 + nrows - 1);
  const bool valid = l < nrows;
  const bool store = (a.flags & F_NO_STORE) == 0;
  LineAcc<T, PS> acc{lds + l * L};
  cx<T> ec[HOIST ? E : 1];
  float gsum[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    if constexpr (HOIST) ec[m] = grad_fold(a, mk(a.scale, (T)0), 0, y, t + TL * m, N);
    gsum[m] = 0.f;
  }
  // block-tiled offset of (y, x = t + TL*m) in the field: lo_t + R*TL*m (TL a multiple
  // of C); any row tile height can read it, storing in place needs whole blocks
  const unsigned lo_t = tiled_off<N>(y, t);
  for (int b = b0; b < b1; ++b) {
    const long long frow = (long long)b * a.fstride + (long long)y * N;
    cx<T> v[E];
    if (tiled) {
      const LP2D_GLB cx<T>* __restrict__ src = glb(a.src + (long long)b * a.fstride);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = gld<NT_GRAD>(src + (lo_t + (unsigned)(tile_rows(N) * TL * m)));
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = a.src[frow + t + TL * m];
    }
    line_fft<T, N, E, true>(v, t, acc, a.tw_inv + (N - 2));
    const cx<T>* __restrict__ us = a.u_saved + (long long)b * a.u_sb + (long long)y * N;
    int yy = y, tt = t;
    if constexpr (!HOIST) asm volatile("" : "+v"(yy), "+v"(tt));  // keep it per field
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = t + TL * m;
      const cx<T> g =
          HOIST ? cmul(v[m], ec[HOIST ? m : 0])
                : cmul(v[m], grad_fold(a, mk(a.scale, (T)0), 0, yy, tt + TL * m, N));
      const cx<T> u = us[x];
      gsum[m] += (float)(u.re * g.im - u.im * g.re);
      if (store && valid) a.dst[frow + x] = g;
    }
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < E; ++m) partial[(long long)y * N + t + TL * m] = gsum[m];
  }
}

// ----------------------------------------------------------------- column tiles
// W adjacent columns [x0, x0+W) of field b.  Threads: (N/E) * W with the column index
// fastest, so every load/store instruction reads W*sizeof(cx) contiguous bytes from
// each of 64/W rows.  lds: W * PaddedLen<N,PS> elements, interleaved by column.
// MIDK selects the spectral stage compiled in: 0 none, 1 transfer function +
// aperture + spectrum store, 2 everything (adds the mask multipliers), 3 as 1 with the
// factor shared by a group of fields.
// The workgroup runs fields [b0, b1) of the same strip in turn.  With MIDK == 3 the
// spectral factor H(kx, ky) * aperture is the same for every field of the call, so it
// is computed once per element before the field loop and kept in registers (2E more
// live registers: a separate schedule for the autotuner to weigh).
template <typename T, int N, int E, int W, int MODE, int PS, int MIDK = 2>
__device__ __forceinline__ void cols_body(const PassArgs<T>& a, int b0, int b1, int x0,
                                          cx<T>* lds, int tid) {
  constexpr int TL = N / E;
  const int c = tid % W, t = tid / W;
  const int x = x0 + c;
  const unsigned lo = (unsigned)(t * N + c);
  InterAcc<T, PS, W> acc{lds + c};
  const bool mid = (a.flags & F_MID) != 0;
  const bool spec_out = (a.flags & F_SPEC_OUT) != 0;
  cx<T> h[MIDK == 3 ? E : 1];
  if constexpr (MIDK == 3) {  // (never used with per-field optics)
#pragma unroll
    for (int m = 0; m < E; ++m)
      h[m] = mid ? spectral_mid<T, false>(a, mk((T)1, (T)0), 0, t + TL * m, x, N)
                 : mk((T)1, (T)0);
  }
  // Only the shared-factor class loops over fields: in the others a loop would let
  // the compiler hoist the per-element factor out of it (all E values live at once).
  const int bend = (MIDK == 3) ? b1 : b0 + 1;
  for (int b = b0; b < bend; ++b) {
    // uniform base (SGPRs) + one 32-bit per-lane offset: every load/store of the
    // strip addresses as base + lane offset + compile-time constant
    const long long ubase = (long long)b * a.fstride + x0;
    cx<T> v[E];
    // the shared-factor class never runs on per-field tables (their optics differ)
    const LP2D_GLB cx<T>* __restrict__ src =
        (MIDK == 3 ? glb(a.src + (long long)b * a.fstride) : field_src(a, b)) + x0;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gld<NT_COL_LD>(src + (lo + (unsigned)(TL * N * m)));
    if constexpr ((MODE & MODE_FWD) != 0) line_fft<T, N, E, false>(v, t, acc, a.tw + (N - 2));
    if constexpr (MIDK == 3) {
      if (mid) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], h[m]);
      }
    } else if constexpr (MIDK == 1) {
      if (mid) {
        const PassArgs<T> ab = field_optics(a, b);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = spectral_mid<T, false>(ab, v[m], b, t + TL * m, x, N);
      }
    } else if constexpr (MIDK == 2) {
      if (mid) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = spectral_mid<T, true>(a, v[m], b, t + TL * m, x, N);
      }
    }
    if constexpr (MIDK > 0) {
      if (spec_out) {
#pragma unroll
        for (int m = 0; m < E; ++m) a.spec_out[ubase + (lo + (unsigned)(TL * N * m))] = v[m];
      }
    }
    // Keep the inverse transform's twiddle loads below the spectral stage: hoisted
    // above it they stay live across it and cost registers.
    asm volatile("" ::: "memory");
    if constexpr ((MODE & MODE_INV) != 0)
      line_fft<T, N, E, true>(v, t, acc, a.tw_inv + (N - 2));
    LP2D_GLB cx<T>* __restrict__ dst =
        (MIDK == 3 ? glb(a.dst + (long long)b * a.fstride) : field_dst(a, b)) + x0;
    const T sc = (a.flags & F_SCALE) ? a.scale : (T)1;
#pragma unroll
    for (int m = 0; m < E; ++m) gst<NT_COL_ST>(dst + (lo + (unsigned)(TL * N * m)), mk(v[m].re * sc, v[m].im * sc));
  }
}

// The gradient pass above, software-pipelined over the group's fields: the rows of
// field b+1 (the plane and the saved forward input) are loaded into registers while
// field b is transformed; twiddles come from LDS (twl) and the first field is peeled
// off the loop, so every wait in the loop covers the prefetch only.  Screen factor
// hoisted (small E).
template <typename T, int N, int E, int LINES, int PS>
__device__ __forceinline__ void rows_grad_pipe_body(const PassArgs<T>& a, int b0, int b1, int y0,
                                                    int nrows, float* __restrict__ partial,
                                                    cx<T>* lds, const cx<T>* twl, int tid) {
  constexpr int TL = N / E;
  constexpr int L = PaddedLen<N, PS>::value;
  const bool tiled = (a.flags & F_TILED) != 0;
  int l, t;
  grad_thread_map<TL, LINES>(tid, tiled, l, t);
  const int y = valid_row(y0 + l, y0 + nrows - 1);
  const bool valid = l < nrows;
  const bool store = (a.flags & F_NO_STORE) == 0;
  LineAcc<T, PS> acc{lds + l * L};
  cx<T> ec[E];
  float gsum[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    ec[m] = grad_fold(a, mk(a.scale, (T)0), 0, y, t + TL * m, N);
    gsum[m] = 0.f;
  }
  const unsigned lo_n = (unsigned)(y * N + t);
  const unsigned lo_t = tiled_off<N>(y, t);
  cx<T> cur[E], nxt[E], ucur[E], unxt[E];
  auto fetch = [&](int bf) {
    const LP2D_GLB cx<T>* __restrict__ src = glb(a.src + (long long)bf * a.fstride);
    const LP2D_GLB cx<T>* __restrict__ us = glb(a.u_saved + (long long)bf * a.u_sb);
    if (tiled) {
#pragma unroll
      for (int m = 0; m < E; ++m) nxt[m] = gld<NT_GRAD>(src + (lo_t + (unsigned)(tile_rows(N) * TL * m)));
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) nxt[m] = gld<NT_GRAD>(src + (lo_n + (unsigned)(TL * m)));
    }
#pragma unroll
    for (int m = 0; m < E; ++m) unxt[m] = gld<NT_GRAD>(us + (lo_n + (unsigned)(TL * m)));
  };
  auto step = [&](int b) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      cur[m] = nxt[m];
      ucur[m] = unxt[m];
    }
    if (b + 1 < b1) fetch(b + 1);  // uniform branch
    line_fft_l<T, N, E, true>(cur, t, acc, twl);
    LP2D_GLB cx<T>* __restrict__ dst = glb(a.dst + (long long)b * a.fstride);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const cx<T> g = cmul(cur[m], ec[m]);
      gsum[m] += (float)(ucur[m].re * g.im - ucur[m].im * g.re);
      if (store && valid) gst<NT_GRAD>(dst + (lo_n + (unsigned)(TL * m)), g);
    }
  };
  if (b0 < b1) {  // an empty group only writes its zero partial sums
    fetch(b0);
    step(b0);
    for (int b = b0 + 1; b < b1; ++b) step(b);
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < E; ++m) partial[(long long)y * N + t + TL * m] = gsum[m];
  }
}

// ----------------------------------------------------------------- pipelined columns
// frac(Kz*z) in cycles for the complex64 path, with Kz = sqrt(max(kz2, 0)).  The square
// root is v_rsq_f64 refined by one Newton step in FP64 (relative error ~1e-14, far
// below what the FP32 phase resolves) instead of the correctly rounded expansion.
// kz2 <= 0 (evanescent, Kz := 0 as beam2d.py:307-308) clamps to 1e-300: then
// Kz*z ~ 1e-150*z and H = 1.  fract() differs from the reference's d - trunc(d) by
// whole cycles only (for z < 0), which leave exp(-2*pi*i*fr) unchanged.
__device__ __forceinline__ float kz_frac_fast(double kz2, double z) {
  const double q = fmax(kz2, 1e-300);
  const double y = __builtin_amdgcn_rsq(q);
  double s = q * y;
  s = fma(fma(-s, s, q), 0.5 * y, s);
  return (float)__builtin_amdgcn_fract(s * z);
}

// Transfer function x aperture of one column, frequency fx: per-column constants.
struct ColMid {
  double c0;   // k0^2 - fx^2
  double fx2;  // fx^2
  double z;    // 0 when the call has no transfer function (then H = 1)
  double ap;   // squared pupil radius, +inf without a pupil
  double inv_ndL;
};
__device__ __forceinline__ ColMid col_mid(const FieldOpt& o, int flags, int sx) {
#pragma clang fp contract(off)
  ColMid m;
  const double fx = (double)sx * o.inv_ndL;
  m.fx2 = fx * fx;
  m.c0 = o.k0sq - m.fx2;
  m.z = (flags & F_H) ? o.z : 0.0;
  m.ap = (flags & F_APERTURE) ? o.ap_r2 : __builtin_huge_val();
  m.inv_ndL = o.inv_ndL;
  return m;
}
// v * H(fy) * pupil(fy) at signed frequency index sy of the column (complex64 path,
// evanescent modes passed through).  Branch-free: same operation order as
// spectral_mid for Kz^2 and the pupil test.
__device__ __forceinline__ cx<float> mid_fast(const ColMid& m, cx<float> v, int sy) {
#pragma clang fp contract(off)
  const double fy = (double)sy * m.inv_ndL;
  const double fy2 = fy * fy;
  const float fr = kz_frac_fast(m.c0 - fy2, m.z);
  const cx<float> h = mk(__builtin_amdgcn_cosf(fr), __builtin_amdgcn_sinf(fr));
  const bool out = m.fx2 + fy2 > m.ap;
  const cx<float> r = cmulc(v, h);  // exp(-2*pi*i*fr)
  return mk(out ? 0.0f : r.re, out ? 0.0f : r.im);
}

// The FP64 part of mid_fast at signed frequency index sy: the fractional phase of
// Kz*z and the pupil test (same operations, same order).
__device__ __forceinline__ float mid_phase(const ColMid& m, int sy, bool& out) {
#pragma clang fp contract(off)
  const double fy = (double)sy * m.inv_ndL;
  const double fy2 = fy * fy;
  out = m.fx2 + fy2 > m.ap;
  return kz_frac_fast(m.c0 - fy2, m.z);
}

// One field of a pipelined column group, staged in LDS at kernel start so that the
// field loop reads its pointers and optics without global loads (a global load there
// would wait behind the prefetch and the previous field's stores).
struct PipeField {
  const void* src;
  void* dst;
  FieldOpt o;
};
constexpr int kPipeMaxPer = 32;

template <typename T>
__device__ __forceinline__ void pipe_stage_fields(const PassArgs<T>& a, int b0, int b1,
                                                  PipeField* pf, int tid) {
  if (tid < b1 - b0) {
    const int b = b0 + tid;
    PipeField f;
    f.src = a.src_tab ? (const void*)a.src_tab[b] : (const void*)(a.src + (long long)b * a.fstride);
    f.dst = a.dst_tab ? (void*)a.dst_tab[b] : (void*)(a.dst + (long long)b * a.fstride);
    if (a.fopt) {
      f.o = a.fopt[b];
    } else {
      f.o.k0sq = a.k0sq;
      f.o.inv_ndL = a.inv_ndL;
      f.o.z = a.z;
      f.o.ap_r2 = a.ap_r2;
    }
    pf[tid] = f;
  }
}

// Column strips of a group of fields [b0, b1) with software pipelining: the strip of
// field b+1 is loaded into registers while field b is transformed, so the memory
// traffic of one field overlaps the arithmetic of the previous one instead of every
// workgroup alternating between the two in step.  Twiddles come from LDS (twl) and
// the fields' pointers and optics from pf, so that nothing in the loop waits for the
// prefetch before it is needed.  Column FFT -> H, pupil -> column IFFT; complex64
// only, evanescent modes pass through (EV_PASS), no mask multipliers.
// Calls the pipelined class takes (host side): the fused forward+inverse pass of the
// complex64 path without mask multipliers, evanescent modes passed through.
template <typename T>
inline bool pipe_eligible(const PassArgs<T>& a, int mode) {
  return sizeof(T) == 4 && mode == (MODE_FWD | MODE_INV) && a.evan == EV_PASS &&
         !(a.flags & (F_SMASK0 | F_SMASK1));
}

// HOISTH: the strip's transfer function is computed once, before the field loop, for
// every field of the group (all fields of the call share one set of optics: a strided
// batch without per-field tables).
//   HOIST_FACTOR: exp(-2*pi*i*fr) x pupil held in registers (2E floats); the product
//                 equals mid_fast's up to the sign of a zero.
//   HOIST_PHASE:  only the FP64 part, the fractional phase fr and the pupil test, held
//                 in registers (E floats and one bit mask); each field evaluates sin/cos
//                 and the product is mid_fast's.
// Measured per call (pipelined 8-column strips): 1024^2 x32 with 32 elements per thread
// 118.9 us, factor 113.9, phase 121.8; 256^2 x64 with 16 elements 22.6 us, phase 20.2,
// factor 22.2; 512^2 x64 with 16: 66.3, phase 62.0, factor 62.7.
enum : int { HOIST_NONE = 0, HOIST_FACTOR = 1, HOIST_PHASE = 2 };
template <typename T, int N, int E, int W, int PS, int HOISTH = HOIST_NONE, int CL = 0>
__device__ __forceinline__ void cols_pipe_body(const PassArgs<T>& a, int b0, int b1, int x0,
                                               cx<T>* lds, const cx<T>* twl,
                                               const PipeField* pf, int tid) {
  static_assert(sizeof(T) == 4, "complex64 path");
  constexpr int TL = N / E;
  const int c = tid % W, t = tid / W;
  const int x = x0 + c;
  const int sx = x < N / 2 ? x : x - N;
  // natural: (y, x) at y*N + x; block-tiled (W == tile_cols(N)): see F_TILED.  Row
  // y = t + TL*m is at lo + TL*N*m in both (TL a multiple of tile_rows(N) when tiled).
  const bool tiled = (a.flags & F_TILED) != 0;
  const unsigned lo_n = (unsigned)(t * N + c);
  const unsigned lo = tiled ? tiled_off<N>(t, c) : lo_n;
  const int xoff = tiled ? (x0 / tile_cols(N)) * (tile_rows(N) * tile_cols(N)) : x0;
  InterAcc<T, PS, W> acc{lds + c};
  const int flags = a.flags;
  const bool mid = (flags & (F_H | F_APERTURE)) != 0;
  const bool spec_out = (flags & F_SPEC_OUT) != 0;
  const bool scale = (flags & F_SCALE) != 0;
  // Step b moves field b's strip (prefetched by step b-1) from nxt into cur, requests
  // field b+1 into nxt, then transforms cur and stores it.  The copy at the head of
  // step b+1 therefore has exactly b's stores behind the loads it waits for, and the
  // compiler's wait count lets those stores stay in flight.  The first step is peeled
  // off the loop so that the loop head is only ever reached in that state (merged
  // with the loop entry the wait would cover the stores as well).
  cx<T> cur[E], nxt[E];
  constexpr bool HF = HOISTH == HOIST_FACTOR, HP = HOISTH == HOIST_PHASE;
  static_assert(!HP || E <= 32, "one mask bit per element");
  cx<T> hm[HF ? E : 1];
  float hfr[HP ? E : 1];
  uint32_t hout = 0;
  if constexpr (HF || HP) {
    if (mid) {
      const ColMid cm = col_mid(pf[0].o, flags, sx);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        bool out;
        const float fr = mid_phase(cm, t + TL * m - (m >= E / 2 ? N : 0), out);
        if constexpr (HF) {
          hm[m] = mk(out ? 0.0f : __builtin_amdgcn_cosf(fr), out ? 0.0f : __builtin_amdgcn_sinf(fr));
        } else {
          hfr[m] = fr;
          hout |= (uint32_t)out << m;
        }
      }
    }
  }
  auto fetch = [&](int bf) {
    const LP2D_GLB cx<T>* __restrict__ src = glb(uniform_ptr((const cx<T>*)pf[bf - b0].src)) + xoff;
#pragma unroll
    for (int m = 0; m < E; ++m) nxt[m] = gld<NT_COL_LD>(src + (lo + (unsigned)(TL * N * m)));
  };
  auto step = [&](int b) {
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = nxt[m];
    if (b + 1 < b1) fetch(b + 1);  // uniform branch
    line_fft_l<T, N, E, false, CL>(cur, t, acc, twl);
    if (HF && mid) {
#pragma unroll
      for (int m = 0; m < E; ++m) cur[m] = cmulc(cur[m], hm[HF ? m : 0]);
    } else if (HP && mid) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const float fr = hfr[HP ? m : 0];
        const cx<float> r = cmulc(cur[m], mk(__builtin_amdgcn_cosf(fr), __builtin_amdgcn_sinf(fr)));
        const bool out = (hout >> m) & 1u;
        cur[m] = mk(out ? 0.0f : r.re, out ? 0.0f : r.im);
      }
    } else if (mid) {
      ColMid cm = col_mid(pf[b - b0].o, flags, sx);
      // opaque per field: otherwise the per-element frequencies (the same for every
      // field of a strip) are hoisted out of the field loop and held for all E
      // elements at once
      int tt = t;
      asm volatile("" : "+v"(cm.c0), "+v"(cm.fx2), "+v"(cm.inv_ndL), "+v"(tt));
      // fftfreq order: index t + TL*m is a negative frequency for m >= E/2
#pragma unroll
      for (int m = 0; m < E; ++m) cur[m] = mid_fast(cm, cur[m], tt + TL * m - (m >= E / 2 ? N : 0));
    }
    if (spec_out) {
      LP2D_GLB cx<T>* __restrict__ so = glb(a.spec_out + (long long)b * a.fstride + x0);
#pragma unroll
      for (int m = 0; m < E; ++m) gst<NT_COL_ST>(so + (lo_n + (unsigned)(TL * N * m)), cur[m]);
    }
    line_fft_l<T, N, E, true, CL>(cur, t, acc, twl);
    if (scale) {
#pragma unroll
      for (int m = 0; m < E; ++m) cur[m] = cscale(cur[m], a.scale);
    }
    LP2D_GLB cx<T>* __restrict__ dst = glb(uniform_ptr((cx<T>*)pf[b - b0].dst)) + xoff;
#pragma unroll
    for (int m = 0; m < E; ++m) gst<NT_COL_ST>(dst + (lo + (unsigned)(TL * N * m)), cur[m]);
  };
  if (b0 >= b1) return;  // (the launchers size the groups so that none is empty)
  fetch(b0);
  step(b0);
  for (int b = b0 + 1; b < b1; ++b) step(b);
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
  const LP2D_GLB cx<T>* fsrc = field_src(a, b);
  LP2D_GLB cx<T>* fdst = field_dst(a, b);
  cx<T> v[E];
  // phase 1: rows (load, spatial prologue, forward row FFT)
  {
    const int y = tid / TL, t = tid % TL;
    const LP2D_GLB cx<T>* __restrict__ src = fsrc + (long long)y * N;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gld<NT_SMALL>(src + (t + TL * m));
    if (a.flags & F_SPATIAL) {
      spatial_pro_all<T, E>(a, v, b, y, t, TL, N);
      if (a.flags & F_FIELD_OUT) {
#pragma unroll
        for (int m = 0; m < E; ++m) a.field_out[fb + (long long)y * N + t + TL * m] = v[m];
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
      const PassArgs<T> ab = field_optics(a, b);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ky = t + TL * m;
        v[m] = spectral_mid(ab, v[m], b, ky, x, N);
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
    LP2D_GLB cx<T>* __restrict__ dst = fdst + (long long)y * N;
    spatial_epi_all<T, E>(a, v, b, y, t, TL, N, fb + (long long)y * N + t, true);
    if ((a.flags & F_NO_STORE) == 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) gst<NT_SMALL>(dst + (t + TL * m), v[m]);
    }
  }
}

// Whole field in LDS, pipelined over a group of fields [b0, b1): the next field is
// loaded into registers while the current one runs its three in-LDS phases (rows ->
// columns with H and pupil -> rows).  Twiddles and the fields' pointers and optics
// come from LDS (twl, pf), so no wait in the loop covers the prefetch before it is
// needed; the first field is peeled off the loop for the same reason as in the
// column pass.  Complex64, evanescent modes passed through, no mask multipliers and
// no spatial prologue/epilogue (flags within F_H, F_APERTURE, F_SCALE, F_SPEC_OUT).
template <typename T, int N, int E, int PS>
__device__ __forceinline__ void small_pipe_body(const PassArgs<T>& a, int b0, int b1,
                                                cx<T>* lds, const cx<T>* twl,
                                                const PipeField* pf, int tid) {
  static_assert(sizeof(T) == 4, "complex64 path");
  constexpr int TL = N / E;
  constexpr int P = SmallPitch<N, PS>::value;
  const int flags = a.flags;
  const bool mid = (flags & (F_H | F_APERTURE)) != 0;
  const bool scale = (flags & F_SCALE) != 0;
  const bool spec_out = (flags & F_SPEC_OUT) != 0;
  // row phases: row y, lane tr (a row's TL lanes are consecutive, within one wave);
  // column phase: column x, lane tc, each wave taking CPW whole columns, so every
  // exchange of a line transform stays inside one wave (no workgroup barrier)
  constexpr int THREADS = N * TL;
  constexpr int CPW = (THREADS < 64 ? THREADS : 64) / TL;
  static_assert(64 % TL == 0 && (THREADS < 64 || THREADS % 64 == 0), "whole lines per wave");
  const int y = tid / TL, tr = tid % TL;
  const int x = (tid / 64) * CPW + tid % CPW, tc = (tid % 64) / CPW;
  const int sx = x < N / 2 ? x : x - N;
  cx<T> cur[E], nxt[E];
  auto fetch = [&](int bf) {
    const LP2D_GLB cx<T>* __restrict__ src =
        glb(uniform_ptr((const cx<T>*)pf[bf - b0].src)) + (long long)y * N;
#pragma unroll
    for (int m = 0; m < E; ++m) nxt[m] = gld<NT_SMALL>(src + (tr + TL * m));
  };
  auto step = [&](int b) {
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = nxt[m];
    if (b + 1 < b1) fetch(b + 1);  // uniform branch
    {  // rows: forward FFT into the tile
      LineAcc<T, PS> acc{lds + y * P};
      line_fft_l<T, N, E, false, 0, true>(cur, tr, acc, twl);
#pragma unroll
      for (int m = 0; m < E; ++m) acc.st(tr + TL * m, cur[m]);
    }
    __syncthreads();
    {  // columns: forward FFT, H and pupil, inverse FFT
      StrideAcc<T, P> acc{lds + pad_idx<PS>(x)};
      cx<T> v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = acc.ld(tc + TL * m);
      wave_sync();  // the column's own lanes read it before its transform rewrites it
      line_fft_l<T, N, E, false, 0, true>(v, tc, acc, twl);
      if (mid) {
        ColMid cm = col_mid(pf[b - b0].o, flags, sx);
        int tt = tc;
        asm volatile("" : "+v"(cm.c0), "+v"(cm.fx2), "+v"(cm.inv_ndL), "+v"(tt));
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = mid_fast(cm, v[m], tt + TL * m - (m >= E / 2 ? N : 0));
      }
      if (spec_out) {
        LP2D_GLB cx<T>* __restrict__ so = glb(a.spec_out + (long long)b * a.fstride + x);
#pragma unroll
        for (int m = 0; m < E; ++m) gst<NT_SMALL>(so + ((long long)(tc + TL * m) * N), v[m]);
      }
      line_fft_l<T, N, E, true, 0, true>(v, tc, acc, twl);
      wave_sync();
#pragma unroll
      for (int m = 0; m < E; ++m) acc.st(tc + TL * m, v[m]);
    }
    __syncthreads();
    {  // rows: inverse FFT, scale, store
      LineAcc<T, PS> acc{lds + y * P};
      cx<T> v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = acc.ld(tr + TL * m);
      wave_sync();  // the row's own lanes read it before its transform rewrites it
      line_fft_l<T, N, E, true, 0, true>(v, tr, acc, twl);
      if (scale) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
      }
      LP2D_GLB cx<T>* __restrict__ dst =
          glb(uniform_ptr((cx<T>*)pf[b - b0].dst)) + (long long)y * N;
#pragma unroll
      for (int m = 0; m < E; ++m) gst<NT_SMALL>(dst + (tr + TL * m), v[m]);
    }
    __syncthreads();  // the next field's row phase rewrites the tile
  };
  if (b0 >= b1) return;  // (the launchers size the groups so that none is empty)
  fetch(b0);
  step(b0);
  for (int b = b0 + 1; b < b1; ++b) step(b);
}

}  // namespace lp2d
