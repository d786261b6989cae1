// Grid resampling of Beam2D.expand / Beam2D.coarse (lightprop2d/beam2d.py:375-404,
// 437-460) behind lp2d_expand / lp2d_coarse, and the bilinear interpolation of
// Beam2D.crop / Beam2D.lens_image (beam2d.py:406-435, 513-536) behind
// lp2d_interp_linear.  Each is one pass over the fields, bound by HBM: expand reads n^2
// and writes n_out^2 elements, coarse reads n^2 and writes (n/m)^2, the interpolation
// gathers at most the window and writes n_out^2.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "../../include/lp2d.h"

extern "C" int lp2d_init(int device);
namespace lp2d {
int set_error(int code, const char* msg);
}

namespace {

// Grid: x = output rows of all fields (batch * rows), y = 256-column chunks.  No
// 64-bit division per element: each workgroup row knows its field and row.
template <typename C>
__global__ void k_expand(const C* __restrict__ in, C* __restrict__ out, int n, int n_out,
                         int d) {
  // edge replication: the d columns left/right repeat the first/last column, then
  // the d rows above/below repeat the first/last (already widened) row
  // (beam2d.py:393-401) -- element (j, i) of the result is in[clamp(j-d), clamp(i-d)]
  const int row = blockIdx.x;
  const int b = row / n_out, j = row - b * n_out;
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  int js = j - d, is = i - d;
  js = js < 0 ? 0 : (js >= n ? n - 1 : js);
  is = is < 0 ? 0 : (is >= n ? n - 1 : is);
  out[(long long)row * n_out + i] = in[((long long)b * n + js) * n + is];
}

template <typename T, typename C>
__global__ void k_coarse(const C* __restrict__ in, C* __restrict__ out, int n, int m) {
  // block mean over m x m (beam2d.py:456-459: reshape (N, m, N, m).mean((1, 3))),
  // accumulated in the field's precision, row of the block by row
  const int no = n / m;
  const int row = blockIdx.x;
  const int b = row / no, J = row - b * no;
  const int I = blockIdx.y * blockDim.x + threadIdx.x;
  if (I >= no) return;
  const C* src = in + ((long long)b * n + (long long)J * m) * n + (long long)I * m;
  T sr = 0, si = 0;
  for (int u = 0; u < m; ++u)
    for (int v = 0; v < m; ++v) {
      const C c = src[(long long)u * n + v];
      sr += c.x;
      si += c.y;
    }
  const T cnt = (T)(m * m);
  C o;
  o.x = sr / cnt;  // numpy's mean: sum / count
  o.y = si / cnt;
  out[(long long)row * no + I] = o;
}

// complex64 fast paths: whole 16-byte vectors per lane.  8 fields 1024^2 <-> 2048^2:
// expand 86.7 -> 83.5 us, coarse 70 -> 49 us (with nontemporal output stores, which
// help the coarse mean and not the expansion).
typedef float v4f __attribute__((ext_vector_type(4)));

// Expand with four output elements (i0 .. i0+3, n_out a multiple of 4) per thread,
// written as two 16-byte stores.
__global__ void k_expand4(const float2* __restrict__ in, float2* __restrict__ out, int n,
                          int n_out, int d) {
  const int row = blockIdx.x;
  const int b = row / n_out, j = row - b * n_out;
  const int i0 = 4 * (blockIdx.y * blockDim.x + threadIdx.x);
  if (i0 >= n_out) return;
  int js = j - d;
  js = js < 0 ? 0 : (js >= n ? n - 1 : js);
  const float2* src = in + ((long long)b * n + js) * n;
  float2 v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int is = i0 + q - d;
    is = is < 0 ? 0 : (is >= n ? n - 1 : is);
    v[q] = src[is];
  }
  v4f* dst = (v4f*)(out + (long long)row * n_out + i0);
  v4f a, c;
  a.x = v[0].x; a.y = v[0].y; a.z = v[1].x; a.w = v[1].y;
  c.x = v[2].x; c.y = v[2].y; c.z = v[3].x; c.w = v[3].y;
  // plain stores: with the nontemporal hint this kernel measured 119.5 us against
  // 83.5 us (8 x 1024^2 -> 2048^2)
  dst[0] = a;
  dst[1] = c;
}

// 2 x 2 block mean with two outputs (I0, I0 + 1; no a multiple of 2) per thread: each
// block row is one 16-byte load; sums in the order of k_coarse (row by row).
__global__ void k_coarse2(const float2* __restrict__ in, float2* __restrict__ out, int n) {
  const int no = n / 2;
  const int row = blockIdx.x;
  const int b = row / no, J = row - b * no;
  const int I0 = 2 * (blockIdx.y * blockDim.x + threadIdx.x);
  if (I0 >= no) return;
  const v4f* r0 = (const v4f*)(in + ((long long)b * n + 2LL * J) * n + 2LL * I0);
  const v4f* r1 = (const v4f*)((const float2*)r0 + n);
  const v4f a0 = r0[0], a1 = r0[1], c0 = r1[0], c1 = r1[1];
  v4f o;
  o.x = (((a0.x + a0.z) + c0.x) + c0.z) / 4.0f;
  o.y = (((a0.y + a0.w) + c0.y) + c0.w) / 4.0f;
  o.z = (((a1.x + a1.z) + c1.x) + c1.z) / 4.0f;
  o.w = (((a1.y + a1.w) + c1.y) + c1.w) / 4.0f;
  __builtin_nontemporal_store(o, (v4f*)(out + (long long)row * no + I0));
}

// ---------------------------------------------------------------- interpolation
// interp2d(kind='linear') on a rectangular grid is FITPACK's degree-1 spline with a
// knot at every sample, evaluated by fpbisp: the argument is clamped to
// [t_0, t_{len-1}], the interval is the last l <= len-2 with t_l <= q, and fpbspl's
// basis values are f*(t_{l+1} - q) and f*(q - t_l) with f = 1/(t_{l+1} - t_l).  The
// knots t_k and queries q_j are formed as integer index times spacing, the way the
// reference builds X (beam2d.py:300-301), so they carry the reference's rounding.
struct AxisTap {
  int l;          // window sample of the left tap
  double h0, h1;  // weights of samples l and l + 1
};

__device__ inline double knot(const lp2d_interp_desc& d, int k) {
  return (double)(k + d.src_i0) * d.src_dx;
}

__device__ inline AxisTap axis_tap(const lp2d_interp_desc& d, int j) {
  double q = ((double)(d.dst_sign * j + d.dst_i0) * d.dst_dx) * d.dst_scale;
  const double tb = knot(d, 0), te = knot(d, d.win_len - 1);
  if (q < tb) q = tb;
  if (q > te) q = te;
  const int last = d.win_len - 2;
  int l = (int)floor((q - tb) / d.src_dx);
  l = l < 0 ? 0 : (l > last ? last : l);
  // the estimate is off by at most one sample where q sits on a knot; settle it on the
  // knots themselves (both loops are bounded by the window)
  while (l < last && knot(d, l + 1) <= q) ++l;
  while (l > 0 && knot(d, l) > q) --l;
  const double tl = knot(d, l), tr = knot(d, l + 1);
  const double f = 1.0 / (tr - tl);
  return AxisTap{l, f * (tr - q), f * (q - tl)};
}

// Chirp factor exp(-i * c * x^2) of one grid coordinate; the chirp of a pixel is the
// product of its column's and its row's factors.
__device__ inline double2 chirp_factor(const lp2d_interp_desc& d, int j) {
  const double x = (double)(j + d.grid_i0) * d.grid_dl;
  double sn, cs;
  sincos(-(x * x) * d.phase_coef, &sn, &cs);
  return make_double2(cs, sn);
}

// One thread per output column, kInterpRows output rows each; grid (column chunks of
// 256, row chunks, fields).  The row taps and row chirp factors are computed once per
// workgroup into LDS.  The value is accumulated in FP64 in fpbisp's order (x tap
// outer, y tap inner), then scaled and chirped in FP64 and rounded once to the field's
// precision.
constexpr int kInterpRows = 16;
constexpr int kInterpThreads = 256;

template <typename C>
__global__ __launch_bounds__(kInterpThreads) void k_interp_linear(
    const C* __restrict__ in, C* __restrict__ out, int n, int n_out, lp2d_interp_desc d) {
  __shared__ AxisTap ty_s[kInterpRows];
  __shared__ double2 cy_s[kInterpRows];
  const bool chirp = d.phase_coef != 0.0;
  const int i0 = blockIdx.y * kInterpRows;
  const int rows = n_out - i0 < kInterpRows ? n_out - i0 : kInterpRows;
  if ((int)threadIdx.x < rows) {
    ty_s[threadIdx.x] = axis_tap(d, i0 + (int)threadIdx.x);
    if (chirp) cy_s[threadIdx.x] = chirp_factor(d, i0 + (int)threadIdx.x);
  }
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out) return;
  const int b = blockIdx.z;
  const AxisTap tx = axis_tap(d, j);
  const double2 cx = chirp ? chirp_factor(d, j) : make_double2(1.0, 0.0);
  const C* win = in + ((long long)b * n + d.win_off) * n + d.win_off + tx.l;
  C* dst = out + ((long long)b * n_out + i0) * n_out + j;
  // Source samples of the current row pair, c<x tap><y tap>.  Consecutive output rows
  // mostly fall in the same or the next source interval (zoom, magnification < 2), so
  // the samples are reused: the row taps are the same in every lane, so the branches
  // are uniform.
  C c00, c01, c10, c11;
  int lprev = -2;
  for (int r = 0; r < rows; ++r) {
    const AxisTap ty = ty_s[r];
    if (ty.l != lprev) {
      const C* r1 = win + (long long)(ty.l + 1) * n;
      if (ty.l == lprev + 1) {
        c00 = c01;
        c10 = c11;
      } else {
        const C* r0 = r1 - n;
        c00 = r0[0];
        c10 = r0[1];
      }
      c01 = r1[0];
      c11 = r1[1];
      lprev = ty.l;
    }
    double vr = ((double)c00.x * tx.h0) * ty.h0;
    double vi = ((double)c00.y * tx.h0) * ty.h0;
    vr += ((double)c01.x * tx.h0) * ty.h1;
    vi += ((double)c01.y * tx.h0) * ty.h1;
    vr += ((double)c10.x * tx.h1) * ty.h0;
    vi += ((double)c10.y * tx.h1) * ty.h0;
    vr += ((double)c11.x * tx.h1) * ty.h1;
    vi += ((double)c11.y * tx.h1) * ty.h1;
    vr *= d.scale;
    vi *= d.scale;
    if (chirp) {
      const double2 cy = cy_s[r];
      const double er = cx.x * cy.x - cx.y * cy.y, ei = cx.x * cy.y + cx.y * cy.x;
      const double t = vr * er - vi * ei;
      vi = vr * ei + vi * er;
      vr = t;
    }
    C o;
    o.x = (decltype(o.x))vr;
    o.y = (decltype(o.y))vi;
    dst[(long long)r * n_out] = o;
  }
}

int prep() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return lp2d::set_error(LP2D_E_HIP, "hipGetDevice failed");
  return lp2d_init(dev);
}


}  // namespace

extern "C" int lp2d_expand(int n, int n_out, int batch, int dtype, const void* in, void* out,
                           void* stream) {
  using lp2d::set_error;
  if (n < 1 || n_out < n || batch < 1 || ((n_out - n) & 1))
    return set_error(LP2D_E_SIZE, "expand needs 1 <= n <= n_out with n_out - n even");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return set_error(LP2D_E_DTYPE, "unknown dtype");
  if (!in || !out) return set_error(LP2D_E_ARG, "null pointer");
  if (int rc = prep()) return rc;
  if ((long long)batch * n_out > 0x7fffffffLL) return set_error(LP2D_E_SIZE, "batch too large");
  const int d = (n_out - n) / 2;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)(batch * n_out), (unsigned)((n_out + 255) / 256));
  if (dtype == LP2D_C64 && n_out % 4 == 0 && ((uintptr_t)out & 15) == 0) {
    const dim3 g4((unsigned)(batch * n_out), (unsigned)((n_out / 4 + 255) / 256));
    k_expand4<<<g4, 256, 0, s>>>((const float2*)in, (float2*)out, n, n_out, d);
  } else if (dtype == LP2D_C64)
    k_expand<float2><<<grid, 256, 0, s>>>((const float2*)in, (float2*)out, n, n_out, d);
  else
    k_expand<double2><<<grid, 256, 0, s>>>((const double2*)in, (double2*)out, n, n_out, d);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : set_error(LP2D_E_HIP, hipGetErrorString(e));
}

extern "C" int lp2d_coarse(int n, int mean_order, int batch, int dtype, const void* in,
                           void* out, void* stream) {
  using lp2d::set_error;
  if (n < 1 || mean_order < 1 || batch < 1 || n % mean_order)
    return set_error(LP2D_E_SIZE, "coarse needs mean_order >= 1 dividing n");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return set_error(LP2D_E_DTYPE, "unknown dtype");
  if (!in || !out) return set_error(LP2D_E_ARG, "null pointer");
  if (int rc = prep()) return rc;
  const int no = n / mean_order;
  if ((long long)batch * no > 0x7fffffffLL) return set_error(LP2D_E_SIZE, "batch too large");
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)(batch * no), (unsigned)((no + 255) / 256));
  if (dtype == LP2D_C64 && mean_order == 2 && no % 2 == 0 && ((uintptr_t)in & 15) == 0 &&
      ((uintptr_t)out & 15) == 0) {
    const dim3 g2((unsigned)(batch * no), (unsigned)((no / 2 + 255) / 256));
    k_coarse2<<<g2, 256, 0, s>>>((const float2*)in, (float2*)out, n);
  } else if (dtype == LP2D_C64)
    k_coarse<float, float2><<<grid, 256, 0, s>>>((const float2*)in, (float2*)out, n, mean_order);
  else
    k_coarse<double, double2><<<grid, 256, 0, s>>>((const double2*)in, (double2*)out, n,
                                                   mean_order);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : set_error(LP2D_E_HIP, hipGetErrorString(e));
}

extern "C" int lp2d_interp_linear(int n, int n_out, int batch, int dtype, const void* in,
                                  void* out, const lp2d_interp_desc* desc, void* stream) {
  using lp2d::set_error;
  if (!desc) return set_error(LP2D_E_ARG, "null interpolation descriptor");
  const lp2d_interp_desc d = *desc;
  if (n < 2 || n_out < 1 || n_out > 65536 || batch < 1 || batch > 65535)
    return set_error(LP2D_E_SIZE, "interp_linear needs n >= 2, 1 <= n_out <= 65536, "
                                  "1 <= batch <= 65535");
  if (d.win_len < 2 || d.win_off < 0 || d.win_off > n - d.win_len)
    return set_error(LP2D_E_SIZE, "interpolation window must hold >= 2 samples inside the field");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return set_error(LP2D_E_DTYPE, "unknown dtype");
  if (!in || !out) return set_error(LP2D_E_ARG, "null pointer");
  if (d.dst_sign != 1 && d.dst_sign != -1) return set_error(LP2D_E_ARG, "dst_sign must be +1 or -1");
  const double params[] = {d.src_dx, d.dst_dx, d.dst_scale, d.scale, d.phase_coef, d.grid_dl};
  for (double v : params)
    if (!(v - v == 0.0)) return set_error(LP2D_E_ARG, "non-finite interpolation parameter");
  if (!(d.src_dx > 0.0)) return set_error(LP2D_E_ARG, "src_dx must be positive");
  // index products stay inside int32: |dst_sign * j + dst_i0|, |k + src_i0|, |j + grid_i0|
  const long long lim = 0x3fffffffLL;
  if (llabs((long long)d.dst_i0) + n_out > lim || llabs((long long)d.src_i0) + d.win_len > lim ||
      llabs((long long)d.grid_i0) + n_out > lim)
    return set_error(LP2D_E_ARG, "index offset out of range");
  const size_t es = dtype == LP2D_C64 ? 8 : 16;
  const char* ib = (const char*)in;
  const char* ob = (const char*)out;
  if (ib < ob + es * (size_t)batch * n_out * n_out && ob < ib + es * (size_t)batch * n * n)
    return set_error(LP2D_E_ARG, "in and out must not overlap");
  if (int rc = prep()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)((n_out + kInterpThreads - 1) / kInterpThreads),
                  (unsigned)((n_out + kInterpRows - 1) / kInterpRows), (unsigned)batch);
  if (dtype == LP2D_C64)
    k_interp_linear<float2><<<grid, kInterpThreads, 0, s>>>((const float2*)in, (float2*)out, n,
                                                            n_out, d);
  else
    k_interp_linear<double2><<<grid, kInterpThreads, 0, s>>>((const double2*)in, (double2*)out,
                                                             n, n_out, d);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : set_error(LP2D_E_HIP, hipGetErrorString(e));
}
