// Mode-basis projection and reconstruction (lightprop2d/beam2d.py:538-667) behind
// lp2d_modes_project / lp2d_modes_construct.  The mode matrix is (nmodes, nb*nb),
// row-major, the field's precision; it meets the central nb x nb window of an n x n
// field (offset dN = (n - nb)/2).  Both kernels read the mode matrix once per group of
// fields and are bound by HBM: a matrix-vector product per field, not a contraction
// wide enough for the matrix cores.
#include <hip/hip_runtime.h>

#include "../../include/lp2d.h"

extern "C" int lp2d_init(int device);
namespace lp2d {
int set_error(int code, const char* msg);
}

namespace {

constexpr int kThreads = 256;
constexpr int kFieldsPerPass = 8;  // fields sharing one read of a mode element

template <typename T>
struct C2 {
  T x, y;
};

// coeffs[b, k] (complex128, pre-zeroed) += sum over a segment of p of
// modes[k, p] * field[b, window(p)]  -- the reference's unconjugated product
// modes_matrix @ flatten_field (beam2d.py:597), accumulated in FP64.
// One workgroup per (pixel segment, kModesPerPass modes, kFieldsPerPass fields): each
// mode element read serves the group's fields and each field element the group's
// modes.  Workgroups go to the 8 XCDs round robin; the linear id is decoded so that
// each XCD takes its own segments and runs all mode groups of a segment back to back,
// the segment of the fields staying in that XCD's L2.
constexpr int kModesPerPass = 4;
template <typename T, int KF>
__global__ __launch_bounds__(kThreads) void k_project(const C2<T>* __restrict__ modes,
                                                      const C2<T>* __restrict__ field,
                                                      double2* __restrict__ coeffs, int n,
                                                      int nb, int nmodes, int batch, int seg,
                                                      int nseg) {
  constexpr int KM = kModesPerPass;
  __shared__ double red[kThreads / 64][2 * KM * KF];
  const int mgroups = (nmodes + KM - 1) / KM;
  const long long id = blockIdx.x;
  const int xcd = (int)(id % 8);
  const long long j = id / 8;
  const int mg = (int)(j % mgroups);
  const long long sg = j / mgroups;      // (segment / 8, field group) on this XCD
  const int per_xcd = (nseg + 7) / 8;    // segment slots per XCD
  const int grp = (int)(sg / per_xcd);
  const int s8 = (int)(sg - (long long)grp * per_xcd);
  const int sgi = s8 * 8 + xcd;          // segment index
  if (sgi >= nseg) return;
  const int k0 = mg * KM;
  const int nk = nmodes - k0 < KM ? nmodes - k0 : KM;
  const int b0 = grp * KF;
  const int nf = batch - b0 < KF ? batch - b0 : KF;
  const long long npix = (long long)nb * nb;
  const long long p0 = (long long)sgi * seg;
  const long long p1 = p0 + seg < npix ? p0 + seg : npix;
  const int dN = (n - nb) / 2;
  double ar[KM][KF] = {}, ai[KM][KF] = {};
  for (long long p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const int jj = (int)(p / nb), ii = (int)(p - (long long)jj * nb);
    const long long off = (long long)(dN + jj) * n + (dN + ii);
    double fx[KF], fy[KF];
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      C2<T> u;
      u.x = 0;
      u.y = 0;
      if (f < nf) u = field[(long long)(b0 + f) * n * n + off];
      fx[f] = (double)u.x;
      fy[f] = (double)u.y;
    }
#pragma unroll
    for (int q = 0; q < KM; ++q) {
      C2<T> m;
      m.x = 0;
      m.y = 0;
      if (q < nk) m = modes[(long long)(k0 + q) * npix + p];
      const double mx = (double)m.x, my = (double)m.y;
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        ar[q][f] += mx * fx[f] - my * fy[f];
        ai[q][f] += mx * fy[f] + my * fx[f];
      }
    }
  }
  // wave reduction, then across the block's waves
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int q = 0; q < KM; ++q) {
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      double r = ar[q][f], i2 = ai[q][f];
      for (int o = 32; o > 0; o >>= 1) {
        r += __shfl_down(r, o, 64);
        i2 += __shfl_down(i2, o, 64);
      }
      if (lane == 0) {
        red[w][2 * (q * KF + f)] = r;
        red[w][2 * (q * KF + f) + 1] = i2;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * KM * KF) {
    const int qf = threadIdx.x / 2, q = qf / KF, f = qf % KF;
    double sum = 0;
    for (int i3 = 0; i3 < kThreads / 64; ++i3) sum += red[i3][threadIdx.x];
    if (q < nk && f < nf) {
      double* dst = reinterpret_cast<double*>(&coeffs[(long long)(b0 + f) * nmodes + k0 + q]);
      atomicAdd(dst + (threadIdx.x & 1), sum);
    }
  }
}

__global__ void k_zero(double2* p, long long count) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

// field[b, j, i] = sum_k coeffs[b, k] * modes[k, window^-1(j, i)] inside the window,
// 0 outside (beam2d.py:664-666 with the zero-padded basis of :555-560).  Accumulated
// in the field's precision, modes in order k = 0, 1, ... as numpy's einsum; each
// thread serves one pixel of up to kFieldsPerPass fields, reading each mode element
// once for all of them.  Grid: x = rows x groups of fields, y = 256-column chunks.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_construct(const C2<T>* __restrict__ modes,
                                                        const double2* __restrict__ coeffs,
                                                        C2<T>* __restrict__ field, int n, int nb,
                                                        int nmodes, int batch) {
  const int row = blockIdx.x;
  const int g = row / n, j = row - g * n;
  const int b0 = g * kFieldsPerPass;
  const int nf = batch - b0 < kFieldsPerPass ? batch - b0 : kFieldsPerPass;
  const int i = blockIdx.y * kThreads + threadIdx.x;
  if (i >= n) return;
  const int dN = (n - nb) / 2;
  const int jj = j - dN, ii = i - dN;
  T sr[kFieldsPerPass] = {}, si[kFieldsPerPass] = {};
  if (jj >= 0 && jj < nb && ii >= 0 && ii < nb) {
    const long long p = (long long)jj * nb + ii;
    const long long npix = (long long)nb * nb;
    for (int k = 0; k < nmodes; ++k) {
      const C2<T> m = modes[(long long)k * npix + p];
#pragma unroll
      for (int f = 0; f < kFieldsPerPass; ++f) {
        if (f < nf) {
          const double2 c = coeffs[(long long)(b0 + f) * nmodes + k];
          const T cr = (T)c.x, ci = (T)c.y;
          sr[f] += m.x * cr - m.y * ci;
          si[f] += m.x * ci + m.y * cr;
        }
      }
    }
  }
#pragma unroll
  for (int f = 0; f < kFieldsPerPass; ++f) {
    if (f < nf) {
      C2<T> o;
      o.x = sr[f];
      o.y = si[f];
      field[((long long)(b0 + f) * n + j) * n + i] = o;
    }
  }
}

int check(int n, int nb, int nmodes, int batch, int dtype, const void* a, const void* b,
          const void* c) {
  using lp2d::set_error;
  if (n < 1 || nb < 1 || nb > n || ((n - nb) & 1) || nmodes < 1 || batch < 1)
    return set_error(LP2D_E_SIZE, "need 1 <= nb <= n, n - nb even, nmodes >= 1, batch >= 1");
  if ((long long)batch * n > 0x7fffffffLL || nmodes > 65535)
    return set_error(LP2D_E_SIZE, "batch or mode count too large");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return set_error(LP2D_E_DTYPE, "unknown dtype");
  if (!a || !b || !c) return set_error(LP2D_E_ARG, "null pointer");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(LP2D_E_HIP, "hipGetDevice failed");
  return lp2d_init(dev);
}

}  // namespace

extern "C" int lp2d_modes_project(int n, int nb, int nmodes, int batch, int dtype,
                                  const void* modes, const void* field, void* coeffs,
                                  void* stream) {
  if (int rc = check(n, nb, nmodes, batch, dtype, modes, field, coeffs)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const long long cnt = (long long)batch * nmodes;
  k_zero<<<(unsigned)((cnt + 255) / 256 < 4096 ? (cnt + 255) / 256 : 4096), 256, 0, s>>>(
      (double2*)coeffs, cnt);
  const long long npix = (long long)nb * nb;
  // fields per workgroup: 8 (each mode element serves 8 fields), or 1 for one field
  const int kf = batch > 1 ? kFieldsPerPass : 1;
  const int groups = (batch + kf - 1) / kf;
  const int mgroups = (nmodes + kModesPerPass - 1) / kModesPerPass;
  // about 1024 workgroups; at least 16 pixels per thread with 8 fields (amortising
  // the reduction of 64 sums), 4 with one
  const int min_px = kf > 1 ? 16 : 4;
  long long nseg = 1024 / ((long long)mgroups * groups);
  const long long maxseg = (npix + (long long)min_px * kThreads - 1) / ((long long)min_px * kThreads);
  if (nseg > maxseg) nseg = maxseg;
  if (nseg < 1) nseg = 1;
  const int seg = (int)((npix + nseg - 1) / nseg);
  nseg = (npix + seg - 1) / seg;
  const long long per_xcd = (nseg + 7) / 8;
  const long long nblocks = 8 * per_xcd * mgroups * groups;
  if (nblocks > 0x7fffffffLL) return lp2d::set_error(LP2D_E_SIZE, "too many modes x fields");
  const unsigned nb_ = (unsigned)nblocks;
  if (dtype == LP2D_C64) {
    if (kf > 1)
      k_project<float, kFieldsPerPass><<<nb_, kThreads, 0, s>>>(
          (const C2<float>*)modes, (const C2<float>*)field, (double2*)coeffs, n, nb, nmodes,
          batch, seg, (int)nseg);
    else
      k_project<float, 1><<<nb_, kThreads, 0, s>>>((const C2<float>*)modes,
                                                   (const C2<float>*)field, (double2*)coeffs, n,
                                                   nb, nmodes, batch, seg, (int)nseg);
  } else {
    if (kf > 1)
      k_project<double, kFieldsPerPass><<<nb_, kThreads, 0, s>>>(
          (const C2<double>*)modes, (const C2<double>*)field, (double2*)coeffs, n, nb, nmodes,
          batch, seg, (int)nseg);
    else
      k_project<double, 1><<<nb_, kThreads, 0, s>>>((const C2<double>*)modes,
                                                    (const C2<double>*)field, (double2*)coeffs, n,
                                                    nb, nmodes, batch, seg, (int)nseg);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : lp2d::set_error(LP2D_E_HIP, hipGetErrorString(e));
}

extern "C" int lp2d_modes_construct(int n, int nb, int nmodes, int batch, int dtype,
                                    const void* modes, const void* coeffs, void* field,
                                    void* stream) {
  if (int rc = check(n, nb, nmodes, batch, dtype, modes, coeffs, field)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int groups = (batch + kFieldsPerPass - 1) / kFieldsPerPass;
  const dim3 grid((unsigned)(groups * n), (unsigned)((n + kThreads - 1) / kThreads));
  if (dtype == LP2D_C64)
    k_construct<float><<<grid, kThreads, 0, s>>>((const C2<float>*)modes, (const double2*)coeffs,
                                                  (C2<float>*)field, n, nb, nmodes, batch);
  else
    k_construct<double><<<grid, kThreads, 0, s>>>((const C2<double>*)modes,
                                                   (const double2*)coeffs, (C2<double>*)field, n,
                                                   nb, nmodes, batch);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : lp2d::set_error(LP2D_E_HIP, hipGetErrorString(e));
}
