// Mode-basis projection and reconstruction (lightprop2d/beam2d.py:538-667) behind
// lp2d_modes_project / lp2d_modes_construct.  The mode matrix is (nmodes, nb*nb),
// row-major, the field's precision; it meets the central nb x nb window of an n x n
// field (offset dN = (n - nb)/2).  The vector kernels read the mode matrix once per
// group of fields (a matrix-vector product per field); the projection of a batch also
// runs as a complex GEMM on the FP64 matrix cores (k_project_mfma).
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
// in the field's precision, modes in order k = 0, 1, ... as numpy's einsum.  A
// workgroup serves PX pixels of one row for FG x KF fields: wave-sized thread groups
// take KF fields each for the same pixels, so the waves of a workgroup read the same
// mode elements (one HBM read, L1 hits after it) and a batch of FG x KF fields reads
// the mode matrix once.  The coefficients of a block of kConstructModes modes are
// staged in LDS in the field's precision (broadcast reads in the inner loop).  Grid:
// x = rows x groups of FG x KF fields, y = PX-column chunks; a row outside the window
// only writes zeros.
constexpr int kConstructModes = 64;
template <typename T, int KF, int FG, int PX>
__global__ __launch_bounds__(PX * FG) void k_construct(const C2<T>* __restrict__ modes,
                                                       const double2* __restrict__ coeffs,
                                                       C2<T>* __restrict__ field, int n, int nb,
                                                       int nmodes, int batch) {
  constexpr int NF = KF * FG;  // fields per workgroup
  constexpr int THREADS = PX * FG;
  __shared__ C2<T> cs[kConstructModes][NF];
  const int row = blockIdx.x;
  const int g = row / n, j = row - g * n;
  const int b0 = g * NF;
  const int nf = batch - b0 < NF ? batch - b0 : NF;
  const int fg = threadIdx.x / PX;           // this thread's group of KF fields
  const int i = blockIdx.y * PX + (threadIdx.x - fg * PX);
  const int dN = (n - nb) / 2;
  const int jj = j - dN, ii = i - dN;
  const bool in_row = jj >= 0 && jj < nb;  // the same for the whole workgroup
  const bool in_px = in_row && i < n && ii >= 0 && ii < nb;
  T sr[KF] = {}, si[KF] = {};
  if (in_row) {
    const long long npix = (long long)nb * nb;
    const long long p = (long long)jj * nb + (in_px ? ii : 0);
    for (int k0 = 0; k0 < nmodes; k0 += kConstructModes) {
      const int nk = nmodes - k0 < kConstructModes ? nmodes - k0 : kConstructModes;
      __syncthreads();  // the previous block's coefficients are no longer read
      for (int e = threadIdx.x; e < kConstructModes * NF; e += THREADS) {
        const int kk = e / NF, f = e - (e / NF) * NF;
        C2<T> c;
        c.x = 0;
        c.y = 0;
        if (kk < nk && f < nf) {
          const double2 d = coeffs[(long long)(b0 + f) * nmodes + k0 + kk];
          c.x = (T)d.x;
          c.y = (T)d.y;
        }
        cs[kk][f] = c;
      }
      __syncthreads();
      if (in_px) {
        for (int kk = 0; kk < nk; ++kk) {
          const C2<T> m = modes[(long long)(k0 + kk) * npix + p];
#pragma unroll
          for (int f = 0; f < KF; ++f) {
            const C2<T> c = cs[kk][fg * KF + f];
            sr[f] += m.x * c.x - m.y * c.y;
            si[f] += m.x * c.y + m.y * c.x;
          }
        }
      }
    }
  }
  if (i < n) {
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      const int fl = fg * KF + f;
      if (fl < nf) {
        C2<T> o;
        o.x = sr[f];
        o.y = si[f];
        field[((long long)(b0 + fl) * n + j) * n + i] = o;
      }
    }
  }
}

// ----------------------------------------------------------------- matrix-core projection
// The projection of a batch is a complex GEMM, C (fields x modes) = F (fields x pixels)
// x M^T, reduced over nb^2 pixels.  With FP64 MFMA (v_mfma_f64_16x16x4f64) every product
// and sum stays in FP64, as the vector kernel's, in a different order.  Real form of the
// unconjugated product: Cr = Mr.Fr - Mi.Fi, Ci = Mr.Fi + Mi.Fr, four real products per
// tile.  Operand layout of the instruction on gfx950 (checked on the device): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; element v of its result is
// D[4v + l / 16][l % 16].
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kMfmaThreads = 512;  // 8 waves, one 16 x 16 (modes x fields) tile each
constexpr int kTileModes = 64, kTileFields = 32;  // per workgroup
constexpr int kChunk = 32;         // pixels staged in LDS per step (8 MFMA k-steps)
// LDS row pitch in doubles: rows of one fragment read (16 rows, lanes r) then start 4
// banks apart, so a half-wave's 32 eight-byte reads fall into distinct bank pairs
constexpr int kPitch = kChunk + 2;
// Workgroups per launch: one per CU.  Measured at 64 modes x 256^2 x 32 fields (one
// tile group): 256 groups 34 us per call, 512 38 us (twice the partial sums to add),
// 128 51 us (half the CUs idle), 1024 50 us; 64-pixel chunks gave the same as 32.
constexpr int kMfmaTargetGroups = 256;

// Workgroup grid of the matrix-core projection: x = pixel groups of px pixels (a
// multiple of kChunk), y = (mode groups of 64) x (field groups of 32).
struct MfmaPlan {
  int gx = 0, mgroups = 0, fgroups = 0, px = 0;
  long long work_bytes = 0;
};
MfmaPlan mfma_plan(int nb, int nmodes, int batch) {
  MfmaPlan p;
  const long long npix = (long long)nb * nb;
  p.mgroups = (nmodes + kTileModes - 1) / kTileModes;
  p.fgroups = (batch + kTileFields - 1) / kTileFields;
  const long long tg = (long long)p.mgroups * p.fgroups;
  if (tg > 65535) return MfmaPlan{};
  const long long chunks = (npix + kChunk - 1) / kChunk;
  long long gx = kMfmaTargetGroups / tg;
  if (gx < 1) gx = 1;
  if (gx > chunks) gx = chunks;
  const long long cpg = (chunks + gx - 1) / gx;  // chunks per pixel group
  p.px = (int)(cpg * kChunk);
  p.gx = (int)((npix + p.px - 1) / p.px);
  p.work_bytes = (long long)p.gx * batch * nmodes * (long long)sizeof(double2);
  return p;
}

// partial[g, b, k] = sum over pixel group g of modes[k, p] * field[b, window(p)] for the
// workgroup's 64 modes x 32 fields.  The next chunk is loaded into registers while the
// current one runs through the matrix cores.
template <typename T>
__global__ __launch_bounds__(kMfmaThreads) void k_project_mfma(
    const C2<T>* __restrict__ modes, const C2<T>* __restrict__ field,
    double2* __restrict__ partial, int n, int nb, int nmodes, int batch, int mgroups, int px) {
  __shared__ double ar[kTileModes][kPitch], ai[kTileModes][kPitch];
  __shared__ double br[kTileFields][kPitch], bi[kTileFields][kPitch];
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int mg = (int)(blockIdx.y % mgroups), fg = (int)(blockIdx.y / mgroups);
  const int k0 = mg * kTileModes, b0 = fg * kTileFields;
  const long long npix = (long long)nb * nb;
  const long long p_begin = (long long)g * px;
  const long long p_end = p_begin + px < npix ? p_begin + px : npix;
  const int dN = (n - nb) / 2;
  constexpr int ML = kTileModes * kChunk / kMfmaThreads;   // mode elements per thread: 4
  constexpr int FL = kTileFields * kChunk / kMfmaThreads;  // field elements per thread: 2
  C2<T> pm[ML], pf[FL];
  auto fetch = [&](long long p0) {
#pragma unroll
    for (int i = 0; i < ML; ++i) {
      const int e = tid + kMfmaThreads * i;
      const int k = k0 + e / kChunk;
      const long long p = p0 + e % kChunk;
      C2<T> v;
      v.x = 0;
      v.y = 0;
      if (k < nmodes && p < p_end) v = modes[(long long)k * npix + p];
      pm[i] = v;
    }
#pragma unroll
    for (int i = 0; i < FL; ++i) {
      const int e = tid + kMfmaThreads * i;
      const int b = b0 + e / kChunk;
      const long long p = p0 + e % kChunk;
      C2<T> v;
      v.x = 0;
      v.y = 0;
      if (b < batch && p < p_end) {
        const int jj = (int)(p / nb), ii = (int)(p - (long long)jj * nb);
        v = field[(long long)b * n * n + (long long)(dN + jj) * n + (dN + ii)];
      }
      pf[i] = v;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < ML; ++i) {
      const int e = tid + kMfmaThreads * i;
      ar[e / kChunk][e % kChunk] = (double)pm[i].x;
      ai[e / kChunk][e % kChunk] = (double)pm[i].y;
    }
#pragma unroll
    for (int i = 0; i < FL; ++i) {
      const int e = tid + kMfmaThreads * i;
      br[e / kChunk][e % kChunk] = (double)pf[i].x;
      bi[e / kChunk][e % kChunk] = (double)pf[i].y;
    }
  };
  const int w = tid / 64, l = tid % 64;
  const int mt = w & 3, ft = w >> 2;  // this wave's mode tile and field tile
  const int r = l & 15, q = l >> 4;
  v4d rr = {0, 0, 0, 0}, ii2 = {0, 0, 0, 0}, ri = {0, 0, 0, 0}, ir = {0, 0, 0, 0};
  fetch(p_begin);
  for (long long p0 = p_begin; p0 < p_end; p0 += kChunk) {
    stage();
    __syncthreads();
    if (p0 + kChunk < p_end) fetch(p0 + kChunk);  // uniform branch
#pragma unroll
    for (int s = 0; s < kChunk / 4; ++s) {
      const double a_r = ar[16 * mt + r][4 * s + q], a_i = ai[16 * mt + r][4 * s + q];
      const double b_r = br[16 * ft + r][4 * s + q], b_i = bi[16 * ft + r][4 * s + q];
      rr = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_r, rr, 0, 0, 0);
      ii2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_i, ii2, 0, 0, 0);
      ri = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_i, ri, 0, 0, 0);
      ir = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_r, ir, 0, 0, 0);
    }
    __syncthreads();  // the next stage() overwrites the tiles
  }
  const int b = b0 + 16 * ft + r;
  if (b < batch) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + 16 * mt + 4 * v + q;
      if (k < nmodes)
        partial[((long long)g * batch + b) * nmodes + k] =
            make_double2(rr[v] - ii2[v], ri[v] + ir[v]);
    }
  }
}

// coeffs[o] = sum over the gx pixel groups of partial[g, o]: kSumOut outputs x
// kSumLanes group lanes per workgroup, so that even a few thousand coefficients spread
// over ~100 workgroups.
constexpr int kSumOut = 16, kSumLanes = 16;
__global__ __launch_bounds__(kSumOut * kSumLanes) void k_sum_groups(
    const double2* __restrict__ partial, double2* __restrict__ coeffs, long long count, int gx) {
  __shared__ double2 red[kSumLanes][kSumOut];
  const int o_l = threadIdx.x % kSumOut, gl = threadIdx.x / kSumOut;
  const long long o = (long long)blockIdx.x * kSumOut + o_l;
  double sr = 0, si = 0;
  if (o < count) {
    for (int g2 = gl; g2 < gx; g2 += kSumLanes) {
      const double2 v = partial[(long long)g2 * count + o];
      sr += v.x;
      si += v.y;
    }
  }
  red[gl][o_l] = make_double2(sr, si);
  __syncthreads();
  if (gl == 0 && o < count) {
    double2 t = red[0][o_l];
    for (int j2 = 1; j2 < kSumLanes; ++j2) {
      t.x += red[j2][o_l].x;
      t.y += red[j2][o_l].y;
    }
    coeffs[o] = t;
  }
}

// Argument checks that need no device.
int check_args(int n, int nb, int nmodes, int batch, int dtype, const void* a, const void* b,
               const void* c) {
  using lp2d::set_error;
  if (n < 1 || nb < 1 || nb > n || ((n - nb) & 1) || nmodes < 1 || batch < 1)
    return set_error(LP2D_E_SIZE, "need 1 <= nb <= n, n - nb even, nmodes >= 1, batch >= 1");
  if ((long long)batch * n > 0x7fffffffLL || nmodes > 65535)
    return set_error(LP2D_E_SIZE, "batch or mode count too large");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return set_error(LP2D_E_DTYPE, "unknown dtype");
  if (!a || !b || !c) return set_error(LP2D_E_ARG, "null pointer");
  return LP2D_OK;
}

int check(int n, int nb, int nmodes, int batch, int dtype, const void* a, const void* b,
          const void* c) {
  using lp2d::set_error;
  if (int rc = check_args(n, nb, nmodes, batch, dtype, a, b, c)) return rc;
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

extern "C" long long lp2d_modes_project_ws_size(int n, int nb, int nmodes, int batch) {
  if (n < 1 || nb < 1 || nb > n || nmodes < 1 || batch < 1) return 0;
  return mfma_plan(nb, nmodes, batch).work_bytes;
}

extern "C" int lp2d_modes_project_ws(int n, int nb, int nmodes, int batch, int dtype,
                                     const void* modes, const void* field, void* coeffs,
                                     void* work, long long work_bytes, void* stream) {
  if (!work) return lp2d_modes_project(n, nb, nmodes, batch, dtype, modes, field, coeffs, stream);
  if (int rc = check_args(n, nb, nmodes, batch, dtype, modes, field, coeffs)) return rc;
  const MfmaPlan pl = mfma_plan(nb, nmodes, batch);
  if (pl.gx < 1) return lp2d::set_error(LP2D_E_SIZE, "too many modes x fields for one launch");
  if (work_bytes < pl.work_bytes)
    return lp2d::set_error(LP2D_E_ARG, "workspace too small: lp2d_modes_project_ws_size bytes needed");
  if (work == coeffs || work == field || work == modes)
    return lp2d::set_error(LP2D_E_ARG, "work must not alias modes, field or coeffs");
  if (int rc = check(n, nb, nmodes, batch, dtype, modes, field, coeffs)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)pl.gx, (unsigned)(pl.mgroups * pl.fgroups));
  if (dtype == LP2D_C64)
    k_project_mfma<float><<<grid, kMfmaThreads, 0, s>>>((const C2<float>*)modes,
                                                         (const C2<float>*)field, (double2*)work,
                                                         n, nb, nmodes, batch, pl.mgroups, pl.px);
  else
    k_project_mfma<double><<<grid, kMfmaThreads, 0, s>>>((const C2<double>*)modes,
                                                          (const C2<double>*)field, (double2*)work,
                                                          n, nb, nmodes, batch, pl.mgroups, pl.px);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lp2d::set_error(LP2D_E_HIP, hipGetErrorString(e));
  const long long cnt = (long long)batch * nmodes;
  const long long sblocks = (cnt + kSumOut - 1) / kSumOut;
  if (sblocks > 0x7fffffffLL) return lp2d::set_error(LP2D_E_SIZE, "too many coefficients");
  k_sum_groups<<<(unsigned)sblocks, kSumOut * kSumLanes, 0, s>>>((const double2*)work,
                                                                (double2*)coeffs, cnt, pl.gx);
  e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : lp2d::set_error(LP2D_E_HIP, hipGetErrorString(e));
}

extern "C" int lp2d_modes_construct(int n, int nb, int nmodes, int batch, int dtype,
                                    const void* modes, const void* coeffs, void* field,
                                    void* stream) {
  if (int rc = check(n, nb, nmodes, batch, dtype, modes, coeffs, field)) return rc;
  hipStream_t s = (hipStream_t)stream;
  // batches of more than 8 fields: 4 groups of 8 per workgroup (64 pixels), the mode
  // matrix read once per 32 fields; otherwise 8 fields for 256 pixels
  const bool wide = batch > kFieldsPerPass;
  const int nf = wide ? 4 * kFieldsPerPass : kFieldsPerPass;
  const int px = wide ? 64 : kThreads;
  const long long rows = (long long)((batch + nf - 1) / nf) * n;
  if (rows > 0x7fffffffLL) return lp2d::set_error(LP2D_E_SIZE, "batch x n too large");
  const dim3 grid((unsigned)rows, (unsigned)((n + px - 1) / px));
  if (dtype == LP2D_C64) {
    if (wide)
      k_construct<float, kFieldsPerPass, 4, 64><<<grid, 256, 0, s>>>(
          (const C2<float>*)modes, (const double2*)coeffs, (C2<float>*)field, n, nb, nmodes, batch);
    else
      k_construct<float, kFieldsPerPass, 1, kThreads><<<grid, kThreads, 0, s>>>(
          (const C2<float>*)modes, (const double2*)coeffs, (C2<float>*)field, n, nb, nmodes, batch);
  } else {
    if (wide)
      k_construct<double, kFieldsPerPass, 4, 64><<<grid, 256, 0, s>>>(
          (const C2<double>*)modes, (const double2*)coeffs, (C2<double>*)field, n, nb, nmodes,
          batch);
    else
      k_construct<double, kFieldsPerPass, 1, kThreads><<<grid, kThreads, 0, s>>>(
          (const C2<double>*)modes, (const double2*)coeffs, (C2<double>*)field, n, nb, nmodes,
          batch);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : lp2d::set_error(LP2D_E_HIP, hipGetErrorString(e));
}
