// C ABI of liblp2d (declared in include/lp2d.h): argument checking, kernel-sequence
// selection for each operation, twiddle tables, and the small elementwise kernels.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lp2d.h"
#include "lp2d_registry.hpp"

using namespace lp2d;

// Scratch of the allocation-free adjoint (lp2d_propagate_adjoint_noalloc): the
// per-group partial sums of a shared screen's phase gradient.  One copy per device,
// held until the process ends; host code reaches it only through the address that
// lp2d_init takes from hipGetSymbolAddress on each device.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

namespace lp2d {
int resident_groups(const void* kernel, int threads) {
  int dev = 0, cus = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1) {
    (void)hipGetLastError();
    cus = 256;
  }
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 1;
  }
  return cus * nb;
}
}  // namespace lp2d

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(LP2D_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

constexpr int kMaxDevices = 64;
struct DeviceTables {
  cx<float>* twf = nullptr;
  cx<double>* twd = nullptr;
  float* scratch = nullptr;  // this device's copy of `scratch`
  bool ready = false;
};
std::mutex g_mu;
DeviceTables g_dev[kMaxDevices];

Registry<float>* g_rf = nullptr;
Registry<double>* g_rd = nullptr;
int g_sel[2][3][kLogMax + 1];  // [dtype][kind: rows, cols, chain][log2 n] -> variant
int g_chunk[2][kLogMax + 1];   // fields per three-sweep chunk (0 = whole batch)

// A plan's copy of the selection for its (dtype, log2 n), taken at lp2d_plan_create;
// during a plan call it is the calling thread's selection for that size, so later
// lp2d_set_variant / lp2d_set_chunk calls do not change what a plan runs.
struct PlanSel {
  int dt = -1, lg = -1;
  int sel[3] = {0, 0, 0};
  int chunk = 0;
};
thread_local const PlanSel* t_plan_sel = nullptr;

int sel_of(int dt, int kind, int lg) {
  const PlanSel* p = t_plan_sel;
  return (p && p->dt == dt && p->lg == lg) ? p->sel[kind] : g_sel[dt][kind][lg];
}
int chunk_of(int dt, int lg) {
  const PlanSel* p = t_plan_sel;
  return (p && p->dt == dt && p->lg == lg) ? p->chunk : g_chunk[dt][lg];
}
struct PlanScope {
  const PlanSel* prev;
  explicit PlanScope(const PlanSel* s) : prev(t_plan_sel) { t_plan_sel = s; }
  ~PlanScope() { t_plan_sel = prev; }
};

void ensure_registry() {
  static std::once_flag once;
  std::call_once(once, [] {
    g_rf = new Registry<float>();
    g_rd = new Registry<double>();
    std::memset(g_rf, 0, sizeof(Registry<float>));
    std::memset(g_rd, 0, sizeof(Registry<double>));
    register_f32_a(*g_rf);
    register_f32_b(*g_rf);
    register_f32_c(*g_rf);
    register_f64_a(*g_rd);
    register_f64_b(*g_rd);
    register_f64_c(*g_rd);
    std::memset(g_sel, 0, sizeof(g_sel));
    std::memset(g_chunk, 0, sizeof(g_chunk));
  });
}

// Forward tables at [0, kTwiddleLen), conjugated ones at [kTwiddleLen, 2*kTwiddleLen).
__global__ void k_init_twiddles(cx<float>* tf, cx<double>* td) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kTwiddleLen - 2) return;
  int N = 2;
  while (e >= 2 * N - 2) N *= 2;
  const int k = e - (N - 2);
  double s, c;
  sincospi(-2.0 * (double)k / (double)N, &s, &c);
  td[e] = mk(c, s);
  tf[e] = mk((float)c, (float)s);
  td[kTwiddleLen + e] = mk(c, -s);
  tf[kTwiddleLen + e] = mk((float)c, (float)-s);
}

int init_device(int device) {
  ensure_registry();
  if (device < 0 || device >= kMaxDevices) return fail(LP2D_E_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceTables& d = g_dev[device];
  if (d.ready) return LP2D_OK;
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t s = nullptr;
  int rc = LP2D_OK;
  if ((e = hipMalloc(&d.twf, 2 * kTwiddleLen * sizeof(cx<float>))) != hipSuccess ||
      (e = hipMalloc(&d.twd, 2 * kTwiddleLen * sizeof(cx<double>))) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
    rc = hip_fail(e, "twiddle table allocation");
  } else {
    k_init_twiddles<<<(kTwiddleLen + 255) / 256, 256, 0, s>>>(d.twf, d.twd);
    void* sp = nullptr;
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess)
      rc = hip_fail(e, "twiddle table init");
    else if ((e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch))) != hipSuccess || !sp)
      rc = hip_fail(e, "scratch address");
    else {
      d.scratch = (float*)sp;
      d.ready = true;
    }
  }
  if (s) hipStreamDestroy(s);
  if (prev != device) hipSetDevice(prev);
  return rc;
}

// The current device's tables, initialising them on first use.  Initialisation
// allocates and synchronises, which a stream under graph capture does not allow: a
// first call on such a stream fails with LP2D_E_INIT instead (lp2d_init beforehand).
int current_tables(DeviceTables** out, hipStream_t s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(LP2D_E_ARG, "device %d out of range", dev);
  if (!g_dev[dev].ready) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
      return fail(LP2D_E_INIT,
                  "liblp2d is not initialised on device %d and the stream is being captured: "
                  "call lp2d_init(%d) before graph capture", dev, dev);
    (void)hipGetLastError();
    int rc = init_device(dev);
    if (rc) return rc;
  }
  *out = &g_dev[dev];
  return LP2D_OK;
}

int log2_exact(int n) {
  if (n < 2 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

int check_n(int n, int batch, int dtype, int* lg) {
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return fail(LP2D_E_DTYPE, "unknown dtype %d", dtype);
  *lg = log2_exact(n);
  if (*lg < 1 || *lg > kLogMax)
    return fail(LP2D_E_SIZE, "n=%d: must be a power of two in [2, %d]", n, kMaxN);
  if (batch < 1) return fail(LP2D_E_SIZE, "batch=%d must be >= 1", batch);
  if ((long long)batch * n > (1LL << 31) / 2)
    return fail(LP2D_E_SIZE, "batch*n too large for one launch");
  return LP2D_OK;
}

int mask_dtype_ok(int t) { return t >= LP2D_MASK_U8 && t <= LP2D_MASK_C128; }

MaskD to_mask(const lp2d_mask& m) {
  MaskD d;
  d.ptr = m.ptr;
  d.dtype = m.dtype;
  d.pad_ = 0;
  d.sb = m.stride_b;
  d.sy = m.stride_y;
  d.sx = m.stride_x;
  return d;
}

template <typename T>
PassArgs<T> base_args(const DeviceTables* t, int n, int batch) {
  PassArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  if constexpr (sizeof(T) == 4)
    a.tw = t->twf;
  else
    a.tw = t->twd;
  a.tw_inv = a.tw + kTwiddleLen;
  a.fstride = (long long)n * n;
  a.batch = batch;
  a.scale = (T)1;
  return a;
}

template <typename T>
void set_optics(PassArgs<T>& a, int n, const lp2d_optics* o) {
  const double dL = o->area_size / n;
  const double k0 = 1.0 / o->wl;
  a.dL = dL;
  a.k0sq = k0 * k0;
  a.inv_ndL = 1.0 / ((double)n * dL);  // numpy.fft.fftfreq: 1.0 / (n * d)
  a.evan = o->evanescent;
  a.lens_rev = (o->lens_f != 0.0) ? k0 / (2.0 * o->lens_f) * 0.15915494309189533577 : 0.0;
  a.ap_r2 = o->aperture_radius * o->aperture_radius;
  a.z = o->z;
}

int check_optics(const lp2d_optics* o) {
  if (!o) return LP2D_OK;
  if (!(o->area_size > 0) || !(o->wl > 0)) return fail(LP2D_E_ARG, "area_size and wl must be > 0");
  if (o->evanescent < 0 || o->evanescent > 2) return fail(LP2D_E_ARG, "bad evanescent mode");
  return LP2D_OK;
}

template <typename T>
Registry<T>* reg();
template <>
Registry<float>* reg<float>() { return g_rf; }
template <>
Registry<double>* reg<double>() { return g_rd; }

template <typename T>
constexpr int dt() { return sizeof(T) == 4 ? 0 : 1; }

template <typename T>
bool use_small(int lg) { return reg<T>()->has_small[lg]; }

// ----------------------------------------------------------------- tracing
struct TraceRec {
  std::string name;
  hipEvent_t start, stop;
};
std::mutex g_trace_mu;
bool g_trace_on = false;
std::vector<TraceRec> g_trace;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

struct TraceScope {
  hipStream_t s;
  bool on;
  TraceRec rec;
  TraceScope(hipStream_t st, const std::string& name) : s(st), on(false) {
    if (!g_trace_on) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    rec.name = name;
    rec.start = pool_event();
    rec.stop = pool_event();
    on = rec.start && rec.stop && hipEventRecord(rec.start, s) == hipSuccess;
  }
  ~TraceScope() {
    if (!on) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    hipEventRecord(rec.stop, s);
    g_trace.push_back(rec);
  }
};

template <typename T>
const char* tname() { return sizeof(T) == 4 ? "c64" : "c128"; }

template <typename T>
int run_rows(const PassArgs<T>& a, int lg, bool inv, hipStream_t s) {
  const RowVar<T>& v = reg<T>()->rows[lg][sel_of(dt<T>(), 0, lg)];
  std::string nm;
  if (g_trace_on) {
    char b[64];
    snprintf(b, sizeof(b), "rows_%s n=%d %s E=%d", inv ? "inv" : "fwd", 1 << lg, tname<T>(), v.E);
    nm = b;
  }
  TraceScope ts(s, nm);
  hipError_t e = v.fn(a, inv, s);
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "row pass launch");
}
template <typename T>
int run_cols(const PassArgs<T>& a, int lg, int mode, hipStream_t s) {
  const ColVar<T>& v = reg<T>()->cols[lg][sel_of(dt<T>(), 1, lg)];
  std::string nm;
  if (g_trace_on) {
    char b[64];
    snprintf(b, sizeof(b), "cols_m%d n=%d %s E=%d W=%d", mode, 1 << lg, tname<T>(), v.E, v.W);
    nm = b;
  }
  TraceScope ts(s, nm);
  hipError_t e = v.fn(a, mode, s);
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "column pass launch");
}
template <typename T>
int run_small(const PassArgs<T>& a, int lg, int mode, hipStream_t s) {
  std::string nm;
  if (g_trace_on) {
    char b[64];
    snprintf(b, sizeof(b), "field_m%d n=%d %s", mode, 1 << lg, tname<T>());
    nm = b;
  }
  TraceScope ts(s, nm);
  hipError_t e = reg<T>()->small[lg].fn(a, mode, s);
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "whole-field launch");
}

// The gradient pass carries the screen factor and the accumulators: it runs the row
// schedule with the fewest elements per thread (lowest register pressure), whose
// factor stays in registers across the fields.  (The selected wide schedule, which
// could read a block-tiled plane in place, must recompute the factor per field and
// measured 1.6x slower, more than the tiled column pass gains.)
template <typename T>
const RowVar<T>& grad_variant(int lg) {
  int vi = 0;
  for (int i = 1; i < reg<T>()->nrows[lg]; ++i)
    if (reg<T>()->rows[lg][i].E < reg<T>()->rows[lg][vi].E) vi = i;
  return reg<T>()->rows[lg][vi];
}

// F_TILED when the selected schedules can hand the intermediate plane between the
// sweeps over in the block-tiled layout: the pipelined complex64 column class with
// strips of tile_cols(n), row tiles of whole tile_rows(n)-row blocks, and (grad) the
// adjoint's gradient pass reading it (whole blocks again, unless it stores nothing).
template <typename T>
int tiled_flag(const PassArgs<T>& cols_a, int lg, bool grad = false, bool grad_store = true) {
  if (sizeof(T) != 4) return 0;
  const int n = 1 << lg;
  const ColVar<T>& cv = reg<T>()->cols[lg][sel_of(dt<T>(), 1, lg)];
  const RowVar<T>& rv = reg<T>()->rows[lg][sel_of(dt<T>(), 0, lg)];
  const int tr = tile_rows(n), tc = tile_cols(n);
  if (!(cv.kind & 2) || cv.W != tc || (n / cv.E) % tr != 0) return 0;
  if (!pipe_eligible(cols_a, MODE_FWD | MODE_INV)) return 0;
  if (rv.lines % tr != 0 || (n / rv.E) % tc != 0) return 0;
  if (grad) {
    const RowVar<T>& gv = grad_variant<T>(lg);
    if ((gv.grad_lines % tr != 0 && grad_store) || (n / gv.E) % tc != 0) return 0;
  }
  return F_TILED;
}

// ----------------------------------------------------------------- elementwise kernels
template <typename T>
__global__ void k_spatial_only(PassArgs<T> a, int n, long long total) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(i / a.fstride);
    const long long r = i - (long long)b * a.fstride;
    const int y = (int)(r / n), x = (int)(r - (long long)y * n);
    a.field_out[i] = spatial_pro(a, a.src[i], b, y, x, n);
  }
}

template <typename T>
__global__ void k_spectral_only(PassArgs<T> a, int n, long long total) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(i / a.fstride);
    const long long r = i - (long long)b * a.fstride;
    const int ky = (int)(r / n), kx = (int)(r - (long long)ky * n);
    a.spec_out[i] = spectral_mid(a, a.src[i], b, ky, kx, n);
  }
}

template <typename T>
__global__ void k_intensity(const cx<T>* __restrict__ f, T* __restrict__ out, long long count) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const cx<T> v = f[i];
    out[i] = v.re * v.re + v.im * v.im;
  }
}

// gphi[b, i] = Im(conj(u[b, i] * exp(i*phase)) * g[b, i]) per field, or its sum over
// the batch (sum != 0: one plane); the screen factor as in the passes (FP64 reduction
// of phase / 2pi to a fraction of a turn).
template <typename T>
__global__ void k_phase_grad(const cx<T>* __restrict__ u, const cx<T>* __restrict__ g,
                             const float* __restrict__ phase, long long phase_sb,
                             float* __restrict__ out, long long count, int batch, int sum) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    float acc = 0.f;
    cx<T> e = mk((T)1, (T)0);
    for (int b = 0; b < batch; ++b) {
      if (phase && (b == 0 || phase_sb != 0)) {
        double r = (double)phase[(long long)b * phase_sb + i] * kInv2Pi;
        r -= rint(r);
        e = exp_i_revolutions<T>(r);
      }
      const cx<T> uu = cmul(u[(long long)b * count + i], e);
      const cx<T> gg = g[(long long)b * count + i];
      const float v = (float)(uu.re * gg.im - uu.im * gg.re);
      if (sum)
        acc += v;
      else
        out[(long long)b * count + i] = v;
    }
    if (sum) out[i] = acc;
  }
}

// out may alias plane 0 of in: each thread reads all planes of its element first.
__global__ void k_sum_planes(const float* in, float* out, long long count, int batch) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < batch; ++b) s += in[(long long)b * count + i];
    out[i] = s;
  }
}

__global__ void k_kz(double* __restrict__ kz, int n, double k0sq, double inv_ndL) {
#pragma clang fp contract(off)
  const long long total = (long long)n * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int ky = (int)(i / n), kx = (int)(i - (long long)ky * n);
    const int sx = kx < (n + 1) / 2 ? kx : kx - n;
    const int sy = ky < (n + 1) / 2 ? ky : ky - n;
    const double fx = (double)sx * inv_ndL, fy = (double)sy * inv_ndL;
    const double kz2 = (k0sq - fx * fx) - fy * fy;
    kz[i] = kz2 >= 0.0 ? sqrt(kz2) : 0.0;
  }
}

// Intensity moments: one workgroup column (blockIdx.y) per field, grid-stride over
// its pixels, workgroup reduction in LDS, one FP64 atomic per workgroup and moment.
template <typename T>
__device__ __forceinline__ double pix_intensity(const cx<T>* f, long long i) {
  const cx<T> v = f[i];
  const T a = v.re * v.re + v.im * v.im;  // |u|^2 in the field's precision
  return (double)a;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;
}

__global__ void k_moments_clear(double* out, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = 0.0;
}

template <typename T>
__global__ void k_moments(const cx<T>* __restrict__ field, int n, double dL, double* out,
                          int pass) {
  __shared__ double sh[16];
  const int b = blockIdx.y;
  const long long total = (long long)n * n;
  const cx<T>* f = field + (long long)b * total;
  double* o = out + 8 * b;
  double xc = 0.0, yc = 0.0;
  if (pass == 1) {
    xc = o[5] / o[0];
    yc = o[6] / o[0];
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int iy = (int)(i / n), ix = (int)(i - (long long)iy * n);
    const double w = pix_intensity(f, i);
    const double X = (double)(ix - n / 2) * dL, Y = (double)(iy - n / 2) * dL;
    if (pass == 0) {
      s0 += w;
      s1 += w * X;
      s2 += w * Y;
    } else {
      s1 += w * (X - xc) * (X - xc);
      s2 += w * (Y - yc) * (Y - yc);
    }
  }
  if (pass == 0) {
    s0 = block_sum(s0, sh);
    if (threadIdx.x == 0) atomicAdd(&o[0], s0);
  }
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) atomicAdd(&o[pass == 0 ? 5 : 3], s1);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) atomicAdd(&o[pass == 0 ? 6 : 4], s2);
}

__global__ void k_moments_final(double* out, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  double* o = out + 8 * b;
  const double s = o[0];
  o[1] = o[5] / s;
  o[2] = o[6] / s;
  o[3] = sqrt(o[3] / s);
  o[4] = sqrt(o[4] / s);
}

unsigned grid_for(long long total) {
  long long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ----------------------------------------------------------------- operations
template <typename T>
int op_fft2(int n, int lg, int batch, const void* in, void* out, int inverse, hipStream_t s) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  if (use_small<T>(lg)) {
    a.src = (const cx<T>*)in;
    a.dst = (cx<T>*)out;
    if (inverse) {
      a.flags = F_SCALE;
      a.scale = (T)(1.0 / ((double)n * n));
    }
    return run_small<T>(a, lg, inverse ? MODE_INV : MODE_FWD, s);
  }
  if (!inverse) {
    a.src = (const cx<T>*)in;
    a.dst = (cx<T>*)out;
    if (int rc = run_rows<T>(a, lg, false, s)) return rc;
    a.src = (const cx<T>*)out;
    return run_cols<T>(a, lg, MODE_FWD, s);
  }
  a.src = (const cx<T>*)in;
  a.dst = (cx<T>*)out;
  if (int rc = run_cols<T>(a, lg, MODE_INV, s)) return rc;
  a.src = (const cx<T>*)out;
  a.flags = F_SCALE;
  a.scale = (T)(1.0 / ((double)n * n));
  return run_rows<T>(a, lg, true, s);
}

template <typename T>
int op_spatial(int n, int lg, int batch, void* field, const lp2d_mask* masks, int nmasks,
               const lp2d_optics* o, const float* phase, int64_t phase_sb, void* spec,
               hipStream_t s) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  int fl = 0;
  if (nmasks > 0) {
    a.m0 = to_mask(masks[0]);
    fl |= F_MASK0;
  }
  if (nmasks > 1) {
    a.m1 = to_mask(masks[1]);
    fl |= F_MASK1;
  }
  if (o) {
    set_optics(a, n, o);
    if (o->lens_f != 0.0) fl |= F_LENS;
  }
  if (phase) {
    a.phase = phase;
    a.phase_sb = phase_sb;
    fl |= F_PHASE;
  }
  a.field_out = (cx<T>*)field;
  fl |= F_FIELD_OUT;
  a.src = (const cx<T>*)field;
  if (!spec) {
    a.flags = fl;
    const long long total = (long long)batch * n * n;
    k_spatial_only<T><<<grid_for(total), 256, 0, s>>>(a, n, total);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LP2D_OK : hip_fail(e, "spatial kernel launch");
  }
  a.dst = (cx<T>*)spec;
  a.flags = fl;
  if (use_small<T>(lg)) return run_small<T>(a, lg, MODE_FWD, s);
  if (int rc = run_rows<T>(a, lg, false, s)) return rc;
  a.flags = 0;
  a.src = (const cx<T>*)spec;
  return run_cols<T>(a, lg, MODE_FWD, s);
}

template <typename T>
int op_spectral(int n, int lg, int batch, void* spec, const lp2d_mask* masks, int nmasks,
                const lp2d_optics* o, void* field_out, hipStream_t s) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  int fl = F_SPEC_OUT;
  if (o) {
    set_optics(a, n, o);
    if (o->z != 0.0) fl |= F_H;
    if (o->aperture_radius > 0.0) fl |= F_APERTURE;
  }
  if (nmasks > 0) {
    a.s0 = to_mask(masks[0]);
    fl |= F_SMASK0;
  }
  if (nmasks > 1) {
    a.s1 = to_mask(masks[1]);
    fl |= F_SMASK1;
  }
  a.src = (const cx<T>*)spec;
  a.spec_out = (cx<T>*)spec;
  if (!field_out) {
    a.flags = fl;
    const long long total = (long long)batch * n * n;
    k_spectral_only<T><<<grid_for(total), 256, 0, s>>>(a, n, total);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LP2D_OK : hip_fail(e, "spectral kernel launch");
  }
  a.dst = (cx<T>*)field_out;
  const T sc = (T)(1.0 / ((double)n * n));
  if (use_small<T>(lg)) {
    a.flags = fl | F_SCALE;
    a.scale = sc;
    return run_small<T>(a, lg, MODE_INV, s);
  }
  a.flags = fl;
  if (int rc = run_cols<T>(a, lg, MODE_INV, s)) return rc;
  PassArgs<T> r = base_args<T>(t, n, batch);
  r.src = (const cx<T>*)field_out;
  r.dst = (cx<T>*)field_out;
  r.flags = F_SCALE;
  r.scale = sc;
  return run_rows<T>(r, lg, true, s);
}

template <typename T>
int op_propagate_chunk(int n, int lg, int batch, const void* in, void* out, const float* phase,
                       int64_t phase_sb, const lp2d_optics* o, void* spec, hipStream_t s);

// Field-to-field propagation.  Large sizes run three sweeps per chunk of fields so
// that a chunk's intermediate planes are still in the Infinity Cache when the next
// sweep reads them (chunk size from lp2d_set_chunk; 0 = the whole batch at once).
template <typename T>
int op_propagate(int n, int lg, int batch, const void* in, void* out, const float* phase,
                 int64_t phase_sb, const lp2d_optics* o, void* spec, hipStream_t s) {
  const int chunk = chunk_of(dt<T>(), lg);
  if (chunk <= 0 || chunk >= batch || use_small<T>(lg))
    return op_propagate_chunk<T>(n, lg, batch, in, out, phase, phase_sb, o, spec, s);
  const long long fs = (long long)n * n;
  for (int c0 = 0; c0 < batch; c0 += chunk) {
    const int cb = batch - c0 < chunk ? batch - c0 : chunk;
    if (int rc = op_propagate_chunk<T>(
            n, lg, cb, (const cx<T>*)in + c0 * fs, (cx<T>*)out + c0 * fs,
            phase ? phase + c0 * phase_sb : nullptr, phase_sb, o,
            spec ? (void*)((cx<T>*)spec + c0 * fs) : nullptr, s))
      return rc;
  }
  return LP2D_OK;
}

template <typename T>
int op_propagate_chunk(int n, int lg, int batch, const void* in, void* out, const float* phase,
                       int64_t phase_sb, const lp2d_optics* o, void* spec, hipStream_t s) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  set_optics(a, n, o);
  int pro = 0, mid = 0;
  if (phase) {
    a.phase = phase;
    a.phase_sb = phase_sb;
    pro |= F_PHASE;
  }
  if (o->lens_f != 0.0) pro |= F_LENS;
  if (o->z != 0.0) mid |= F_H;
  if (o->aperture_radius > 0.0) mid |= F_APERTURE;
  if (spec) {
    a.spec_out = (cx<T>*)spec;
    mid |= F_SPEC_OUT;
  }
  const T sc = (T)(1.0 / ((double)n * n));
  a.src = (const cx<T>*)in;
  a.dst = (cx<T>*)out;
  if (use_small<T>(lg)) {
    a.flags = pro | mid | F_SCALE;
    a.scale = sc;
    return run_small<T>(a, lg, MODE_FWD | MODE_INV, s);
  }
  a.flags = mid;
  const int tl = tiled_flag<T>(a, lg);
  a.flags = pro | tl;
  if (int rc = run_rows<T>(a, lg, false, s)) return rc;
  a.src = (const cx<T>*)out;
  a.flags = mid | tl;
  if (int rc = run_cols<T>(a, lg, MODE_FWD | MODE_INV, s)) return rc;
  a.flags = F_SCALE | tl;
  a.scale = sc;
  return run_rows<T>(a, lg, true, s);
}

// Field groups of the gradient pass over a batch: about 512K threads in flight,
// whatever the tile height; the kernel gives group g the fields [g*per,
// min(batch, (g+1)*per)), so the count is recomputed from per and no group is empty
// (batch 5 in 4 groups of 2 would leave group 3 at field 6).
template <typename T>
int grad_groups(int lg, int batch) {
  const RowVar<T>& v = grad_variant<T>(lg);
  const int n = 1 << lg;
  const int tpf = (n + v.grad_lines - 1) / v.grad_lines;
  int groups = (2048 * 256 / ((n / v.E) * v.grad_lines)) / tpf;
  if (groups < 1) groups = 1;
  if (groups > batch) groups = batch;
  const int per = (batch + groups - 1) / groups;
  return (batch + per - 1) / per;
}

// The gradient pass with a shared screen: each group of fields sums its phase
// gradients into one n x n plane of `partial` (planes 0 .. groups-1), and a second
// kernel adds the planes into `out` (which may be plane 0 of partial).  With a single
// group the pass writes `out` directly.
// Multi-plane stack u_{k+1} = P(z)[u_k exp(i screens[k])], k < K, screens shared by the
// batch.  With a block-tiled intermediate (complex64 three-sweep sizes) it runs
//   rows(screen 0) -> [cols -> chain(inverse rows, screen k+1, forward rows)] x (K-1)
//   -> cols -> inverse rows,
// 2K + 1 sweeps instead of 3K, in the output buffer; the chain passes store the plane
// inputs u_k that planes[k] asks for.  Otherwise one fused propagation per plane.
template <typename T>
int op_stack(int n, int lg, int batch, const void* in, void* out, const float* screens, int K,
             const lp2d_optics* o, void* const* planes, hipStream_t s) {
  const long long nn = (long long)n * n;
  auto plane = [&](int k) -> void* { return (planes && k > 0 && k < K) ? planes[k] : nullptr; };
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  set_optics(a, n, o);
  int mid = 0;
  if (o->z != 0.0) mid |= F_H;
  if (o->aperture_radius > 0.0) mid |= F_APERTURE;
  a.flags = mid;
  const int tl = use_small<T>(lg) ? 0 : tiled_flag<T>(a, lg);
  const int cv = sel_of(dt<T>(), 2, lg);
  const RowChainLaunch<T> chain = reg<T>()->nchain[lg] > cv ? reg<T>()->chain[lg][cv] : nullptr;
  if (!tl || !chain || K < 2 || o->lens_f != 0.0) {
    const void* cur = in;
    for (int k = 0; k < K; ++k) {
      void* dst = (k + 1 < K && plane(k + 1)) ? plane(k + 1) : out;
      if (int rc = op_propagate_chunk<T>(n, lg, batch, cur, dst, screens + k * nn, 0, o,
                                         nullptr, s))
        return rc;
      cur = dst;
    }
    return LP2D_OK;
  }
  const T sc = (T)(1.0 / ((double)n * n));
  a.src = (const cx<T>*)in;
  a.dst = (cx<T>*)out;
  a.phase = screens;
  a.phase_sb = 0;
  a.flags = F_PHASE | tl;
  if (int rc = run_rows<T>(a, lg, false, s)) return rc;
  for (int k = 0; k < K; ++k) {
    PassArgs<T> c = base_args<T>(t, n, batch);
    set_optics(c, n, o);
    c.src = (const cx<T>*)out;
    c.dst = (cx<T>*)out;
    c.flags = mid | tl;
    if (int rc = run_cols<T>(c, lg, MODE_FWD | MODE_INV, s)) return rc;
    PassArgs<T> r = base_args<T>(t, n, batch);
    r.src = (const cx<T>*)out;
    r.dst = (cx<T>*)out;
    r.scale = sc;
    if (k + 1 == K) {
      r.flags = F_SCALE | tl;
      return run_rows<T>(r, lg, true, s);
    }
    r.phase = screens + (k + 1) * nn;
    r.phase_sb = 0;
    r.field_out = (cx<T>*)plane(k + 1);
    r.flags = F_SCALE | F_PHASE | tl | (r.field_out ? F_FIELD_OUT : 0);
    std::string nm;
    if (g_trace_on) {
      char b[64];
      snprintf(b, sizeof(b), "rows_chain n=%d %s E=%d", n, tname<T>(), reg<T>()->chain_E[lg][cv]);
      nm = b;
    }
    hipError_t e;
    {
      TraceScope ts(s, nm);
      e = chain(r, s);
    }
    if (e != hipSuccess) return hip_fail(e, "chained row pass launch");
  }
  return LP2D_OK;
}

template <typename T>
int run_rows_grad(const PassArgs<T>& a, int lg, float* partial, float* out, hipStream_t s) {
  const RowVar<T>& v = grad_variant<T>(lg);
  const int n = 1 << lg;
  const int groups = grad_groups<T>(lg, a.batch);
  std::string nm;
  if (g_trace_on) {
    char b[64];
    snprintf(b, sizeof(b), "rows_grad n=%d %s E=%d", n, tname<T>(), v.E);
    nm = b;
  }
  hipError_t e;
  {
    TraceScope ts(s, nm);
    e = v.grad_fn(a, groups, groups > 1 ? partial : out, s);
  }
  if (e != hipSuccess) return hip_fail(e, "adjoint gradient pass launch");
  if (groups > 1) {
    const long long cnt = (long long)n * n;
    k_sum_planes<<<grid_for(cnt), 256, 0, s>>>(partial, out, cnt, groups);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "sum kernel launch");
  }
  return LP2D_OK;
}

template <typename T>
// scratch (allocation-free form): with a shared screen over a batch, gphase is one
// n x n plane and the per-field or per-group gradient planes go to scratch; otherwise
// (scratch == nullptr) they go to gphase's planes and the sum to its plane 0.
int op_adjoint(int n, int lg, int batch, const void* gout, const void* u, const float* phase,
               int64_t phase_sb, const lp2d_optics* o, void* gin, float* gphase, int aflags,
               void* work, hipStream_t s, float* scratch = nullptr) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  set_optics(a, n, o);
  a.z = -o->z;  // conj(H(z)) = H(-z) under the reference's trunc rule; evanescent stays 1
  const int no_store = (aflags & LP2D_ADJ_NO_GRAD_IN) ? F_NO_STORE : 0;
  int mid = 0, epi = F_SCALE | no_store;
  if (o->z != 0.0) mid |= F_H;
  if (o->aperture_radius > 0.0) mid |= F_APERTURE;
  if (phase || o->lens_f != 0.0) {
    a.phase = phase;
    a.phase_sb = phase_sb;
    a.u_saved = (const cx<T>*)u;
    a.u_sb = (long long)n * n;
    a.grad_phase = gphase;
    epi |= F_GRAD;
  }
  const bool shared_sum = gphase && phase_sb == 0 && batch > 1;
  float* planes = (shared_sum && scratch) ? scratch : gphase;  // per-field / per-group planes
  const T sc = (T)(1.0 / ((double)n * n));
  a.scale = sc;
  a.src = (const cx<T>*)gout;
  a.dst = (cx<T>*)gin;
  if (use_small<T>(lg)) {
    a.flags = mid | epi;
    if (shared_sum) a.grad_phase = planes;
    if (int rc = run_small<T>(a, lg, MODE_FWD | MODE_INV, s)) return rc;
  } else {
    a.flags = mid;
    // With a workspace the column pass writes there and the last pass reads it: no
    // in-place conversion of a tiled plane, so any gradient schedule can read one.
    int tl = tiled_flag<T>(a, lg, shared_sum, no_store == 0 && !work);
    const bool oop = work && tl;
    a.flags = tl;
    if (int rc = run_rows<T>(a, lg, false, s)) return rc;
    a.src = (const cx<T>*)gin;
    if (oop) a.dst = (cx<T>*)work;
    a.flags = mid | tl;
    if (int rc = run_cols<T>(a, lg, MODE_FWD | MODE_INV, s)) return rc;
    if (oop) {
      a.src = (const cx<T>*)work;
      a.dst = (cx<T>*)gin;
    }
    if (shared_sum) {
      a.flags = F_SCALE | no_store | tl;
      return run_rows_grad<T>(a, lg, planes, gphase, s);
    }
    a.flags = epi | tl;
    if (int rc = run_rows<T>(a, lg, true, s)) return rc;
  }
  if (shared_sum) {  // per-field planes from the whole-field kernel: reduce into gphase
    const long long cnt = (long long)n * n;
    k_sum_planes<<<grid_for(cnt), 256, 0, s>>>(planes, gphase, cnt, batch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sum kernel launch");
  }
  return LP2D_OK;
}

// Field groups of the stack adjoint's chained pass (about 512K threads in flight, no
// empty group, as grad_groups).
template <typename T>
int adj_chain_groups(int lg, int batch) {
  const Registry<T>* r = reg<T>();
  if (!r->adj_chain[lg][0]) return 1;
  const int tpf = (1 << lg) / r->adj_lines[lg];
  int groups = (2048 * 256 / r->adj_threads[lg]) / tpf;
  if (groups < 1) groups = 1;
  if (groups > batch) groups = batch;
  const int per = (batch + groups - 1) / groups;
  return (batch + per - 1) / per;
}

// Float workspace of op_stack_adjoint: the per-group (chained pass) or per-field /
// per-group (plane-by-plane adjoint) phase-gradient planes of the shared screens.
template <typename T>
size_t stack_adjoint_ws(int lg, int batch) {
  if (batch < 2) return 0;
  const size_t plane = ((size_t)1 << (2 * lg)) * sizeof(float);
  const int gc = adj_chain_groups<T>(lg, batch);
  const size_t chain = gc > 1 ? (size_t)gc * plane : 0;
  const int gg = grad_groups<T>(lg, batch);
  const size_t per = use_small<T>(lg) ? (size_t)batch * plane : (gg > 1 ? (size_t)gg * plane : 0);
  return chain > per ? chain : per;
}

// Adjoint of op_stack: from grad_out = dL/du_K back to dL/du_0 (grad_in) and the
// batch-summed screen gradients.  With a block-tiled intermediate it runs
//   forward rows -> [cols(-z) -> adjoint chain(inverse rows, fold screen k, gradient
//   of screen k, forward rows)] x (K-1) -> cols(-z) -> last chain (natural store),
// 2K + 1 sweeps in grad_in; otherwise one adjoint per plane.  uk[k] = u_k, the input
// of plane k (the chained form reads it only for the screen gradients).
template <typename T>
int op_stack_adjoint(int n, int lg, int batch, const void* gout, const void* const* uk,
                     const float* screens, int K, const lp2d_optics* o, void* gin,
                     float* gscreens, int aflags, float* ws, hipStream_t s) {
  const long long nn = (long long)n * n;
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, n, batch);
  set_optics(a, n, o);
  a.z = -o->z;
  int mid = 0;
  if (o->z != 0.0) mid |= F_H;
  if (o->aperture_radius > 0.0) mid |= F_APERTURE;
  a.flags = mid;
  const int tl = use_small<T>(lg) ? 0 : tiled_flag<T>(a, lg);
  const RowAdjChainLaunch<T> mid_fn = reg<T>()->adj_chain[lg][0];
  const RowAdjChainLaunch<T> last_fn = reg<T>()->adj_chain[lg][1];
  if (!tl || !mid_fn || !last_fn || o->lens_f != 0.0) {
    const void* cur = gout;
    for (int k = K - 1; k >= 0; --k) {
      const int fl = k > 0 ? 0 : (aflags & LP2D_ADJ_NO_GRAD_IN);
      if (int rc = op_adjoint<T>(n, lg, batch, cur, uk ? uk[k] : nullptr, screens + k * nn, 0, o,
                                 gin, gscreens ? gscreens + k * nn : nullptr, fl, nullptr, s,
                                 ws))
        return rc;
      cur = gin;
    }
    return LP2D_OK;
  }
  const int groups = adj_chain_groups<T>(lg, batch);
  a.src = (const cx<T>*)gout;
  a.dst = (cx<T>*)gin;
  a.flags = tl;
  if (int rc = run_rows<T>(a, lg, false, s)) return rc;
  const T sc = (T)(1.0 / ((double)n * n));
  for (int k = K - 1; k >= 0; --k) {
    PassArgs<T> c = base_args<T>(t, n, batch);
    set_optics(c, n, o);
    c.z = -o->z;
    c.src = (const cx<T>*)gin;
    c.dst = (cx<T>*)gin;
    c.flags = mid | tl;
    if (int rc = run_cols<T>(c, lg, MODE_FWD | MODE_INV, s)) return rc;
    PassArgs<T> r = base_args<T>(t, n, batch);
    r.src = (const cx<T>*)gin;
    r.dst = (cx<T>*)gin;
    r.scale = sc;
    r.phase = screens + k * nn;
    r.phase_sb = 0;
    r.u_saved = gscreens ? (const cx<T>*)uk[k] : nullptr;
    r.u_sb = nn;
    r.flags = F_SCALE | tl | (k == 0 && (aflags & LP2D_ADJ_NO_GRAD_IN) ? F_NO_STORE : 0);
    float* gk = gscreens ? gscreens + k * nn : nullptr;
    float* part = gk && groups > 1 ? ws : gk;
    std::string nm;
    if (g_trace_on) {
      char b[64];
      snprintf(b, sizeof(b), "rows_adj_chain n=%d %s%s", n, tname<T>(), k ? "" : " last");
      nm = b;
    }
    hipError_t e;
    {
      TraceScope ts(s, nm);
      e = (k ? mid_fn : last_fn)(r, groups, part, s);
    }
    if (e != hipSuccess) return hip_fail(e, "adjoint chained row pass launch");
    if (part && part != gk) {
      k_sum_planes<<<grid_for(nn), 256, 0, s>>>(part, gk, nn, groups);
      if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "sum kernel launch");
    }
  }
  return LP2D_OK;
}

// ----------------------------------------------------------------- uneven batches
// Fields are grouped by size; each size bucket keeps device tables of its fields'
// source / destination pointers and optics, and runs through the tuned per-size
// kernels (one launch for n <= 128 c64, three sweeps otherwise).
struct Bucket {
  int n = 0, lg = 0, count = 0;
  int flags = 0;
  void* d_src = nullptr;
  void* d_dst = nullptr;
  FieldOpt* d_opt = nullptr;
};

template <typename T>
int run_bucket(const Bucket& bk, hipStream_t s) {
  DeviceTables* t;
  if (int rc = current_tables(&t, s)) return rc;
  PassArgs<T> a = base_args<T>(t, bk.n, bk.count);
  a.fstride = 0;
  a.fopt = bk.d_opt;
  a.src_tab = (const cx<T>* const*)bk.d_src;
  a.dst_tab = (cx<T>* const*)bk.d_dst;
  const T sc = (T)(1.0 / ((double)bk.n * bk.n));
  if (use_small<T>(bk.lg)) {
    a.flags = bk.flags | F_SCALE;
    a.scale = sc;
    return run_small<T>(a, bk.lg, MODE_FWD | MODE_INV, s);
  }
  a.flags = bk.flags;
  const int tl = tiled_flag<T>(a, bk.lg);
  a.flags = tl;
  if (int rc = run_rows<T>(a, bk.lg, false, s)) return rc;
  a.src_tab = (const cx<T>* const*)bk.d_dst;  // later sweeps work in the destinations
  a.flags = bk.flags | tl;
  if (int rc = run_cols<T>(a, bk.lg, MODE_FWD | MODE_INV, s)) return rc;
  a.flags = F_SCALE | tl;
  a.scale = sc;
  return run_rows<T>(a, bk.lg, true, s);
}

}  // namespace

// A plan: one field geometry on one device, with the selection it was created under.
struct lp2d_plan {
  int n = 0, lg = 0, batch = 0, dtype = 0, device = 0;
  PlanSel sel;
};

namespace {

// The plan's device must be current (its tables and selection belong to it).
int plan_enter(const lp2d_plan* p) {
  if (!p) return fail(LP2D_E_ARG, "null plan");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev != p->device)
    return fail(LP2D_E_ARG, "plan made on device %d used with device %d current", p->device, dev);
  return LP2D_OK;
}

// Bytes of the adjoint workspace: the column pass's output plane (complex64 sizes run
// in three sweeps), then the per-group (or per-field) gradient planes of a shared
// screen, each part 256-byte aligned.
template <typename T>
void adjoint_ws_parts(int n, int lg, int batch, size_t* plane_bytes, size_t* grad_bytes) {
  const size_t nn = (size_t)n * n;
  const bool small = use_small<T>(lg);
  *plane_bytes = (!small && sizeof(T) == 4) ? (size_t)batch * nn * sizeof(cx<T>) : 0;
  size_t g = 0;
  if (batch > 1) {
    const int gg = grad_groups<T>(lg, batch);
    g = (small ? (size_t)batch : (gg > 1 ? (size_t)gg : 0)) * nn * sizeof(float);
  }
  *grad_bytes = g;
}
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct lp2d_batch {
  int device = 0;
  int dtype = 0;
  void* d_mem = nullptr;
  std::vector<Bucket> buckets;
};

// ===================================================================== C ABI
extern "C" {

int lp2d_init(int device) { return init_device(device); }

const char* lp2d_last_error(void) { return g_err.c_str(); }

int lp2d_version(void) { return 10000; }

int lp2d_max_n(int dtype) { return dtype == LP2D_C64 || dtype == LP2D_C128 ? kMaxN : 0; }

int lp2d_kz_grid(int n, const lp2d_optics* o, double* kz, void* stream) {
  int lg;
  if (int rc = check_n(n, 1, LP2D_C64, &lg)) return rc;
  if (!o || !kz) return fail(LP2D_E_ARG, "null optics or output");
  if (int rc = check_optics(o)) return rc;
  DeviceTables* t;
  if (int rc = current_tables(&t, (hipStream_t)stream)) return rc;
  const double dL = o->area_size / n, k0 = 1.0 / o->wl;
  k_kz<<<grid_for((long long)n * n), 256, 0, (hipStream_t)stream>>>(kz, n, k0 * k0,
                                                                      1.0 / ((double)n * dL));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "kz kernel launch");
}

int lp2d_moments(int n, int batch, int dtype, const void* field, double dL, double* out,
                 void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!field || !out) return fail(LP2D_E_ARG, "null pointer");
  if (!(dL > 0)) return fail(LP2D_E_ARG, "dL must be > 0");
  hipStream_t s = (hipStream_t)stream;
  k_moments_clear<<<(8 * batch + 255) / 256, 256, 0, s>>>(out, 8 * batch);
  long long per = ((long long)n * n + 255) / 256;
  const unsigned gx = (unsigned)(per < 1024 ? per : 1024);
  const dim3 grid(gx, (unsigned)batch);
  for (int pass = 0; pass < 2; ++pass) {
    if (dtype == LP2D_C64)
      k_moments<float><<<grid, 256, 0, s>>>((const cx<float>*)field, n, dL, out, pass);
    else
      k_moments<double><<<grid, 256, 0, s>>>((const cx<double>*)field, n, dL, out, pass);
  }
  k_moments_final<<<(batch + 63) / 64, 64, 0, s>>>(out, batch);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "moments kernel launch");
}

int lp2d_fft2(int n, int batch, int dtype, const void* in, void* out, int inverse, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!in || !out) return fail(LP2D_E_ARG, "null field pointer");
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64 ? op_fft2<float>(n, lg, batch, in, out, inverse, s)
                           : op_fft2<double>(n, lg, batch, in, out, inverse, s);
}

int lp2d_spatial_apply(int n, int batch, int dtype, void* field, const lp2d_mask* masks, int nmasks,
                       const lp2d_optics* optics, const float* phase, int64_t phase_stride_b,
                       void* spectrum_out, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!field) return fail(LP2D_E_ARG, "null field pointer");
  if (nmasks < 0 || nmasks > 2 || (nmasks > 0 && !masks)) return fail(LP2D_E_ARG, "nmasks must be 0..2");
  for (int i = 0; i < nmasks; ++i) {
    if (!masks[i].ptr) return fail(LP2D_E_ARG, "mask %d has a null pointer", i);
    if (!mask_dtype_ok(masks[i].dtype)) return fail(LP2D_E_DTYPE, "mask %d: unknown dtype %d", i, masks[i].dtype);
  }
  if (int rc = check_optics(optics)) return rc;
  if (spectrum_out == field) return fail(LP2D_E_ARG, "spectrum_out must not alias field");
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_spatial<float>(n, lg, batch, field, masks, nmasks, optics, phase, phase_stride_b, spectrum_out, s)
             : op_spatial<double>(n, lg, batch, field, masks, nmasks, optics, phase, phase_stride_b, spectrum_out, s);
}

int lp2d_spectral_apply(int n, int batch, int dtype, void* spectrum, const lp2d_mask* masks,
                        int nmasks, const lp2d_optics* optics, void* field_out, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!spectrum) return fail(LP2D_E_ARG, "null spectrum pointer");
  if (nmasks < 0 || nmasks > 2 || (nmasks > 0 && !masks)) return fail(LP2D_E_ARG, "nmasks must be 0..2");
  for (int i = 0; i < nmasks; ++i) {
    if (!masks[i].ptr) return fail(LP2D_E_ARG, "mask %d has a null pointer", i);
    if (!mask_dtype_ok(masks[i].dtype)) return fail(LP2D_E_DTYPE, "mask %d: unknown dtype %d", i, masks[i].dtype);
  }
  if (int rc = check_optics(optics)) return rc;
  if (field_out == spectrum) return fail(LP2D_E_ARG, "field_out must not alias spectrum");
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64 ? op_spectral<float>(n, lg, batch, spectrum, masks, nmasks, optics, field_out, s)
                           : op_spectral<double>(n, lg, batch, spectrum, masks, nmasks, optics, field_out, s);
}

int lp2d_phase_grad(int n, int batch, int dtype, const void* field, const void* grad,
                    const float* phase, int64_t phase_stride_b, float* grad_phase,
                    int sum_batch, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!field || !grad || !grad_phase) return fail(LP2D_E_ARG, "null pointer");
  const long long nn = (long long)n * n;
  if (phase_stride_b != 0 && phase_stride_b != nn)
    return fail(LP2D_E_ARG, "phase_stride_b must be 0 or n*n");
  if ((const void*)grad_phase == field || (const void*)grad_phase == grad)
    return fail(LP2D_E_ARG, "grad_phase must not alias the fields");
  hipStream_t s = (hipStream_t)stream;
  const int sum = sum_batch ? 1 : 0;
  if (dtype == LP2D_C64)
    k_phase_grad<float><<<grid_for(nn), 256, 0, s>>>((const cx<float>*)field, (const cx<float>*)grad, phase, phase_stride_b, grad_phase, nn, batch, sum);
  else
    k_phase_grad<double><<<grid_for(nn), 256, 0, s>>>((const cx<double>*)field, (const cx<double>*)grad, phase, phase_stride_b, grad_phase, nn, batch, sum);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "phase gradient kernel launch");
}

int lp2d_intensity(int64_t count, int dtype, const void* field, void* out, void* stream) {
  if (count < 0) return fail(LP2D_E_SIZE, "negative count");
  if (!field || !out) return fail(LP2D_E_ARG, "null pointer");
  if (count == 0) return LP2D_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == LP2D_C64)
    k_intensity<float><<<grid_for(count), 256, 0, s>>>((const cx<float>*)field, (float*)out, count);
  else if (dtype == LP2D_C128)
    k_intensity<double><<<grid_for(count), 256, 0, s>>>((const cx<double>*)field, (double*)out, count);
  else
    return fail(LP2D_E_DTYPE, "unknown dtype %d", dtype);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "intensity kernel launch");
}

int lp2d_propagate(int n, int batch, int dtype, const void* field_in, void* field_out,
                   const float* phase, int64_t phase_stride_b, const lp2d_optics* optics,
                   void* spectrum_out, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!field_in || !field_out || !optics) return fail(LP2D_E_ARG, "null pointer");
  if (int rc = check_optics(optics)) return rc;
  if (spectrum_out && (spectrum_out == field_out || spectrum_out == field_in))
    return fail(LP2D_E_ARG, "spectrum_out must not alias the fields");
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_propagate<float>(n, lg, batch, field_in, field_out, phase, phase_stride_b, optics, spectrum_out, s)
             : op_propagate<double>(n, lg, batch, field_in, field_out, phase, phase_stride_b, optics, spectrum_out, s);
}

int lp2d_propagate_adjoint(int n, int batch, int dtype, const void* grad_out, const void* field_in,
                           const float* phase, int64_t phase_stride_b, const lp2d_optics* optics,
                           void* grad_in, float* grad_phase, int adj_flags, void* stream) {
  return lp2d_propagate_adjoint_ws(n, batch, dtype, grad_out, field_in, phase, phase_stride_b,
                                   optics, grad_in, grad_phase, adj_flags, nullptr, stream);
}

int lp2d_propagate_adjoint_ws(int n, int batch, int dtype, const void* grad_out,
                              const void* field_in, const float* phase, int64_t phase_stride_b,
                              const lp2d_optics* optics, void* grad_in, float* grad_phase,
                              int adj_flags, void* work, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!grad_out || !grad_in || !optics) return fail(LP2D_E_ARG, "null pointer");
  if (int rc = check_optics(optics)) return rc;
  if (grad_phase && !phase) return fail(LP2D_E_ARG, "grad_phase requires phase");
  if ((phase || optics->lens_f != 0.0) && !field_in)
    return fail(LP2D_E_ARG, "field_in is required with a phase screen or lens");
  if (field_in && grad_in == field_in) return fail(LP2D_E_ARG, "grad_in must not alias field_in");
  if (adj_flags & ~LP2D_ADJ_NO_GRAD_IN) return fail(LP2D_E_ARG, "unknown adj_flags bits");
  if (work && (work == grad_in || work == grad_out || work == field_in))
    return fail(LP2D_E_ARG, "work must not alias grad_out, field_in or grad_in");
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_adjoint<float>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, work, s)
             : op_adjoint<double>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, work, s);
}

int lp2d_propagate_stack(int n, int batch, int dtype, const void* field_in, void* field_out,
                         const float* screens, int nscreens, const lp2d_optics* optics,
                         void* const* planes_out, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!field_in || !field_out || !optics || !screens) return fail(LP2D_E_ARG, "null pointer");
  if (nscreens < 1) return fail(LP2D_E_ARG, "nscreens must be >= 1");
  if (int rc = check_optics(optics)) return rc;
  if (planes_out)
    for (int k = 1; k < nscreens; ++k)
      if (planes_out[k] && (planes_out[k] == field_out || planes_out[k] == field_in))
        return fail(LP2D_E_ARG, "planes_out[%d] must not alias the fields", k);
  ensure_registry();
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_stack<float>(n, lg, batch, field_in, field_out, screens, nscreens, optics, planes_out, s)
             : op_stack<double>(n, lg, batch, field_in, field_out, screens, nscreens, optics, planes_out, s);
}

int lp2d_propagate_adjoint_noalloc(int n, int batch, int dtype, const void* grad_out,
                                   const void* field_in, const float* phase,
                                   int64_t phase_stride_b, const lp2d_optics* optics,
                                   void* grad_in, float* grad_phase, int adj_flags,
                                   void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!grad_out || !grad_in || !optics) return fail(LP2D_E_ARG, "null pointer");
  if (int rc = check_optics(optics)) return rc;
  if (grad_phase && !phase) return fail(LP2D_E_ARG, "grad_phase requires phase");
  if ((phase || optics->lens_f != 0.0) && !field_in)
    return fail(LP2D_E_ARG, "field_in is required with a phase screen or lens");
  if (field_in && grad_in == field_in) return fail(LP2D_E_ARG, "grad_in must not alias field_in");
  if (adj_flags & ~LP2D_ADJ_NO_GRAD_IN) return fail(LP2D_E_ARG, "unknown adj_flags bits");
  ensure_registry();
  // scratch needed: the per-field planes of the whole-field kernel, or the per-group
  // planes of the gradient pass, of a shared screen over a batch
  size_t need = 0;
  if (grad_phase && phase_stride_b == 0 && batch > 1) {
    const size_t plane = (size_t)n * n * sizeof(float);
    const bool small = dtype == LP2D_C64 ? use_small<float>(lg) : use_small<double>(lg);
    const int groups = dtype == LP2D_C64 ? grad_groups<float>(lg, batch) : grad_groups<double>(lg, batch);
    need = small ? (size_t)batch * plane : (groups > 1 ? (size_t)groups * plane : 0);
  }
  if (need > kScratchBytes)
    return fail(LP2D_E_SIZE,
                "n=%d batch=%d: the shared-screen gradient needs %zu bytes of scratch, the "
                "allocation-free adjoint holds %zu (use lp2d_propagate_adjoint)",
                n, batch, need, kScratchBytes);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices || !g_dev[dev].ready || !g_dev[dev].scratch)
    return fail(LP2D_E_INIT, "call lp2d_init(%d) before lp2d_propagate_adjoint_noalloc", dev);
  float* sp = g_dev[dev].scratch;
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_adjoint<float>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, nullptr, s, sp)
             : op_adjoint<double>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, nullptr, s, sp);
}

int64_t lp2d_propagate_stack_adjoint_ws_size(int n, int batch, int dtype) {
  int lg;
  if (check_n(n, batch, dtype, &lg)) return -1;
  ensure_registry();
  return (int64_t)(dtype == LP2D_C64 ? stack_adjoint_ws<float>(lg, batch)
                                     : stack_adjoint_ws<double>(lg, batch));
}

int lp2d_propagate_stack_adjoint(int n, int batch, int dtype, const void* grad_out,
                                 const void* const* planes_in, const float* screens,
                                 int nscreens, const lp2d_optics* optics, void* grad_in,
                                 float* grad_screens, int adj_flags, void* work,
                                 int64_t work_bytes, void* stream) {
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  if (!grad_out || !grad_in || !optics || !screens) return fail(LP2D_E_ARG, "null pointer");
  if (nscreens < 1) return fail(LP2D_E_ARG, "nscreens must be >= 1");
  if (int rc = check_optics(optics)) return rc;
  if (adj_flags & ~LP2D_ADJ_NO_GRAD_IN) return fail(LP2D_E_ARG, "unknown adj_flags bits");
  if (!planes_in) return fail(LP2D_E_ARG, "null planes_in");
  for (int k = 0; k < nscreens; ++k) {
    if (!planes_in[k]) return fail(LP2D_E_ARG, "planes_in[%d] is null", k);
    if (planes_in[k] == grad_in) return fail(LP2D_E_ARG, "grad_in must not alias planes_in[%d]", k);
  }
  ensure_registry();
  const size_t need = dtype == LP2D_C64 ? stack_adjoint_ws<float>(lg, batch)
                                        : stack_adjoint_ws<double>(lg, batch);
  if (grad_screens && need > 0) {
    if (!work || work_bytes < (int64_t)need)
      return fail(LP2D_E_SIZE, "n=%d batch=%d: work needs %zu bytes, got %lld", n, batch, need,
                  (long long)work_bytes);
    if (work == grad_in || work == grad_out || work == grad_screens || work == screens)
      return fail(LP2D_E_ARG, "work must not alias grad_out, grad_in, screens or grad_screens");
    for (int k = 0; k < nscreens; ++k)
      if (work == planes_in[k]) return fail(LP2D_E_ARG, "work must not alias planes_in[%d]", k);
  }
  hipStream_t s = (hipStream_t)stream;
  float* ws = grad_screens && need > 0 ? (float*)work : nullptr;
  return dtype == LP2D_C64
             ? op_stack_adjoint<float>(n, lg, batch, grad_out, planes_in, screens, nscreens, optics, grad_in, grad_screens, adj_flags, ws, s)
             : op_stack_adjoint<double>(n, lg, batch, grad_out, planes_in, screens, nscreens, optics, grad_in, grad_screens, adj_flags, ws, s);
}

int lp2d_plan_create(int n, int batch, int dtype, lp2d_plan** out) {
  if (!out) return fail(LP2D_E_ARG, "null output handle");
  *out = nullptr;
  int lg;
  if (int rc = check_n(n, batch, dtype, &lg)) return rc;
  ensure_registry();
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (int rc = init_device(dev)) return rc;
  lp2d_plan* p = new (std::nothrow) lp2d_plan();
  if (!p) return fail(LP2D_E_ARG, "out of host memory");
  p->n = n;
  p->lg = lg;
  p->batch = batch;
  p->dtype = dtype;
  p->device = dev;
  p->sel.dt = dtype;
  p->sel.lg = lg;
  for (int k = 0; k < 3; ++k) p->sel.sel[k] = g_sel[dtype][k][lg];
  p->sel.chunk = g_chunk[dtype][lg];
  *out = p;
  return LP2D_OK;
}

int lp2d_plan_destroy(lp2d_plan* plan) {
  delete plan;
  return LP2D_OK;
}

int64_t lp2d_plan_workspace_size(const lp2d_plan* p, int op) {
  if (!p) return -1;
  size_t a = 0, b = 0;
  switch (op) {
    case LP2D_OP_FFT2:
    case LP2D_OP_PROPAGATE:
    case LP2D_OP_STACK:
      return 0;
    case LP2D_OP_ADJOINT:
      if (p->dtype == LP2D_C64)
        adjoint_ws_parts<float>(p->n, p->lg, p->batch, &a, &b);
      else
        adjoint_ws_parts<double>(p->n, p->lg, p->batch, &a, &b);
      return (int64_t)(align256(a) + b);
    case LP2D_OP_STACK_ADJOINT:
      return (int64_t)(p->dtype == LP2D_C64 ? stack_adjoint_ws<float>(p->lg, p->batch)
                                            : stack_adjoint_ws<double>(p->lg, p->batch));
    default:
      return -1;
  }
}

int lp2d_plan_fft2(const lp2d_plan* p, const void* in, void* out, int inverse, void* stream) {
  if (int rc = plan_enter(p)) return rc;
  PlanScope ps(&p->sel);
  return lp2d_fft2(p->n, p->batch, p->dtype, in, out, inverse, stream);
}

int lp2d_plan_propagate(const lp2d_plan* p, const void* field_in, void* field_out,
                        const float* phase, int64_t phase_stride_b, const lp2d_optics* optics,
                        void* spectrum_out, void* stream) {
  if (int rc = plan_enter(p)) return rc;
  PlanScope ps(&p->sel);
  return lp2d_propagate(p->n, p->batch, p->dtype, field_in, field_out, phase, phase_stride_b,
                        optics, spectrum_out, stream);
}

int lp2d_plan_propagate_adjoint(const lp2d_plan* p, const void* grad_out, const void* field_in,
                                const float* phase, int64_t phase_stride_b,
                                const lp2d_optics* optics, void* grad_in, float* grad_phase,
                                int adj_flags, void* work, int64_t work_bytes, void* stream) {
  if (int rc = plan_enter(p)) return rc;
  const int n = p->n, batch = p->batch, dtype = p->dtype, lg = p->lg;
  if (!grad_out || !grad_in || !optics) return fail(LP2D_E_ARG, "null pointer");
  if (int rc = check_optics(optics)) return rc;
  if (grad_phase && !phase) return fail(LP2D_E_ARG, "grad_phase requires phase");
  if ((phase || optics->lens_f != 0.0) && !field_in)
    return fail(LP2D_E_ARG, "field_in is required with a phase screen or lens");
  if (field_in && grad_in == field_in) return fail(LP2D_E_ARG, "grad_in must not alias field_in");
  if (adj_flags & ~LP2D_ADJ_NO_GRAD_IN) return fail(LP2D_E_ARG, "unknown adj_flags bits");
  size_t plane_b = 0, grad_b = 0;
  if (dtype == LP2D_C64)
    adjoint_ws_parts<float>(n, lg, batch, &plane_b, &grad_b);
  else
    adjoint_ws_parts<double>(n, lg, batch, &plane_b, &grad_b);
  const size_t need = align256(plane_b) + grad_b;
  if (need > 0 && (!work || work_bytes < (int64_t)need))
    return fail(LP2D_E_SIZE, "n=%d batch=%d: work needs %zu bytes, got %lld", n, batch, need,
                (long long)work_bytes);
  if (work && (work == grad_in || work == grad_out || work == field_in ||
               (grad_phase && work == (void*)grad_phase)))
    return fail(LP2D_E_ARG, "work must not alias the other arrays");
  char* w = (char*)work;
  void* plane = plane_b ? (void*)w : nullptr;
  float* gplanes = grad_b ? (float*)(w + align256(plane_b)) : nullptr;
  PlanScope ps(&p->sel);
  hipStream_t s = (hipStream_t)stream;
  return dtype == LP2D_C64
             ? op_adjoint<float>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, plane, s, gplanes)
             : op_adjoint<double>(n, lg, batch, grad_out, field_in, phase, phase_stride_b, optics, grad_in, grad_phase, adj_flags, plane, s, gplanes);
}

int lp2d_plan_propagate_stack(const lp2d_plan* p, const void* field_in, void* field_out,
                              const float* screens, int nscreens, const lp2d_optics* optics,
                              void* const* planes_out, void* stream) {
  if (int rc = plan_enter(p)) return rc;
  PlanScope ps(&p->sel);
  return lp2d_propagate_stack(p->n, p->batch, p->dtype, field_in, field_out, screens, nscreens,
                              optics, planes_out, stream);
}

int lp2d_plan_propagate_stack_adjoint(const lp2d_plan* p, const void* grad_out,
                                      const void* const* planes_in, const float* screens,
                                      int nscreens, const lp2d_optics* optics, void* grad_in,
                                      float* grad_screens, int adj_flags, void* work,
                                      int64_t work_bytes, void* stream) {
  if (int rc = plan_enter(p)) return rc;
  PlanScope ps(&p->sel);
  return lp2d_propagate_stack_adjoint(p->n, p->batch, p->dtype, grad_out, planes_in, screens,
                                      nscreens, optics, grad_in, grad_screens, adj_flags, work,
                                      work_bytes, stream);
}

int lp2d_batch_create(const lp2d_field_desc* descs, int count, int dtype, lp2d_batch** out) {
  if (!out) return fail(LP2D_E_ARG, "null output handle");
  *out = nullptr;
  if (!descs || count < 1) return fail(LP2D_E_ARG, "need at least one field descriptor");
  if (dtype != LP2D_C64 && dtype != LP2D_C128) return fail(LP2D_E_DTYPE, "unknown dtype %d", dtype);
  std::vector<std::vector<int>> by_lg(kLogMax + 1);
  for (int i = 0; i < count; ++i) {
    const lp2d_field_desc& d = descs[i];
    const int lg = log2_exact(d.n);
    if (lg < 1 || lg > kLogMax)
      return fail(LP2D_E_SIZE, "field %d: n=%d must be a power of two in [2, %d]", i, d.n, kMaxN);
    if (!d.src || !d.dst) return fail(LP2D_E_ARG, "field %d: null pointer", i);
    if (!(d.area_size > 0) || !(d.wl > 0)) return fail(LP2D_E_ARG, "field %d: area_size and wl must be > 0", i);
    by_lg[lg].push_back(i);
  }
  DeviceTables* t;
  if (int rc = current_tables(&t, nullptr)) return rc;
  lp2d_batch* b = new lp2d_batch();
  hipGetDevice(&b->device);
  b->dtype = dtype;
  // one allocation: per bucket [src ptrs | dst ptrs | optics]
  size_t bytes = 0;
  for (int lg = kLogMax; lg >= 1; --lg)
    if (!by_lg[lg].empty()) bytes += by_lg[lg].size() * (2 * sizeof(void*) + sizeof(FieldOpt));
  std::vector<unsigned char> host(bytes);
  size_t off = 0;
  for (int lg = kLogMax; lg >= 1; --lg) {  // largest fields first
    const std::vector<int>& ids = by_lg[lg];
    if (ids.empty()) continue;
    Bucket bk;
    bk.n = 1 << lg;
    bk.lg = lg;
    bk.count = (int)ids.size();
    bool any_ap = false;
    void** sp = (void**)(host.data() + off);
    void** dp = sp + ids.size();
    FieldOpt* op = (FieldOpt*)(dp + ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
      const lp2d_field_desc& d = descs[ids[k]];
      const double dL = d.area_size / d.n, k0 = 1.0 / d.wl;
      sp[k] = const_cast<void*>(d.src);
      dp[k] = d.dst;
      op[k].k0sq = k0 * k0;
      op[k].inv_ndL = 1.0 / ((double)d.n * dL);
      op[k].z = d.z;
      op[k].ap_r2 = d.aperture_radius > 0.0 ? d.aperture_radius * d.aperture_radius : HUGE_VAL;
      any_ap = any_ap || d.aperture_radius > 0.0;
    }
    bk.flags = F_H | (any_ap ? F_APERTURE : 0);
    bk.d_src = (void*)(off);
    bk.d_dst = (void*)(off + ids.size() * sizeof(void*));
    bk.d_opt = (FieldOpt*)(off + 2 * ids.size() * sizeof(void*));
    off += ids.size() * (2 * sizeof(void*) + sizeof(FieldOpt));
    b->buckets.push_back(bk);
  }
  hipError_t e = hipMalloc(&b->d_mem, bytes);
  if (e == hipSuccess) e = hipMemcpy(b->d_mem, host.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(b->d_mem);
    delete b;
    return hip_fail(e, "batch table upload");
  }
  for (Bucket& bk : b->buckets) {  // offsets -> device addresses
    bk.d_src = (char*)b->d_mem + (size_t)bk.d_src;
    bk.d_dst = (char*)b->d_mem + (size_t)bk.d_dst;
    bk.d_opt = (FieldOpt*)((char*)b->d_mem + (size_t)bk.d_opt);
  }
  *out = b;
  return LP2D_OK;
}

int lp2d_batch_run(lp2d_batch* b, void* stream) {
  if (!b) return fail(LP2D_E_ARG, "null batch");
  int dev = 0;
  hipGetDevice(&dev);
  if (dev != b->device) return fail(LP2D_E_ARG, "batch was created on device %d, current is %d", b->device, dev);
  hipStream_t s = (hipStream_t)stream;
  for (const Bucket& bk : b->buckets) {
    const int rc = b->dtype == LP2D_C64 ? run_bucket<float>(bk, s) : run_bucket<double>(bk, s);
    if (rc) return rc;
  }
  return LP2D_OK;
}

int lp2d_batch_destroy(lp2d_batch* b) {
  if (!b) return LP2D_OK;
  hipFree(b->d_mem);
  delete b;
  return LP2D_OK;
}

int lp2d_sum_planes(int64_t count, int batch, const float* in, float* out, void* stream) {
  if (count < 0 || batch < 1) return fail(LP2D_E_SIZE, "bad count or batch");
  if (!in || !out) return fail(LP2D_E_ARG, "null pointer");
  if (count == 0) return LP2D_OK;
  k_sum_planes<<<grid_for(count), 256, 0, (hipStream_t)stream>>>(in, out, count, batch);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LP2D_OK : hip_fail(e, "sum kernel launch");
}

int lp2d_variant_count(int n, int dtype, int kind) {
  ensure_registry();
  const int lg = log2_exact(n);
  if (lg < 1 || lg > kLogMax || (dtype != 0 && dtype != 1) || kind < 0 || kind > 3) return 0;
  if (kind == 2) return dtype == 0 ? (int)g_rf->has_small[lg] : (int)g_rd->has_small[lg];
  if (kind == 3) return dtype == 0 ? g_rf->nchain[lg] : g_rd->nchain[lg];
  if (dtype == 0) return kind == 0 ? g_rf->nrows[lg] : g_rf->ncols[lg];
  return kind == 0 ? g_rd->nrows[lg] : g_rd->ncols[lg];
}

int lp2d_set_variant(int n, int dtype, int kind, int variant) {
  if (kind != 0 && kind != 1 && kind != 3)
    return fail(LP2D_E_ARG, "kind must be 0 (rows), 1 (columns) or 3 (chained rows)");
  const int cnt = lp2d_variant_count(n, dtype, kind);
  if (cnt == 0) return fail(LP2D_E_SIZE, "no variants for n=%d dtype=%d kind=%d", n, dtype, kind);
  if (variant < 0 || variant >= cnt) return fail(LP2D_E_ARG, "variant %d out of range [0,%d)", variant, cnt);
  g_sel[dtype][kind == 3 ? 2 : kind][log2_exact(n)] = variant;
  return LP2D_OK;
}

int lp2d_set_chunk(int n, int dtype, int fields) {
  ensure_registry();
  const int lg = log2_exact(n);
  if (lg < 1 || lg > kLogMax || (dtype != 0 && dtype != 1)) return fail(LP2D_E_SIZE, "bad n or dtype");
  if (fields < 0) return fail(LP2D_E_ARG, "fields must be >= 0");
  g_chunk[dtype][lg] = fields;
  return LP2D_OK;
}

int lp2d_get_chunk(int n, int dtype) {
  ensure_registry();
  const int lg = log2_exact(n);
  if (lg < 1 || lg > kLogMax || (dtype != 0 && dtype != 1)) return fail(LP2D_E_SIZE, "bad n or dtype");
  return g_chunk[dtype][lg];
}

int lp2d_get_variant(int n, int dtype, int kind) {
  if ((kind != 0 && kind != 1 && kind != 3) || lp2d_variant_count(n, dtype, kind) == 0)
    return fail(LP2D_E_SIZE, "no variants");
  return g_sel[dtype][kind == 3 ? 2 : kind][log2_exact(n)];
}

int lp2d_variant_name(int n, int dtype, int kind, int variant, char* buf, int buflen) {
  const int cnt = lp2d_variant_count(n, dtype, kind);
  if (kind == 2 || variant < 0 || variant >= cnt || !buf || buflen < 1)
    return fail(LP2D_E_ARG, "bad variant query");
  const int lg = log2_exact(n);
  if (kind == 3) {
    const Registry<float>* rf = g_rf;
    const Registry<double>* rd = g_rd;
    const int E = dtype ? rd->chain_E[lg][variant] : rf->chain_E[lg][variant];
    const int L = dtype ? rd->chain_lines[lg][variant] : rf->chain_lines[lg][variant];
    const int tl = dtype ? rd->chain_twl[lg][variant] : rf->chain_twl[lg][variant];
    return snprintf(buf, buflen, "chain E=%d lines=%d threads=%d%s", E, L, (n / E) * L,
                    tl ? " lds-twiddles" : "");
  }
  if (kind == 0) {
    const int E = dtype ? g_rd->rows[lg][variant].E : g_rf->rows[lg][variant].E;
    const int L = dtype ? g_rd->rows[lg][variant].lines : g_rf->rows[lg][variant].lines;
    const int th = dtype ? g_rd->rows[lg][variant].threads : g_rf->rows[lg][variant].threads;
    const int lds = dtype ? g_rd->rows[lg][variant].lds : g_rf->rows[lg][variant].lds;
    const int rk = dtype ? g_rd->rows[lg][variant].kind : g_rf->rows[lg][variant].kind;
    return snprintf(buf, buflen, "rows E=%d lines=%d threads=%d lds=%d%s", E, L, th, lds,
                    (rk & 2) ? " pipelined" : (rk & 4) ? " lds-twiddles" : "");
  }
  const int E = dtype ? g_rd->cols[lg][variant].E : g_rf->cols[lg][variant].E;
  const int W = dtype ? g_rd->cols[lg][variant].W : g_rf->cols[lg][variant].W;
  const int th = dtype ? g_rd->cols[lg][variant].threads : g_rf->cols[lg][variant].threads;
  const int lds = dtype ? g_rd->cols[lg][variant].lds : g_rf->cols[lg][variant].lds;
  const int ck = dtype ? g_rd->cols[lg][variant].kind : g_rf->cols[lg][variant].kind;
  return snprintf(buf, buflen, "cols E=%d W=%d threads=%d lds=%d%s%s", E, W, th, lds,
                  (ck & 2) ? " pipelined" : "", (ck & 1) ? " shared-H" : "");
}

}  // extern "C"

extern "C" {

int lp2d_trace_enable(int on) {
  std::lock_guard<std::mutex> lk(g_trace_mu);
  g_trace_on = on != 0;
  return LP2D_OK;
}

int lp2d_trace_collect(lp2d_trace_stat* stats, int max_stats, int* n_stats) {
  if (!n_stats || (max_stats > 0 && !stats)) return fail(LP2D_E_ARG, "null output");
  std::lock_guard<std::mutex> lk(g_trace_mu);
  std::vector<lp2d_trace_stat> agg;
  int rc = LP2D_OK;
  for (TraceRec& r : g_trace) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(r.stop);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.start, r.stop);
    if (e != hipSuccess) rc = hip_fail(e, "trace event");
    size_t k = 0;
    for (; k < agg.size(); ++k)
      if (r.name == agg[k].name) break;
    if (k == agg.size()) {
      lp2d_trace_stat st;
      std::memset(&st, 0, sizeof(st));
      std::snprintf(st.name, sizeof(st.name), "%s", r.name.c_str());
      agg.push_back(st);
    }
    agg[k].launches += 1;
    agg[k].total_ms += ms;
    g_event_pool.push_back(r.start);
    g_event_pool.push_back(r.stop);
  }
  g_trace.clear();
  *n_stats = (int)agg.size();
  for (int i = 0; i < (int)agg.size() && i < max_stats; ++i) stats[i] = agg[i];
  return rc;
}

}  // extern "C"

namespace lp2d {
// Status + message for entry points defined in other translation units.
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace lp2d
