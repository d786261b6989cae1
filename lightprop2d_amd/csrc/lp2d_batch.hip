// Uneven batches: many independent fields of different sizes, distances and optics
// in one launch per sweep.
//
// Work is cut into tiles of ~16K complex elements each (a whole field for n <= 128,
// a row tile or a column strip for larger n), all handled by 1024-thread workgroups
// with one common LDS footprint.  Three persistent launches run the three sweeps:
//   sweep 1: whole-field propagations (n <= 128) + forward row tiles (n >= 256)
//   sweep 2: column strips (forward FFT, H, aperture, inverse FFT)
//   sweep 3: inverse row tiles (1/n^2 folded in)
// Each launch has a fixed grid (one workgroup per CU by default) that strides over
// its tile list, ordered largest field first.  lp2d_batch_run allocates nothing, so a
// caller can capture it in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lp2d.h"
#include "lp2d_kernels.hpp"

using namespace lp2d;

namespace {

constexpr int kThreads = 1024;
constexpr int kEb = 8;  // elements per thread in the tile bodies

// n -> kind index; kinds are (sweep, n) pairs handled by the switch below.
enum Kind : int {
  K_SMALL64 = 0,
  K_SMALL128,
  K_ROWF256, K_ROWF512, K_ROWF1024, K_ROWF2048, K_ROWF4096,
  K_COL256, K_COL512, K_COL1024, K_COL2048, K_COL4096,
  K_ROWI256, K_ROWI512, K_ROWI1024, K_ROWI2048, K_ROWI4096,
};

struct FieldDev {
  const cx<float>* src;
  cx<float>* dst;
  double k0sq, inv_ndL, z, ap_r2;
  int n;
  int pad_;
};

struct Item {
  int field;
  int kind;
  int tile;
  int pad_;
};

template <int N>
struct TileShape {
  static constexpr int TL = N / kEb;
  static constexpr int LINES = kThreads / TL;  // rows per row tile / columns per strip
  static constexpr int PS = 3;
  static constexpr int LDS = LINES * PaddedLen<N, PS>::value * (int)sizeof(cx<float>);
};
constexpr int kSmall128Lds = 128 * SmallPitch<128, 4>::value * (int)sizeof(cx<float>);
constexpr int kSmall64Lds = 64 * SmallPitch<64, 2>::value * (int)sizeof(cx<float>);
constexpr int cmaxi(int a, int b) { return a > b ? a : b; }
constexpr int kLdsBytes = cmaxi(kSmall128Lds, cmaxi(kSmall64Lds, TileShape<256>::LDS));

struct BatchArgs {
  const FieldDev* fields;
  const Item* items;
  const cx<float>* tw;
  int count;
};

__device__ __forceinline__ PassArgs<float> field_args(const FieldDev& f, const cx<float>* tw) {
  PassArgs<float> a;
  memset(&a, 0, sizeof(a));
  a.tw = tw;
  a.tw_inv = tw + kTwiddleLen;
  a.src = f.src;
  a.dst = f.dst;
  a.fstride = 0;
  a.batch = 1;
  a.k0sq = f.k0sq;
  a.inv_ndL = f.inv_ndL;
  a.z = f.z;
  a.ap_r2 = f.ap_r2;
  a.scale = 1.0f;
  return a;
}

template <int N>
__device__ __forceinline__ void row_tile(PassArgs<float>& a, int tile, bool inv, cx<float>* lds) {
  using S = TileShape<N>;
  if (inv) {
    a.src = a.dst;
    a.flags = F_SCALE;
    a.scale = 1.0f / ((float)N * (float)N);
    rows_body<float, N, kEb, S::LINES, true, S::PS, false>(a, 0, tile * S::LINES, S::LINES,
                                                          lds, threadIdx.x);
  } else {
    a.flags = 0;
    rows_body<float, N, kEb, S::LINES, false, S::PS, false>(a, 0, tile * S::LINES, S::LINES,
                                                           lds, threadIdx.x);
  }
}

template <int N>
__device__ __forceinline__ void col_tile(PassArgs<float>& a, int tile, int mid,
                                         cx<float>* lds) {
  using S = TileShape<N>;
  a.src = a.dst;
  a.flags = mid;
  cols_body<float, N, kEb, S::LINES, MODE_FWD | MODE_INV, S::PS, 1>(a, 0, 1, tile * S::LINES,
                                                                   lds, threadIdx.x);
}

__global__ __launch_bounds__(kThreads) void k_batch_sweep(BatchArgs ba) {
  __shared__ cx<float> lds[kLdsBytes / sizeof(cx<float>)];
  for (int i = blockIdx.x; i < ba.count; i += gridDim.x) {
    const Item it = ba.items[i];
    const FieldDev f = ba.fields[it.field];
    PassArgs<float> a = field_args(f, ba.tw);
    const int mid = (f.z != 0.0 ? F_H : 0) | (f.ap_r2 > 0.0 ? F_APERTURE : 0);
    switch (it.kind) {
      case K_SMALL64:
        a.flags = mid | F_SCALE;
        a.scale = 1.0f / (64.0f * 64.0f);
        small_body<float, 64, 4, 2>(a, 0, MODE_FWD | MODE_INV, lds, threadIdx.x);
        break;
      case K_SMALL128:
        a.flags = mid | F_SCALE;
        a.scale = 1.0f / (128.0f * 128.0f);
        small_body<float, 128, 16, 4>(a, 0, MODE_FWD | MODE_INV, lds, threadIdx.x);
        break;
      case K_ROWF256: row_tile<256>(a, it.tile, false, lds); break;
      case K_ROWF512: row_tile<512>(a, it.tile, false, lds); break;
      case K_ROWF1024: row_tile<1024>(a, it.tile, false, lds); break;
      case K_ROWF2048: row_tile<2048>(a, it.tile, false, lds); break;
      case K_ROWF4096: row_tile<4096>(a, it.tile, false, lds); break;
      case K_COL256: col_tile<256>(a, it.tile, mid, lds); break;
      case K_COL512: col_tile<512>(a, it.tile, mid, lds); break;
      case K_COL1024: col_tile<1024>(a, it.tile, mid, lds); break;
      case K_COL2048: col_tile<2048>(a, it.tile, mid, lds); break;
      case K_COL4096: col_tile<4096>(a, it.tile, mid, lds); break;
      case K_ROWI256: row_tile<256>(a, it.tile, true, lds); break;
      case K_ROWI512: row_tile<512>(a, it.tile, true, lds); break;
      case K_ROWI1024: row_tile<1024>(a, it.tile, true, lds); break;
      case K_ROWI2048: row_tile<2048>(a, it.tile, true, lds); break;
      case K_ROWI4096: row_tile<4096>(a, it.tile, true, lds); break;
      default: break;
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

int lg_of(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

int tiles_per_sweep(int n) { return n * n / (kThreads * kEb); }  // 8192 elements per tile

thread_local std::string g_batch_err;

}  // namespace

struct lp2d_batch {
  int device = 0;
  FieldDev* d_fields = nullptr;
  Item* d_items = nullptr;
  cx<float>* tw = nullptr;
  int count[3] = {0, 0, 0};
  int offset[3] = {0, 0, 0};
  int grid = 0;
  int default_grid = 0;
};

extern "C" {

int lp2d_twiddles_f32(const void** out);  // lp2d_api.hip
void* lp2d_trace_begin(void* stream, const char* name);
void lp2d_trace_end(void* token);

int lp2d_batch_create(const lp2d_field_desc* descs, int count, int dtype, lp2d_batch** out) {
  if (!out || !descs || count < 1) return LP2D_E_ARG;
  *out = nullptr;
  if (dtype != LP2D_C64) return LP2D_E_DTYPE;
  std::vector<FieldDev> fields(count);
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) {
    const lp2d_field_desc& d = descs[i];
    const int lg = lg_of(d.n);
    if (lg < 6 || lg > 12) return LP2D_E_SIZE;  // 64 .. 4096
    if (!d.src || !d.dst || !(d.area_size > 0) || !(d.wl > 0)) return LP2D_E_ARG;
    const double dL = d.area_size / d.n, k0 = 1.0 / d.wl;
    fields[i] = FieldDev{(const cx<float>*)d.src, (cx<float>*)d.dst, k0 * k0,
                         1.0 / ((double)d.n * dL), d.z, d.aperture_radius * d.aperture_radius,
                         d.n, 0};
    order[i] = i;
  }
  // largest fields first, so the longest tiles start early in every sweep
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return fields[a].n > fields[b].n; });
  std::vector<Item> sweep[3];
  for (int i : order) {
    const int n = fields[i].n;
    const int lg = lg_of(n);
    if (n <= 128) {
      sweep[0].push_back(Item{i, n == 64 ? K_SMALL64 : K_SMALL128, 0, 0});
      continue;
    }
    const int k = lg - 8;  // 256 -> 0
    const int tiles = tiles_per_sweep(n);
    for (int t = 0; t < tiles; ++t) sweep[0].push_back(Item{i, K_ROWF256 + k, t, 0});
    for (int t = 0; t < tiles; ++t) sweep[1].push_back(Item{i, K_COL256 + k, t, 0});
    for (int t = 0; t < tiles; ++t) sweep[2].push_back(Item{i, K_ROWI256 + k, t, 0});
  }
  lp2d_batch* b = new lp2d_batch();
  hipGetDevice(&b->device);
  std::vector<Item> all;
  for (int s = 0; s < 3; ++s) {
    b->offset[s] = (int)all.size();
    b->count[s] = (int)sweep[s].size();
    all.insert(all.end(), sweep[s].begin(), sweep[s].end());
  }
  const void* tw = nullptr;
  int rc = lp2d_twiddles_f32(&tw);
  if (rc != LP2D_OK) {
    delete b;
    return rc;
  }
  b->tw = (cx<float>*)tw;
  if (hipMalloc(&b->d_fields, sizeof(FieldDev) * fields.size()) != hipSuccess ||
      hipMalloc(&b->d_items, sizeof(Item) * std::max<size_t>(1, all.size())) != hipSuccess ||
      hipMemcpy(b->d_fields, fields.data(), sizeof(FieldDev) * fields.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_items, all.data(), sizeof(Item) * all.size(), hipMemcpyHostToDevice) !=
          hipSuccess) {
    hipFree(b->d_fields);
    hipFree(b->d_items);
    delete b;
    return LP2D_E_HIP;
  }
  int cus = 0, per_cu = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_sweep, kThreads, 0);
  b->default_grid = std::max(1, cus * std::max(1, per_cu));
  b->grid = b->default_grid;
  *out = b;
  return LP2D_OK;
}

int lp2d_batch_set_grid(lp2d_batch* b, int workgroups) {
  if (!b || workgroups < 0) return LP2D_E_ARG;
  b->grid = workgroups == 0 ? b->default_grid : workgroups;
  return LP2D_OK;
}

int lp2d_batch_run(lp2d_batch* b, void* stream) {
  if (!b) return LP2D_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int sw = 0; sw < 3; ++sw) {
    if (b->count[sw] == 0) continue;
    BatchArgs ba{b->d_fields, b->d_items + b->offset[sw], b->tw, b->count[sw]};
    const int grid = std::min(b->grid, b->count[sw]);
    static const char* names[3] = {"batch_sweep1", "batch_sweep2", "batch_sweep3"};
    void* tok = lp2d_trace_begin(stream, names[sw]);
    k_batch_sweep<<<grid, kThreads, 0, s>>>(ba);
    const hipError_t e = hipGetLastError();
    lp2d_trace_end(tok);
    if (e != hipSuccess) return LP2D_E_HIP;
  }
  return LP2D_OK;
}

int lp2d_batch_destroy(lp2d_batch* b) {
  if (!b) return LP2D_OK;
  hipFree(b->d_fields);
  hipFree(b->d_items);
  delete b;
  return LP2D_OK;
}

}  // extern "C"
