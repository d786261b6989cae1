// Kernel templates, their launchers and the per-size variant lists.  Included by the
// per-precision translation units (lp2d_f32_*.hip, lp2d_f64_*.hip) so the large set
// of instantiations compiles in parallel.
#pragma once
#include <atomic>

#include "lp2d_registry.hpp"

namespace lp2d {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int floor_pow2(int n) { return n < 2 ? 1 : 2 * floor_pow2(n >> 1); }
constexpr int kLdsBudget = 160 * 1024 - 2048;

// ----------------------------------------------------------------- kernels
// Every pass kernel asks for at least two waves per SIMD: the compiler then budgets at
// most 256 registers per lane (one more and a 256-thread workgroup drops to a single
// wave per SIMD, halving the latency hiding).
#define LP2D_OCC __attribute__((amdgpu_waves_per_eu(2)))
template <typename T, int N, int E, int LINES, bool INV, int PS, bool FEAT>
__global__ __launch_bounds__((N / E) * LINES) LP2D_OCC void k_rows(PassArgs<T> a) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  constexpr int TPF = (N + LINES - 1) / LINES;  // row tiles per field
  const int tile = blockIdx.x;
  const int b = tile / TPF;
  const int y0 = (tile - b * TPF) * LINES;
  rows_body<T, N, E, LINES, INV, PS, FEAT>(a, b, y0, cmin(LINES, N - y0), lds, threadIdx.x);
}

// The row pass with its twiddles staged once in LDS (compact mode CL), each workgroup
// taking every gridDim.x-th row tile of the batch (one resident round).
template <typename T, int N, int E, int LINES, bool INV, int PS, bool FEAT, int CL,
          bool PHO = false>
__global__ __launch_bounds__((N / E) * LINES) LP2D_OCC void k_rows_l(PassArgs<T> a, int nunits) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  __shared__ cx<T> twl[TwlLen<N, E, CL>::value];
  twl_fill<T, N, E, (N / E) * LINES, 1, CL>(twl, a.tw, threadIdx.x);
  __syncthreads();
  constexpr int TPF = N / LINES;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int b = u / TPF;
    rows_body<T, N, E, LINES, INV, PS, FEAT, CL, PHO>(a, b, (u - b * TPF) * LINES, LINES, lds,
                                                      threadIdx.x, twl);
  }
}

// Pipelined row pass: about one resident round of workgroups, each taking every
// gridDim.x-th row tile of the batch; the twiddle tables are staged in LDS once.
template <typename T, int N, int E, int LINES, bool INV, int PS, bool TILED, bool PH = false,
          bool WSK = true>
__global__ __launch_bounds__((N / E) * LINES) LP2D_OCC void k_rows_pipe(PassArgs<T> a, int nunits) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  __shared__ cx<T> twl[TwlLen<N, E>::value];
  twl_fill<T, N, E, (N / E) * LINES>(twl, a.tw, threadIdx.x);
  __syncthreads();
  rows_pipe_body<T, N, E, LINES, INV, PS, TILED, PH, WSK>(a, blockIdx.x, gridDim.x, nunits, lds,
                                                          twl, threadIdx.x);
}

// The pipelined gradient body when its screen factor is hoisted and the LDS twiddle
// tables fit beside the row tiles.
template <typename T, int N, int E, int LINES, int PS>
struct GradPipe {
  static constexpr bool value =
      LINES * PaddedLen<N, PS>::value * (int)sizeof(cx<T>) +
          TwlLen<N, E>::value * (int)sizeof(cx<T>) <= kLdsBudget;
};

template <typename T, int N, int E, int LINES, int PS, bool HOIST>
__global__ __launch_bounds__((N / E) * LINES) LP2D_OCC void k_rows_grad(PassArgs<T> a, int groups,
                                                                float* partial) {
  constexpr bool PIPE = HOIST && GradPipe<T, N, E, LINES, PS>::value;
  __shared__ cx<T> lds[LINES * PaddedLen<N, PS>::value];
  __shared__ cx<T> twl[PIPE ? TwlLen<N, E>::value : 1];
  constexpr int TPF = (N + LINES - 1) / LINES;
  const int tile = blockIdx.x % TPF, g = blockIdx.x / TPF;
  const int per = (a.batch + groups - 1) / groups;
  const int b0 = g * per, b1 = cmin(a.batch, b0 + per);
  const int y0 = tile * LINES;
  if constexpr (PIPE) {
    twl_fill<T, N, E, (N / E) * LINES>(twl, a.tw, threadIdx.x);
    __syncthreads();
    rows_grad_pipe_body<T, N, E, LINES, PS>(a, b0, b1, y0, cmin(LINES, N - y0),
                                            partial + (long long)g * N * N, lds, twl,
                                            threadIdx.x);
  } else {
    rows_grad_body<T, N, E, LINES, PS, HOIST>(a, b0, b1, y0, cmin(LINES, N - y0),
                                              partial + (long long)g * N * N, lds, threadIdx.x);
  }
}

// Chained-plane row pass: one workgroup per row tile of LINES rows of one field.
// WPE: waves per SIMD the register budget is sized for.
template <typename T, int N, int E, int LINES, int PS, int WPE>
__global__ __launch_bounds__((N / E) * LINES) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_rows_chain(PassArgs<T> a) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  constexpr int TPF = N / LINES;
  const int tile = blockIdx.x;
  const int b = tile / TPF;
  rows_chain_body<T, N, E, LINES, PS>(a, b, (tile - b * TPF) * LINES, lds, nullptr,
                                      threadIdx.x);
}

// The same with the twiddles staged once in LDS (compact last pass when CL), each
// workgroup taking every gridDim.x-th row tile of the batch (one resident round).
template <typename T, int N, int E, int LINES, int PS, int WPE, int CL>
__global__ __launch_bounds__((N / E) * LINES) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_rows_chain_l(PassArgs<T> a, int nunits) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  __shared__ cx<T> twl[TwlLen<N, E, CL>::value];
  twl_fill<T, N, E, (N / E) * LINES, 1, CL>(twl, a.tw, threadIdx.x);
  __syncthreads();
  constexpr int TPF = N / LINES;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int b = u / TPF;
    rows_chain_body<T, N, E, LINES, PS, CL>(a, b, (u - b * TPF) * LINES, lds, twl, threadIdx.x);
  }
}

// Grid: (N/W strips) x (field groups of `per` fields).
template <typename T, int N, int E, int W, int MODE, int PS, int MIDK>
__global__ __launch_bounds__((N / E) * W) LP2D_OCC void k_cols(PassArgs<T> a, int per) {
  __shared__ cx<T> lds[W * PaddedLen<N, PS>::value];
  constexpr int SPF = N / W;  // column strips per field
  const int g = gridDim.x, i = blockIdx.x;
  // Workgroups go to the 8 XCDs round robin.  Give each XCD a contiguous run of
  // strips so that narrow strips sharing 128-byte lines meet in the same L2.
  const int blk = (g % 8 == 0) ? (i % 8) * (g / 8) + i / 8 : i;
  const int grp = blk / SPF;
  const int x0 = (blk - grp * SPF) * W;
  const int b0 = grp * per;
  const int b1 = cmin(a.batch, b0 + per);
  cols_body<T, N, E, W, MODE, PS, MIDK>(a, b0, b1, x0, lds, threadIdx.x);
}

// Pipelined column pass: grid (N/W strips) x (field groups of `per` <= kPipeMaxPer
// fields); the twiddle tables of the strip's passes and the group's field pointers
// and optics are staged in LDS once per workgroup.
template <typename T, int N, int E, int W, int PS, int HOISTH = HOIST_NONE, int CL = 0>
__global__ __launch_bounds__((N / E) * W) LP2D_OCC void k_cols_pipe(PassArgs<T> a, int per) {
  __shared__ cx<T> lds[W * PaddedLen<N, PS>::value];
  __shared__ cx<T> twl[TwlLen<N, E, CL>::value];
  __shared__ PipeField pf[kPipeMaxPer];
  constexpr int THREADS = (N / E) * W;
  static_assert(THREADS >= kPipeMaxPer, "one thread per staged field");
  constexpr int SPF = N / W;
  const int g = gridDim.x, i = blockIdx.x;
  const int blk = (g % 8 == 0) ? (i % 8) * (g / 8) + i / 8 : i;
  const int grp = blk / SPF;
  const int x0 = (blk - grp * SPF) * W;
  const int b0 = grp * per;
  const int b1 = cmin(a.batch, b0 + per);
  twl_fill<T, N, E, THREADS, 1, CL>(twl, a.tw, threadIdx.x);
  pipe_stage_fields(a, b0, b1, pf, threadIdx.x);
  __syncthreads();
  cols_pipe_body<T, N, E, W, PS, HOISTH, CL>(a, b0, b1, x0, lds, twl, pf, threadIdx.x);
}

// Whole field in LDS, pipelined over groups of `per` <= kPipeMaxPer fields.
template <typename T, int N, int E, int PS>
__global__ __launch_bounds__(N * (N / E)) void k_small_pipe(PassArgs<T> a, int per) {
  __shared__ cx<T> lds[N * SmallPitch<N, PS>::value];
  __shared__ cx<T> twl[TwlLen<N, E>::value];
  __shared__ PipeField pf[kPipeMaxPer];
  constexpr int THREADS = N * (N / E);
  static_assert(THREADS >= kPipeMaxPer, "one thread per staged field");
  const int b0 = blockIdx.x * per;
  const int b1 = cmin(a.batch, b0 + per);
  twl_fill<T, N, E, THREADS>(twl, a.tw, threadIdx.x);
  pipe_stage_fields(a, b0, b1, pf, threadIdx.x);
  __syncthreads();
  small_pipe_body<T, N, E, PS>(a, b0, b1, lds, twl, pf, threadIdx.x);
}

template <typename T, int N, int E, int PS>
__global__ __launch_bounds__(N * (N / E)) void k_small(PassArgs<T> a, int mode) {
  __shared__ cx<T> lds[N * SmallPitch<N, PS>::value];
  small_body<T, N, E, PS>(a, blockIdx.x, mode, lds, threadIdx.x);
}

// ----------------------------------------------------------------- shapes
template <typename T, int N, int E>
struct RowShape {
  static constexpr int TL = N / E;
  static constexpr int PS = cmax(1, ilog2c(E));
  static constexpr int LINES = cmax(1, cmin(N, 256 / TL));
  static constexpr int THREADS = TL * LINES;
  static constexpr int LDS = LINES * RowPitch<N, PS, LINES>::value * (int)sizeof(cx<T>);
};

template <typename T, int N, int E, int WT>
struct ColShape {
  static constexpr int TL = N / E;
  static constexpr int PS = cmax(1, ilog2c(E));
  static constexpr int L = PaddedLen<N, PS>::value;
  static constexpr int W = floor_pow2(
      cmax(1, cmin(cmin(WT, N), cmin(1024 / TL, kLdsBudget / (L * (int)sizeof(cx<T>))))));
  static constexpr int THREADS = TL * W;
  static constexpr int LDS = W * L * (int)sizeof(cx<T>);
};

template <typename T, int N, int EMAX>
struct SmallShape {
  static constexpr int E = cmin(N, EMAX);
  static constexpr int PS = cmax(1, ilog2c(E));
  static constexpr int THREADS = N * (N / E);
  static constexpr int LDS = N * SmallPitch<N, PS>::value * (int)sizeof(cx<T>);
  static constexpr bool OK = THREADS <= 1024 && LDS <= kLdsBudget;
};

// ----------------------------------------------------------------- launchers
// Workgroups of `kernel` resident at once on the current device: measured on first
// use and kept per kernel (every caller computes the same value).
inline int cached_slots(std::atomic<int>& slot, const void* kernel, int threads) {
  int v = slot.load(std::memory_order_relaxed);
  if (v == 0) {
    v = resident_groups(kernel, threads);
    slot.store(v, std::memory_order_relaxed);
  }
  return v;
}

template <typename T, int N, int E, int LINES, int PS>
hipError_t launch_rows(const PassArgs<T>& a, bool inv, hipStream_t s) {
  if ((a.flags & F_TILED) && (LINES % tile_rows(N) != 0 || (N / E) % tile_cols(N) != 0))
    return hipErrorInvalidValue;
  constexpr int threads = (N / E) * LINES;
  constexpr int TPF = (N + LINES - 1) / LINES;
  const unsigned grid = (unsigned)((long long)a.batch * TPF);
  const bool feat = (a.flags & (F_SPATIAL | F_GRAD | F_INTENSITY | F_NO_STORE)) != 0;
  if (inv) {
    if (feat)
      k_rows<T, N, E, LINES, true, PS, true><<<grid, threads, 0, s>>>(a);
    else
      k_rows<T, N, E, LINES, true, PS, false><<<grid, threads, 0, s>>>(a);
  } else {
    if (feat)
      k_rows<T, N, E, LINES, false, PS, true><<<grid, threads, 0, s>>>(a);
    else
      k_rows<T, N, E, LINES, false, PS, false><<<grid, threads, 0, s>>>(a);
  }
  return hipGetLastError();
}

template <typename T, int N, int E, int LINES, bool INV, int PS, bool FEAT, int CL, bool PHO>
hipError_t run_rows_l(const PassArgs<T>& a, int nu, hipStream_t s) {
  constexpr int threads = (N / E) * LINES;
  static std::atomic<int> slot{0};
  const int sl = cached_slots(
      slot, (const void*)&k_rows_l<T, N, E, LINES, INV, PS, FEAT, CL, PHO>, threads);
  const int grid = nu < sl ? nu : sl;
  k_rows_l<T, N, E, LINES, INV, PS, FEAT, CL, PHO><<<grid, threads, 0, s>>>(a, nu);
  return hipGetLastError();
}
template <typename T, int N, int E, int LINES, int PS, int CL>
hipError_t launch_rows_l(const PassArgs<T>& a, bool inv, hipStream_t s) {
  if ((a.flags & F_TILED) && (LINES % tile_rows(N) != 0 || (N / E) % tile_cols(N) != 0))
    return hipErrorInvalidValue;
  // passes with spatial features keep the plain class (in the persistent loop their
  // extra state spills: 636-748 bytes per lane at 4096 points), except a forward pass
  // whose only feature is a phase screen
  const long long units = (long long)a.batch * (N / LINES);
  const int fl = a.flags & (F_SPATIAL | F_GRAD | F_INTENSITY | F_NO_STORE);
  const bool pho = !inv && fl == F_PHASE && a.phase;
  if ((fl != 0 && !pho) || N % LINES != 0 || units > 0x7fffffff)
    return launch_rows<T, N, E, LINES, PS>(a, inv, s);
  const int nu = (int)units;
  if (pho) return run_rows_l<T, N, E, LINES, false, PS, false, CL, true>(a, nu, s);
  return inv ? run_rows_l<T, N, E, LINES, true, PS, false, CL, false>(a, nu, s)
             : run_rows_l<T, N, E, LINES, false, PS, false, CL, false>(a, nu, s);
}

// The pipelined class covers the row passes of strided complex64 batches without
// spatial features (the passes of a plain propagation); anything else runs the plain
// class of the same shape.  (Over an uneven batch's per-field pointer tables, read one
// tile ahead, it measured slower than the plain class: 2.43 against 2.36 ms for
// 2048 x 128^2 + 1536 x 256^2 + 512 x 512^2.)
template <typename T, int N, int E, int LINES, bool INV, int PS, bool TILED, bool PH,
          bool WSK = true>
hipError_t run_rows_pipe_k(const PassArgs<T>& a, int nu, hipStream_t s) {
  constexpr int threads = (N / E) * LINES;
  // same value on every thread that races to set it
  static std::atomic<int> slot{0};
  const int sl = cached_slots(
      slot, (const void*)&k_rows_pipe<T, N, E, LINES, INV, PS, TILED, PH, WSK>, threads);
  const int grid = nu < sl ? nu : sl;
  k_rows_pipe<T, N, E, LINES, INV, PS, TILED, PH, WSK><<<grid, threads, 0, s>>>(a, nu);
  return hipGetLastError();
}
// A pass converting between the natural and the block-tiled layout in place keeps the
// workgroup-synchronised transform (the wave-local one would let a wave store over
// data another wave has yet to load).
template <typename T, int N, int E, int LINES, bool INV, int PS, bool TILED, bool PH>
hipError_t run_rows_pipe(const PassArgs<T>& a, int nu, hipStream_t s) {
  if constexpr (TILED) {
    if ((const void*)a.src == (const void*)a.dst)
      return run_rows_pipe_k<T, N, E, LINES, INV, PS, TILED, PH, false>(a, nu, s);
  }
  return run_rows_pipe_k<T, N, E, LINES, INV, PS, TILED, PH, true>(a, nu, s);
}

template <typename T, int N, int E, int LINES, int PS>
hipError_t launch_rows_pipe(const PassArgs<T>& a, bool inv, hipStream_t s) {
  constexpr bool TILE_OK = LINES % tile_rows(N) == 0 && (N / E) % tile_cols(N) == 0;
  if ((a.flags & F_TILED) && !TILE_OK) return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 4 && N % LINES == 0) {
    const int fl = a.flags & (F_SPATIAL | F_GRAD | F_INTENSITY | F_NO_STORE);
    // the forward pass may carry a phase screen alone (shared or per field); at 512
    // points its prefetched screen values spill (136 bytes per lane), and the plain
    // pass with the screen prologue runs instead
    constexpr bool PH_OK = N != 512;
    const bool ph = PH_OK && !inv && fl == F_PHASE && a.phase;
    const long long units = (long long)a.batch * (N / LINES);
    if ((fl == 0 || ph) && !a.src_tab && !a.dst_tab && units <= 0x7fffffff) {
      const bool tiled = (a.flags & F_TILED) != 0;
      const int nu = (int)units;
      if constexpr (TILE_OK) {
        if (tiled) {
          if (inv) return run_rows_pipe<T, N, E, LINES, true, PS, true, false>(a, nu, s);
          if constexpr (PH_OK)
            if (ph) return run_rows_pipe<T, N, E, LINES, false, PS, true, true>(a, nu, s);
          return run_rows_pipe<T, N, E, LINES, false, PS, true, false>(a, nu, s);
        }
      }
      if (inv) return run_rows_pipe<T, N, E, LINES, true, PS, false, false>(a, nu, s);
      if constexpr (PH_OK)
        if (ph) return run_rows_pipe<T, N, E, LINES, false, PS, false, true>(a, nu, s);
      return run_rows_pipe<T, N, E, LINES, false, PS, false, false>(a, nu, s);
    }
  }
  return launch_rows<T, N, E, LINES, PS>(a, inv, s);
}

template <typename T, int N, int E, int LINES, int PS, bool HOIST>
hipError_t launch_rows_grad(const PassArgs<T>& a, int groups, float* partial, hipStream_t s) {
  // reading a tiled plane in place needs whole row blocks unless nothing is stored
  const bool in_place = (const void*)a.src == (const void*)a.dst;
  if ((a.flags & F_TILED) &&
      (((LINES % tile_rows(N)) != 0 && !(a.flags & F_NO_STORE) && in_place) ||
       (N / E) % tile_cols(N) != 0))
    return hipErrorInvalidValue;
  constexpr int threads = (N / E) * LINES;
  constexpr int TPF = (N + LINES - 1) / LINES;
  k_rows_grad<T, N, E, LINES, PS, HOIST><<<(unsigned)(TPF * groups), threads, 0, s>>>(a, groups,
                                                                                   partial);
  return hipGetLastError();
}

// The chained-plane pass runs on block-tiled planes of whole tiles only (complex64
// schedules whose row tile is a whole number of row blocks and whose threads cover
// whole tile widths); other shapes report hipErrorInvalidValue and the caller runs
// the plain sweeps.
template <typename T, int N, int E, int LINES, int PS>
constexpr bool chain_ok() {
  return sizeof(T) == 4 && N % LINES == 0 && TiledRowMap<N, N / E, LINES>::ok;
}
template <typename T, int N, int E, int LINES, int PS, int WPE = 2>
hipError_t launch_rows_chain(const PassArgs<T>& a, hipStream_t s) {
  if constexpr (chain_ok<T, N, E, LINES, PS>()) {
    if (!(a.flags & F_TILED) || !a.phase || a.src_tab || a.dst_tab) return hipErrorInvalidValue;
    constexpr int threads = (N / E) * LINES;
    const unsigned grid = (unsigned)((long long)a.batch * (N / LINES));
    k_rows_chain<T, N, E, LINES, PS, WPE><<<grid, threads, 0, s>>>(a);
    return hipGetLastError();
  } else {
    (void)a;
    (void)s;
    return hipErrorInvalidValue;
  }
}

// Adjoint chained pass: grid (N/LINES row tiles) x (field groups); twiddles in LDS
// (TWL) or from the global tables.
template <typename T, int N, int E, int LINES, int PS, int CL, bool TWL, bool LAST, int WPE = 2>
__global__ __launch_bounds__((N / E) * LINES) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_rows_adj_chain(PassArgs<T> a, int groups, float* partial) {
  __shared__ cx<T> lds[LINES * RowPitch<N, PS, LINES>::value];
  __shared__ cx<T> twl[TWL ? TwlLen<N, E, CL>::value : 1];
  constexpr int TPF = N / LINES;
  const int tile = blockIdx.x % TPF, g = blockIdx.x / TPF;
  const int per = (a.batch + groups - 1) / groups;
  const int b0 = g * per, b1 = cmin(a.batch, b0 + per);
  if constexpr (TWL) {
    twl_fill<T, N, E, (N / E) * LINES, 1, CL>(twl, a.tw, threadIdx.x);
    __syncthreads();
  }
  rows_adj_chain_body<T, N, E, LINES, PS, CL, LAST>(
      a, b0, b1, tile * LINES, partial ? partial + (long long)g * N * N : nullptr, lds,
      TWL ? twl : nullptr, threadIdx.x);
}

template <typename T, int N, int E, int LINES, int CL>
constexpr int chain_l_lds() {
  return (LINES * RowPitch<N, cmax(1, ilog2c(E)), LINES>::value + TwlLen<N, E, CL>::value) *
         (int)sizeof(cx<T>);
}
template <typename T, int N, int E, int LINES, int PS, int WPE, int CL>
hipError_t launch_rows_chain_l(const PassArgs<T>& a, hipStream_t s) {
  if constexpr (chain_ok<T, N, E, LINES, PS>() && chain_l_lds<T, N, E, LINES, CL>() <= kLdsBudget) {
    if (!(a.flags & F_TILED) || !a.phase || a.src_tab || a.dst_tab) return hipErrorInvalidValue;
    constexpr int threads = (N / E) * LINES;
    const long long units = (long long)a.batch * (N / LINES);
    if (units > 0x7fffffff) return hipErrorInvalidValue;
    static std::atomic<int> slot{0};
    const int slots = cached_slots(slot, (const void*)&k_rows_chain_l<T, N, E, LINES, PS, WPE, CL>,
                                   threads);
    const int grid = (int)(units < slots ? units : slots);
    k_rows_chain_l<T, N, E, LINES, PS, WPE, CL><<<grid, threads, 0, s>>>(a, (int)units);
    return hipGetLastError();
  } else {
    (void)a;
    (void)s;
    return hipErrorInvalidValue;
  }
}

template <typename T, int N, int E, int LINES, int CL, bool TWL>
constexpr int adj_chain_lds() {
  return (LINES * RowPitch<N, cmax(1, ilog2c(E)), LINES>::value +
          (TWL ? TwlLen<N, E, CL>::value : 0)) *
         (int)sizeof(cx<T>);
}
template <typename T, int N, int E, int LINES, int PS, int CL, bool TWL, bool LAST, int WPE = 2>
hipError_t launch_rows_adj_chain(const PassArgs<T>& a, int groups, float* partial, hipStream_t s) {
  if constexpr (chain_ok<T, N, E, LINES, PS>() &&
                adj_chain_lds<T, N, E, LINES, CL, TWL>() <= kLdsBudget) {
    if (!(a.flags & F_TILED) || !a.phase || (partial && !a.u_saved) || a.src_tab || a.dst_tab ||
        groups < 1)
      return hipErrorInvalidValue;
    constexpr int threads = (N / E) * LINES;
    const long long grid = (long long)(N / LINES) * groups;
    k_rows_adj_chain<T, N, E, LINES, PS, CL, TWL, LAST, WPE><<<(unsigned)grid, threads, 0, s>>>(
        a, groups, partial);
    return hipGetLastError();
  } else {
    (void)a;
    (void)groups;
    (void)partial;
    (void)s;
    return hipErrorInvalidValue;
  }
}

// Fields per column workgroup: share the transfer function across a group of fields
// while keeping at least ~kColBlocks workgroups in flight.
constexpr int kColBlocks = 1024;

template <typename T, int N, int E, int W, int PS, bool HOIST>
hipError_t launch_cols(const PassArgs<T>& a, int mode, hipStream_t s) {
  if (a.flags & F_TILED) return hipErrorInvalidValue;  // natural layout only
  constexpr int threads = (N / E) * W;
  constexpr int SPF = N / W;
  // the shared-factor class needs one set of optics for the whole call
  const bool hoist = HOIST && a.fopt == nullptr;
  const bool shared_h = hoist && (a.flags & F_MID) && !(a.flags & (F_SMASK0 | F_SMASK1));
  int per = shared_h ? (int)((long long)a.batch * SPF / kColBlocks) : 1;
  if (per < 1) per = 1;
  if (per > 8) per = 8;
  const int groups = (a.batch + per - 1) / per;
  const unsigned grid = (unsigned)((long long)groups * SPF);
  // spectral stage class: none / transfer function + aperture + store / with masks
  const int midk = (a.flags & (F_SMASK0 | F_SMASK1)) ? 2 : (a.flags & F_MID) ? 1 : 0;
  constexpr int FI = MODE_FWD | MODE_INV;
  switch (mode * 4 + midk) {
    case MODE_FWD * 4 + 0:
      k_cols<T, N, E, W, MODE_FWD, PS, 0><<<grid, threads, 0, s>>>(a, per);
      break;
    case MODE_INV * 4 + 0:
      k_cols<T, N, E, W, MODE_INV, PS, 0><<<grid, threads, 0, s>>>(a, per);
      break;
    case MODE_INV * 4 + 1:
      if constexpr (HOIST) {
        if (hoist) {
          k_cols<T, N, E, W, MODE_INV, PS, 3><<<grid, threads, 0, s>>>(a, per);
          break;
        }
      }
      k_cols<T, N, E, W, MODE_INV, PS, 1><<<grid, threads, 0, s>>>(a, per);
      break;
    case FI * 4 + 0:
    case FI * 4 + 1:
      if constexpr (HOIST) {
        if (hoist) {
          k_cols<T, N, E, W, FI, PS, 3><<<grid, threads, 0, s>>>(a, per);
          break;
        }
      }
      k_cols<T, N, E, W, FI, PS, 1><<<grid, threads, 0, s>>>(a, per);
      break;
    case FI * 4 + 2:
      k_cols<T, N, E, W, FI, PS, 2><<<grid, threads, 0, s>>>(a, per);
      break;
    default:  // forward-only passes with a spectral stage, and inverse ones with masks
      if (mode == MODE_FWD)
        k_cols<T, N, E, W, MODE_FWD, PS, 2><<<grid, threads, 0, s>>>(a, per);
      else
        k_cols<T, N, E, W, MODE_INV, PS, 2><<<grid, threads, 0, s>>>(a, per);
      break;
  }
  return hipGetLastError();
}

// The pipelined class covers the fused forward+inverse column pass of the complex64
// path with transfer function / pupil / spectrum output; anything else (mask
// multipliers, other evanescent modes, single-direction passes) runs the plain class
// of the same shape.  Each workgroup takes `per` fields so that the grid is about one
// round of resident workgroups.
template <typename T, int N, int E, int W, int PS, int HOISTH = HOIST_NONE, int CL = 0>
hipError_t launch_cols_pipe(const PassArgs<T>& a, int mode, hipStream_t s) {
  const bool ok = pipe_eligible(a, mode);
  // the block-tiled layout is only defined for strips of tile_cols(N) columns, and only
  // this class reads it
  if ((a.flags & F_TILED) && (!ok || W != tile_cols(N) || (N / E) % tile_rows(N) != 0))
    return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 4) {
    if (ok) {
      constexpr int threads = (N / E) * W;
      constexpr int SPF = N / W;
      // the hoisted factor needs one set of optics for the whole call
      const bool hoist = HOISTH != HOIST_NONE && a.fopt == nullptr;
      static std::atomic<int> slot[2];
      int slots = 0;
      if constexpr (HOISTH != HOIST_NONE) {
        if (hoist)
          slots = cached_slots(slot[1], (const void*)&k_cols_pipe<T, N, E, W, PS, HOISTH, CL>, threads);
      }
      if (!hoist)
        slots = cached_slots(slot[0], (const void*)&k_cols_pipe<T, N, E, W, PS, HOIST_NONE, CL>, threads);
      int groups = slots / SPF;
      if (groups < 1) groups = 1;
      if (groups > a.batch) groups = a.batch;
      int per = (a.batch + groups - 1) / groups;
      if (per > kPipeMaxPer) per = kPipeMaxPer;
      // at 4096 a field's 1024 strips already take four rounds of the resident
      // workgroups: groups of at most 4 fields measured 800-811 us per 8 fields against
      // 833-841 us for one group of 8 (2 fields: 797-813 us, 1 field: 830-841 us).  At
      // 1024, where the groups of 8 fill one round, caps of 4 and 2 took 125 and 131 us
      // against 115 us.
      if constexpr (N >= 4096) {
        if (per > 4) per = 4;
      }
      groups = (a.batch + per - 1) / per;
      if constexpr (HOISTH != HOIST_NONE) {
        if (hoist) {
          k_cols_pipe<T, N, E, W, PS, HOISTH, CL><<<(unsigned)(groups * SPF), threads, 0, s>>>(a, per);
          return hipGetLastError();
        }
      }
      k_cols_pipe<T, N, E, W, PS, HOIST_NONE, CL><<<(unsigned)(groups * SPF), threads, 0, s>>>(a, per);
      return hipGetLastError();
    }
  }
  return launch_cols<T, N, E, W, PS, false>(a, mode, s);
}

// The fused complex64 propagation (transfer function / pupil / scale / spectrum
// output only) runs the pipelined whole-field kernel; every other combination the
// general one.
template <typename T, int N, int E, int PS>
hipError_t launch_small(const PassArgs<T>& a, int mode, hipStream_t s) {
  constexpr int PIPE_OK_FLAGS = F_H | F_APERTURE | F_SCALE | F_SPEC_OUT;
  if constexpr (sizeof(T) == 4 && N >= 32 && E == 16) {
    constexpr int threads = N * (N / E);
    constexpr int lds = N * SmallPitch<N, PS>::value * (int)sizeof(cx<T>) +
                        TwlLen<N, E>::value * (int)sizeof(cx<T>) +
                        kPipeMaxPer * (int)sizeof(PipeField);
    if constexpr (lds <= kLdsBudget) {
      if (mode == (MODE_FWD | MODE_INV) && a.evan == EV_PASS &&
          (a.flags & ~PIPE_OK_FLAGS) == 0) {
        static std::atomic<int> slot{0};
        const int slots = cached_slots(slot, (const void*)&k_small_pipe<T, N, E, PS>, threads);
        int groups = slots < a.batch ? slots : a.batch;
        if (groups < 1) groups = 1;
        int per = (a.batch + groups - 1) / groups;
        if (per > kPipeMaxPer) per = kPipeMaxPer;
        groups = (a.batch + per - 1) / per;
        k_small_pipe<T, N, E, PS><<<(unsigned)groups, threads, 0, s>>>(a, per);
        return hipGetLastError();
      }
    }
  }
  k_small<T, N, E, PS><<<(unsigned)a.batch, N * (N / E), 0, s>>>(a, mode);
  return hipGetLastError();
}

// ----------------------------------------------------------------- registration
template <typename T, int N, int E>
void add_row(Registry<T>& r) {
  using S = RowShape<T, N, E>;
  constexpr int lg = ilog2c(N);
  int& k = r.nrows[lg];
  if (k >= kMaxVariants) return;
  // (taller gradient tiles measured slower: 8 rows, for reading a block-tiled plane in
  // place, because a 1024-thread workgroup per CU stalls the CU at every exchange
  // barrier; 4 rows at 1024 (512 threads), reading the adjoint's workspace in 256-byte
  // pieces, by 5% on the gradient pass)
  // (at 4096 points, 2-row gradient tiles over the 2 x 4 block-tiled workspace, 1024
  // threads, measured 1156 us against 749 us for one row per workgroup)
  r.rows[lg][k++] = RowVar<T>{&launch_rows<T, N, E, S::LINES, S::PS>,
                              &launch_rows_grad<T, N, E, S::LINES, S::PS, (E <= 8)>, E,
                              S::LINES, S::THREADS, S::LDS, S::LINES, 0};
}

template <typename T, int N, int E>
void add_row_pipe(Registry<T>& r) {
  using S = RowShape<T, N, E>;
  constexpr int lg = ilog2c(N);
  int& k = r.nrows[lg];
  if (k >= kMaxVariants) return;
  constexpr int lds = S::LDS + TwlLen<N, E>::value * (int)sizeof(cx<T>);
  if constexpr (lds <= kLdsBudget)
    r.rows[lg][k++] = RowVar<T>{&launch_rows_pipe<T, N, E, S::LINES, S::PS>,
                                &launch_rows_grad<T, N, E, S::LINES, S::PS, (E <= 8)>, E,
                                S::LINES, S::THREADS, lds, S::LINES, 2};
}

// The plain row pass with LDS twiddles (kind bit 2), compact mode CL.
template <typename T, int N, int E, int CL>
void add_row_l(Registry<T>& r) {
  using S = RowShape<T, N, E>;
  constexpr int lg = ilog2c(N);
  int& k = r.nrows[lg];
  if (k >= kMaxVariants) return;
  constexpr int lds = S::LDS + TwlLen<N, E, CL>::value * (int)sizeof(cx<T>);
  if constexpr (lds <= kLdsBudget)
    r.rows[lg][k++] = RowVar<T>{&launch_rows_l<T, N, E, S::LINES, S::PS, CL>,
                                &launch_rows_grad<T, N, E, S::LINES, S::PS, (E <= 8)>, E,
                                S::LINES, S::THREADS, lds, S::LINES, 4};
}

template <typename T, int N, int E, int WT, bool HOIST = false>
void add_col(Registry<T>& r) {
  using S = ColShape<T, N, E, WT>;
  constexpr int lg = ilog2c(N);
  int& k = r.ncols[lg];
  if (k >= kMaxVariants) return;
  r.cols[lg][k++] = ColVar<T>{&launch_cols<T, N, E, S::W, S::PS, HOIST>, E, S::W, S::THREADS,
                              S::LDS, HOIST ? 1 : 0};
}

template <typename T, int N, int E, int WT, int HOISTH = HOIST_NONE, int CL = 0>
void add_col_pipe(Registry<T>& r) {
  using S = ColShape<T, N, E, WT>;
  constexpr int lg = ilog2c(N);
  int& k = r.ncols[lg];
  if (k >= kMaxVariants) return;
  constexpr int lds = S::LDS + TwlLen<N, E, CL>::value * (int)sizeof(cx<T>) +
                      kPipeMaxPer * (int)sizeof(PipeField);
  if constexpr (lds <= kLdsBudget)
    r.cols[lg][k++] = ColVar<T>{&launch_cols_pipe<T, N, E, S::W, S::PS, HOISTH, CL>, E, S::W,
                                S::THREADS, lds, HOISTH != HOIST_NONE ? 3 : 2};
}

template <typename T, int N>
void register_n(Registry<T>& r) {
  constexpr int lg = ilog2c(N);
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int EMAX = F32 ? 16 : 8;
  constexpr int E0 = cmin(N, EMAX);
  constexpr int E1 = cmin(N, EMAX / 2);
  constexpr int WT = F32 ? 16 : 8;  // 128-byte row segments
  r.nrows[lg] = 0;
  r.ncols[lg] = 0;
  add_row<T, N, E0>(r);
  add_col<T, N, E0, WT>(r);
  if constexpr (E1 != E0 && N >= 16) {
    add_row<T, N, E1>(r);
    add_col<T, N, E1, WT>(r);
  }
  if constexpr (N >= 256) {
    add_col<T, N, E0, WT / 2>(r);
    if constexpr (F32) {
      add_row<T, N, 32>(r);
      add_col<T, N, 32, 16>(r);
      add_col<T, N, 32, 8>(r);
      add_col<T, N, 32, 8, true>(r);
      add_col<T, N, 16, 8, true>(r);
      add_col_pipe<T, N, 32, 8>(r);
      add_col_pipe<T, N, 16, 8>(r);
      add_col_pipe<T, N, 32, 4>(r);
      add_col_pipe<T, N, 32, 8, HOIST_FACTOR>(r);
      add_col_pipe<T, N, 16, 8, HOIST_PHASE>(r);
      // the compact twiddle table lets 4-column strips of 4096 points pipeline (the
      // hoisted factor does not fit in the registers beside them); at 2048 the split
      // compact table brings 4-column strips to 77 KiB of LDS, two workgroups per CU
      // (163 against 223 us per 8 fields)
      if constexpr (N >= 4096) add_col_pipe<T, N, 32, 4, HOIST_NONE, true>(r);
      if constexpr (N == 2048) add_col_pipe<T, N, 32, 4, HOIST_NONE, 2>(r);
      add_row_pipe<T, N, 32>(r);
      if constexpr (N == 4096) add_row_l<T, N, 32, 2>(r);
    } else {
      add_row<T, N, 16>(r);
      add_col<T, N, 16, 8>(r);
    }
  }
  // chained-plane row pass: 32 elements per thread in the row pass's shape (2 waves
  // per SIMD) with global or LDS-staged twiddles, and 16 elements per thread in row
  // tiles of at least one tile height (4 waves per SIMD).  Variant 0 is the fastest
  // measured per size (stack of 5 screens, per pass; 32 global / 32 LDS / 16):
  // 4096 x8: 680 / 627 / 846 us; 1024 x32: 142 / 117 / 148; 512 x64: 67 / 63 / 81;
  // 256 x128: 36 / 34 / 33.  At 4096 the 2-row tiles with the split compact table
  // (75 KiB of LDS, two workgroups per CU) take 529 us and come first.
  r.nchain[lg] = 0;
  for (int k = 0; k < 4; ++k) r.chain_twl[lg][k] = 0;
  r.adj_chain[lg][0] = r.adj_chain[lg][1] = nullptr;
  r.adj_threads[lg] = r.adj_lines[lg] = 0;
  if constexpr (F32 && N >= 256) {
    using S = RowShape<T, N, 32>;
    constexpr int L16 = cmax(tile_rows(N), 256 / (N / 16));
    constexpr int PS16 = cmax(1, ilog2c(16));
    constexpr bool ok32 = chain_ok<T, N, 32, S::LINES, S::PS>();
    constexpr bool ok16 = (N / 16) * L16 <= 1024 && chain_ok<T, N, 16, L16, PS16>();
    auto add32 = [&] {
      if constexpr (ok32) {
        const int k = r.nchain[lg]++;
        r.chain[lg][k] = &launch_rows_chain<T, N, 32, S::LINES, S::PS, 2>;
        r.chain_E[lg][k] = 32;
        r.chain_lines[lg][k] = S::LINES;
      }
    };
    auto add16 = [&] {
      if constexpr (ok16) {
        const int k = r.nchain[lg]++;
        r.chain[lg][k] = &launch_rows_chain<T, N, 16, L16, PS16, 4>;
        r.chain_E[lg][k] = 16;
        r.chain_lines[lg][k] = L16;
      }
    };
    // twiddles in LDS: the 32-element shape (rows of the row pass, or 4 rows at 4096
    // with the compact last-pass table: one 512-thread workgroup per CU)
    constexpr int LL = N >= 4096 ? 4 : S::LINES;
    constexpr int CLL = N >= 4096 ? 1 : 0;
    constexpr bool okl = chain_ok<T, N, 32, LL, S::PS>() &&
                         chain_l_lds<T, N, 32, LL, CLL>() <= kLdsBudget &&
                         (N / 32) * LL <= 1024;
    auto add32l = [&] {
      if constexpr (okl) {
        const int k = r.nchain[lg]++;
        r.chain[lg][k] = &launch_rows_chain_l<T, N, 32, LL, S::PS, 2, CLL>;
        r.chain_E[lg][k] = 32;
        r.chain_lines[lg][k] = LL;
        r.chain_twl[lg][k] = 1;
      }
    };
    // 2-row tiles of 4096 points with the split compact table: two workgroups per CU
    constexpr bool oks = N == 4096 && chain_ok<T, N, 32, S::LINES, S::PS>() &&
                         chain_l_lds<T, N, 32, S::LINES, 2>() <= 80 * 1024;
    auto add32s = [&] {
      if constexpr (oks) {
        const int k = r.nchain[lg]++;
        r.chain[lg][k] = &launch_rows_chain_l<T, N, 32, S::LINES, S::PS, 2, 2>;
        r.chain_E[lg][k] = 32;
        r.chain_lines[lg][k] = S::LINES;
        r.chain_twl[lg][k] = 1;
      }
    };
    if constexpr (N == 4096) {
      add32s();
      add32l();
      add32();
      add16();
    } else if constexpr (N >= 512) {
      add32l();
      add32();
      add16();
    } else {
      add16();
      add32();
      add32l();
    }
    // the stack's adjoint: 16 elements per thread in row tiles of whole blocks (the
    // group's gradient sums stay in registers beside two transforms: with 32 elements
    // the chained form spilled 280-400 bytes per lane, and so does a 1024-thread
    // workgroup held to 128 registers, so 2048 and 8192 keep the plane-by-plane
    // adjoint), compact last-pass twiddles in LDS where two workgroups per CU keep
    // room for them (at 4096 the split table with the radix-16 last pass, and the
    // passes held to 128 registers for four waves per SIMD: 957 -> 617 us and 689 ->
    // 519 us per 8 fields against the global tables), the global tables otherwise
    constexpr int LA = L16;
    constexpr int CLA = N == 4096 ? 2 : 1;
    constexpr bool TWA = adj_chain_lds<T, N, 16, LA, CLA, true>() <= 80 * 1024;
    if constexpr (ok16 && (N / 16) * LA <= 512 &&
                  adj_chain_lds<T, N, 16, LA, CLA, TWA>() <= kLdsBudget) {
      constexpr int WPA = N == 4096 ? 4 : 2;
      r.adj_chain[lg][0] = &launch_rows_adj_chain<T, N, 16, LA, PS16, CLA, TWA, false, WPA>;
      r.adj_chain[lg][1] = &launch_rows_adj_chain<T, N, 16, LA, PS16, CLA, TWA, true, WPA>;
      r.adj_threads[lg] = (N / 16) * LA;
      r.adj_lines[lg] = LA;
    }
  }
  using SS = SmallShape<T, N, EMAX>;
  if constexpr (SS::OK) {
    r.small[lg] = SmallVar<T>{&launch_small<T, N, SS::E, SS::PS>, SS::E, SS::THREADS, SS::LDS};
    r.has_small[lg] = true;
  } else {
    r.has_small[lg] = false;
  }
}

}  // namespace lp2d
