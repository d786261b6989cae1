// Kernel-variant registry: for every (precision, n) the launchable schedules of the
// row pass, the column pass and the whole-field-in-LDS kernel.  The autotuner picks
// one variant per (n, precision, kind) through lp2d_set_variant.
#pragma once
#include <hip/hip_runtime.h>

#include "lp2d_kernels.hpp"

namespace lp2d {

template <typename T>
using RowLaunch = hipError_t (*)(const PassArgs<T>&, bool inv, hipStream_t);
template <typename T>
using ColLaunch = hipError_t (*)(const PassArgs<T>&, int mode, hipStream_t);
template <typename T>
using SmallLaunch = hipError_t (*)(const PassArgs<T>&, int mode, hipStream_t);

template <typename T>
using RowGradLaunch = hipError_t (*)(const PassArgs<T>&, int groups, float* partial,
                                     hipStream_t);
template <typename T>
using RowChainLaunch = hipError_t (*)(const PassArgs<T>&, hipStream_t);
template <typename T>
using RowAdjChainLaunch = hipError_t (*)(const PassArgs<T>&, int groups, float* partial,
                                         hipStream_t);

template <typename T>
struct RowVar {
  RowLaunch<T> fn;
  RowGradLaunch<T> grad_fn;  // adjoint last pass with the batch-summed phase gradient
  int E, lines, threads, lds;
  int grad_lines;  // row tile of grad_fn: whole 8-row blocks where the threads allow
  int kind;        // bit 1: software-pipelined over the row tiles of the batch
                   // bit 2: twiddles staged in LDS, one resident round of workgroups
};
template <typename T>
struct ColVar {
  ColLaunch<T> fn;
  int E, W, threads, lds;
  int kind;  // bit 0: transfer function computed once per group of fields (shared-H)
             // bit 1: software-pipelined over the fields of a group
};
template <typename T>
struct SmallVar {
  SmallLaunch<T> fn;
  int E, threads, lds;
};

constexpr int kMaxVariants = 16;
constexpr int kLogMax = 13;  // n <= 8192

template <typename T>
struct Registry {
  RowVar<T> rows[kLogMax + 1][kMaxVariants];
  int nrows[kLogMax + 1];
  ColVar<T> cols[kLogMax + 1][kMaxVariants];
  int ncols[kLogMax + 1];
  SmallVar<T> small[kLogMax + 1];
  bool has_small[kLogMax + 1];
  // chained-plane row pass (inverse rows + next screen + forward rows over the
  // block-tiled plane): its shapes per size (none where the plane is never tiled)
  RowChainLaunch<T> chain[kLogMax + 1][4];
  int chain_E[kLogMax + 1][4], chain_lines[kLogMax + 1][4];
  int chain_twl[kLogMax + 1][4];  // 1: twiddles staged in LDS
  int nchain[kLogMax + 1];
  // adjoint of a stack: [0] chained (inverse rows, fold, gradient, forward rows),
  // [1] last plane (inverse rows, fold, gradient, natural store); threads per group
  RowAdjChainLaunch<T> adj_chain[kLogMax + 1][2];
  int adj_threads[kLogMax + 1], adj_lines[kLogMax + 1];
};

// Workgroups of a column-pass kernel resident at once on the current device.
int resident_groups(const void* kernel, int threads);

void register_f32_a(Registry<float>& r);   // n = 2 .. 256
void register_f32_b(Registry<float>& r);   // n = 512, 1024
void register_f32_c(Registry<float>& r);   // n = 2048 .. 8192
void register_f64_a(Registry<double>& r);  // n = 2 .. 256
void register_f64_b(Registry<double>& r);  // n = 512, 1024
void register_f64_c(Registry<double>& r);  // n = 2048 .. 8192

}  // namespace lp2d
