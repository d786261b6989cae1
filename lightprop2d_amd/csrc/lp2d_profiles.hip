// Device-side field and mask generators (lightprop2d/profiles.py:21-136) behind
// lp2d_profile.  One thread per output element over the broadcast grid (y[j], x[i]);
// FP64 arithmetic in the reference's operation order, so the float outputs agree with
// numpy to the last bit or two of exp/sin/cos and the masks compare the same values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/lp2d.h"

extern "C" int lp2d_init(int device);
namespace lp2d {
int set_error(int code, const char* msg);
}

namespace {

struct Prof {
  int kind, nx, ny, b, out_dtype;
  double a0, rho0, r2, thresh, x0, y0, dx, dy, slits;
};

__device__ __forceinline__ bool round_hole(const Prof& p, double x, double y) {
#pragma clang fp contract(off)
  // gaussian_beam(x - x0, y - y0, 1, R) >= 1/exp(0.5)   (profiles.py:79-80)
  const double u = x - p.x0, v = y - p.y0;
  return exp(-(u * u + v * v) / p.r2) >= p.thresh;
}

__device__ __forceinline__ bool square(double x, double y, double x0, double y0, double d) {
  return fabs(x - x0) < d / 2 && fabs(y - y0) < d / 2;
}

__global__ void k_profile(Prof p, const double* __restrict__ xs, const double* __restrict__ ys,
                          const double* __restrict__ rnd, void* __restrict__ out) {
#pragma clang fp contract(off)
  const long long total = (long long)p.nx * p.ny;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % p.nx), j = (int)(e / p.nx);
    const double x = xs[i], y = ys[j];
    double re = 0.0, im = 0.0;
    bool mask = false;
    switch (p.kind) {
      case LP2D_PROF_PLANE:
        re = 1.0;
        break;
      case LP2D_PROF_GAUSSIAN: {
        const double u = x - p.x0, v = y - p.y0;
        re = p.a0 * exp(-(u * u + v * v) / (p.rho0 * p.rho0));
        break;
      }
      case LP2D_PROF_ROUND_HOLE:
        mask = round_hole(p, x, y);
        break;
      case LP2D_PROF_RECT_HOLE:
        mask = fabs(x - p.x0) < (p.dx / 2) && fabs(y - p.y0) < (p.dy / 2);
        break;
      case LP2D_PROF_SLITS: {
        const double dl = p.slits / 2;
        mask = square(x, y, p.x0 - dl, p.y0, p.dx) || square(x, y, p.x0 + dl, p.y0, p.dx);
        break;
      }
      default: {  // random kinds: binned host draw, (ny/b) x (nx/b), row-major
        const int nxb = p.nx / p.b;
        const int ib = i / p.b, jb = j / p.b;
        // np.kron over a (ny/b, nx/b) array yields (ny/b*b, nx/b*b): outside it (n not
        // a multiple of b) the reference would fail to broadcast; those elements are 0
        const bool inside = ib < nxb && jb < p.ny / p.b;
        double u = inside ? rnd[(long long)jb * nxb + ib] : 0.0;
        if (p.kind == LP2D_PROF_RANDOM) {
          re = u;
        } else {
          const bool hole = round_hole(p, x, y);
          u = hole ? u : 0.0;
          if (p.kind == LP2D_PROF_RANDOM_HOLE) {
            re = u;
          } else {  // hole * (cos(2 pi u) + i sin(2 pi u))   (profiles.py:110-112)
            const double angle = u * 2.0 * 3.141592653589793;
            re = hole ? cos(angle) : 0.0;
            im = hole ? sin(angle) : 0.0;
          }
        }
        break;
      }
    }
    if (p.out_dtype == LP2D_MASK_U8) {
      static_cast<uint8_t*>(out)[e] = mask ? 1 : 0;
    } else if (p.out_dtype == LP2D_MASK_F64) {
      static_cast<double*>(out)[e] = re;
    } else {
      static_cast<double2*>(out)[e] = make_double2(re, im);
    }
  }
}

int out_kind(int kind) {
  switch (kind) {
    case LP2D_PROF_ROUND_HOLE:
    case LP2D_PROF_RECT_HOLE:
    case LP2D_PROF_SLITS:
      return LP2D_MASK_U8;
    case LP2D_PROF_RANDOM_PHASE:
      return LP2D_MASK_C128;
    default:
      return LP2D_MASK_F64;
  }
}

}  // namespace

extern "C" int lp2d_profile(int kind, int nx, int ny, const double* x, const double* y,
                            const lp2d_profile_params* params, const double* rand, void* out,
                            int out_dtype, void* stream) {
  using lp2d::set_error;
  if (kind < LP2D_PROF_PLANE || kind > LP2D_PROF_RANDOM_PHASE)
    return set_error(LP2D_E_ARG, "unknown generator kind");
  if (nx < 1 || ny < 1) return set_error(LP2D_E_SIZE, "grid sizes must be >= 1");
  if (!x || !y || !out || !params) return set_error(LP2D_E_ARG, "null pointer");
  if (out_dtype != out_kind(kind))
    return set_error(LP2D_E_DTYPE, "output dtype does not match the generator");
  const bool random = kind >= LP2D_PROF_RANDOM;
  if (random && (!rand || params->binning < 1 || params->binning > nx || params->binning > ny))
    return set_error(LP2D_E_ARG, "random generators need the host draw and 1 <= binning <= n");
  if ((kind == LP2D_PROF_ROUND_HOLE || kind >= LP2D_PROF_RANDOM_HOLE) && !(params->r != 0.0))
    return set_error(LP2D_E_ARG, "round hole radius must be non-zero");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(LP2D_E_HIP, "hipGetDevice failed");
  if (int rc = lp2d_init(dev)) return rc;
  Prof p;
  p.kind = kind;
  p.nx = nx;
  p.ny = ny;
  p.b = random ? params->binning : 1;
  p.out_dtype = out_dtype;
  p.a0 = params->a0;
  p.rho0 = params->rho0;
  p.r2 = params->r * params->r;
  p.thresh = params->thresh;
  p.x0 = params->x0;
  p.y0 = params->y0;
  p.dx = params->dx;
  p.dy = params->dy;
  p.slits = params->slits;
  const long long total = (long long)nx * ny;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_profile<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(p, x, y, rand, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(LP2D_E_HIP, hipGetErrorString(e));
  return LP2D_OK;
}
