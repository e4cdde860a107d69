// Objective-function kernels: the quadratic neighbour prior, a fixed-order
// FP64 reduction of per-workgroup partial sums, and the element-wise pieces
// of the OSEM vector-Jacobian product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <mutex>

#include "pmrt_device.hpp"

namespace mi {

// The FP64 partial sums of the allocation-free prior entry point: a fixed
// array in every GPU's global memory, held until the process ends. Host code
// reaches it only through the per-device address from hipGetSymbolAddress
// (noalloc_partials); the array's name is not a device address on the host.
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {

constexpr int kPriorTileX = 64;
constexpr int kPriorTileY = 4;
constexpr int kPriorThreads = kPriorTileX * kPriorTileY;

// Neighbourhood priors, one voxel per thread, accumulated in double in
// (dz, dy, dx) order; neighbours outside the image are skipped.
//  KIND 1, QuadraticPrior::compute_value / compute_gradient
//    (QuadraticPrior.cxx:229-367): float terms w * (a-b)^2 / 4 and w * (a-b).
//  KIND 2, RelativeDifferencePrior::compute_value / compute_gradient
//    (RelativeDifferencePrior.cxx:280-433): value term
//    w * 0.5 * (double)(a-b)^2 / (float)(a + b + gamma|a-b| + eps), gradient term
//    float w * derivative_10(a, b); both 0 at a = b = 0 when eps = 0.
// ZFAST: the device layout [y][x][zp] (threads along z), else [z][y][x]
// (threads along x). Workgroup = 64 x 4 voxels; its value partial (without
// beta) is written to partial[block]; grad = beta * gradient.
template <int KIND, bool KAPPA, bool ZFAST>
__global__ void __launch_bounds__(kPriorThreads) prior_kernel(const float* __restrict__ img,
                                                              const float* __restrict__ kappa, PriorImage im,
                                                              PriorWeights W, int wz0, int wz1, float beta, float gamma,
                                                              float eps, float* __restrict__ grad,
                                                              double* __restrict__ partial) {
  const int c = blockIdx.x * kPriorTileX + static_cast<int>(threadIdx.x % kPriorTileX);
  const int b = blockIdx.y * kPriorTileY + static_cast<int>(threadIdx.x / kPriorTileX);
  const int a = blockIdx.z;
  const int nz = im.nz, ny = im.ny, nx = im.nx;
  const int z = ZFAST ? c : a, y = ZFAST ? a : b, x = ZFAST ? b : c;
  // strides of z, y, x
  const int64_t sz = ZFAST ? 1 : static_cast<int64_t>(ny) * nx;
  const int64_t sy = ZFAST ? static_cast<int64_t>(nx) * im.nzp : nx;
  const int64_t sx = ZFAST ? im.nzp : 1;
  double value = 0.;
  if (x < nx && y < ny && z < nz) {
    const int64_t v = z * sz + y * sy + x * sx;
    const float xv = img[v];
    const float kv = KAPPA ? kappa[v] : 1.F;
    const int min_dz = max(wz0, -z), max_dz = min(wz1, nz - 1 - z);
    const int min_dy = max(-1, -y), max_dy = min(1, ny - 1 - y);
    const int min_dx = max(-1, -x), max_dx = min(1, nx - 1 - x);
    double g = 0.;
    for (int dz = min_dz; dz <= max_dz; ++dz)
      for (int dy = min_dy; dy <= max_dy; ++dy)
        for (int dx = min_dx; dx <= max_dx; ++dx) {
          const int64_t n = v + dz * sz + dy * sy + dx * sx;
          const float w = W.w[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
          const float xn = img[n];
          double cv, cg;
          if constexpr (KIND == 1) {
            const float d = xv - xn;
            cv = static_cast<double>(w * (d * d) / 4);
            cg = static_cast<double>(w * d);
          } else {
            const float ad = fabsf(xv - xn);
            if (eps == 0.F && xv == 0.F && xn == 0.F) {
              cv = 0.0;
              cg = 0.0;
            } else {
              const float den = xv + xn + gamma * ad + eps;
              const double dd = static_cast<double>(xv - xn);
              cv = static_cast<double>(w) * (0.5 * (dd * dd / den));
              const float t = gamma * ad + xv + 3 * xn + 2 * eps;
              const double num = dd * t;
              const double ds = static_cast<double>(xv + xn) + gamma * ad + eps;
              cg = static_cast<double>(w * static_cast<float>(num / (ds * ds)));
            }
          }
          if (KAPPA) {
            const float kk = kv * kappa[n];
            cv *= kk;
            cg *= kk;
          }
          value += cv;
          g += cg;
        }
    if (grad) grad[v] = static_cast<float>(g * beta);
  }
  // workgroup sum of the value terms
  __shared__ double red[kPriorThreads / 64];
  for (int off = 32; off > 0; off >>= 1) value += __shfl_xor(value, off, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = value;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.;
    for (int i = 0; i < kPriorThreads / 64; ++i) s += red[i];
    partial[(static_cast<int64_t>(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
  }
}

// derivative_20 / derivative_11 of the relative difference prior
// (RelativeDifferencePrior.cxx:617-641): the reference's pow of float
// arguments is the double pow; a * a of a float is exact in double and
// (a * a) * a is a^3 correctly rounded.
__device__ inline float rdp_d20(float xj, float xk, float gamma, float eps) {
  if (!(xj > 0.F || xk > 0.F || eps > 0.F)) return INFINITY;
  const double a = static_cast<double>(2 * xk + eps);
  const double d = static_cast<double>(xj + xk + gamma * fabsf(xj - xk) + eps);
  return static_cast<float>(2 * (a * a) / ((d * d) * d));
}
__device__ inline float rdp_d11(float xj, float xk, float gamma, float eps) {
  if (!(xj > 0.F || xk > 0.F || eps > 0.F)) return INFINITY;
  const float num = -2 * (2 * xj + eps) * (2 * xk + eps);
  const double d = static_cast<double>(xj + xk + gamma * fabsf(xj - xk) + eps);
  return static_cast<float>(static_cast<double>(num) / ((d * d) * d));
}

// out[j] += beta * sum_k w (f20(x_j, x_k) in_j + f11(x_j, x_k) in_k) kappa_j kappa_k
// (accumulate_Hessian_times_input, QuadraticPrior.cxx:604-690,
// RelativeDifferencePrior.cxx:535-614; KIND 1 quadratic f20 = 1, f11 = -1; 2
// relative difference), or KIND 3: out[j] += beta * sum_k w in_k kappa_j
// kappa_k (the quadratic prior's add_multiplication_with_approximate_Hessian,
// :537-602). Float arithmetic in (dz, dy, dx) order, [z][y][x] images.
template <int KIND, bool KAPPA>
__global__ void __launch_bounds__(kPriorThreads) prior_hessian_kernel(const float* __restrict__ x,
                                                                      const float* __restrict__ in,
                                                                      const float* __restrict__ kappa, PriorImage im,
                                                                      PriorWeights W, int wz0, int wz1, float beta,
                                                                      float gamma, float eps, float* __restrict__ out) {
  const int xx = blockIdx.x * kPriorTileX + static_cast<int>(threadIdx.x % kPriorTileX);
  const int y = blockIdx.y * kPriorTileY + static_cast<int>(threadIdx.x / kPriorTileX);
  const int z = blockIdx.z;
  const int nz = im.nz, ny = im.ny, nx = im.nx;
  if (xx >= nx || y >= ny) return;
  const int64_t plane = static_cast<int64_t>(ny) * nx;
  const int64_t v = z * plane + static_cast<int64_t>(y) * nx + xx;
  const float xv = x[v], iv = in[v];
  const float kv = KAPPA ? kappa[v] : 1.F;
  const int min_dz = max(wz0, -z), max_dz = min(wz1, nz - 1 - z);
  const int min_dy = max(-1, -y), max_dy = min(1, ny - 1 - y);
  const int min_dx = max(-1, -xx), max_dx = min(1, nx - 1 - xx);
  float result = 0.F;
  for (int dz = min_dz; dz <= max_dz; ++dz)
    for (int dy = min_dy; dy <= max_dy; ++dy)
      for (int dx = min_dx; dx <= max_dx; ++dx) {
        const int64_t n = v + dz * plane + static_cast<int64_t>(dy) * nx + dx;
        float current = W.w[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
        if constexpr (KIND == 3) {
          current = current * in[n];
        } else {
          if (current == 0.F) continue;
          const float xn = x[n];
          const float f20 = KIND == 1 ? 1.F : rdp_d20(xv, xn, gamma, eps);
          const float f11 = KIND == 1 ? -1.F : rdp_d11(xv, xn, gamma, eps);
          if (dz == 0 && dy == 0 && dx == 0) current *= f20 * iv;
          else current *= (f20 * iv + f11 * in[n]);
        }
        if (KAPPA) current *= kv * kappa[n];
        result += current;
      }
  out[v] = out[v] + result * beta;
}

// Maximum of n floats: per-workgroup maxima into part[blockIdx.x].
__global__ void __launch_bounds__(256) max_partials_kernel(const float* __restrict__ v, int64_t n,
                                                           float* __restrict__ part) {
  float m = -FLT_MAX;
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
    m = fmaxf(m, v[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float red[4];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// *small = max(max(part) * small_num, 0) -- the threshold of divide()
// (numerics/divide.inl:33-34).
__global__ void __launch_bounds__(256) small_value_kernel(const float* __restrict__ part, int n, float small_num,
                                                          float* __restrict__ small) {
  float m = -FLT_MAX;
  for (int i = threadIdx.x; i < n; i += 256) m = fmaxf(m, part[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float red[4];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float sv = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3])) * small_num;
    *small = sv > 0 ? sv : 0.F;
  }
}

// One-step-late update of OSMAPOSL with a prior (OSMAPOSLReconstruction.cxx:403-516)
// on the device layout: denominator from the prior gradient g and the subset
// sensitivity s -- "additive": min/max-clamped g / num_subsets + s into
// [s/10, 10 s]; "multiplicative": s * clamp(1 + g, 0.1, 10) -- then
// divide(u, d, 1e-6) with the precomputed *small, threshold_upper_lower(lo, hi)
// unless the first subiteration, and x *= u. Clamps use std::min / std::max
// semantics.
__global__ void osl_update_kernel(float* __restrict__ x, const float* __restrict__ upd,
                                  const float* __restrict__ pgrad, const float* __restrict__ sens,
                                  const float* __restrict__ small, int nsub, int multiplicative, int threshold,
                                  float lo, float hi, int nz, int nzp, int64_t n) {
  const float sm = *small;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (static_cast<int>(i % nzp) >= nz) continue;
    const float s = sens[i];
    float d = pgrad[i];
    if (!multiplicative) {
      d = d / nsub + s;
      const float hi_d = s * 10, lo_d = s / 10;
      d = (hi_d < d) ? hi_d : d;
      d = (d < lo_d) ? lo_d : d;
    } else {
      d += 1;
      d = (10.F < d) ? 10.F : d;
      d = (d < 1 / 10.F) ? 1 / 10.F : d;
      d *= s;
    }
    float u = upd[i];
    if (fabsf(d) <= sm && fabsf(u) <= sm) u = 0;
    else u = u / d;
    if (threshold) {
      if (u > hi) u = hi;
      else if (lo > u) u = lo;
    }
    x[i] = x[i] * u;
  }
}

// *out = scale * sum(p[0..n)), one workgroup, fixed order: thread t sums
// p[t], p[t+1024], ...; then a pairwise tree over the 1024 thread sums.
__global__ void __launch_bounds__(1024) sum_partials_kernel(const double* __restrict__ p, int n, double scale,
                                                            double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0.;
  for (int i = threadIdx.x; i < n; i += 1024) s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] * scale;
}

// v may alias g or x (same-index element-wise update).
__global__ void osem_vjp_scale_kernel(const float* g, const float* x, const float* __restrict__ sens, float* v,
                                      int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float s = sens[i];
    v[i] = s != 0.F ? g[i] * x[i] / s : 0.F;
  }
}

__global__ void osem_vjp_mask_kernel(const float* __restrict__ g, const float* __restrict__ u, float* __restrict__ out,
                                     int64_t n, float lo, float hi) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float ui = u[i];
    out[i] = (ui > lo && ui < hi) ? g[i] : 0.F;
  }
}

__global__ void add_product_kernel(float* __restrict__ out, const float* __restrict__ g, const float* __restrict__ u,
                                   int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = out[i] + g[i] * u[i];
}

__global__ void accumulate_kernel(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = dst[i] + src[i];
}

__global__ void subtract_kernel(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = dst[i] - src[i];
}

__global__ void divide_scalar_kernel(float* __restrict__ dst, float d, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = dst[i] / d;
}

int elementwise_blocks(int64_t n) { return static_cast<int>(std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace

namespace {
dim3 prior_grid(const PriorImage& im) {
  const int nc = im.zfast ? im.nz : im.nx, nb = im.zfast ? im.nx : im.ny, na = im.zfast ? im.ny : im.nz;
  return dim3((nc + kPriorTileX - 1) / kPriorTileX, (nb + kPriorTileY - 1) / kPriorTileY, na);
}
}  // namespace

int prior_num_partials(const PriorImage& im) {
  const dim3 g = prior_grid(im);
  return static_cast<int>(g.x * g.y * g.z);
}

void launch_prior(int kind, const float* d_img, const float* d_kappa, const PriorImage& im, const PriorWeights& w,
                  int only_2D, float beta, float gamma, float eps, float* d_grad, double* d_partial,
                  hipStream_t stream) {
  const dim3 grid = prior_grid(im);
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
#define MI_PRIOR(K, KA, ZF)                                                                                         \
  hipLaunchKernelGGL((prior_kernel<K, KA, ZF>), grid, dim3(kPriorThreads), 0, stream, d_img, d_kappa, im, w, wz0, wz1, \
                     beta, gamma, eps, d_grad, d_partial)
#define MI_PRIOR_Z(K, KA)                 \
  if (im.zfast) { MI_PRIOR(K, KA, true); } \
  else { MI_PRIOR(K, KA, false); }
#define MI_PRIOR_K(K)                        \
  if (d_kappa) { MI_PRIOR_Z(K, true) }        \
  else { MI_PRIOR_Z(K, false) }
  if (kind == 2) { MI_PRIOR_K(2) }
  else { MI_PRIOR_K(1) }
#undef MI_PRIOR_K
#undef MI_PRIOR_Z
#undef MI_PRIOR
}

void launch_prior_hessian(int kind, const float* d_x, const float* d_in, const float* d_kappa, const PriorImage& im,
                          const PriorWeights& w, int only_2D, float beta, float gamma, float eps, float* d_out,
                          hipStream_t stream) {
  const dim3 grid((im.nx + kPriorTileX - 1) / kPriorTileX, (im.ny + kPriorTileY - 1) / kPriorTileY, im.nz);
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
#define MI_PH(K, KA)                                                                                              \
  hipLaunchKernelGGL((prior_hessian_kernel<K, KA>), grid, dim3(kPriorThreads), 0, stream, d_x, d_in, d_kappa, im, \
                     w, wz0, wz1, beta, gamma, eps, d_out)
#define MI_PH_K(K)                   \
  if (d_kappa) { MI_PH(K, true); }   \
  else { MI_PH(K, false); }
  if (kind == 2) { MI_PH_K(2) }
  else if (kind == 3) { MI_PH_K(3) }
  else { MI_PH_K(1) }
#undef MI_PH_K
#undef MI_PH
}

int max_num_partials() { return 1024; }

void launch_small_value(const float* d_v, int64_t n, float small_num, float* d_part, float* d_small,
                        hipStream_t stream) {
  const int blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, max_num_partials())));
  hipLaunchKernelGGL(max_partials_kernel, dim3(blocks), dim3(256), 0, stream, d_v, n, d_part);
  hipLaunchKernelGGL(small_value_kernel, dim3(1), dim3(256), 0, stream, d_part, blocks, small_num, d_small);
}

void launch_osl_update(float* d_x, const float* d_upd, const float* d_pgrad, const float* d_sens, const float* d_small,
                       int nsub, int multiplicative, int threshold, float lo, float hi, int nz, int nzp, int64_t n,
                       hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(osl_update_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_x, d_upd, d_pgrad, d_sens,
                     d_small, nsub, multiplicative, threshold, lo, hi, nz, nzp, n);
}

void launch_accumulate(float* d_dst, const float* d_src, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(accumulate_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_dst, d_src, n);
}

void launch_subtract(float* d_dst, const float* d_src, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(subtract_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_dst, d_src, n);
}

void launch_divide_scalar(float* d_dst, float d, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(divide_scalar_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_dst, d, n);
}

void launch_sum_partials(const double* d_partial, int n, double scale, double* d_out, hipStream_t stream) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, stream, d_partial, n, scale, d_out);
}

void launch_osem_vjp_scale(const float* d_g, const float* d_x, const float* d_sens, float* d_v, int64_t n,
                           hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(osem_vjp_scale_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_g, d_x, d_sens, d_v, n);
}

void launch_osem_vjp_mask(const float* d_g, const float* d_u, float* d_out, int64_t n, float lo, float hi,
                          hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(osem_vjp_mask_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_g, d_u, d_out, n, lo,
                     hi);
}

void launch_add_product(float* d_out, const float* d_g, const float* d_u, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(add_product_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_out, d_g, d_u, n);
}

// the device address of `scratch` on `device` (the current device), looked
// up on the first call for that device and kept
double* noalloc_partials(int device) {
  static std::mutex m;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(m);
  if (!addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    addr[device] = static_cast<double*>(p);
  }
  return addr[device];
}

}  // namespace mi
