// Objective-function kernels: the quadratic neighbour prior, a fixed-order
// FP64 reduction of per-workgroup partial sums, and the element-wise pieces
// of the OSEM vector-Jacobian product.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pmrt_device.hpp"

namespace mi {
namespace {

constexpr int kPriorTileX = 64;
constexpr int kPriorTileY = 4;
constexpr int kPriorThreads = kPriorTileX * kPriorTileY;

// QuadraticPrior::compute_value / compute_gradient (QuadraticPrior.cxx:229-367)
// for one voxel per thread, image [nz][ny][nx]. Neighbours outside the image
// are skipped; each term is evaluated in float as the reference does
// (w * (a-b)^2 / 4 and w * (a-b)) and accumulated in double in (dz, dy, dx)
// order. Workgroup = 64 x 4 voxels of one plane; its value partial is written
// to partial[block].
template <bool KAPPA>
__global__ void __launch_bounds__(kPriorThreads) quadratic_prior_kernel(const float* __restrict__ img,
                                                                        const float* __restrict__ kappa, int nz, int ny,
                                                                        int nx, PriorWeights W, int wz0, int wz1,
                                                                        float beta, float* __restrict__ grad,
                                                                        double* __restrict__ partial) {
  const int x = blockIdx.x * kPriorTileX + static_cast<int>(threadIdx.x % kPriorTileX);
  const int y = blockIdx.y * kPriorTileY + static_cast<int>(threadIdx.x / kPriorTileX);
  const int z = blockIdx.z;
  double value = 0.;
  if (x < nx && y < ny) {
    const int64_t plane = static_cast<int64_t>(ny) * nx;
    const int64_t v = z * plane + static_cast<int64_t>(y) * nx + x;
    const float xv = img[v];
    const float kv = KAPPA ? kappa[v] : 1.F;
    const int min_dz = max(wz0, -z), max_dz = min(wz1, nz - 1 - z);
    const int min_dy = max(-1, -y), max_dy = min(1, ny - 1 - y);
    const int min_dx = max(-1, -x), max_dx = min(1, nx - 1 - x);
    double g = 0.;
    for (int dz = min_dz; dz <= max_dz; ++dz)
      for (int dy = min_dy; dy <= max_dy; ++dy)
        for (int dx = min_dx; dx <= max_dx; ++dx) {
          const int64_t n = v + dz * plane + static_cast<int64_t>(dy) * nx + dx;
          const float w = W.w[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
          const float d = xv - img[n];
          double cv = static_cast<double>(w * (d * d) / 4);
          double cg = static_cast<double>(w * d);
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

__global__ void add_product_kernel(float* __restrict__ out, const float* __restrict__ g, const float* __restrict__ u,
                                   int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = out[i] + g[i] * u[i];
}

int elementwise_blocks(int64_t n) { return static_cast<int>(std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace

int prior_num_partials(int nz, int ny, int nx) {
  return ((nx + kPriorTileX - 1) / kPriorTileX) * ((ny + kPriorTileY - 1) / kPriorTileY) * nz;
}

void launch_quadratic_prior(const float* d_img, const float* d_kappa, int nz, int ny, int nx, const PriorWeights& w,
                            int only_2D, float beta, float* d_grad, double* d_partial, hipStream_t stream) {
  const dim3 grid((nx + kPriorTileX - 1) / kPriorTileX, (ny + kPriorTileY - 1) / kPriorTileY, nz);
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
  if (d_kappa)
    hipLaunchKernelGGL(quadratic_prior_kernel<true>, grid, dim3(kPriorThreads), 0, stream, d_img, d_kappa, nz, ny, nx, w,
                       wz0, wz1, beta, d_grad, d_partial);
  else
    hipLaunchKernelGGL(quadratic_prior_kernel<false>, grid, dim3(kPriorThreads), 0, stream, d_img, d_kappa, nz, ny, nx,
                       w, wz0, wz1, beta, d_grad, d_partial);
}

void launch_sum_partials(const double* d_partial, int n, double scale, double* d_out, hipStream_t stream) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, stream, d_partial, n, scale, d_out);
}

void launch_osem_vjp_scale(const float* d_g, const float* d_x, const float* d_sens, float* d_v, int64_t n,
                           hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(osem_vjp_scale_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_g, d_x, d_sens, d_v, n);
}

void launch_add_product(float* d_out, const float* d_g, const float* d_u, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(add_product_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_out, d_g, d_u, n);
}

}  // namespace mi
