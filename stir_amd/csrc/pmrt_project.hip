// Forward and back projection kernels of the ray-tracing-matrix pair.
//
// Data layout in HBM (device layout):
//   image       [y][x][z]  with z padded to nzp (a column of z values is contiguous)
//   proj data   segment blocks in storage order 0,+1,-1,..., then [view][tang][axial]
//               (the axial positions of one (segment, view, tang) are contiguous)
// Both projections are wave-per-unit gathers with lanes along z / axial
// position, so every image or projection-data access of a wave is one
// contiguous run (coalesced); the matrix elements a wave consumes are
// wave-uniform and come through scalar loads. Summation order per bin is
// the row order (forward) and the table order (back), both fixed, so
// results are deterministic; no float atomics are used.
//
// Compiled with -ffp-contract=off: the forward projection sums
// img * w in row order exactly as ProjMatrixElemsForOneBin::forward_project
// (ProjMatrixElemsForOneBin.cxx:354-371).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "pmrt_device.hpp"

namespace mi {

namespace {

constexpr int kWave = 64;
constexpr int kNumXcd = 8;

// blockIdx -> work index so that each XCD (block i runs on XCD i % 8) gets a
// contiguous range of work: neighbouring families / columns share L2 lines.
__device__ inline int xcd_remap(int bid, int nblocks) {
  const int per = (nblocks + kNumXcd - 1) / kNumXcd;
  const int xcd = bid % kNumXcd;
  const int idx = bid / kNumXcd;
  const int full = nblocks % kNumXcd == 0 ? kNumXcd : nblocks % kNumXcd;  // XCDs that get `per` blocks
  // XCDs [0, full) own `per` blocks, the others `per - 1`
  const int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return start + idx;
}

template <int IP>
__device__ inline void inplane_op(int x, int y, int& xo, int& yo) {
  if constexpr (IP == IP_ID) { xo = x; yo = y; }
  else if constexpr (IP == IP_XMX) { xo = -x; yo = y; }
  else if constexpr (IP == IP_YMY) { xo = x; yo = -y; }
  else if constexpr (IP == IP_XMX_YMY) { xo = -x; yo = -y; }
  else if constexpr (IP == IP_XY_YX) { xo = y; yo = x; }
  else if constexpr (IP == IP_XMY_YX) { xo = -y; yo = x; }
  else if constexpr (IP == IP_XY_YMX) { xo = y; yo = -x; }
  else { xo = -y; yo = -x; }
}

__device__ inline void decode(int c, int& x, int& y, int& z) {
  x = (c & 1023) - 512;
  y = ((c >> 10) & 1023) - 512;
  z = static_cast<int>(static_cast<unsigned>(c) >> 20) - 2048;
}

template <int IP>
__device__ inline float fp_row(const int2* __restrict__ row, int n, int zsign, int zl, bool valid,
                               const float* __restrict__ img, const ImageDims& d) {
  float acc = 0.F;
  int e = 0;
  for (; e < n; ++e) {
    const int2 el = row[e];
    int x, y, z, xo, yo;
    decode(el.x, x, y, z);
    inplane_op<IP>(x, y, xo, yo);
    const int col = (yo - d.min_y) * d.nx + (xo - d.min_x);
    const int zz = zsign * z + zl;
    if (valid && static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz)) {
      const float v = img[static_cast<int64_t>(col) * d.nzp + zz];
      acc = acc + v * __int_as_float(el.y);
    }
  }
  return acc;
}

template <int MODE>
__global__ void __launch_bounds__(kWave) fp_kernel(const FpTask* __restrict__ tasks, int ntasks,
                                                   const Family* __restrict__ fams, const int2* __restrict__ pool,
                                                   const float* __restrict__ img, float* __restrict__ out, ImageDims d,
                                                   FpData data) {
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  if (t >= ntasks) return;
  const int lane = threadIdx.x;
  const FpTask task = tasks[t];
  const Family f = fams[task.family];
  const int a = task.a0 + lane;
  const bool valid = a < f.nax;
  const int zl = f.zadd0 + f.ppa * a - d.min_z;
  const int2* row = pool + f.row_start;
  float acc;
  switch (f.inplane) {
    case IP_ID: acc = fp_row<IP_ID>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_XMX: acc = fp_row<IP_XMX>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_YMY: acc = fp_row<IP_YMY>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_XMX_YMY: acc = fp_row<IP_XMX_YMY>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_XY_YX: acc = fp_row<IP_XY_YX>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_XMY_YX: acc = fp_row<IP_XMY_YX>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    case IP_XY_YMX: acc = fp_row<IP_XY_YMX>(row, f.row_len, f.zsign, zl, valid, img, d); break;
    default: acc = fp_row<IP_XMY_YMX>(row, f.row_len, f.zsign, zl, valid, img, d); break;
  }
  const int64_t i = static_cast<int64_t>(f.sino_base) + a;
  if constexpr (MODE == FP_PLAIN) {
    if (valid) out[i] = acc;
  } else if constexpr (MODE == FP_RATIO || MODE == FP_GRADIENT) {
    if (valid) {
      // estimated += additive; divide_and_truncate (recon_array_functions.cxx:172-266)
      const float ybar = acc + (data.additive ? data.additive[i] : data.additive_constant);
      float num = data.measured[i];
      const float small_value = data.vg_small[f.viewgram];
      if (num <= small_value) {
        num = 0;
      } else {
        const float max_quotient = 10000.F;
        if (num > max_quotient * ybar) num = max_quotient;
        else num = num / ybar;
      }
      if constexpr (MODE == FP_GRADIENT) num = num - (data.mult ? data.mult[i] : 1.F);
      out[i] = num;
    }
  } else {  // FP_LOGLIK: accumulate_loglikelihood (recon_array_functions.cxx:336-377)
    double v = 0.0;
    if (valid) {
      float est = acc + (data.additive ? data.additive[i] : data.additive_constant);
      if (data.mult) est = est * data.mult[i];
      const float y = data.measured[i];
      const float small_value = data.vg_small[f.viewgram];
      const float ne = (est < y / 10000.F) ? y / 10000.F : est;
      if (y <= small_value) v = -static_cast<double>(ne);
      else v = y * log(static_cast<double>(ne)) - static_cast<double>(ne);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (lane == 0) data.partial[t] = v;
  }
}

// Voxel-driven back projection: one wave per (image column (y,x), chunk of
// 64 z values), lanes along z. The entries of the column's table are
// wave-uniform and the projection data each entry reads are one contiguous run.
template <int MODE>
__global__ void __launch_bounds__(kWave) bp_kernel(const int64_t* __restrict__ list_off, const BpEntry* __restrict__ ent,
                                                   int sb, int se, const float* __restrict__ proj,
                                                   float* __restrict__ img, ImageDims d, int nzch, BpData data) {
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int col = w / nzch;
  if (col >= d.ncols) return;
  const int z = (w - col * nzch) * kWave + static_cast<int>(threadIdx.x);
  const int zi = d.min_z + z;
  float acc = 0.F;
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    for (int64_t e = beg; e < end; ++e) {
      const BpEntry E = ent[e];
      const int nax = E.nax_ppa & 0xffff;
      const int ppa = E.nax_ppa >> 16;
      const int diff = zi - E.k;
      int a;
      bool ok;
      if (ppa == 1) {
        a = diff;
        ok = static_cast<unsigned>(a) < static_cast<unsigned>(nax);
      } else if (ppa == 2) {
        a = diff >> 1;
        ok = diff >= 0 && (diff & 1) == 0 && a < nax;
      } else {
        a = diff / ppa;
        ok = diff >= 0 && a * ppa == diff && a < nax;
      }
      if (ok) acc = acc + E.w * proj[static_cast<int64_t>(E.sino_base) + a];
    }
  }
  if (z >= d.nzp) return;
  const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
  if constexpr (MODE == BP_PLAIN) {
    img[idx] = z < d.nz ? acc : 0.F;
  } else {
    if (z < d.nz) {
      // divide(update, sens, 0) (numerics/divide.inl:26-47), threshold, multiply
      float u = acc;
      const float sv = data.sens[idx];
      if (fabsf(sv) <= 0.F && fabsf(u) <= 0.F) u = 0;
      else u = u / sv;
      if (data.threshold) {
        if (u > FLT_MAX) u = FLT_MAX;
        else if (0.F > u) u = 0.F;
      }
      data.estimate[idx] = data.estimate[idx] * u;
    }
  }
}

// dst[c * dst_ld + r] = src[r * src_ld + c], 32x32 tiles through LDS
__global__ void __launch_bounds__(256) transpose_kernel(const float* __restrict__ src, float* __restrict__ dst, int rows,
                                                        int cols, int src_ld, int dst_ld) {
  __shared__ float tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) tile[k][threadIdx.x] = src[static_cast<int64_t>(r) * src_ld + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) dst[static_cast<int64_t>(c) * dst_ld + r] = tile[threadIdx.x][k];
  }
}

__global__ void fill_kernel(float* d, int64_t n, float v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    d[i] = v;
}

// small_value = max(max(viewgram) * 1e-6, 0) per viewgram (recon_array_functions.cxx:184)
__global__ void __launch_bounds__(256) viewgram_max_kernel(const float* __restrict__ proj, const int64_t* __restrict__ off,
                                                           const int32_t* __restrict__ len, float* __restrict__ small) {
  __shared__ float red[256];
  const int vg = blockIdx.x;
  const float* p = proj + off[vg];
  const int n = len[vg];
  float m = -FLT_MAX;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, p[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float v = red[0] * 0.000001F;
    small[vg] = (0.F < v) ? v : 0.F;
  }
}

// zero the first and last axial position of `count` runs of length `run`
__global__ void zero_end_planes_kernel(float* proj, const int64_t* off, int count, int run) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  proj[off[0] + static_cast<int64_t>(i) * run] = 0.F;
  proj[off[0] + static_cast<int64_t>(i) * run + run - 1] = 0.F;
}

}  // namespace

void launch_forward_project(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams, const int2* d_pool,
                            const float* d_img, float* d_out, const ImageDims& dims, const FpData& data,
                            hipStream_t stream) {
  if (ntasks <= 0) return;
  const dim3 grid(ntasks), block(kWave);
  switch (mode) {
    case FP_PLAIN:
      hipLaunchKernelGGL(fp_kernel<FP_PLAIN>, grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img, d_out, dims, data);
      break;
    case FP_RATIO:
      hipLaunchKernelGGL(fp_kernel<FP_RATIO>, grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img, d_out, dims, data);
      break;
    case FP_GRADIENT:
      hipLaunchKernelGGL(fp_kernel<FP_GRADIENT>, grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img, d_out, dims,
                         data);
      break;
    case FP_LOGLIK:
      hipLaunchKernelGGL(fp_kernel<FP_LOGLIK>, grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img, d_out, dims, data);
      break;
  }
}

void launch_back_project(BpMode mode, const int64_t* d_list_off, const BpEntry* d_entries, int subset_begin,
                         int subset_end, const float* d_proj, float* d_img, const ImageDims& dims, const BpData& data,
                         hipStream_t stream) {
  const int nzch = (dims.nzp + kWave - 1) / kWave;
  const dim3 grid(static_cast<unsigned>(dims.ncols) * nzch), block(kWave);
  if (mode == BP_PLAIN)
    hipLaunchKernelGGL(bp_kernel<BP_PLAIN>, grid, block, 0, stream, d_list_off, d_entries, subset_begin, subset_end, d_proj,
                       d_img, dims, nzch, data);
  else
    hipLaunchKernelGGL(bp_kernel<BP_OSEM_UPDATE>, grid, block, 0, stream, d_list_off, d_entries, subset_begin, subset_end,
                       d_proj, d_img, dims, nzch, data);
}

void launch_transpose(const float* src, float* dst, int rows, int cols, int src_ld, int dst_ld, hipStream_t stream) {
  if (rows <= 0 || cols <= 0) return;
  const dim3 grid((cols + 31) / 32, (rows + 31) / 32), block(32, 8);
  hipLaunchKernelGGL(transpose_kernel, grid, block, 0, stream, src, dst, rows, cols, src_ld, dst_ld);
}

void launch_fill(float* d, int64_t n, float v, hipStream_t stream) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, d, n, v);
}

void launch_viewgram_max(const float* d_proj, const int64_t* d_vg_off, const int32_t* d_vg_len, int nvg, float* d_small,
                         hipStream_t stream) {
  if (nvg <= 0) return;
  hipLaunchKernelGGL(viewgram_max_kernel, dim3(nvg), dim3(256), 0, stream, d_proj, d_vg_off, d_vg_len, d_small);
}

void launch_zero_end_planes(float* d_proj, const int64_t* d_off, int count, int run, int stride, hipStream_t stream) {
  (void)stride;
  if (count <= 0) return;
  hipLaunchKernelGGL(zero_end_planes_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, d_proj, d_off, count, run);
}

}  // namespace mi
