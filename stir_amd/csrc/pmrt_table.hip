// Builds the voxel-driven back-projection table: for every (subset, image
// column) the list of (family run, z offset, weight) entries whose matrix
// elements fall in that column. The entries are generated in family order,
// element order, and then stably sorted by (subset, column); the resulting
// per-column order is therefore fixed, which makes the back projection
// deterministic run to run.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "pmrt_device.hpp"

namespace mi {

namespace {

__device__ inline void decode(int c, int& x, int& y, int& z) {
  x = (c & 1023) - 512;
  y = ((c >> 10) & 1023) - 512;
  z = static_cast<int>(static_cast<unsigned>(c) >> 20) - 2048;
}

__device__ inline void inplane_any(int ip, int x, int y, int& xo, int& yo) {
  switch (ip) {
    case IP_ID: xo = x; yo = y; break;
    case IP_XMX: xo = -x; yo = y; break;
    case IP_YMY: xo = x; yo = -y; break;
    case IP_XMX_YMY: xo = -x; yo = -y; break;
    case IP_XY_YX: xo = y; yo = x; break;
    case IP_XMY_YX: xo = -y; yo = x; break;
    case IP_XY_YMX: xo = y; yo = -x; break;
    default: xo = -y; yo = -x; break;
  }
}

// one wave per family, lanes over the elements of its basic row
__global__ void __launch_bounds__(64) count_entries_kernel(const Family* __restrict__ fams, int fam_begin, int fam_end,
                                                           const int2* __restrict__ pool, int subset, ImageDims d,
                                                           int32_t* __restrict__ counts) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    int x, y, z, xo, yo;
    decode(pool[f.row_start + e].x, x, y, z);
    inplane_any(f.inplane, x, y, xo, yo);
    const int col = (yo - d.min_y) * d.nx + (xo - d.min_x);
    atomicAdd(&counts[static_cast<int64_t>(subset) * d.ncols + col], 1);
  }
}

__global__ void __launch_bounds__(64) make_entries_kernel(const Family* __restrict__ fams, int fam_begin, int fam_end,
                                                          const int64_t* __restrict__ fam_entry_off,
                                                          const int2* __restrict__ pool, int subset, ImageDims d,
                                                          uint32_t* __restrict__ keys, BpEntry* __restrict__ ent) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  const int64_t base = fam_entry_off[fi - fam_begin];
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    const int2 el = pool[f.row_start + e];
    int x, y, z, xo, yo;
    decode(el.x, x, y, z);
    inplane_any(f.inplane, x, y, xo, yo);
    const int col = (yo - d.min_y) * d.nx + (xo - d.min_x);
    keys[base + e] = static_cast<uint32_t>(subset) * static_cast<uint32_t>(d.ncols) + static_cast<uint32_t>(col);
    BpEntry b;
    b.sino_base = f.sino_base;
    b.k = f.zsign * z + f.zadd0;
    b.nax_ppa = static_cast<int32_t>(f.nax) | (static_cast<int32_t>(f.ppa) << 16);
    b.w = __int_as_float(el.y);
    ent[base + e] = b;
  }
}

// number of (bin, element) pairs whose voxel z lies inside the image:
// a in [0, nax) with 0 <= zsign*z + zadd0 + ppa*a - min_z < nz
__global__ void __launch_bounds__(64) count_nonzeros_kernel(const Family* __restrict__ fams, int nfam,
                                                            const int2* __restrict__ pool, ImageDims d,
                                                            unsigned long long* __restrict__ total) {
  const int fi = blockIdx.x;
  if (fi >= nfam) return;
  const Family f = fams[fi];
  unsigned long long cnt = 0;
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    int x, y, z;
    decode(pool[f.row_start + e].x, x, y, z);
    const int k = f.zsign * z + f.zadd0 - d.min_z;  // image z of axial position 0
    // ppa*a in [-k, nz-1-k]
    const int lo_num = -k, hi_num = d.nz - 1 - k;
    int lo = lo_num <= 0 ? 0 : (lo_num + f.ppa - 1) / f.ppa;
    int hi = hi_num < 0 ? -1 : hi_num / f.ppa;
    if (hi > f.nax - 1) hi = f.nax - 1;
    if (hi >= lo) cnt += static_cast<unsigned long long>(hi - lo + 1);
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if (threadIdx.x == 0) atomicAdd(total, cnt);
}

__global__ void iota_kernel(uint32_t* v, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    v[i] = static_cast<uint32_t>(i);
}

__global__ void gather_kernel(const BpEntry* __restrict__ src, const uint32_t* __restrict__ idx, BpEntry* __restrict__ dst,
                              int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[idx[i]];
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

}  // namespace

void launch_count_entries(const Family* d_fams, int fam_begin, int fam_end, const int2* d_pool, int subset,
                          const ImageDims& dims, int32_t* d_counts, hipStream_t stream) {
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(count_entries_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, fam_begin, fam_end, d_pool,
                     subset, dims, d_counts);
}

void launch_make_entries(const Family* d_fams, int fam_begin, int fam_end, const int64_t* d_fam_entry_off,
                         const int2* d_pool, int subset, const ImageDims& dims, uint32_t* d_keys, BpEntry* d_entries,
                         hipStream_t stream) {
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(make_entries_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, fam_begin, fam_end,
                     d_fam_entry_off, d_pool, subset, dims, d_keys, d_entries);
}

void launch_count_nonzeros(const Family* d_fams, int nfam, const int2* d_pool, const ImageDims& dims,
                           unsigned long long* d_total, hipStream_t stream) {
  if (nfam <= 0) return;
  hipLaunchKernelGGL(count_nonzeros_kernel, dim3(nfam), dim3(64), 0, stream, d_fams, nfam, d_pool, dims, d_total);
}

void launch_iota(uint32_t* d_vals, int64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(256), 0, stream, d_vals, n);
}

void launch_gather_entries(const BpEntry* d_src, const uint32_t* d_idx, BpEntry* d_dst, int64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n)), dim3(256), 0, stream, d_src, d_idx, d_dst, n);
}

// Stable key/value radix sort (deterministic) and exclusive scan, from hipCUB.
hipError_t radix_sort_pairs(void* d_temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, int n, int end_bit, hipStream_t stream) {
  return hipcub::DeviceRadixSort::SortPairs(d_temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, end_bit, stream);
}

hipError_t exclusive_scan_counts(void* d_temp, size_t& temp_bytes, const int32_t* in, int64_t* out, int n,
                                 hipStream_t stream) {
  return hipcub::DeviceScan::ExclusiveSum(d_temp, temp_bytes, in, out, n, stream);
}

hipError_t exclusive_scan_u32(void* d_temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, int n,
                              hipStream_t stream) {
  return hipcub::DeviceScan::ExclusiveSum(d_temp, temp_bytes, in, out, n, stream);
}

namespace {
__global__ void fill_int2_kernel(int2* d, int64_t n, int2 v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    d[i] = v;
}
}  // namespace

void launch_fill_int2(int2* d, int64_t n, int2 v, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_int2_kernel, dim3(grid_for(n)), dim3(256), 0, stream, d, n, v);
}

}  // namespace mi
