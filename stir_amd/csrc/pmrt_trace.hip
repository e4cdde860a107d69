// Device row builder: traces the basic rows of the ray-tracing matrix.
//
// One thread per basic bin. Each thread owns three scratch buffers of `cap`
// elements (key = packed z,y,x sorted lexicographically; value = weight) and
// reproduces, in the reference's float evaluation order:
//   RayTraceVoxelsOnCartesianGrid      RayTraceVoxelsOnCartesianGrid.cxx:52-269
//   ray_trace_one_lor                  ProjMatrixByBinUsingRayTracing.cxx:438-579
//   multiple tangential LORs + merge   ProjMatrixByBinUsingRayTracing.cxx:755-782,
//                                      ProjMatrixElemsForOneBin.cxx:162-277
//   add_adjacent_z / z+1 copies        ProjMatrixByBinUsingRayTracing.cxx:785-954
// This file must be compiled with -ffp-contract=off and without fast-math.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmrt_device.hpp"

namespace mi {

namespace {

struct Buf {
  uint64_t* k;
  float* w;
};

struct Ctx {
  int cap;
  int overflow;
};

__device__ inline uint64_t pack_key(int z, int y, int x) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(z + 32768)) << 32) |
         (static_cast<uint64_t>(static_cast<uint32_t>(y + 32768)) << 16) | static_cast<uint64_t>(static_cast<uint32_t>(x + 32768));
}
__device__ inline int key_z(uint64_t k) { return static_cast<int>((k >> 32) & 0xffff) - 32768; }
__device__ inline int key_y(uint64_t k) { return static_cast<int>((k >> 16) & 0xffff) - 32768; }
__device__ inline int key_x(uint64_t k) { return static_cast<int>(k & 0xffff) - 32768; }

__device__ inline int stir_round(float x) { return x >= 0 ? static_cast<int>(x + 0.5F) : -static_cast<int>(-x + 0.5F); }
__device__ inline float fmin_first(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ inline float fmax_first(float a, float b) { return (a < b) ? b : a; }  // std::max

// float inner product (z,y,x order), then double sqrt (BasicCoordinate.inl:408-429)
__device__ inline double norm3(float z, float y, float x) {
  float acc = 0.F;
  acc = acc + z * z;
  acc = acc + y * y;
  acc = acc + x * x;
  return sqrt(static_cast<double>(acc));
}
__device__ inline bool half_integer(float a) { return fabsf(floorf(a) + .5F - a) < .0001F; }

__device__ inline void push(Buf b, int& n, Ctx& c, int z, int y, int x, float w) {
  if (n >= c.cap) {
    c.overflow = 1;
    return;
  }
  b.k[n] = pack_key(z, y, x);
  b.w[n] = w;
  ++n;
}

// bottom-up merge sort of a[0,n) by key, tmp has room for n elements
__device__ void sort_rows(Buf a, int n, Buf tmp) {
  if (n < 2) return;
  Buf src = a, dst = tmp;
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      const int mid = min(lo + width, n), hi = min(lo + 2 * width, n);
      int i = lo, j = mid, o = lo;
      while (i < mid && j < hi) {
        if (src.k[j] < src.k[i]) { dst.k[o] = src.k[j]; dst.w[o++] = src.w[j++]; }
        else { dst.k[o] = src.k[i]; dst.w[o++] = src.w[i++]; }
      }
      while (i < mid) { dst.k[o] = src.k[i]; dst.w[o++] = src.w[i++]; }
      while (j < hi) { dst.k[o] = src.k[j]; dst.w[o++] = src.w[j++]; }
    }
    Buf t = src; src = dst; dst = t;
  }
  if (src.k != a.k)
    for (int i = 0; i < n; ++i) { a.k[i] = src.k[i]; a.w[i] = src.w[i]; }
}

// ProjMatrixElemsForOneBin::merge: sort both, sorted union with duplicates summed
// (this + other). Result in out. Returns early (no sort) when b is empty.
__device__ int merge_rows(Buf a, int na, Buf b, int nb, Buf out, Ctx& c) {
  if (nb == 0) {
    for (int i = 0; i < na; ++i) { out.k[i] = a.k[i]; out.w[i] = a.w[i]; }
    return na;
  }
  sort_rows(a, na, out);
  sort_rows(b, nb, out);
  int i = 0, j = 0, o = 0;
  while (j < nb) {
    if (o >= c.cap) { c.overflow = 1; return 0; }
    if (i == na) { out.k[o] = b.k[j]; out.w[o++] = b.w[j++]; continue; }
    if (b.k[j] == a.k[i]) { out.k[o] = a.k[i]; out.w[o++] = a.w[i++] + b.w[j++]; }
    else if (b.k[j] < a.k[i]) { out.k[o] = b.k[j]; out.w[o++] = b.w[j++]; }
    else { out.k[o] = a.k[i]; out.w[o++] = a.w[i++]; }
  }
  while (i < na) {
    if (o >= c.cap) { c.overflow = 1; return 0; }
    out.k[o] = a.k[i]; out.w[o++] = a.w[i++];
  }
  return o;
}

// one Siddon trace without the in-plane split (RayTraceVoxelsOnCartesianGrid.cxx:59-101,138-268)
__device__ void siddon(Buf out, int& n, Ctx& c, float sz, float sy, float sx, float ez, float ey, float ex, float vz,
                       float vy, float vx, float nc) {
  const float dz = ez - sz, dy = ey - sy, dx = ex - sx;
  if (norm3(dz, dy, dx) <= .00001F) return;
  const float d12 = static_cast<float>(norm3(dz * vz, dy * vy, dx * vx) * static_cast<double>(nc));
  const int sign_x = dx >= 0 ? 1 : -1;
  const int sign_y = dy >= 0 ? 1 : -1;
  const int sign_z = dz >= 0 ? 1 : -1;
  const float small_difference = 1.E-4F;
  const bool zx = fabsf(dx) <= small_difference;
  const bool zy = fabsf(dy) <= small_difference;
  const bool zz = fabsf(dz) <= small_difference;
  const float inc_x = zx ? d12 * 1000000.F : d12 / fabsf(dx);
  const float inc_y = zy ? d12 * 1000000.F : d12 / fabsf(dy);
  const float inc_z = zz ? d12 * 1000000.F : d12 / fabsf(dz);
  const float xmax = stir_round(ex) + sign_x * 0.5F;
  const float ymax = stir_round(ey) + sign_y * 0.5F;
  const float zmax = stir_round(ez) + sign_z * 0.5F;
  const float axend = zx ? d12 * 1000000.F : (xmax - sx) * inc_x * sign_x * .9999F;
  const float ayend = zy ? d12 * 1000000.F : (ymax - sy) * inc_y * sign_y * .9999F;
  const float azend = zz ? d12 * 1000000.F : (zmax - sz) * inc_z * sign_z * .9999F;
  const float amax = fmin_first(axend, fmin_first(ayend, azend));
  int cz = stir_round(sz), cy = stir_round(sy), cx = stir_round(sx);
  float az = zz ? -inc_z : ((cz - sz) - sign_z * 0.5F) * inc_z * sign_z;
  float ax = zx ? -inc_x : ((cx - sx) - sign_x * 0.5F) * inc_x * sign_x;
  float ay = zy ? -inc_y : ((cy - sy) - sign_y * 0.5F) * inc_y * sign_y;
  float a = fmax_first(ax, fmax_first(ay, az));
  if (zx) ax = axend; else ax += inc_x;
  if (zy) ay = ayend; else ay += inc_y;
  if (zz) az = azend; else az += inc_z;
  while (a < amax && !c.overflow) {
    if (ax < ay) {
      if (ax < az) { push(out, n, c, cz, cy, cx, ax - a); a = ax; ax += inc_x; cx += sign_x; }
      else { push(out, n, c, cz, cy, cx, az - a); a = az; az += inc_z; cz += sign_z; }
    } else if (ay < az) { push(out, n, c, cz, cy, cx, ay - a); a = ay; ay += inc_y; cy += sign_y; }
    else { push(out, n, c, cz, cy, cx, az - a); a = az; az += inc_z; cz += sign_z; }
  }
}

// RayTraceVoxelsOnCartesianGrid with its split for rays lying in a plane
// between voxels (:103-131). The recursion splits at most once per axis, in
// the order z, y, x; each leaf traces the points shifted by +-0.5 in the split
// axes with the normalisation halved per split, and the union is sorted.
__device__ void ray_trace(Buf out, int& n, Buf tmp, Ctx& c, float sz, float sy, float sx, float ez, float ey, float ex,
                          float vz, float vy, float vx, float nc) {
  const float dz = ez - sz, dy = ey - sy, dx = ex - sx;
  if (norm3(dz, dy, dx) <= .00001F) return;
  const float small_difference = 1.E-4F;
  const bool split_z = fabsf(dz) <= small_difference && half_integer(sz);
  const bool split_y = fabsf(dy) <= small_difference && half_integer(sy);
  const bool split_x = fabsf(dx) <= small_difference && half_integer(sx);
  if (!split_z && !split_y && !split_x) {
    siddon(out, n, c, sz, sy, sx, ez, ey, ex, vz, vy, vx, nc);
    return;
  }
  const int nzs = split_z ? 2 : 1, nys = split_y ? 2 : 1, nxs = split_x ? 2 : 1;
  float ncl = nc;
  if (split_z) ncl = ncl / 2;
  if (split_y) ncl = ncl / 2;
  if (split_x) ncl = ncl / 2;
  for (int iz = 0; iz < nzs; ++iz)
    for (int iy = 0; iy < nys; ++iy)
      for (int ix = 0; ix < nxs; ++ix) {
        const float oz = split_z ? (iz ? .5F : -.5F) : 0.F;
        const float oy = split_y ? (iy ? .5F : -.5F) : 0.F;
        const float ox = split_x ? (ix ? .5F : -.5F) : 0.F;
        siddon(out, n, c, sz + oz, sy + oy, sx + ox, ez + oz, ey + oy, ex + ox, vz, vy, vx, ncl);
      }
  sort_rows(out, n, tmp);
}

// ray_trace_one_lor (ProjMatrixByBinUsingRayTracing.cxx:438-579)
__device__ void trace_one_lor(Buf out, int& n, Buf tmp, Ctx& c, float s, float t, float cphi, float sphi, float costheta,
                              float tantheta, float offset_in_z, const TraceConsts& tc, int num_lors) {
  float max_a, min_a;
  const float fovrad = tc.fovrad;
  if (tc.restrict_cyl) {
    if (fabsf(s) > fovrad) return;
    if (fabsf(s) == fovrad) { max_a = min_a = 0; }
    else { max_a = sqrtf(fovrad * fovrad - s * s); min_a = -max_a; }
  } else {
    if (fabsf(cphi) < 1.E-3 || fabsf(sphi) < 1.E-3) {
      if (fovrad < fabsf(s)) return;
      max_a = fovrad;
      min_a = -fovrad;
    } else {
      const int sgs = sphi < 0 ? -1 : 1, sgc = cphi < 0 ? -1 : 1;
      max_a = fmin_first((fovrad * sgs - s * cphi) / sphi, (fovrad * sgc + s * sphi) / cphi);
      min_a = fmax_first((-fovrad * sgs - s * cphi) / sphi, (-fovrad * sgc + s * sphi) / cphi);
      if (min_a > max_a - 1.E-3 * tc.vx) return;
    }
  }
  const float st_x = (s * cphi + max_a * sphi) / tc.vx;
  const float st_y = (s * sphi - max_a * cphi) / tc.vy;
  const float st_z = (t / costheta + offset_in_z - max_a * tantheta) / tc.vz;
  const float sp_x = (s * cphi + min_a * sphi) / tc.vx;
  const float sp_y = (s * sphi - min_a * cphi) / tc.vy;
  const float sp_z = (t / costheta + offset_in_z - min_a * tantheta) / tc.vz;
  const bool fwd = st_z < sp_z || (st_z == sp_z && (st_y < sp_y || (st_y == sp_y && st_x <= sp_x)));
  const float nc = 1 / tc.vx / num_lors;
  if (fwd) ray_trace(out, n, tmp, c, st_z, st_y, st_x, sp_z, sp_y, sp_x, tc.vz, tc.vy, tc.vx, nc);
  else ray_trace(out, n, tmp, c, sp_z, sp_y, sp_x, st_z, st_y, st_x, tc.vz, tc.vy, tc.vx, nc);
}

// add_adjacent_z (ProjMatrixByBinUsingRayTracing.cxx:832-954), in place
__device__ int add_adjacent_z(Buf lor, int n, Ctx& c, float zf, float right) {
  const int org = n;
  const int z0 = key_z(lor.k[0]);
  for (int zi = 1;; ++zi) {
    const float ov = fmin_first(right, zf + zi + .5F) - fmax_first(0.F, zf + zi - .5F);
    if (ov <= 0.0001) break;
    const uint64_t zkey = static_cast<uint64_t>(static_cast<uint32_t>(z0 + zi + 32768)) << 32;
    for (int k = 0; k < org; ++k) {
      if (n >= c.cap) { c.overflow = 1; return n; }
      lor.k[n] = (lor.k[k] & 0xffffffffull) | zkey;
      lor.w[n++] = ov > .9999 ? lor.w[k] : lor.w[k] * ov;
    }
  }
  if (zf > 0.5) {
    const float ov = zf - .5F;
    const uint64_t zkey = static_cast<uint64_t>(static_cast<uint32_t>(z0 - 1 + 32768)) << 32;
    for (int k = 0; k < org; ++k) {
      if (n >= c.cap) { c.overflow = 1; return n; }
      lor.k[n] = (lor.k[k] & 0xffffffffull) | zkey;
      lor.w[n++] = ov > .9999 ? lor.w[k] : lor.w[k] * ov;
    }
  }
  {
    const float ov = fmin_first(right, zf + .5F) - fmax_first(0.F, zf - .5F);
    if (ov < .9999)
      for (int k = 0; k < org; ++k) lor.w[k] *= ov;
  }
  return n;
}

}  // namespace

// calculate_proj_matrix_elems_for_one_bin (ProjMatrixByBinUsingRayTracing.cxx:582-829)
__global__ void __launch_bounds__(64) trace_basic_rows_kernel(const BasicRay* __restrict__ rays, int first, int count,
                                                              TraceConsts tc, uint64_t* scratch_k, float* scratch_w,
                                                              int cap, int* out_len, int* out_buf, int* overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const BasicRay r = rays[first + i];
  const size_t base = static_cast<size_t>(i) * 3 * cap;
  Buf buf[3] = {{scratch_k + base, scratch_w + base},
                {scratch_k + base + cap, scratch_w + base + cap},
                {scratch_k + base + 2 * static_cast<size_t>(cap), scratch_w + base + 2 * static_cast<size_t>(cap)}};
  int ia = 0, ib = 1, ic = 2;  // lor, second operand, free
  Ctx c{cap, 0};
  int n = 0;
  if (tc.num_tlors == 1) {
    trace_one_lor(buf[ia], n, buf[ic], c, r.s, r.t, r.cphi, r.sphi, r.costheta, r.tantheta, r.offset_in_z, tc, r.num_lors);
  } else {
    const int nt = tc.num_tlors;
    float cur = r.s - r.s_inc * (nt - 1) / 2.F;
    for (int k = 1; k <= nt && !c.overflow; ++k, cur += r.s_inc) {
      int m = 0;
      trace_one_lor(buf[ib], m, buf[ic], c, cur, r.t, r.cphi, r.sphi, r.costheta, r.tantheta, r.offset_in_z, tc,
                    r.num_lors * nt);
      if (m == 0) continue;  // merge returns early without touching lor
      n = merge_rows(buf[ia], n, buf[ib], m, buf[ic], c);
      const int t = ia; ia = ic; ic = t;
    }
  }
  if (n > 0 && !c.overflow) {
    if (r.tantheta == 0) {
      const float zf = key_z(buf[ia].k[0]) + tc.origin_z_over_vz - tc.zmid;
      const float left = (r.t - r.samp_z / 2) / tc.vz;
      const float right = (r.t + r.samp_z / 2) / tc.vz;
      n = add_adjacent_z(buf[ia], n, c, zf - left, right - left);
    } else if (r.num_lors > 1) {
      for (int k = 0; k < n; ++k) { buf[ib].k[k] = buf[ia].k[k]; buf[ib].w[k] = buf[ia].w[k]; }
      const int m = n;
      for (int zi = 1; zi < r.num_lors && !c.overflow; ++zi) {
        for (int k = 0; k < m; ++k) buf[ib].k[k] += static_cast<uint64_t>(1) << 32;
        n = merge_rows(buf[ia], n, buf[ib], m, buf[ic], c);
        const int t = ia; ia = ic; ic = t;
      }
    }
  }
  if (c.overflow) {
    atomicMax(overflow, 1);
    out_len[first + i] = 0;
    out_buf[first + i] = ia;
    return;
  }
  out_len[first + i] = n;
  out_buf[first + i] = ia;
}

// Packs the traced rows into the element pool: {coords, weight} with
// coords = (x+512) | (y+512)<<10 | (z+2048)<<20.
__global__ void pack_rows_kernel(int first, int count, const uint64_t* scratch_k, const float* scratch_w, int cap,
                                 const int* lens, const int* bufs, const int64_t* row_start, int2* pool, int* range_error,
                                 int4 lim) {
  const int i = blockIdx.x;
  if (i >= count) return;
  const int n = lens[first + i];
  const size_t base = (static_cast<size_t>(i) * 3 + bufs[first + i]) * cap;
  const int64_t out = row_start[first + i];
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const uint64_t key = scratch_k[base + k];
    const int z = key_z(key), y = key_y(key), x = key_x(key);
    // every symmetry image of (x,y) must be a voxel of the image: lim = (xlo, xhi, ylo, yhi)
    if (x < lim.x || x > lim.y || y < lim.z || y > lim.w || z < -2048 || z > 2047) atomicMax(range_error, 1);
    const int coords = (x + 512) | ((y + 512) << 10) | ((z + 2048) << 20);
    pool[out + k] = make_int2(coords, __float_as_int(scratch_w[base + k]));
  }
}

void launch_trace_basic_rows(const BasicRay* d_rays, int first, int count, const TraceConsts& tc, uint64_t* scratch_k,
                             float* scratch_w, int cap, int* d_len, int* d_buf, int* d_overflow, hipStream_t stream) {
  const int block = 64;
  const int grid = (count + block - 1) / block;
  hipLaunchKernelGGL(trace_basic_rows_kernel, dim3(grid), dim3(block), 0, stream, d_rays, first, count, tc, scratch_k,
                     scratch_w, cap, d_len, d_buf, d_overflow);
}

void launch_pack_rows(int first, int count, const uint64_t* scratch_k, const float* scratch_w, int cap, const int* d_len,
                      const int* d_buf, const int64_t* d_row_start, int2* d_pool, int* d_range_error, int4 lim,
                      hipStream_t stream) {
  hipLaunchKernelGGL(pack_rows_kernel, dim3(count), dim3(256), 0, stream, first, count, scratch_k, scratch_w, cap, d_len,
                     d_buf, d_row_start, d_pool, d_range_error, lim);
}

}  // namespace mi
