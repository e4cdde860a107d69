// Grouped projection kernels: a wave is split into 64/G lane groups of G
// lanes, and each group runs one work item -- one family (forward
// projection, lanes over its axial positions) or one image column (back
// projection, lanes over z) -- of any problem of a batch. Small scanners
// have families of 8-32 axial positions and images of 15-31 planes; packing
// several of them into one 64-wide wave keeps the lanes busy, and one launch
// covers every problem of a batch.
//
// Per-bin sums run in row order and per-voxel sums in table order, exactly
// as in fp_kernel / bp_kernel, so results are identical to those kernels.
// Row elements and table entries are loaded per lane (the members of a
// group load the same address).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "pmrt_device.hpp"
#include "pmrt_kernels.hpp"

namespace mi {

namespace {

using namespace dev;

constexpr int kGroupUnroll = 8;

// (x, y) -> (a x + b y, c x + e y) for the in-plane operation codes
__device__ inline void inplane_coeffs(int ip, int& a, int& b, int& c, int& e) {
  a = b = c = e = 0;
  switch (ip) {
    case IP_ID: a = 1; e = 1; break;
    case IP_XMX: a = -1; e = 1; break;
    case IP_YMY: a = 1; e = -1; break;
    case IP_XMX_YMY: a = -1; e = -1; break;
    case IP_XY_YX: b = 1; c = 1; break;
    case IP_XMY_YX: b = -1; c = 1; break;
    case IP_XY_YMX: b = 1; c = -1; break;
    default: b = -1; c = -1; break;
  }
}

__device__ inline int wave_max(int v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
  return v;
}

template <int MODE, int G>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_group_kernel(const ProblemDesc* __restrict__ descs,
                                                                         const WorkItem* __restrict__ work,
                                                                         int nwork) {
  constexpr int F = kWave / G;
  const int wave = static_cast<int>(threadIdx.x) / kWave;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int a = lane % G;
  const int wi = (batch_block() * kWavesPerBlock + wave) * F + lane / G;
  WorkItem W{0, -1};
  if (wi < nwork) W = work[wi];
  const bool item = W.index >= 0;
  const ProblemDesc& P = descs[W.problem];
  const ImageDims d = P.d;
  Family f{};
  if (item) f = as_const(P.fams)[W.index];
  const bool valid = item && a < f.nax;
  int ca, cb, cc, ce;
  inplane_coeffs(f.inplane, ca, cb, cc, ce);
  const int zs = f.zsign;
  const int zb = f.zadd0 + f.ppa * a - d.min_z;
  cptr<int2> row = as_const(P.pool) + f.row_start;
  const int n = item ? f.row_len : 0;
  const int nmax = wave_max(n);
  cptr<float> img = as_const(P.fp_in);
  float acc = 0.F;
  for (int e0 = 0; e0 < nmax; e0 += kGroupUnroll) {
    float v[kGroupUnroll], w[kGroupUnroll];
    bool ok[kGroupUnroll];
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j) {
      const int e = e0 + j;
      const bool has = e < n;
      int2 el = make_int2(0, 0);
      if (has) el = row[e];
      int x, y, z;
      decode(el.x, x, y, z);
      const int xo = ca * x + cb * y, yo = cc * x + ce * y;
      const int col = has ? (yo - d.min_y) * d.nx + (xo - d.min_x) : 0;
      const int zz = __mul24(zs, z) + zb;  // zs = +-1, |z| < 2^11
      ok[j] = valid && has && static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz);
      w[j] = __int_as_float(el.y);
      v[j] = img[static_cast<int64_t>(col) * d.nzp + clamp_index(zz, d.nz)];
    }
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j)
      if (ok[j]) acc = acc + v[j] * w[j];
  }
  if (valid) {
    double ll = 0.0;
    fp_epilogue<MODE>(acc, static_cast<int64_t>(f.sino_base) + a, f.viewgram, P.fd, P.fp_out, ll);
  }
}

template <int MODE, int G>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_group_kernel(const ProblemDesc* __restrict__ descs,
                                                                         const WorkItem* __restrict__ work, int nwork,
                                                                         int subset, int threshold) {
  constexpr int F = kWave / G;
  const int wave = static_cast<int>(threadIdx.x) / kWave;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int zl = lane % G;
  const int wi = (batch_block() * kWavesPerBlock + wave) * F + lane / G;
  WorkItem W{0, -1};
  if (wi < nwork) W = work[wi];
  const bool item = W.index >= 0;
  const ProblemDesc& P = descs[W.problem];
  const ImageDims d = P.d;
  int64_t beg = 0, n = 0;
  if (item) {
    cptr<int64_t> lo = as_const(P.list_off);
    beg = lo[static_cast<int64_t>(subset) * d.ncols + W.index];
    n = lo[static_cast<int64_t>(subset) * d.ncols + W.index + 1] - beg;
  }
  const int nmax = wave_max(static_cast<int>(n));
  const int zi = d.min_z + zl;
  const bool zin = item && zl < d.nz;
  cptr<BpEntry> ent = as_const(P.entries) + beg;
  cptr<float> proj = as_const(P.bp_in);
  float acc = 0.F;
  for (int e0 = 0; e0 < nmax; e0 += kGroupUnroll) {
    float p[kGroupUnroll], w[kGroupUnroll];
    bool ok[kGroupUnroll];
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j) {
      const int e = e0 + j;
      const bool has = e < n;
      BpEntry E{0, 0, 1, 0.F};
      if (has) E = ent[e];
      const int nax = E.nax_ppa & 0xffff, ppa = E.nax_ppa >> 16;
      const int diff = zi - E.k;
      int ax;
      bool in;
      if (ppa == 1) {
        ax = diff;
        in = static_cast<unsigned>(ax) < static_cast<unsigned>(nax);
      } else if (ppa == 2) {
        ax = diff >> 1;
        in = (diff & 1) == 0 && static_cast<unsigned>(ax) < static_cast<unsigned>(nax);
      } else {
        ax = diff / max(ppa, 1);
        in = diff >= 0 && ax * ppa == diff && ax < nax;
      }
      ok[j] = has && zin && in;
      w[j] = E.w;
      p[j] = proj[static_cast<int64_t>(E.sino_base) + min(max(ax, 0), max(nax - 1, 0))];
    }
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j)
      if (ok[j]) acc = acc + w[j] * p[j];
  }
  if (!item || zl >= d.nzp) return;
  const int64_t idx = static_cast<int64_t>(W.index) * d.nzp + zl;
  if constexpr (MODE == BP_PLAIN) {
    as_global(P.bp_out)[idx] = zl < d.nz ? acc : 0.F;
  } else {
    if (zl < d.nz) {
      const float u = osem_factor(acc, as_const(as_const(P.sens)[subset])[idx], threshold, P.thr_lo, P.thr_hi);
      as_global(P.estimate)[idx] = as_global(P.estimate)[idx] * u;
    }
  }
}

// Row groups: the families of one basic row (its symmetry-related
// (segment, view, tangential) positions, all in the same subset; the host
// groups them by row start and length) share every row element; a wave runs up to 64 axial positions of several such families,
// lanes laid out family after family. Row elements are wave-uniform (scalar
// loads) and each lane applies its own family's in-plane operation, so a
// short family no longer leaves most of a wave idle. Each bin sums its row in
// row order, as fp_kernel does, so results are identical to it.
template <int MODE>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_rowgroup_kernel(const ProblemDesc* __restrict__ descs,
                                                                            const RowGroupItem* __restrict__ items,
                                                                            const int32_t* __restrict__ item_fams,
                                                                            int nitems) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int wi = batch_block() * kWavesPerBlock + wave;
  if (wi >= nitems) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const RowGroupItem it = items[wi];
  const int p = __builtin_amdgcn_readfirstlane(it.problem);
  const ProblemDesc& P = descs[p];
  const ImageDims d = P.d;
  cptr<Family> fams = as_const(P.fams);
  cptr<int32_t> list = as_const(item_fams) + it.first;
  // this lane's family and axial position
  // byte offset of the voxel column of element (x, y) for this lane's
  // family: A4 x + B4 y + C4 (the in-plane operation folded into the
  // column index); lanes without a bin keep 0 and a z offset no element reaches
  int A4 = 0, B4 = 0, C4 = 0, zs = 1, zb = 1 << 20, vg = 0;
  int64_t out = 0;
  bool valid = false;
  int off = 0, row_start = 0, row_len = 0;
  for (int k = 0; k < it.count; ++k) {
    const Family f = fams[list[k]];
    if (k == 0) {
      row_start = f.row_start;
      row_len = f.row_len;
    }
    const int a = lane - off;
    if (a >= 0 && a < f.nax) {
      int ca, cb, cc, ce;  // xo = ca x + cb y, yo = cc x + ce y
      inplane_coeffs(f.inplane, ca, cb, cc, ce);
      const int nzp4 = 4 * d.nzp;
      A4 = (cc * d.nx + ca) * nzp4;
      B4 = (ce * d.nx + cb) * nzp4;
      C4 = (-d.min_y * d.nx - d.min_x) * nzp4;
      zs = f.zsign;
      zb = f.zadd0 + f.ppa * a - d.min_z;
      out = static_cast<int64_t>(f.sino_base) + a;
      vg = f.viewgram;
      valid = true;
    }
    off += f.nax;
  }
  cptr<int2> row = as_const(P.pool) + row_start;
  cptr<float> img = as_const(P.fp_in);
  float acc = 0.F;
  int e = 0;
  for (; e + kGroupUnroll <= row_len; e += kGroupUnroll) {
    float v[kGroupUnroll], w[kGroupUnroll];
    bool ok[kGroupUnroll];
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j) {
      const int2 el = row[e + j];
      int x, y, z;
      decode(el.x, x, y, z);
      // |A4|, |B4| <= 4 nx nzp < 2^23 and |x|, |y| < 2^9: 24-bit multiplies
      const unsigned col4 = static_cast<unsigned>(__mul24(A4, x) + __mul24(B4, y) + C4);
      const int zz = __mul24(zs, z) + zb;  // zs = +-1, |z| < 2^11
      ok[j] = static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz);
      w[j] = __int_as_float(el.y);
      v[j] = load_byteoff(img, col4 + (static_cast<unsigned>(clamp_index(zz, d.nz)) << 2));
    }
#pragma unroll
    for (int j = 0; j < kGroupUnroll; ++j)
      if (ok[j]) acc = acc + v[j] * w[j];
  }
  for (; e < row_len; ++e) {
    const int2 el = row[e];
    int x, y, z;
    decode(el.x, x, y, z);
    const unsigned col4 = static_cast<unsigned>(__mul24(A4, x) + __mul24(B4, y) + C4);
    const int zz = __mul24(zs, z) + zb;  // zs = +-1, |z| < 2^11
    if (static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz))
      acc = acc + load_byteoff(img, col4 + (static_cast<unsigned>(zz) << 2)) * __int_as_float(el.y);
  }
  if (valid) {
    double ll = 0.0;
    fp_epilogue<MODE>(acc, out, vg, P.fd, P.fp_out, ll);
  }
}

unsigned group_blocks(int nwork, int G) {
  const int per_block = kWavesPerBlock * (kWave / G);
  return batch_grid((nwork + per_block - 1) / per_block);
}

}  // namespace

void launch_fp_group(FpMode mode, int G, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork,
                     hipStream_t stream) {
  if (nwork <= 0) return;
  const dim3 grid(group_blocks(nwork, G)), block(kWave * kWavesPerBlock);
#define MI_FPG(M, GG) hipLaunchKernelGGL((fp_group_kernel<M, GG>), grid, block, 0, stream, d_descs, d_work, nwork)
#define MI_FPG_G(M)                 \
  switch (G) {                      \
    case 8: MI_FPG(M, 8); break;    \
    case 16: MI_FPG(M, 16); break;  \
    case 32: MI_FPG(M, 32); break;  \
    default: MI_FPG(M, 64); break;  \
  }
  if (mode == FP_RATIO) { MI_FPG_G(FP_RATIO) }
  else { MI_FPG_G(FP_PLAIN) }
#undef MI_FPG_G
#undef MI_FPG
}

void launch_fp_rowgroup(FpMode mode, const ProblemDesc* d_descs, const RowGroupItem* d_items, const int32_t* d_item_fams,
                        int nitems, hipStream_t stream) {
  if (nitems <= 0) return;
  const dim3 grid(batch_grid((nitems + kWavesPerBlock - 1) / kWavesPerBlock)), block(kWave * kWavesPerBlock);
  if (mode == FP_RATIO)
    hipLaunchKernelGGL(fp_rowgroup_kernel<FP_RATIO>, grid, block, 0, stream, d_descs, d_items, d_item_fams, nitems);
  else
    hipLaunchKernelGGL(fp_rowgroup_kernel<FP_PLAIN>, grid, block, 0, stream, d_descs, d_items, d_item_fams, nitems);
}

void launch_bp_group(BpMode mode, int G, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork, int subset,
                     int threshold, hipStream_t stream) {
  if (nwork <= 0) return;
  const dim3 grid(group_blocks(nwork, G)), block(kWave * kWavesPerBlock);
#define MI_BPG(M, GG)                                                                                             \
  hipLaunchKernelGGL((bp_group_kernel<M, GG>), grid, block, 0, stream, d_descs, d_work, nwork, subset, threshold)
#define MI_BPG_G(M)                 \
  switch (G) {                      \
    case 16: MI_BPG(M, 16); break;  \
    case 32: MI_BPG(M, 32); break;  \
    default: MI_BPG(M, 64); break;  \
  }
  if (mode == BP_OSEM_UPDATE) { MI_BPG_G(BP_OSEM_UPDATE) }
  else { MI_BPG_G(BP_PLAIN) }
#undef MI_BPG_G
#undef MI_BPG
}

}  // namespace mi
