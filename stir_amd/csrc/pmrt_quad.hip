// Back projection over 2 x 2 blocks of image columns.
//
// Neighbouring columns meet largely the same families: at mMR size a 2 x 2
// block's columns meet only about twice the (family, plane) pairs of one
// column (2.2-2.6 times in the oblique segments). The block table merges the four
// columns' lists of the column table into one list per (subset, block) whose
// entries carry the family's run once and one weight per column (0 where a
// column has no element of that family at that plane). One wave back-projects
// a block: each merged entry's run is gathered once and added into the four
// columns' sums, so the projection-data gathers fall to about half.
//
// Order: the merged list is ordered by (family, z of the element relative to
// the first element of its basic row, taken as unsigned). Within one column's
// list (family, element order) that order is the same, because a basic row
// holds the elements of one (x, y) at ascending z (the tracer runs towards +z,
// sorted unions are z-major) except add_adjacent_z's copies at z0 - 1, which
// come last (and whose relative z is the largest unsigned value). Each column
// therefore sums exactly its column-table entries in column-table order: the
// block back projection equals the column back projection bit for bit. A
// merged entry in which a column has no element adds mul_legacy(0, p) = +0.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmrt_device.hpp"
#include "pmrt_kernels.hpp"

namespace mi {

namespace {

using namespace dev;

// a * b, except that a zero operand gives +0 whatever the other one is
// (v_mul_legacy_f32)
__device__ float mul_legacy(float a, float b) __asm("llvm.amdgcn.fmul.legacy");

__device__ inline int basic_z(int c) { return static_cast<int>(static_cast<unsigned>(c) >> 20) - 2048; }

// merge key of every entry of the sorted column table: (family << 12) |
// ((z - z of the row's first element) & 0xfff); perm[i] is entry i's index in
// generation order (family offsets famoff, element e = perm[i] - famoff[f])
__global__ void quad_keys_kernel(const uint32_t* __restrict__ perm, int64_t n, const int64_t* __restrict__ famoff,
                                 int nfam, const Family* __restrict__ fams, const int2* __restrict__ pool,
                                 uint64_t* __restrict__ key) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t g = perm[i];
    int lo = 0, hi = nfam - 1;  // last family whose first entry is at or before g
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (famoff[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    const Family f = fams[lo];
    const int e = static_cast<int>(g - famoff[lo]);
    const int z = basic_z(pool[f.row_start + e].x), z0 = basic_z(pool[f.row_start].x);
    key[i] = (static_cast<uint64_t>(lo) << 12) | (static_cast<unsigned>(z - z0) & 0xfffu);
  }
}

// One thread per (subset, block): four-way merge of the block's column lists
// by key. COUNT: number of merged entries; otherwise the entries, written
// from qoff. *bad is set if a column's keys decrease or one key names two
// different runs (then the table is not used).
template <bool COUNT>
__global__ void quad_merge_kernel(const int64_t* __restrict__ list_off, const BpEntry* __restrict__ ent,
                                  const uint64_t* __restrict__ key, ImageDims d, int nqx, int nq, int nsub,
                                  const int64_t* __restrict__ qoff, int32_t* __restrict__ qcount,
                                  QuadEntry* __restrict__ qent, int* __restrict__ bad) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(nsub) * nq) return;
  const int s = static_cast<int>(t / nq), qi = static_cast<int>(t % nq);
  const int qx = qi % nqx, qy = qi / nqx;
  int64_t pos[4], end[4];
  uint64_t head[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cx = 2 * qx + (j & 1), cy = 2 * qy + (j >> 1);
    pos[j] = end[j] = 0;
    if (cx < d.nx && cy < d.ny) {
      const int64_t c = static_cast<int64_t>(s) * d.ncols + static_cast<int64_t>(cy) * d.nx + cx;
      pos[j] = list_off[c];
      end[j] = list_off[c + 1];
    }
    head[j] = pos[j] < end[j] ? key[pos[j]] : ~0ull;
  }
  int64_t out = COUNT ? 0 : qoff[t];
  int32_t cnt = 0;
  int err = 0;
  for (;;) {
    uint64_t kmin = ~0ull;
#pragma unroll
    for (int j = 0; j < 4; ++j) kmin = head[j] < kmin ? head[j] : kmin;
    if (kmin == ~0ull) break;
    QuadEntry q;
    q.sino_base = 0; q.k = 0; q.nax_ppa = 0; q.pad = 0;
    bool first = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      q.w[j] = 0.F;
      if (head[j] != kmin) continue;
      if (!COUNT) {
        const BpEntry E = ent[pos[j]];
        if (first) {
          q.sino_base = E.sino_base; q.k = E.k; q.nax_ppa = E.nax_ppa;
        } else if (E.sino_base != q.sino_base || E.k != q.k || E.nax_ppa != q.nax_ppa) {
          err = 1;
        }
        q.w[j] = E.w;
      }
      first = false;
      ++pos[j];
      head[j] = pos[j] < end[j] ? key[pos[j]] : ~0ull;
      if (head[j] < kmin) err = 1;
    }
    if (!COUNT) qent[out + cnt] = q;
    ++cnt;
  }
  if (COUNT) qcount[t] = cnt;
  if (err) atomicMax(bad, 1);
}

// block index of work item w: tiles of tile.x x tile.y blocks in row-major
// tile order; -1 past the image
__device__ inline int tiled_quad(int w, int nqx, int nqy, int2 tile) {
  const int tiles_x = (nqx + tile.x - 1) / tile.x;
  const int t = w / (tile.x * tile.y), r = w - t * (tile.x * tile.y);
  const int y = (t / tiles_x) * tile.y + r / tile.x;
  const int x = (t % tiles_x) * tile.x + r % tile.x;
  return (y < nqy && x < nqx) ? y * nqx + x : -1;
}

// One wave per 2 x 2 block of columns of a one-plane table (one plane per
// axial position), lanes over z (NSLOT slots of 64 planes). Entries are
// wave-uniform; each entry's run is a range-checked raw buffer of nax values
// (planes outside the run read 0, as in bp_column). U merged entries have
// their gathers in flight together.
template <int MODE, int NSLOT, bool LO32, int U>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_quad_kernel(const int64_t* __restrict__ qoff,
                                                                        const QuadEntry* __restrict__ qent, int sb,
                                                                        int se, const float* __restrict__ proj,
                                                                        float* __restrict__ img, ImageDims d, int nqx,
                                                                        int nqy, int nwork, int2 tile, int remap,
                                                                        BpData data, int accumulate) {
#if defined(__HIP_DEVICE_COMPILE__)  // the raw-buffer helpers exist in device code only
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int b = remap > 1 ? xcd_chunk_remap(blockIdx.x, remap) : static_cast<int>(blockIdx.x);
  const int w = b * kWavesPerBlock + wave;
  if (w >= nwork) return;
  const int qi = tiled_quad(w, nqx, nqy, tile);
  if (qi < 0) return;
  const int nq = nqx * nqy;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  cptr<QuadEntry> ent = as_const(qent);
  const int qx = qi % nqx, qy = qi / nqx;
  float acc[4][NSLOT];
  int zi4[NSLOT];  // 4 * plane of each slot
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int z = lane + kWave * q;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // accumulate: continue the sums the previous launch left in the image
      const int cx = 2 * qx + (c & 1), cy = 2 * qy + (c >> 1);
      acc[c][q] = accumulate && z < d.nzp && cx < d.nx && cy < d.ny
                      ? as_global(img)[(static_cast<int64_t>(cy) * d.nx + cx) * d.nzp + z]
                      : 0.F;
    }
    zi4[q] = 4 * (d.min_z + lane + kWave * q);
    asm volatile("" : "+v"(zi4[q]));  // opaque: one subtraction per entry and slot, no folded offsets
  }
  for (int s = sb; s < se; ++s) {
    const int64_t beg = as_const(qoff)[static_cast<int64_t>(s) * nq + qi];
    const int64_t end = as_const(qoff)[static_cast<int64_t>(s) * nq + qi + 1];
    int64_t e = beg;
    for (; e + U <= end; e += U) {
      QuadEntry Es[U];  // all requested before any is decoded
#pragma unroll
      for (int j = 0; j < U; ++j) Es[j] = ent[e + j];
      float p[U][NSLOT];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const __amdgpu_buffer_rsrc_t run =
            LO32 ? run_buffer_lo32(as_const(proj), static_cast<unsigned>(Es[j].sino_base), Es[j].nax_ppa & 0xffff)
                 : run_buffer(as_const(proj), static_cast<unsigned>(Es[j].sino_base), Es[j].nax_ppa & 0xffff);
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) p[j][q] = load_run(run, zi4[q] - 4 * Es[j].k);
      }
#pragma unroll
      for (int j = 0; j < U; ++j)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int q = 0; q < NSLOT; ++q) acc[c][q] = acc[c][q] + mul_legacy(Es[j].w[c], p[j][q]);
    }
    for (; e < end; ++e) {
      const QuadEntry E = ent[e];
      const __amdgpu_buffer_rsrc_t run =
          LO32 ? run_buffer_lo32(as_const(proj), static_cast<unsigned>(E.sino_base), E.nax_ppa & 0xffff)
               : run_buffer(as_const(proj), static_cast<unsigned>(E.sino_base), E.nax_ppa & 0xffff);
      float p[NSLOT];
#pragma unroll
      for (int q = 0; q < NSLOT; ++q) p[q] = load_run(run, zi4[q] - 4 * E.k);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) acc[c][q] = acc[c][q] + mul_legacy(E.w[c], p[q]);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int cx = 2 * qx + (c & 1), cy = 2 * qy + (c >> 1);
    if (cx >= d.nx || cy >= d.ny) continue;
    const int64_t col = static_cast<int64_t>(cy) * d.nx + cx;
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) {
      const int z = lane + kWave * q;
      if (z >= d.nzp) continue;
      const int64_t idx = col * d.nzp + z;
      if constexpr (MODE == BP_PLAIN) {
        as_global(img)[idx] = z < d.nz ? acc[c][q] : 0.F;
      } else {
        if (z < d.nz) {
          const float u = osem_factor(acc[c][q], as_const(data.sens)[idx], data.threshold, data.thr_lo, data.thr_hi);
          as_global(data.estimate)[idx] = as_global(data.estimate)[idx] * u;
          if (data.update) as_global(data.update)[idx] = u;
        }
      }
    }
  }
#endif
}

unsigned grid_of(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

}  // namespace

void launch_quad_keys(const uint32_t* d_perm, int64_t n, const int64_t* d_famoff, int nfam, const Family* d_fams,
                      const int2* d_pool, uint64_t* d_key, hipStream_t stream) {
  if (n <= 0 || nfam <= 0) return;
  hipLaunchKernelGGL(quad_keys_kernel, dim3(grid_of(n, 256)), dim3(256), 0, stream, d_perm, n, d_famoff, nfam, d_fams,
                     d_pool, d_key);
}

void launch_quad_merge(bool count, const int64_t* d_list_off, const BpEntry* d_ent, const uint64_t* d_key,
                       const ImageDims& dims, int nqx, int nq, int nsub, const int64_t* d_qoff, int32_t* d_qcount,
                       QuadEntry* d_qent, int* d_bad, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(nsub) * nq;
  if (n <= 0) return;
  const dim3 grid(static_cast<unsigned>((n + 127) / 128)), block(128);
  if (count)
    hipLaunchKernelGGL(quad_merge_kernel<true>, grid, block, 0, stream, d_list_off, d_ent, d_key, dims, nqx, nq, nsub,
                       d_qoff, d_qcount, d_qent, d_bad);
  else
    hipLaunchKernelGGL(quad_merge_kernel<false>, grid, block, 0, stream, d_list_off, d_ent, d_key, dims, nqx, nq, nsub,
                       d_qoff, d_qcount, d_qent, d_bad);
}

void launch_back_project_quad(BpMode mode, const int64_t* d_qoff, const QuadEntry* d_qent, int subset_begin,
                              int subset_end, const float* d_proj, float* d_img, const ImageDims& dims,
                              const BpData& data, const BpTuning& tune, hipStream_t stream, int64_t proj_values) {
  const int nslot = (dims.nzp + kWave - 1) / kWave;
  const int nqx = (dims.nx + 1) / 2, nqy = (dims.ny + 1) / 2;
  // tiles in blocks: half the column tile's extent (at least one block)
  const int2 tile = make_int2(tune.tile_x > 1 ? tune.tile_x / 2 : 1, tune.tile_y > 1 ? tune.tile_y / 2 : 1);
  const int nwork = ((nqx + tile.x - 1) / tile.x) * ((nqy + tile.y - 1) / tile.y) * tile.x * tile.y;
  int nblocks = (nwork + kWavesPerBlock - 1) / kWavesPerBlock;
  const int remap = tune.xcd_remap > 1 ? tune.xcd_remap : 0;
  if (remap > 1) {  // chunked remap: whole rounds of 8 chunks
    const int round = kNumXcd * remap;
    nblocks = (nblocks + round - 1) / round * round;
  }
  const dim3 grid(nblocks), block(kWave * kWavesPerBlock);
  const uint64_t first = reinterpret_cast<uint64_t>(d_proj);
  const bool lo32 = proj_values > 0 && (first >> 32) == ((first + static_cast<uint64_t>(proj_values) * 4 - 1) >> 32);
  // a plain projection of several subsets: one launch per subset, each
  // continuing the sums in the image (as launch_back_project)
  const bool split = tune.split_subsets && mode == BP_PLAIN && subset_end - subset_begin > 1;
  const int nlaunch = split ? subset_end - subset_begin : 1;
  for (int li = 0; li < nlaunch; ++li) {
  const int sb = split ? subset_begin + li : subset_begin, se = split ? sb + 1 : subset_end, acc = split && li > 0;
#define MI_QBP_U(M, S, L, U)                                                                                     \
  hipLaunchKernelGGL((bp_quad_kernel<M, S, L, U>), grid, block, 0, stream, d_qoff, d_qent, sb, se, d_proj, d_img,  \
                     dims, nqx, nqy, nwork, tile, remap, data, acc)
#define MI_QBP_L(M, S, L)                              \
  if (tune.quad_unroll <= 4) MI_QBP_U(M, S, L, 4);     \
  else if (tune.quad_unroll <= 6) MI_QBP_U(M, S, L, 6); \
  else MI_QBP_U(M, S, L, 8)
#define MI_QBP(M, S)               \
  if (lo32) { MI_QBP_L(M, S, true); } \
  else { MI_QBP_L(M, S, false); }
#define MI_QBP_S(M)              \
  switch (nslot) {               \
    case 1: MI_QBP(M, 1); break; \
    case 2: MI_QBP(M, 2); break; \
    case 3: MI_QBP(M, 3); break; \
    default: MI_QBP(M, 4); break; \
  }
  if (mode == BP_PLAIN) { MI_QBP_S(BP_PLAIN) }
  else { MI_QBP_S(BP_OSEM_UPDATE) }
  }
#undef MI_QBP_S
#undef MI_QBP
#undef MI_QBP_L
#undef MI_QBP_U
}

}  // namespace mi
