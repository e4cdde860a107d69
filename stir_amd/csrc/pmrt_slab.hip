// Back projection through projection-data slabs staged in LDS.
//
// The image is cut into tiles of tx x ty columns; one workgroup owns a tile
// and keeps the sums of all its voxels in registers: wave w holds columns
// w*CPW .. w*CPW + CPW - 1 of the tile, 2 slots of 64 planes each (lane l of
// slot q is plane 64 q + l). The projection data are visited one phase at a
// time: a phase is a (subset, view, group of segments). For a phase, the
// tile's columns read only the runs of that view whose tangential positions
// cross the tile, in the group's segments -- a slab that the workgroup
// copies into LDS once (one contiguous range per segment) and then reads as
// often as its columns' table entries ask. Global traffic per projection is
// the table (streamed once) plus the slabs (each run read by the tiles along
// its LOR), instead of one gather per matrix element.
//
// Slab layout: kSlabLead zeros, then per segment of the group 64 zeros and,
// for every tangential position of the tile's range, its run of nax values
// followed by 64 zeros; kSlabLead zeros at the end. An entry = (byte offset
// of its run's axial index -k, where k is the plane of axial position 0;
// weight for slot 0; weight for slot 1). Lane l of slot q reads axial index
// 64 q + l - k. When the entry reaches slot q (k <= 64 q + 63 and
// k + nax - 1 >= 64 q) that index lies in [-63, nax + 62]: the run or its
// zeros, so the lane adds exactly its term or w * 0. When it does not, the
// slot's weight is 0 and the read stays inside the slab (the leading and
// trailing zeros cover the reach of any plane): it adds w * 0 = 0 too. No
// per-lane test, clamp or select; one address per entry serves both slots
// (the second slot's read at the instruction's +256 byte offset).
//
// Table layout: per (subset, tile, phase, wave) a dense block of `steps` x
// CPW entries, entry (t, i) = the t-th term of the wave's column i, in
// family order then row order (the entries are generated in that order and
// stably radix-sorted). Columns with fewer terms are padded with entries of
// weight 0 that read the slab's leading zeros. The wave walks its block one
// step at a time, all CPW columns in lockstep: the entries come in long
// scalar loads and every voxel sums its terms in a fixed order
// (deterministic run to run).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmrt_device.hpp"
#include "pmrt_kernels.hpp"

namespace mi {

namespace {

using namespace dev;

constexpr int kRunsInFlight = 4;  // slab runs whose loads a wave issues together
static_assert(kSlabPad == kWave && kSlabLead == 2 * kWave, "slab zeros are written 64 lanes at a time");

__device__ inline void tile_of(const SlabArgs& A, int col_x, int col_y, int& t, int& c) {
  const int tx = col_x / A.tx, ty = col_y / A.ty;
  t = ty * A.ntx + tx;
  c = (col_y - ty * A.ty) * A.tx + (col_x - tx * A.tx);
}

// The matrix element e of family fi, with its column, tile and planes; false
// when it has no voxel in the image (it then adds nothing anywhere).
struct SlabElem {
  int t, c, k, zlo, zhi;
  float w;
};
__device__ inline bool slab_elem(const Family& f, int2 el, const ImageDims& d, const SlabArgs& A, SlabElem& E) {
  int x, y, z, xo, yo;
  decode(el.x, x, y, z);
  inplane_any_op(f.inplane, x, y, xo, yo);
  const int cx = xo - d.min_x, cy = yo - d.min_y;
  E.k = f.zsign * z + f.zadd0 - d.min_z;  // plane of axial position 0
  E.zlo = max(E.k, 0);
  E.zhi = min(E.k + f.nax - 1, d.nz - 1);
  if (cx < 0 || cx >= d.nx || cy < 0 || cy >= d.ny || E.zhi < E.zlo) return false;
  tile_of(A, cx, cy, E.t, E.c);
  E.w = __int_as_float(el.y);
  return true;
}

// pass 1: the tangential range of every (subset, tile, view slot)
__global__ void __launch_bounds__(64) slab_range_kernel(const Family* __restrict__ fams,
                                                        const int4* __restrict__ fam_slab, int fam_begin, int fam_end,
                                                        const int2* __restrict__ pool, int subset, ImageDims d,
                                                        SlabArgs A, int2* __restrict__ rng) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  const int4 fs = fam_slab[fi];  // view slot, tangential index, segment slot
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    SlabElem E;
    if (!slab_elem(f, pool[f.row_start + e], d, A, E)) continue;
    const int64_t g = (static_cast<int64_t>(subset) * A.ntiles + E.t) * A.nv_max + fs.x;
    atomicMin(&rng[g].x, fs.y);
    atomicMax(&rng[g].y, fs.y);
  }
}

__device__ inline int64_t slab_key(const SlabArgs& A, int subset, int t, int vslot, int group, int c) {
  const int64_t phase = static_cast<int64_t>(vslot) * A.ngroups + group;
  return ((static_cast<int64_t>(subset) * A.ntiles + t) * (A.nv_max * A.ngroups) + phase) * (A.tx * A.ty) + c;
}

// pass 2: the number of terms of every (subset, tile, phase, column)
__global__ void __launch_bounds__(64) slab_count_kernel(const Family* __restrict__ fams,
                                                        const int4* __restrict__ fam_slab, int fam_begin, int fam_end,
                                                        const int2* __restrict__ pool, int subset, ImageDims d,
                                                        SlabArgs A, uint32_t* __restrict__ counts) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  const int4 fs = fam_slab[fi];
  const int group = as_const(A.segs)[fs.z].w;
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    SlabElem E;
    if (!slab_elem(f, pool[f.row_start + e], d, A, E)) continue;
    atomicAdd(&counts[slab_key(A, subset, E.t, fs.x, group, E.c)], 1U);
  }
}

// pass 3: every element's key and entry, in family order, element order
__global__ void __launch_bounds__(64) slab_make_kernel(const Family* __restrict__ fams,
                                                       const int4* __restrict__ fam_slab, int fam_begin, int fam_end,
                                                       const int64_t* __restrict__ fam_entry_off,
                                                       const int2* __restrict__ pool, int subset, ImageDims d,
                                                       SlabArgs A, uint32_t* __restrict__ keys,
                                                       SlabEntry* __restrict__ ent) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  const int4 fs = fam_slab[fi];
  const int4 sg = as_const(A.segs)[fs.z];  // device offset, nax, -, group
  const int64_t base = fam_entry_off[fi - fam_begin];
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    SlabElem E;
    SlabEntry b{0, 0.F, 0.F};
    uint32_t key = 0xffffffffU;  // sorts after every list; never read
    if (slab_elem(f, pool[f.row_start + e], d, A, E)) {
      const int2 r = as_const(A.rng)[(static_cast<int64_t>(subset) * A.ntiles + E.t) * A.nv_max + fs.x];
      const int T = r.y - r.x + 1;
      const int2 pos = as_const(A.seg_pos)[fs.z];  // nax and segments before it in its group
      const int run = slab_segment_start(pos.y, pos.x, T) + kSlabPad + (fs.y - r.x) * (sg.y + kSlabPad);
      b.off4 = 4 * (run - E.k);
      b.w0 = E.zlo < kWave ? E.w : 0.F;
      b.w1 = E.zhi >= kWave ? E.w : 0.F;
      key = static_cast<uint32_t>(slab_key(A, subset, E.t, fs.x, sg.w, E.c));
    }
    keys[base + e] = key;
    ent[base + e] = b;
  }
}

// steps of every (subset, tile, phase, wave) block: the most terms of its columns
__global__ void slab_steps_kernel(const uint32_t* __restrict__ counts, int64_t nblocks, int cpw, int tc,
                                  uint32_t* __restrict__ steps) {
  const int nw = (tc + cpw - 1) / cpw;
  for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; b < nblocks;
       b += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t tp = b / nw;
    const int w = static_cast<int>(b - tp * nw);
    uint32_t m = 0;
    for (int i = 0; i < cpw; ++i) {
      const int c = w * cpw + i;
      if (c < tc) m = max(m, counts[tp * tc + c]);
    }
    steps[b] = m;
  }
}

// pass 4: sorted element r -> its block slot (first entry of the wave's
// block + rank within its column's list * CPW + column within the wave)
__global__ void slab_place_kernel(const uint32_t* __restrict__ keys_sorted, const uint32_t* __restrict__ idx_sorted,
                                  const SlabEntry* __restrict__ raw, const uint32_t* __restrict__ list_start,
                                  const uint2* __restrict__ hdr, int64_t n, int cpw, int tc,
                                  SlabEntry* __restrict__ table) {
  const int nw = (tc + cpw - 1) / cpw;
  for (int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint32_t key = keys_sorted[r];
    const int64_t tp = key / tc;
    const int c = static_cast<int>(key - tp * tc);
    const int w = c / cpw, i = c - w * cpw;
    const uint32_t rank = static_cast<uint32_t>(r) - list_start[key];
    table[static_cast<int64_t>(hdr[tp * nw + w].x) + static_cast<int64_t>(rank) * cpw + i] = raw[idx_sorted[r]];
  }
}

__global__ void slab_fill_kernel(SlabEntry* __restrict__ t, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    t[i] = SlabEntry{0, 0.F, 0.F};
}

// Copies the slab of segment group `group` of view `view`, tangential
// indices [lo, lo + T), into LDS: wave w takes runs w, w + nwaves, ... of the
// group's (segment, tangential) runs, kRunsInFlight at a time, and writes
// each run's values and the 64 zeros after it (and the zeros that open its
// segment); the first and the last wave write the lead and tail zeros.
__device__ inline void load_slab(const SlabArgs& A, cptr<float> proj, int view, int group, int lo, int T, float* slab,
                                 int wave, int nwaves, int lane) {
  const int2 gr = as_const(A.groups)[group];  // first segment slot, end slot
  const int nruns = T * (gr.y - gr.x);
  const int2 last = as_const(A.seg_pos)[gr.y - 1];
  const int end = slab_segment_start(last.y, last.x, T) + kSlabPad + T * (as_const(A.segs)[gr.y - 1].y + kSlabPad);
  if (wave == 0) {
    slab[lane] = 0.F;
    slab[kWave + lane] = 0.F;
  }
  if (wave == nwaves - 1) {
    slab[end + lane] = 0.F;
    slab[end + kWave + lane] = 0.F;
  }
  for (int r0 = wave; r0 < nruns; r0 += kRunsInFlight * nwaves) {
    float v[kRunsInFlight][2];
    int dst[kRunsInFlight], nax[kRunsInFlight];
#pragma unroll
    for (int u = 0; u < kRunsInFlight; ++u) {
      const int ri = r0 + u * nwaves;  // wave-uniform
      nax[u] = 0;
      dst[u] = -1;
      v[u][0] = v[u][1] = 0.F;
      if (ri < nruns) {
        const int j = gr.x + ri / T, r = ri % T;
        const int4 sg = as_const(A.segs)[j];
        const int2 pos = as_const(A.seg_pos)[j];
        nax[u] = sg.y;
        const int start = slab_segment_start(pos.y, pos.x, T);
        dst[u] = start + kSlabPad + r * (sg.y + kSlabPad);
        cptr<float> src = proj + sg.x + (static_cast<int64_t>(view) * A.ntang + lo + r) * sg.y;
        if (lane < sg.y) v[u][0] = src[lane];
        if (lane + kWave < sg.y) v[u][1] = src[lane + kWave];
        if (r == 0) slab[start + lane] = 0.F;  // the segment's opening zeros
      }
    }
#pragma unroll
    for (int u = 0; u < kRunsInFlight; ++u) {
      if (dst[u] < 0) continue;
      float* run = slab + dst[u];
      if (lane < nax[u]) run[lane] = v[u][0];
      if (lane + kWave < nax[u]) run[lane + kWave] = v[u][1];
      run[nax[u] + lane] = 0.F;  // the zeros after the run
    }
  }
}

// one step of the wave's block: entry i is the next term of column i
template <int CPW>
__device__ inline void slab_step(cptr<SlabEntry> pe, const char* __restrict__ lds, int lane4, float (&tot)[CPW][2]) {
  SlabEntry E[CPW];
#pragma unroll
  for (int i = 0; i < CPW; ++i) E[i] = pe[i];
  float v[CPW][2];
#pragma unroll
  for (int i = 0; i < CPW; ++i) {
    const char* a = lds + (lane4 + E[i].off4);
    v[i][0] = *reinterpret_cast<const float*>(a);
    v[i][1] = *reinterpret_cast<const float*>(a + 256);
  }
#pragma unroll
  for (int i = 0; i < CPW; ++i) {
    tot[i][0] = __builtin_fmaf(E[i].w0, v[i][0], tot[i][0]);
    tot[i][1] = __builtin_fmaf(E[i].w1, v[i][1], tot[i][1]);
  }
}

template <int MODE, int CPW>
__global__ void __launch_bounds__(1024) bp_slab_kernel(SlabArgs A, int sb, int se, const float* __restrict__ proj,
                                                        float* __restrict__ img, ImageDims d, BpData data) {
  extern __shared__ float slab[];
  const int nwaves = A.nwaves;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int lane4 = 4 * lane;
  const int t = as_const(A.tile_order)[blockIdx.x];
  const int TC = A.tx * A.ty;
  const int nph_max = A.nv_max * A.ngroups;
  float tot[CPW][2];
#pragma unroll
  for (int i = 0; i < CPW; ++i) tot[i][0] = tot[i][1] = 0.F;
  cptr<SlabEntry> ent = as_const(A.ent);
  const char* lds = reinterpret_cast<const char*>(slab);
  for (int s = sb; s < se; ++s) {
    const int nv = as_const(A.nviews)[s];
    for (int vs = 0; vs < nv; ++vs) {
      const int2 r = as_const(A.rng)[(static_cast<int64_t>(s) * A.ntiles + t) * A.nv_max + vs];
      if (r.y < r.x) continue;  // no entries of this tile for this view (same for the whole workgroup)
      const int view = as_const(A.view_of)[s * A.nv_max + vs];
      for (int gidx = 0; gidx < A.ngroups; ++gidx) {
        const int64_t blk = ((static_cast<int64_t>(s) * A.ntiles + t) * nph_max + vs * A.ngroups + gidx) * nwaves + wave;
        const uint2 h = as_const(A.hdr)[blk];  // first entry, steps
        // touch every 128-byte line of this phase's table block (all waves):
        // the lines reach L2 while the slab loads, and the scalar loads of the
        // entries below hit there instead of waiting on HBM
        int touched = 0;
        {
          const int64_t blk0 = blk - wave;
          const uint2 h0 = as_const(A.hdr)[blk0], hl = as_const(A.hdr)[blk0 + nwaves - 1];
          cptr<int> words = (cptr<int>)A.ent;
          const int64_t w0 = static_cast<int64_t>(h0.x) * 3, w1 = (static_cast<int64_t>(hl.x) + int64_t{hl.y} * CPW) * 3;
          for (int64_t o = w0 + 32 * static_cast<int64_t>(threadIdx.x); o < w1; o += 32 * static_cast<int64_t>(blockDim.x))
            touched += words[o];
        }
        __syncthreads();  // the previous phase's reads are done
        load_slab(A, as_const(proj), view, gidx, r.x, r.y - r.x + 1, slab, wave, nwaves, lane);
        __syncthreads();
        cptr<SlabEntry> pe = ent + h.x;
        uint32_t st = h.y;
        for (; st >= 2; st -= 2, pe += 2 * CPW) {
          slab_step<CPW>(pe, lds, lane4, tot);
          slab_step<CPW>(pe + CPW, lds, lane4, tot);
        }
        if (st) slab_step<CPW>(pe, lds, lane4, tot);
        asm volatile("" ::"v"(touched));  // the touching loads are complete by now
      }
    }
  }
  // the tile's voxels
  const int tx0 = (t % A.ntx) * A.tx, ty0 = (t / A.ntx) * A.ty;
#pragma unroll
  for (int i = 0; i < CPW; ++i) {
    const int c = wave * CPW + i;
    if (c >= TC) break;
    const int cx = tx0 + c % A.tx, cy = ty0 + c / A.tx;
    if (cx >= d.nx || cy >= d.ny) continue;
    const int col = cy * d.nx + cx;
    gptr<float> out = as_global(img) + static_cast<int64_t>(col) * d.nzp;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int z = lane + kWave * q;
      if (z >= d.nzp) continue;
      if constexpr (MODE == BP_PLAIN) {
        out[z] = z < d.nz ? tot[i][q] : 0.F;
      } else {
        if (z < d.nz) {
          const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
          const float u = osem_factor(tot[i][q], as_const(data.sens)[idx], data.threshold, data.thr_lo, data.thr_hi);
          as_global(data.estimate)[idx] = as_global(data.estimate)[idx] * u;
          if (data.update) as_global(data.update)[idx] = u;
        }
      }
    }
  }
}

// the columns of tiles without any entry: zero (or the OSEM update of a zero
// back projection)
template <int MODE>
__global__ void __launch_bounds__(256) slab_empty_tiles_kernel(SlabArgs A, const int32_t* __restrict__ tiles,
                                                               int ntiles, float* __restrict__ img, ImageDims d,
                                                               BpData data) {
  const int TC = A.tx * A.ty;
  const int64_t n = static_cast<int64_t>(ntiles) * TC * d.nzp;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int z = static_cast<int>(i % d.nzp);
    const int64_t tc = i / d.nzp;
    const int t = tiles[tc / TC], c = static_cast<int>(tc % TC);
    const int cx = (t % A.ntx) * A.tx + c % A.tx, cy = (t / A.ntx) * A.ty + c / A.tx;
    if (cx >= d.nx || cy >= d.ny) continue;
    const int64_t idx = (static_cast<int64_t>(cy) * d.nx + cx) * d.nzp + z;
    if constexpr (MODE == BP_PLAIN) {
      img[idx] = 0.F;
    } else {
      if (z < d.nz) {
        const float u = osem_factor(0.F, data.sens[idx], data.threshold, data.thr_lo, data.thr_hi);
        data.estimate[idx] = data.estimate[idx] * u;
        if (data.update) data.update[idx] = u;
      }
    }
  }
}

unsigned grid_of(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

}  // namespace

void launch_slab_range(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end, const int2* d_pool,
                       int subset, const ImageDims& dims, const SlabArgs& A, int2* d_rng, hipStream_t stream) {
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(slab_range_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, d_fam_slab, fam_begin,
                     fam_end, d_pool, subset, dims, A, d_rng);
}

void launch_slab_count(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end, const int2* d_pool,
                       int subset, const ImageDims& dims, const SlabArgs& A, uint32_t* d_counts, hipStream_t stream) {
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(slab_count_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, d_fam_slab, fam_begin,
                     fam_end, d_pool, subset, dims, A, d_counts);
}

void launch_slab_make(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end,
                      const int64_t* d_fam_entry_off, const int2* d_pool, int subset, const ImageDims& dims,
                      const SlabArgs& A, uint32_t* d_keys, SlabEntry* d_entries, hipStream_t stream) {
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(slab_make_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, d_fam_slab, fam_begin,
                     fam_end, d_fam_entry_off, d_pool, subset, dims, A, d_keys, d_entries);
}

void launch_slab_steps(const uint32_t* d_counts, int64_t nblocks, int cpw, int tc, uint32_t* d_steps,
                       hipStream_t stream) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(slab_steps_kernel, dim3(grid_of(nblocks)), dim3(256), 0, stream, d_counts, nblocks, cpw, tc, d_steps);
}

void launch_slab_place(const uint32_t* d_keys_sorted, const uint32_t* d_idx_sorted, const SlabEntry* d_raw,
                       const uint32_t* d_list_start, const uint2* d_hdr, int64_t n, int cpw, int tc,
                       SlabEntry* d_table, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(slab_place_kernel, dim3(grid_of(n)), dim3(256), 0, stream, d_keys_sorted, d_idx_sorted, d_raw,
                     d_list_start, d_hdr, n, cpw, tc, d_table);
}

void launch_slab_fill(SlabEntry* d_table, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(slab_fill_kernel, dim3(grid_of(n)), dim3(256), 0, stream, d_table, n);
}

void launch_back_project_slab(BpMode mode, const SlabArgs& A, int subset_begin, int subset_end, const float* d_proj,
                              float* d_img, const ImageDims& dims, const BpData& data, hipStream_t stream) {
  const int TC = A.tx * A.ty;
  const dim3 block(kWave * A.nwaves);
  const size_t lds = static_cast<size_t>(A.slab_floats) * sizeof(float);
  if (A.nbusy > 0) {
    const dim3 grid(A.nbusy);
    // slabs above 64 KB need the kernel's dynamic LDS limit raised (once per instantiation)
#define MI_SLAB(M, C)                                                                                               \
  do {                                                                                                              \
    static bool attr_set = false;                                                                                   \
    if (!attr_set) {                                                                                                \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bp_slab_kernel<M, C>),                               \
                                hipFuncAttributeMaxDynamicSharedMemorySize, kSlabMaxBytes);                         \
      attr_set = true;                                                                                              \
    }                                                                                                               \
    hipLaunchKernelGGL((bp_slab_kernel<M, C>), grid, block, lds, stream, A, subset_begin, subset_end, d_proj, d_img, \
                       dims, data);                                                                                 \
  } while (0)
#define MI_SLAB_C(M)                              \
  switch (A.cpw) {                                \
    case 2: MI_SLAB(M, 2); break;                 \
    case 4: MI_SLAB(M, 4); break;                 \
    default: MI_SLAB(M, 8); break;                \
  }
    if (mode == BP_PLAIN) { MI_SLAB_C(BP_PLAIN) }
    else { MI_SLAB_C(BP_OSEM_UPDATE) }
#undef MI_SLAB_C
#undef MI_SLAB
  }
  const int nempty = A.ntiles - A.nbusy;
  if (nempty > 0) {
    const int64_t n = static_cast<int64_t>(nempty) * TC * dims.nzp;
    if (mode == BP_PLAIN)
      hipLaunchKernelGGL(slab_empty_tiles_kernel<BP_PLAIN>, dim3(grid_of(n)), dim3(256), 0, stream, A,
                         A.tile_order + A.nbusy, nempty, d_img, dims, data);
    else
      hipLaunchKernelGGL(slab_empty_tiles_kernel<BP_OSEM_UPDATE>, dim3(grid_of(n)), dim3(256), 0, stream, A,
                         A.tile_order + A.nbusy, nempty, d_img, dims, data);
  }
}

}  // namespace mi
