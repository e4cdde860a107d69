// Time-of-flight forward and back projection.
//
// TOF rows are the non-TOF basic rows with each value multiplied by the TOF
// kernel of the element's position along the LOR
// (ProjMatrixByBin::apply_tof_kernel, ProjMatrixByBin.inl:115-197). TOF keeps
// only the z-shift symmetry, so every family is one (segment, view,
// tangential position) with its own row, and the kernel value of an element
// is the same for all axial positions of the family. The weights are not
// stored: a wave computes them for 64 row elements at a time (one element
// per lane, FastErf interpolation in FP64) and broadcasts them with
// readlane while it walks the row, so projection data are read or written
// once per TOF bin and the image once per matrix element, as without TOF.
//
// Projection data: TOF bin outermost, each TOF block in the non-TOF device
// layout. Arithmetic follows the oracle's restatement in float order
// (compiled with -ffp-contract=off), so forward projections match it bit for
// bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "pmrt_device.hpp"
#include "pmrt_kernels.hpp"

namespace mi {

namespace {

using namespace dev;

// d = -<voxel - mid, unit> with the voxel's scanner coordinates
// (ProjMatrixByBin.inl:159-167)
__device__ inline float tof_distance(const FamilyLor& L, int z, int y, int x, const TofConsts& T) {
  const float pz = (T.vz * static_cast<float>(z) + T.oz) - T.z_off;
  const float py = (T.vy * static_cast<float>(y) + T.oy) - 0.F;
  const float px = (T.vx * static_cast<float>(x) + T.ox) - 0.F;
  float ip = 0.F;
  ip = ip + (pz - L.mz) * L.uz;
  ip = ip + (py - L.my) * L.uy;
  ip = ip + (px - L.mx) * L.ux;
  return -ip;
}

// FastErf::get_erf_linear_interpolation (FastErf.inl:46-61)
__device__ inline double erf_interp(const double* __restrict__ tab, double xp) {
  xp = fmax(fmin(kErfMax, xp), -kErfMax);
  const double xi = (xp + kErfMax) / ((2 * kErfMax) / kErfSamples);
  const int lo = static_cast<int>(floor(xi));
  return tab[lo] + (xi - lo) * (tab[lo + 1] - tab[lo]);
}

// value * get_tof_value(low - d, high - d) (ProjMatrixByBin.inl:163-172,186-195);
// 0 where the reference drops the element
__device__ inline float tof_weight(float w, float d, int k, const TofConsts& T) {
  const float d1n = (T.low[k] - d) * T.r;
  const float d2n = (T.high[k] - d) * T.r;
  if ((d1n >= 4.F && d2n >= 4.F) || (d1n <= -4.F && d2n <= -4.F)) return 0.F;
  const float v = static_cast<float>(0.5 * (erf_interp(T.erf, d2n) - erf_interp(T.erf, d1n)));
  return (v > 0.F && w != 0.F) ? w * v : 0.F;
}

// TOF weights of one element for all TOF bins. With shared bin boundaries
// (high[k] == low[k+1], as set_tof_mash_factor produces them) the erf of each
// boundary is evaluated once, and only for bins that pass the zero test;
// the values are those of tof_weight bit for bit.
template <int NT>
__device__ inline void tof_weights(float w, float d, const TofConsts& T, float (&tw)[NT]) {
  if (!T.shared_bounds) {
#pragma unroll
    for (int k = 0; k < NT; ++k) tw[k] = k < T.ntof ? tof_weight(w, d, k, T) : 0.F;
    return;
  }
  float xb[NT + 1];
  xb[0] = (T.low[0] - d) * T.r;
#pragma unroll
  for (int j = 1; j <= NT; ++j) xb[j] = j <= T.ntof ? (T.high[j - 1] - d) * T.r : 0.F;
  bool live[NT];
  bool need[NT + 1];
#pragma unroll
  for (int j = 0; j <= NT; ++j) need[j] = false;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    live[k] = k < T.ntof && !((xb[k] >= 4.F && xb[k + 1] >= 4.F) || (xb[k] <= -4.F && xb[k + 1] <= -4.F));
    if (live[k]) {
      need[k] = true;
      need[k + 1] = true;
    }
  }
  double e[NT + 1];
#pragma unroll
  for (int j = 0; j <= NT; ++j) e[j] = need[j] ? erf_interp(T.erf, xb[j]) : 0.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const float v = static_cast<float>(0.5 * (e[k + 1] - e[k]));
    tw[k] = (live[k] && v > 0.F && w != 0.F) ? w * v : 0.F;
  }
}

__device__ inline float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

constexpr int kTofUnroll = 2;

// One family per wave; lanes over axial positions (NSLOT per lane), NT >= ntof.
template <int MODE, int NT, int NSLOT>
__device__ void fp_tof_task(const Family& f, const FamilyLor& lor, const int2* __restrict__ pool,
                            const float* __restrict__ img, float* __restrict__ out, const ImageDims& d,
                            const TofConsts& T, const FpData& data, int t, int lane) {
  int zb[NSLOT];
  bool valid[NSLOT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int a = lane + kWave * q;
    valid[q] = a < f.nax;
    zb[q] = f.zadd0 + f.ppa * a - d.min_z;
  }
  float acc[NSLOT][NT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q)
#pragma unroll
    for (int k = 0; k < NT; ++k) acc[q][k] = 0.F;
  const int2* row = pool + f.row_start;
  const int n = f.row_len;
  const int zs = f.zsign;
  for (int c0 = 0; c0 < n; c0 += kWave) {
    // TOF weights of element c0 + lane
    float tw[NT];
    {
      const int e = c0 + lane;
      float dist = 0.F, w = 0.F;
      if (e < n) {
        const int2 el = row[e];
        int x, y, z;
        decode(el.x, x, y, z);
        dist = tof_distance(lor, z, y, x, T);
        w = __int_as_float(el.y);
      }
      tof_weights<NT>(w, dist, T, tw);
    }
    const int m = min(kWave, n - c0);
    int j = 0;
#if defined(__HIP_DEVICE_COMPILE__)
    // each element's image column as a range-checked buffer of nz values: a
    // voxel outside the image reads 0 and adds nothing (0 * w == 0 for the
    // finite, non-zero weights used), so no per-lane clamp, test or select;
    // lanes without a bin sum values that are never stored
    for (; j + kTofUnroll <= m; j += kTofUnroll) {
      float v[kTofUnroll][NSLOT];
#pragma unroll
      for (int u = 0; u < kTofUnroll; ++u) {
        int x, y, z;
        decode(row[c0 + j + u].x, x, y, z);
        const int col = (y - d.min_y) * d.nx + (x - d.min_x);
        const __amdgpu_buffer_rsrc_t column =
            run_buffer(as_const(img), static_cast<unsigned>(col * d.nzp), d.nz);
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) v[u][q] = load_run(column, 4 * (__mul24(zs, z) + zb[q]));
      }
#pragma unroll
      for (int u = 0; u < kTofUnroll; ++u)
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          if (k < T.ntof) {
            const float wk = bcast(tw[k], j + u);
            if (wk != 0.F) {
#pragma unroll
              for (int q = 0; q < NSLOT; ++q) acc[q][k] = acc[q][k] + v[u][q] * wk;
            }
          }
        }
    }
#endif
    for (; j + kTofUnroll <= m; j += kTofUnroll) {
      float v[kTofUnroll][NSLOT];
      bool ok[kTofUnroll][NSLOT];
#pragma unroll
      for (int u = 0; u < kTofUnroll; ++u) {
        int x, y, z;
        decode(row[c0 + j + u].x, x, y, z);
        const float* colp = img + static_cast<int64_t>((y - d.min_y) * d.nx + (x - d.min_x)) * d.nzp;
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) {
          const int zz = __mul24(zs, z) + zb[q];  // zs = +-1, |z| < 2^11
          ok[u][q] = valid[q] && static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz);
          v[u][q] = load_u32off(colp, static_cast<unsigned>(clamp_index(zz, d.nz)));
        }
      }
#pragma unroll
      for (int u = 0; u < kTofUnroll; ++u)
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          // a predicate, not a break: the loop must stay fully unrolled so that
          // acc and tw remain in registers
          if (k < T.ntof) {
            // elements whose TOF weight is 0 are not in the reference's TOF row:
            // skip them (wave-uniform), so an infinite image value does not give 0 * inf
            const float wk = bcast(tw[k], j + u);
            if (wk != 0.F) {
#pragma unroll
              for (int q = 0; q < NSLOT; ++q)
                if (ok[u][q]) acc[q][k] = acc[q][k] + v[u][q] * wk;
            }
          }
        }
    }
    for (; j < m; ++j) {
      int x, y, z;
      decode(row[c0 + j].x, x, y, z);
      const float* colp = img + static_cast<int64_t>((y - d.min_y) * d.nx + (x - d.min_x)) * d.nzp;
      float v[NSLOT];
      bool ok[NSLOT];
#pragma unroll
      for (int q = 0; q < NSLOT; ++q) {
        const int zz = __mul24(zs, z) + zb[q];  // zs = +-1, |z| < 2^11
        ok[q] = valid[q] && static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz);
        v[q] = load_u32off(colp, static_cast<unsigned>(clamp_index(zz, d.nz)));
      }
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        if (k < T.ntof) {
          const float wk = bcast(tw[k], j);
          if (wk != 0.F) {
#pragma unroll
            for (int q = 0; q < NSLOT; ++q)
              if (ok[q]) acc[q][k] = acc[q][k] + v[q] * wk;
          }
        }
      }
    }
  }
  double ll = 0.0;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    if (!valid[q]) continue;
    const int64_t i0 = static_cast<int64_t>(f.sino_base) + lane + kWave * q;
#pragma unroll
    for (int k = 0; k < NT; ++k)
      if (k < T.ntof)
        fp_epilogue<MODE>(acc[q][k], static_cast<int64_t>(k) * T.block + i0, k * T.nvg + f.viewgram, data, out, ll);
  }
  if constexpr (fp_accumulates_loglik(MODE)) {
    for (int off = kWave / 2; off > 0; off >>= 1) ll += __shfl_xor(ll, off, kWave);
    if (lane == 0) data.partial[t] = ll;
  }
}

template <int MODE, int NT>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_tof_kernel(const FpTask* __restrict__ tasks, int ntasks,
                                                                       const Family* __restrict__ fams,
                                                                       const FamilyLor* __restrict__ lors,
                                                                       const int2* __restrict__ pool,
                                                                       const float* __restrict__ img,
                                                                       float* __restrict__ out, ImageDims d, TofConsts T,
                                                                       FpData data) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int t = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (t >= ntasks) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int fi = tasks[t].first;
  const Family f = fams[fi];
  const FamilyLor lor = lors[fi];
  // TOF families hold at most 2 x 64 axial positions (checked at set-up)
  if (f.nax <= kWave) fp_tof_task<MODE, NT, 1>(f, lor, pool, img, out, d, T, data, t, lane);
  else fp_tof_task<MODE, NT, 2>(f, lor, pool, img, out, d, T, data, t, lane);
}

// Voxel-driven TOF back projection: one wave per image column, as bp_kernel;
// for each table entry only the TOF bins with a non-zero weight are read.
template <int MODE, int PPA, int NSLOT, int NT>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_tof_kernel(const int64_t* __restrict__ list_off,
                                                                       const BpEntry* __restrict__ ent,
                                                                       const float* __restrict__ ent_d, int sb, int se,
                                                                       const float* __restrict__ proj,
                                                                       float* __restrict__ img, ImageDims d, int nwork,
                                                                       int2 tile, int remap, TofConsts T, BpData data) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int w = (remap ? xcd_remap(blockIdx.x, gridDim.x) : static_cast<int>(blockIdx.x)) * kWavesPerBlock + wave;
  if (w >= nwork) return;
  const int col = tiled_column(w, d, tile);
  if (col < 0) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  float acc[NSLOT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) acc[q] = 0.F;
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    for (int64_t c0 = beg; c0 < end; c0 += kWave) {
      // TOF weights and the range of TOF bins with a non-zero weight, of entry c0 + lane
      float tw[NT];
      int klo = NT, khi = -1;
      {
        const int64_t e = c0 + lane;
        float wv = 0.F, dist = 0.F;
        if (e < end) {
          wv = ent[e].w;
          dist = ent_d[e];
        }
        tof_weights<NT>(wv, dist, T, tw);
#pragma unroll
        for (int k = 0; k < NT; ++k)
          if (tw[k] != 0.F) {
            klo = min(klo, k);
            khi = k;
          }
      }
      const int m = static_cast<int>(min(static_cast<int64_t>(kWave), end - c0));
#if defined(__HIP_DEVICE_COMPILE__)
      if constexpr (PPA == 1) {
        // one plane per axial position: the chunk's entries come from one
        // vector load (lane j holds entry c0 + j) and each (entry, TOF bin)
        // run is read as a range-checked buffer of nax values, so a plane
        // the entry does not reach reads 0 (x + w*0 == x); two entries per
        // step keep their gathers in flight together
        int4 mine = make_int4(0, 0, 0, 0);
        if (lane < m) mine = reinterpret_cast<cptr<int4>>(as_const(ent))[c0 + lane];
        int zi4[NSLOT];
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) zi4[q] = 4 * (d.min_z + lane + kWave * q);
        for (int j = 0; j < m; j += 2) {
          float pv[2][NT][NSLOT];
          float wv[2][NT];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int jj = j + u < m ? j + u : j;  // a repeated entry adds nothing (its weights are zeroed)
            const int kl = __builtin_amdgcn_readlane(klo, jj), kh = __builtin_amdgcn_readlane(khi, jj);
            const int base = __builtin_amdgcn_readlane(mine.x, jj);
            const int k4 = 4 * __builtin_amdgcn_readlane(mine.y, jj);
            const int nax = __builtin_amdgcn_readlane(mine.z, jj) & 0xffff;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              const float wk = (j + u < m && k >= kl && k <= kh) ? bcast(tw[k], jj) : 0.F;
              wv[u][k] = wk;
#pragma unroll
              for (int q = 0; q < NSLOT; ++q) pv[u][k][q] = 0.F;
              if (wk != 0.F) {  // wave-uniform: only the TOF bins of the reference's TOF row
                const __amdgpu_buffer_rsrc_t run =
                    run_buffer(as_const(proj), static_cast<unsigned>(static_cast<int64_t>(k) * T.block + base), nax);
#pragma unroll
                for (int q = 0; q < NSLOT; ++q) pv[u][k][q] = load_run(run, zi4[q] - k4);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int k = 0; k < NT; ++k)
              if (wv[u][k] != 0.F) {
#pragma unroll
                for (int q = 0; q < NSLOT; ++q) acc[q] = acc[q] + wv[u][k] * pv[u][k][q];
              }
        }
        continue;
      }
#endif
      for (int j = 0; j < m; ++j) {
        const int kl = __builtin_amdgcn_readlane(klo, j), kh = __builtin_amdgcn_readlane(khi, j);
        if (kh < 0) continue;
        const BpEntry E = ent[c0 + j];
        unsigned a[NSLOT];
        bool ok[NSLOT];
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) a[q] = entry_axial<PPA>(E.k, E.nax_ppa, d.min_z + lane + kWave * q, ok[q]);
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          if (k < kl || k > kh) continue;
          const float wk = bcast(tw[k], j);
          if (wk == 0.F) continue;  // not in the reference's TOF row
          const float* run = proj + static_cast<int64_t>(k) * T.block + E.sino_base;
#pragma unroll
          for (int q = 0; q < NSLOT; ++q) {
            const float p = load_u32off(run, a[q]);
            if (ok[q]) acc[q] = acc[q] + wk * p;
          }
        }
      }
    }
  }
  float* outc = img + static_cast<int64_t>(col) * d.nzp;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int z = lane + kWave * q;
    if (z >= d.nzp) continue;
    if constexpr (MODE == BP_PLAIN) {
      outc[z] = z < d.nz ? acc[q] : 0.F;
    } else {
      if (z < d.nz) {
        const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
        const float u = osem_factor(acc[q], data.sens[idx], data.threshold, data.thr_lo, data.thr_hi);
        data.estimate[idx] = data.estimate[idx] * u;
        if (data.update) data.update[idx] = u;
      }
    }
  }
}

__global__ void __launch_bounds__(64) entry_tof_distance_kernel(const Family* __restrict__ fams, int fam_begin,
                                                                int fam_end, const int64_t* __restrict__ fam_entry_off,
                                                                const FamilyLor* __restrict__ lors,
                                                                const int2* __restrict__ pool, TofConsts T,
                                                                float* __restrict__ out) {
  const int fi = fam_begin + blockIdx.x;
  if (fi >= fam_end) return;
  const Family f = fams[fi];
  const FamilyLor L = lors[fi];
  const int64_t base = fam_entry_off[fi - fam_begin];
  for (int e = threadIdx.x; e < f.row_len; e += blockDim.x) {
    int x, y, z;
    decode(pool[f.row_start + e].x, x, y, z);
    out[base + e] = tof_distance(L, z, y, x, T);
  }
}

__global__ void gather_floats_kernel(const float* __restrict__ src, const uint32_t* __restrict__ idx,
                                     float* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[idx[i]];
}

// ---- stored TOF rows

// the first and last TOF bin with a non-zero weight (klo > khi when none)
template <int NT>
__device__ inline void live_range(const float (&tw)[NT], int& klo, int& khi) {
  klo = NT;
  khi = -1;
#pragma unroll
  for (int k = 0; k < NT; ++k)
    if (tw[k] != 0.F) {
      klo = min(klo, k);
      khi = k;
    }
}

__global__ void tof_live_span_kernel(const BpEntry* __restrict__ ent, const float* __restrict__ ent_d, int64_t n,
                                     TofConsts T, int* __restrict__ span) {
  int best = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    float tw[kMaxTofBins];
    tof_weights<kMaxTofBins>(ent[i].w, ent_d[i], T, tw);
    int klo, khi;
    live_range(tw, klo, khi);
    best = max(best, khi - klo + 1);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, kWave));
  if (threadIdx.x % kWave == 0) atomicMax(span, best);
}

// one wave per family, lanes over its row; the same weights as fp_tof_task
__global__ void __launch_bounds__(kWave) tof_fp_records_kernel(const Family* __restrict__ fams, int nfam,
                                                               const FamilyLor* __restrict__ lors,
                                                               const int2* __restrict__ pool, ImageDims d,
                                                               TofConsts T, int nt, int w, int32_t* __restrict__ rec) {
  const int fi = blockIdx.x;
  if (fi >= nfam) return;
  const Family f = fams[fi];
  const FamilyLor lor = lors[fi];
  const int words = tof_fp_words(w);
  for (int e = threadIdx.x; e < f.row_len; e += kWave) {
    const int2 el = pool[f.row_start + e];
    int x, y, z;
    decode(el.x, x, y, z);
    float tw[kMaxTofBins];
    tof_weights<kMaxTofBins>(__int_as_float(el.y), tof_distance(lor, z, y, x, T), T, tw);
    int klo, khi;
    live_range(tw, klo, khi);
    const int c = khi < klo ? 0 : min(klo, nt - w);
    int32_t* r = rec + (static_cast<int64_t>(f.row_start) + e) * words;
    r[0] = ((y - d.min_y) * d.nx + (x - d.min_x)) | (c << 20);
    r[1] = static_cast<int32_t>(static_cast<uint32_t>(f.zsign * z) << 16);
    for (int j = 0; j < w; ++j) r[2 + j] = __float_as_int(c + j < kMaxTofBins ? tw[c + j] : 0.F);
  }
}

// run lengths of equal window starts, from the end of each row (one thread
// per family)
__global__ void tof_fp_runs_kernel(const Family* __restrict__ fams, int nfam, int w, int32_t* __restrict__ rec) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= nfam) return;
  const Family f = fams[fi];
  const int words = tof_fp_words(w);
  int32_t* r = rec + static_cast<int64_t>(f.row_start) * words;
  int run = 0, next_c = -1;
  for (int e = f.row_len - 1; e >= 0; --e) {
    const int c = r[static_cast<int64_t>(e) * words] >> 20;
    run = (c == next_c && run < 0xffff) ? run + 1 : 1;
    next_c = c;
    int32_t& h = r[static_cast<int64_t>(e) * words + 1];
    h = (h & static_cast<int32_t>(0xffff0000u)) | run;
  }
}

__global__ void tof_bp_records_kernel(const BpEntry* __restrict__ ent, const float* __restrict__ ent_d, int64_t n,
                                      TofConsts T, int w, int32_t* __restrict__ rec) {
  const int words = tof_bp_words(w);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const BpEntry E = ent[i];
    float tw[kMaxTofBins];
    tof_weights<kMaxTofBins>(E.w, ent_d[i], T, tw);
    int klo, khi;
    live_range(tw, klo, khi);
    const int nk = khi < klo ? 0 : khi - klo + 1;
    if (nk == 0) klo = 0;
    int32_t* r = rec + i * words;
    r[0] = static_cast<int32_t>(klo * T.block + E.sino_base);  // the run of the window's first bin
    r[1] = E.k;
    r[2] = (E.nax_ppa & 0xfffff) | (klo << 20) | (nk << 25);
    for (int j = 0; j < w; ++j) r[3 + j] = __float_as_int(j < nk ? tw[klo + j] : 0.F);
  }
}

// a * b, except that a zero operand gives +0 whatever the other one is
// (v_mul_legacy_f32)
__device__ float mul_legacy(float a, float b) __asm("llvm.amdgcn.fmul.legacy");

template <int N>
struct alignas(16) Words {
  int32_t v[N];
};

// acc[q][C + j] += v[q] * w[j] for the W bins of a window starting at C; the
// legacy multiply gives +0 for a zero weight whatever the image value (an
// element outside the reference's TOF row adds nothing, as it is skipped
// there), and x + (+0) == x for every accumulator value other than -0, which
// a sum started at +0 never takes
template <int C, int NT, int W, int NSLOT>
__device__ inline void window_add(float (&acc)[NSLOT][NT], const float (&v)[NSLOT], const int32_t* wb) {
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const float wj = __int_as_float(wb[j]);
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) acc[q][C + j] = acc[q][C + j] + mul_legacy(v[q], wj);
  }
}

constexpr int kRowGroup = 4;  // elements whose image gathers are issued together

// the elements [e, stop) of a run whose window starts at bin C; LO32: the
// image lies inside one 4 GB-aligned window (a column's buffer start is one
// 32-bit add)
template <int C, int NT, int W, int NSLOT, bool LO32>
__device__ inline void fp_tof_run(cptr<int32_t> row, int e, int stop, const float* __restrict__ img,
                                  const ImageDims& d, const int (&zb4)[NSLOT], float (&acc)[NSLOT][NT]) {
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr int R = tof_fp_words(W);
  const cptr<Words<R>> rec = reinterpret_cast<cptr<Words<R>>>(row);
  // an image voxel outside the column's nz planes (or a negative plane)
  // reads 0 from the range-checked column buffer and adds +0
  for (; e + kRowGroup <= stop; e += kRowGroup) {
    Words<R> E[kRowGroup];
    float v[kRowGroup][NSLOT];
#pragma unroll
    for (int u = 0; u < kRowGroup; ++u) {
      E[u] = rec[e + u];
      const unsigned colf = static_cast<unsigned>(__mul24(E[u].v[0] & 0xfffff, d.nzp));
      const __amdgpu_buffer_rsrc_t column =
          LO32 ? run_buffer_lo32(as_const(img), colf, d.nz) : run_buffer(as_const(img), colf, d.nz);
      const int z4 = (E[u].v[1] >> 16) * 4;
#pragma unroll
      for (int q = 0; q < NSLOT; ++q) v[u][q] = load_run(column, z4 + zb4[q]);
    }
#pragma unroll
    for (int u = 0; u < kRowGroup; ++u) window_add<C, NT, W, NSLOT>(acc, v[u], E[u].v + 2);
  }
  for (; e < stop; ++e) {
    const Words<R> E = rec[e];
    const unsigned colf = static_cast<unsigned>(__mul24(E.v[0] & 0xfffff, d.nzp));
    const __amdgpu_buffer_rsrc_t column =
        LO32 ? run_buffer_lo32(as_const(img), colf, d.nz) : run_buffer(as_const(img), colf, d.nz);
    const int z4 = (E.v[1] >> 16) * 4;
    float v[NSLOT];
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) v[q] = load_run(column, z4 + zb4[q]);
    window_add<C, NT, W, NSLOT>(acc, v, E.v + 2);
  }
#endif
}

// One family per wave over stored records: lanes over axial positions, each
// element's image column gathered once and added to the accumulators of its
// window, in row order (the reference's per-bin sum order). The window start
// is dispatched once per run of elements that share it.
template <int MODE, int NT, int W, int NSLOT, bool LO32>
__device__ void fp_tof_rows_task(const Family& f, const int32_t* __restrict__ rec, const float* __restrict__ img,
                                 float* __restrict__ out, const ImageDims& d, const TofConsts& T, const FpData& data,
                                 int t, int lane) {
  constexpr int R = tof_fp_words(W);
  static_assert(NT - W < 16, "window start beyond the dispatch");
  int zb4[NSLOT];
  bool valid[NSLOT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int a = lane + kWave * q;
    valid[q] = a < f.nax;
    zb4[q] = 4 * (f.zadd0 + f.ppa * a - d.min_z);
  }
  float acc[NSLOT][NT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q)
#pragma unroll
    for (int k = 0; k < NT; ++k) acc[q][k] = 0.F;
#if defined(__HIP_DEVICE_COMPILE__)
  const cptr<int32_t> row = as_const(rec) + static_cast<int64_t>(f.row_start) * R;
  const int n = f.row_len;
  for (int e = 0; e < n;) {
    const int c = row[static_cast<int64_t>(e) * R] >> 20;
    const int stop = e + (row[static_cast<int64_t>(e) * R + 1] & 0xffff);
    switch (c) {
#define MI_TW(C)                                                                               \
  case C:                                                                                      \
    if constexpr (C <= NT - W) fp_tof_run<C, NT, W, NSLOT, LO32>(row, e, stop, img, d, zb4, acc); \
    break;
      MI_TW(0) MI_TW(1) MI_TW(2) MI_TW(3) MI_TW(4) MI_TW(5) MI_TW(6) MI_TW(7)
      MI_TW(8) MI_TW(9) MI_TW(10) MI_TW(11) MI_TW(12) MI_TW(13) MI_TW(14) MI_TW(15)
#undef MI_TW
      default: break;
    }
    e = stop;
  }
#endif
  double ll = 0.0;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    if (!valid[q]) continue;
    const int64_t i0 = static_cast<int64_t>(f.sino_base) + lane + kWave * q;
#pragma unroll
    for (int k = 0; k < NT; ++k)
      if (k < T.ntof)
        fp_epilogue<MODE>(acc[q][k], static_cast<int64_t>(k) * T.block + i0, k * T.nvg + f.viewgram, data, out, ll);
  }
  if constexpr (fp_accumulates_loglik(MODE)) {
    for (int off = kWave / 2; off > 0; off >>= 1) ll += __shfl_xor(ll, off, kWave);
    if (lane == 0) data.partial[t] = ll;
  }
}

template <int MODE, int NT, int W, bool LO32>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_tof_rows_kernel(const FpTask* __restrict__ tasks,
                                                                            int ntasks,
                                                                            const Family* __restrict__ fams,
                                                                            const int32_t* __restrict__ rec,
                                                                            const float* __restrict__ img,
                                                                            float* __restrict__ out, ImageDims d,
                                                                            TofConsts T, FpData data) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int t = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (t >= ntasks) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const Family f = fams[tasks[t].first];
  if (f.nax <= kWave) fp_tof_rows_task<MODE, NT, W, 1, LO32>(f, rec, img, out, d, T, data, t, lane);
  else fp_tof_rows_task<MODE, NT, W, 2, LO32>(f, rec, img, out, d, T, data, t, lane);
}

constexpr int kEntryGroup = 2;  // table entries whose projection-data gathers are issued together

// Voxel-driven back projection over stored records: one wave per image
// column, as bp_tof_kernel. Each entry gathers the runs of its W window bins
// at once; bins past its live range get an empty buffer (no read, value 0)
// and weight 0, and the legacy multiply makes a zero weight add +0 whatever
// the projection value.
template <int MODE, int PPA, int NSLOT, int W>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_tof_rows_kernel(const int64_t* __restrict__ list_off,
                                                                            const int32_t* __restrict__ rec, int sb,
                                                                            int se, const float* __restrict__ proj,
                                                                            float* __restrict__ img, ImageDims d,
                                                                            int nwork, int2 tile, int remap,
                                                                            int64_t block, int accumulate,
                                                                            int64_t sentinel, BpData data) {
  constexpr int R = tof_bp_words(W);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int b = remap == 1 ? xcd_remap(blockIdx.x, gridDim.x)
                : remap > 1  ? xcd_chunk_remap(blockIdx.x, remap)
                             : static_cast<int>(blockIdx.x);
  const int w = b * kWavesPerBlock + wave;
  if (w >= nwork) return;
  const int col = tiled_column(w, d, tile);
  if (col < 0) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  float acc[NSLOT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) acc[q] = 0.F;
#if defined(__HIP_DEVICE_COMPILE__)
  const cptr<Words<R>> table = reinterpret_cast<cptr<Words<R>>>(as_const(rec));
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    int64_t e = beg;
    for (; e < end; e += kEntryGroup) {
      float p[kEntryGroup][W][NSLOT];
      float wt[kEntryGroup][W];
#pragma unroll
      for (int u = 0; u < kEntryGroup; ++u) {
        // past the list's end: the all-zero record after the table (no live
        // bin, weights 0), so it adds nothing
        const Words<R> E = table[e + u < end ? e + u : sentinel];
        const int base = E.v[0], k = E.v[1], np = E.v[2] & 0xfffff;
        const int nk = static_cast<int>(static_cast<unsigned>(E.v[2]) >> 25);
        const int nax = np & 0xffff;
        int off[NSLOT];
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) {
          const int zi = d.min_z + lane + kWave * q;
          if constexpr (PPA == 1) {
            off[q] = whole_offset(4 * (zi - k));  // outside [0, nax) the run buffer reads 0
          } else {
            bool ok;
            const unsigned a = entry_axial<PPA>(k, np, zi, ok);
            off[q] = ok ? static_cast<int>(a << 2) : -4;
          }
        }
        // the record holds the run of bin klo; the run of bin klo + j starts
        // one TOF block after that of bin klo + j - 1: one 64-bit step per bin
        uint64_t run_addr = reinterpret_cast<uint64_t>(proj) + 4 * static_cast<uint64_t>(static_cast<uint32_t>(base));
#pragma unroll
        for (int j = 0; j < W; ++j) {
          wt[u][j] = __int_as_float(E.v[3 + j]);  // stored as 0 past the live range
          const __amdgpu_buffer_rsrc_t run = __builtin_amdgcn_make_buffer_rsrc(
              reinterpret_cast<void*>(run_addr), static_cast<short>(0), j < nk ? 4 * nax : 0, 0x00020000);
#pragma unroll
          for (int q = 0; q < NSLOT; ++q) p[u][j][q] = load_run(run, off[q]);
          run_addr += 4 * static_cast<uint64_t>(block);
        }
      }
#pragma unroll
      for (int u = 0; u < kEntryGroup; ++u)
#pragma unroll
        for (int j = 0; j < W; ++j)
#pragma unroll
          for (int q = 0; q < NSLOT; ++q) acc[q] = acc[q] + mul_legacy(wt[u][j], p[u][j][q]);
    }
  }
#endif
  float* outc = img + static_cast<int64_t>(col) * d.nzp;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int z = lane + kWave * q;
    if (z >= d.nzp) continue;
    if constexpr (MODE == BP_PLAIN) {
      outc[z] = z < d.nz ? (accumulate ? outc[z] + acc[q] : acc[q]) : 0.F;
    } else {
      if (z < d.nz) {
        const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
        const float u = osem_factor(acc[q], data.sens[idx], data.threshold, data.thr_lo, data.thr_hi);
        data.estimate[idx] = data.estimate[idx] * u;
        if (data.update) data.update[idx] = u;
      }
    }
  }
}

}  // namespace

void launch_forward_project_tof(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams,
                                const FamilyLor* d_lors, const int2* d_pool, const float* d_img, float* d_out,
                                const ImageDims& dims, const TofConsts& tof, const FpData& data, hipStream_t stream) {
  if (ntasks <= 0) return;
  const dim3 grid((ntasks + kWavesPerBlock - 1) / kWavesPerBlock), block(kWave * kWavesPerBlock);
#define MI_FPT(M, NT)                                                                                              \
  hipLaunchKernelGGL((fp_tof_kernel<M, NT>), grid, block, 0, stream, d_tasks, ntasks, d_fams, d_lors, d_pool, d_img, \
                     d_out, dims, tof, data)
#define MI_FPT_CASE(M)                      \
  case M:                                   \
    if (tof.ntof <= 12) MI_FPT(M, 12);      \
    else MI_FPT(M, 16);                     \
    break;
  switch (mode) {
    MI_FPT_CASE(FP_PLAIN)
    MI_FPT_CASE(FP_RATIO)
    MI_FPT_CASE(FP_GRADIENT)
    MI_FPT_CASE(FP_LOGLIK)
    MI_FPT_CASE(FP_GRADIENT_LOGLIK)
    MI_FPT_CASE(FP_RATIO_LOGLIK)
    MI_FPT_CASE(FP_DRATIO)
    MI_FPT_CASE(FP_SCALE)
    MI_FPT_CASE(FP_LM_RATIO)
    MI_FPT_CASE(FP_LM_VALUE)
  }
#undef MI_FPT_CASE
#undef MI_FPT
}

void launch_back_project_tof(BpMode mode, int ppa, const int64_t* d_list_off, const BpEntry* d_entries,
                             const float* d_entry_d, int subset_begin, int subset_end, const float* d_proj,
                             float* d_img, const ImageDims& dims, const TofConsts& tof, const BpData& data,
                             const BpTuning& tune, hipStream_t stream) {
  const int nslot = (dims.nzp + kWave - 1) / kWave;
  const int2 tile = make_int2(tune.tile_x > 0 ? tune.tile_x : 1, tune.tile_y > 0 ? tune.tile_y : 1);
  const int nwork = ((dims.nx + tile.x - 1) / tile.x) * ((dims.ny + tile.y - 1) / tile.y) * tile.x * tile.y;
  const dim3 grid((nwork + kWavesPerBlock - 1) / kWavesPerBlock), block(kWave * kWavesPerBlock);
#define MI_BPT(M, P, S, NT)                                                                                         \
  hipLaunchKernelGGL((bp_tof_kernel<M, P, S, NT>), grid, block, 0, stream, d_list_off, d_entries, d_entry_d,       \
                     subset_begin, subset_end, d_proj, d_img, dims, nwork, tile, tune.xcd_remap, tof, data)
#define MI_BPT_NT(M, P, S)                 \
  if (tof.ntof <= 12) MI_BPT(M, P, S, 12); \
  else MI_BPT(M, P, S, 16);
#define MI_BPT_SLOTS(M, P)            \
  switch (nslot) {                    \
    case 1: MI_BPT_NT(M, P, 1) break; \
    case 2: MI_BPT_NT(M, P, 2) break; \
    case 3: MI_BPT_NT(M, P, 3) break; \
    default: MI_BPT_NT(M, P, 4) break; \
  }
  if (mode == BP_PLAIN) {
    if (ppa == 1) { MI_BPT_SLOTS(BP_PLAIN, 1) }
    else if (ppa == 2) { MI_BPT_SLOTS(BP_PLAIN, 2) }
    else { MI_BPT_SLOTS(BP_PLAIN, 0) }
  } else {
    if (ppa == 1) { MI_BPT_SLOTS(BP_OSEM_UPDATE, 1) }
    else if (ppa == 2) { MI_BPT_SLOTS(BP_OSEM_UPDATE, 2) }
    else { MI_BPT_SLOTS(BP_OSEM_UPDATE, 0) }
  }
#undef MI_BPT_SLOTS
#undef MI_BPT_NT
#undef MI_BPT
}

void launch_entry_tof_distance(const Family* d_fams, int fam_begin, int fam_end, const int64_t* d_fam_entry_off,
                               const FamilyLor* d_lors, const int2* d_pool, const ImageDims& dims,
                               const TofConsts& tof, float* d_entry_d, hipStream_t stream) {
  (void)dims;
  if (fam_end <= fam_begin) return;
  hipLaunchKernelGGL(entry_tof_distance_kernel, dim3(fam_end - fam_begin), dim3(64), 0, stream, d_fams, fam_begin,
                     fam_end, d_fam_entry_off, d_lors, d_pool, tof, d_entry_d);
}

void launch_gather_floats(const float* d_src, const uint32_t* d_idx, float* d_dst, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(gather_floats_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, d_src, d_idx,
                     d_dst, n);
}

namespace {
unsigned elementwise_blocks(int64_t n) { return static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 65536)); }
}  // namespace

void launch_tof_live_span(const BpEntry* d_entries, const float* d_entry_d, int64_t n, const TofConsts& tof,
                          int* d_span, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tof_live_span_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_entries, d_entry_d, n,
                     tof, d_span);
}

void launch_tof_fp_records(const Family* d_fams, int nfam, const FamilyLor* d_lors, const int2* d_pool,
                           const ImageDims& dims, const TofConsts& tof, int nt, int w, int32_t* d_rec,
                           hipStream_t stream) {
  if (nfam <= 0) return;
  hipLaunchKernelGGL(tof_fp_records_kernel, dim3(nfam), dim3(kWave), 0, stream, d_fams, nfam, d_lors, d_pool, dims, tof,
                     nt, w, d_rec);
  hipLaunchKernelGGL(tof_fp_runs_kernel, dim3((nfam + 255) / 256), dim3(256), 0, stream, d_fams, nfam, w, d_rec);
}

void launch_tof_bp_records(const BpEntry* d_entries, const float* d_entry_d, int64_t n, const TofConsts& tof, int w,
                           int32_t* d_rec, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tof_bp_records_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, stream, d_entries, d_entry_d, n,
                     tof, w, d_rec);
}

void launch_forward_project_tof_rows(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams,
                                     const int32_t* d_rec, int nt, int w, const float* d_img, float* d_out,
                                     const ImageDims& dims, const TofConsts& tof, const FpData& data,
                                     hipStream_t stream) {
  if (ntasks <= 0) return;
  const dim3 grid((ntasks + kWavesPerBlock - 1) / kWavesPerBlock), block(kWave * kWavesPerBlock);
  // the image inside one 4 GB-aligned window of the address space
  const uint64_t first = reinterpret_cast<uint64_t>(d_img);
  const bool lo32 = (first >> 32) == ((first + static_cast<uint64_t>(dims.ncols) * dims.nzp * 4 - 1) >> 32);
#define MI_FPR(M, NT, W)                                                                                           \
  if (lo32)                                                                                                         \
    hipLaunchKernelGGL((fp_tof_rows_kernel<M, NT, W, true>), grid, block, 0, stream, d_tasks, ntasks, d_fams, d_rec, \
                       d_img, d_out, dims, tof, data);                                                              \
  else                                                                                                              \
    hipLaunchKernelGGL((fp_tof_rows_kernel<M, NT, W, false>), grid, block, 0, stream, d_tasks, ntasks, d_fams,      \
                       d_rec, d_img, d_out, dims, tof, data)
#define MI_FPR_CASE(M)                                            \
  case M:                                                         \
    if (nt == 12 && w == kTofFpNarrow) { MI_FPR(M, 12, kTofFpNarrow); } \
    else { MI_FPR(M, 16, kTofFpWide); }                               \
    break;
  switch (mode) {
    MI_FPR_CASE(FP_PLAIN)
    MI_FPR_CASE(FP_RATIO)
    MI_FPR_CASE(FP_GRADIENT)
    MI_FPR_CASE(FP_LOGLIK)
    MI_FPR_CASE(FP_GRADIENT_LOGLIK)
    MI_FPR_CASE(FP_RATIO_LOGLIK)
    MI_FPR_CASE(FP_DRATIO)
    MI_FPR_CASE(FP_SCALE)
    MI_FPR_CASE(FP_LM_RATIO)
    MI_FPR_CASE(FP_LM_VALUE)
  }
#undef MI_FPR_CASE
#undef MI_FPR
}

void launch_back_project_tof_rows(BpMode mode, int ppa, const int64_t* d_list_off, const int32_t* d_rec, int w,
                                  int64_t num_entries, int subset_begin, int subset_end, const float* d_proj,
                                  float* d_img, const ImageDims& dims, const TofConsts& tof, const BpData& data,
                                  const BpTuning& tune, hipStream_t stream) {
  const int nslot = (dims.nzp + kWave - 1) / kWave;
  const int2 tile = make_int2(tune.tile_x > 0 ? tune.tile_x : 1, tune.tile_y > 0 ? tune.tile_y : 1);
  const int nwork = ((dims.nx + tile.x - 1) / tile.x) * ((dims.ny + tile.y - 1) / tile.y) * tile.x * tile.y;
  int nblocks = (nwork + kWavesPerBlock - 1) / kWavesPerBlock;
  if (tune.xcd_remap > 1) {  // chunked remap: whole rounds of 8 chunks
    const int round = 8 * tune.xcd_remap;
    nblocks = (nblocks + round - 1) / round * round;
  }
  const dim3 grid(nblocks), block(kWave * kWavesPerBlock);
  // a plain projection of several subsets runs one launch per subset, each
  // adding to the image: the waves in flight then share one subset's
  // projection data in L2 (11 TOF blocks make the whole data set many times
  // the L2 and MALL)
  const bool split = mode == BP_PLAIN && subset_end - subset_begin > 1;
  for (int s = subset_begin; s < subset_end; ++s) {
    const int sb = split ? s : subset_begin, se = split ? s + 1 : subset_end, accumulate = split && s > subset_begin;
#define MI_BPR(M, P, S, W)                                                                                    \
  hipLaunchKernelGGL((bp_tof_rows_kernel<M, P, S, W>), grid, block, 0, stream, d_list_off, d_rec, sb, se, d_proj, \
                     d_img, dims, nwork, tile, tune.xcd_remap, tof.block, accumulate, num_entries, data)
#define MI_BPR_W(M, P, S)                         \
  if (w == kTofBpNarrow) MI_BPR(M, P, S, kTofBpNarrow); \
  else MI_BPR(M, P, S, kTofBpWide);
#define MI_BPR_SLOTS(M, P)            \
  switch (nslot) {                    \
    case 1: MI_BPR_W(M, P, 1) break;  \
    case 2: MI_BPR_W(M, P, 2) break;  \
    case 3: MI_BPR_W(M, P, 3) break;  \
    default: MI_BPR_W(M, P, 4) break; \
  }
  if (mode == BP_PLAIN) {
    if (ppa == 1) { MI_BPR_SLOTS(BP_PLAIN, 1) }
    else if (ppa == 2) { MI_BPR_SLOTS(BP_PLAIN, 2) }
    else { MI_BPR_SLOTS(BP_PLAIN, 0) }
  } else {
    if (ppa == 1) { MI_BPR_SLOTS(BP_OSEM_UPDATE, 1) }
    else if (ppa == 2) { MI_BPR_SLOTS(BP_OSEM_UPDATE, 2) }
    else { MI_BPR_SLOTS(BP_OSEM_UPDATE, 0) }
  }
    if (!split) break;
  }
#undef MI_BPR_SLOTS
#undef MI_BPR_W
#undef MI_BPR
}

}  // namespace mi
