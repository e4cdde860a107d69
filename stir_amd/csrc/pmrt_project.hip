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
#include "pmrt_kernels.hpp"

namespace mi {

namespace {

using namespace dev;

constexpr int kUnroll = 8;      // matrix elements whose gathers are in flight together
constexpr int kRowSlack = 16;   // elements after the last row copy (read-ahead of the forward row loop)

// Sum of img * w over one row in row order for every bin of the lane
// (ProjMatrixElemsForOneBin.cxx:354-371), from the row's copy for the family's
// in-plane operation (see expand_rows_kernel): element = ((column << 12) |
// (z + 2048), weight), so the column needs no decoding or symmetry arithmetic,
// only its byte offset. Each element's image column is a range-checked raw
// buffer of nz values read at the lane's z: a voxel outside the image reads 0,
// and x + 0*w == x, so every bin still sums exactly the reference's terms. The
// elements of the next kUnroll group are requested before the current
// group's gathers are issued (the copies end in kUnroll elements of slack, so
// reading a group past the row's end stays inside the allocation; those
// elements are never used). LO32: the image lies inside one 4 GB-aligned
// window, so a column's buffer start is one 32-bit add (4 scalar instructions
// per element instead of 7; the CU's one scalar unit is shared by its waves).
template <int NSLOT, bool LO32>
__device__ inline void fp_row_ops(cptr<int2> row, int n, const LaneBins<NSLOT>& L, cptr<float> img, const ImageDims& d,
                                  float (&acc)[NSLOT]) {
  // elements per group: 16 gathers in flight whatever the slot count, so
  // that no task shape holds more registers than another needs
  constexpr int U = NSLOT >= 4 ? 4 : (NSLOT == 3 ? 5 : (NSLOT == 2 ? 8 : 16));
  static_assert(U <= kRowSlack, "read-ahead beyond the row copies' slack");
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) acc[q] = 0.F;
  int e = 0;
#if defined(__HIP_DEVICE_COMPILE__)
  if (n >= U) {
    // image z of element z + 2048 is zs (z + 2048) + zb - 2048 zs
    int zs4[NSLOT], zb4[NSLOT];
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) {
      zs4[q] = 4 * L.zs[q];
      zb4[q] = 4 * L.zb[q] - 2048 * zs4[q];
    }
    int2 cur[U];
#pragma unroll
    for (int j = 0; j < U; ++j) cur[j] = row[j];
    for (; e + U <= n; e += U) {
      int2 nxt[U];
#pragma unroll
      for (int j = 0; j < U; ++j) nxt[j] = row[e + U + j];
      float v[U][NSLOT], w[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const unsigned col = static_cast<unsigned>(cur[j].x) >> 12;
        const int zu = cur[j].x & 0xfff;
        const __amdgpu_buffer_rsrc_t column = LO32 ? run_buffer_lo32(img, col * static_cast<unsigned>(d.nzp), d.nz)
                                                   : run_buffer(img, col * static_cast<unsigned>(d.nzp), d.nz);
        w[j] = __int_as_float(cur[j].y);
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) v[j][q] = load_run(column, __mul24(zs4[q], zu) + zb4[q]);
      }
#pragma unroll
      for (int j = 0; j < U; ++j)
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) acc[q] = acc[q] + v[j][q] * w[j];
#pragma unroll
      for (int j = 0; j < U; ++j) cur[j] = nxt[j];
    }
  }
#endif
  for (; e < n; ++e) {
    const int2 el = row[e];
    const int col = static_cast<int>(static_cast<unsigned>(el.x) >> 12);
    const int z = (el.x & 0xfff) - 2048;
    const unsigned col4 = static_cast<unsigned>(col * d.nzp) << 2;
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) {
      const int zz = __mul24(L.zs[q], z) + L.zb[q];  // zs = +-1, |z| < 2^11
      if (L.valid[q] && static_cast<unsigned>(zz) < static_cast<unsigned>(d.nz))
        acc[q] = acc[q] + load_byteoff(img, col4 + (static_cast<unsigned>(zz) << 2)) * __int_as_float(el.y);
    }
  }
}

template <int MODE, int NSLOT, bool LO32 = false>
__device__ inline void fp_task(const Family& f1, const Family& f2, bool pair, cptr<int2> pool, cptr<float> img,
                               float* __restrict__ out, const ImageDims& d, const FpData& data, int t, int lane,
                               cptr<int2> ops, int64_t nelem) {
  LaneBins<NSLOT> L;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int g = lane + kWave * q;
    const bool in1 = g < f1.nax;
    const bool in2 = pair && !in1 && g < f1.nax + f2.nax;
    const int a = in1 ? g : g - f1.nax;
    L.valid[q] = in1 || in2;
    const Family& f = in1 ? f1 : f2;
    L.zs[q] = in1 ? f1.zsign : f2.zsign;
    L.zb[q] = (in1 ? f1.zadd0 + f1.ppa * a : f2.zadd0 + f2.ppa * a) - d.min_z;
    L.out[q] = static_cast<int64_t>(in1 ? f1.sino_base : f2.sino_base) + a;
    L.vg[q] = in1 ? f1.viewgram : f2.viewgram;
    (void)f;
  }
  float acc[NSLOT];
  fp_row_ops<NSLOT, LO32>(ops + static_cast<int64_t>(f1.inplane) * nelem + f1.row_start, f1.row_len, L, img, d, acc);
  double ll = 0.0;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q)
    if (L.valid[q]) fp_epilogue<MODE>(acc[q], L.out[q], L.vg[q], data, out, ll);
  if constexpr (fp_accumulates_loglik(MODE)) {
    for (int off = kWave / 2; off > 0; off >>= 1) ll += __shfl_xor(ll, off, kWave);
    if (lane == 0) as_global(data.partial)[t] = ll;
  }
}

// One wave per task; kWavesPerBlock tasks per workgroup. The number of axial
// slots per lane (1..4) follows the task's total number of bins.
template <int MODE, bool LO32>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_kernel(const FpTask* __restrict__ tasks, int ntasks,
                                                                   const Family* __restrict__ fams,
                                                                   const int2* __restrict__ pool,
                                                                   const float* __restrict__ img, float* __restrict__ out,
                                                                   ImageDims d, FpData data, int xcd_chunk,
                                                                   const int2* __restrict__ ops, int64_t nelem) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int b = xcd_chunk > 0 ? xcd_chunk_remap(blockIdx.x, xcd_chunk) : xcd_remap(blockIdx.x, gridDim.x);
  const int t = b * kWavesPerBlock + wave;
  if (t >= ntasks) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const FpTask task = tasks[t];
  const Family f1 = fams[task.first];
  const bool pair = task.second >= 0;
  const Family f2 = fams[pair ? task.second : task.first];
  const int total = f1.nax + (pair ? f2.nax : 0);
  const int nslot = (total + kWave - 1) / kWave;
  cptr<int2> rows = ops ? as_const(ops) : nullptr;
  switch (nslot) {
    case 1: fp_task<MODE, 1, LO32>(f1, f2, pair, as_const(pool), as_const(img), out, d, data, t, lane, rows, nelem); break;
    case 2: fp_task<MODE, 2, LO32>(f1, f2, pair, as_const(pool), as_const(img), out, d, data, t, lane, rows, nelem); break;
    case 3: fp_task<MODE, 3, LO32>(f1, f2, pair, as_const(pool), as_const(img), out, d, data, t, lane, rows, nelem); break;
    default: fp_task<MODE, 4, LO32>(f1, f2, pair, as_const(pool), as_const(img), out, d, data, t, lane, rows, nelem); break;
  }
}

// Voxel-driven back projection: one wave per image column (y,x), NSLOT
// z values per lane (z = lane + 64*slot), kWavesPerBlock columns per
// workgroup. Each table entry (wave-uniform, scalar loads) is decoded once
// for all slots; kUnroll entries have their projection-data reads in flight
// together and are accumulated in table order.
// Back projection of one image column (wave-uniform col) over subsets [sb, se).
// LO32: the projection data lie inside one 4 GB-aligned window (a run's
// buffer start is one 32-bit add).
// accumulate (BP_PLAIN): the sums start from the image's values (the sums of
// the subsets before sb, left by the previous launch), so a projection of
// several subsets in one launch per subset adds the same terms in the same
// order as one launch.
template <int MODE, int PPA, int NSLOT, bool SKIP, bool LO32 = false>
__device__ inline void bp_column(int col, cptr<int64_t> list_off, cptr<BpEntry> ent, int sb, int se, cptr<float> proj,
                                 float* __restrict__ img, const ImageDims& d, const BpData& data, int lane,
                                 bool accumulate = false) {
  float acc[NSLOT];
  int zi4[NSLOT];  // 4 * plane of each slot
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int z = lane + kWave * q;
    acc[q] = accumulate && z < d.nzp ? as_global(img)[static_cast<int64_t>(col) * d.nzp + z] : 0.F;
    zi4[q] = 4 * (d.min_z + lane + kWave * q);
    asm volatile("" : "+v"(zi4[q]));  // opaque: keeps zi4 - 4k one subtraction per entry
  }
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    int64_t e = beg;
    for (; e + kUnroll <= end; e += kUnroll) {
      float p[kUnroll][NSLOT], wt[kUnroll];
      bool ok[kUnroll][NSLOT];
      BpEntry Es[kUnroll];  // all kUnroll entries requested before any is decoded
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) Es[j] = ent[e + j];
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) {
        const BpEntry& E = Es[j];
        // the run's byte offset joins the lane's offset (one 32-bit add), so
        // the gather addresses from the kernel-wide base: the projection
        // data hold fewer than 2^30 values (checked at set-up)
        const unsigned run4 = static_cast<unsigned>(E.sino_base) << 2;
        wt[j] = E.w;
        // z range the entry's run can reach (wave-uniform): [k, k + ppa*(nax-1)]
        const int ppa_e = PPA > 0 ? PPA : (E.nax_ppa >> 16);
        const int zlo = E.k - d.min_z, zhi = zlo + ppa_e * ((E.nax_ppa & 0xffff) - 1);
#if defined(__HIP_DEVICE_COMPILE__)
        // one plane per axial position: the run as a buffer of nax values,
        // whose range check returns 0 for the planes the entry does not
        // reach (x + w*0 == x: the same sums as skipping them). (Taking this
        // path for the one-plane entries of mixed tables, behind a per-entry
        // test, made those tables' back projection slower.)
        if constexpr (PPA == 1 && !SKIP) {
          const __amdgpu_buffer_rsrc_t run =
              LO32 ? run_buffer_lo32(proj, static_cast<unsigned>(E.sino_base), E.nax_ppa & 0xffff)
                   : run_buffer(proj, static_cast<unsigned>(E.sino_base), E.nax_ppa & 0xffff);
#pragma unroll
          for (int q = 0; q < NSLOT; ++q) {
            ok[j][q] = true;
            p[j][q] = load_run(run, zi4[q] - 4 * E.k);
          }
          continue;
        }
#endif
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) {
          if (!SKIP || (zhi >= kWave * q && zlo < kWave * (q + 1))) {
            const unsigned off = entry_offset<PPA>(E.k, E.nax_ppa, d.min_z + lane + kWave * q, zi4[q], ok[j][q]);
            p[j][q] = load_byteoff(proj, run4 + off);
          } else {
            ok[j][q] = false;
            p[j][q] = 0.F;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kUnroll; ++j)
#pragma unroll
        for (int q = 0; q < NSLOT; ++q)
          if (ok[j][q]) acc[q] = acc[q] + wt[j] * p[j][q];
    }
    for (; e < end; ++e) {
      const BpEntry E = ent[e];
      const unsigned run4 = static_cast<unsigned>(E.sino_base) << 2;
#pragma unroll
      for (int q = 0; q < NSLOT; ++q) {
        bool ok;
        const unsigned off = entry_offset<PPA>(E.k, E.nax_ppa, d.min_z + lane + kWave * q, zi4[q], ok);
        if (ok) acc[q] = acc[q] + E.w * load_byteoff(proj, run4 + off);
      }
    }
  }
  gptr<float> out = as_global(img) + static_cast<int64_t>(col) * d.nzp;
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int z = lane + kWave * q;
    if (z >= d.nzp) continue;
    if constexpr (MODE == BP_PLAIN) {
      out[z] = z < d.nz ? acc[q] : 0.F;
    } else {
      if (z < d.nz) {
        const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
        const float u = osem_factor(acc[q], as_const(data.sens)[idx], data.threshold, data.thr_lo, data.thr_hi);
        as_global(data.estimate)[idx] = as_global(data.estimate)[idx] * u;
        if (data.update) as_global(data.update)[idx] = u;
      }
    }
  }
}

// Back projection of one short image column (nzp <= 64 / NPART planes) with
// the wave split into NPART parts of 64 / NPART lanes: part p takes the
// entries p, p + NPART, ... of the column's list (wave-uniform loads, each
// lane selecting its part's entry), so a column of 16 planes keeps all 64
// lanes busy. Each part sums its entries in table order; the parts' sums are
// added in part order (a fixed order, not the single-accumulator order of
// bp_column).
template <int MODE, int PPA, int NPART>
__device__ inline void bp_column_parts(int col, cptr<int64_t> list_off, cptr<BpEntry> ent, int sb, int se,
                                       cptr<float> proj, float* __restrict__ img, const ImageDims& d,
                                       const BpData& data, int lane) {
  constexpr int ZL = kWave / NPART;
  constexpr int U = kUnroll / NPART > 0 ? kUnroll / NPART : 1;  // entry groups in flight
  const int part = lane / ZL;
  const int zi = d.min_z + lane % ZL;
  float acc = 0.F;
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    int64_t e = beg;
    for (; e + NPART * U <= end; e += NPART * U) {
      float p[U], wt[U];
      bool ok[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        BpEntry E = load_entry_uniform(ent + e + j * NPART);
#pragma unroll
        for (int q = 1; q < NPART; ++q) {
          const BpEntry Eq = load_entry_uniform(ent + e + j * NPART + q);
          if (part == q) E = Eq;
        }
        wt[j] = E.w;
        const unsigned a = entry_axial<PPA>(E.k, E.nax_ppa, zi, ok[j]);
        p[j] = proj[static_cast<int64_t>(E.sino_base) + a];
      }
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (ok[j]) acc = acc + wt[j] * p[j];
    }
    for (; e < end; e += NPART) {
      const int64_t n = end - e;  // entries left (wave-uniform)
      BpEntry E = load_entry_uniform(ent + e);
#pragma unroll
      for (int q = 1; q < NPART; ++q) {
        if (q < n) {
          const BpEntry Eq = load_entry_uniform(ent + e + q);
          if (part == q) E = Eq;
        }
      }
      bool ok;
      const unsigned a = entry_axial<PPA>(E.k, E.nax_ppa, zi, ok);
      const float pv = proj[static_cast<int64_t>(E.sino_base) + a];
      if (ok && part < n) acc = acc + E.w * pv;
    }
  }
  // parts' sums in part order, on the lanes of part 0
  float total = acc;
#pragma unroll
  for (int q = 1; q < NPART; ++q) total = total + __shfl(acc, lane + q * ZL, kWave);
  const int z = lane;
  if (part != 0 || z >= d.nzp) return;
  gptr<float> out = as_global(img) + static_cast<int64_t>(col) * d.nzp;
  if constexpr (MODE == BP_PLAIN) {
    out[z] = z < d.nz ? total : 0.F;
  } else {
    if (z < d.nz) {
      const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
      const float u = osem_factor(total, as_const(data.sens)[idx], data.threshold, data.thr_lo, data.thr_hi);
      as_global(data.estimate)[idx] = as_global(data.estimate)[idx] * u;
      if (data.update) as_global(data.update)[idx] = u;
    }
  }
}

// Back projection of one image column of a table whose families all have 2
// planes per axial position (span 1), nzp <= 64. An entry with z offset k
// reaches the planes k + 2a only, all of the parity of k - min_z. The wave is
// split into groups of GL lanes: group g takes the parity g & 1 and is part
// g >> 1 of that parity's P = 32 / GL parts; its lane i holds the plane
// 2i + parity. The column's list is read in chunks of 64 entries, one entry
// per lane (one vector load), and written to the wave's 64 LDS slots
// compacted by parity (the chunk's even entries in table order, then its odd
// ones); step t gives part q of a parity that parity's entry t*P + q (one LDS
// read per lane). With P == 1 every plane adds its entries in table order, as
// bp_column does (identical sums); with P > 1 each part adds every P-th entry
// of its parity and the parts' sums are added in part order.
template <int MODE, int GL>
__device__ inline void bp_column_parity(int col, cptr<int64_t> list_off, cptr<BpEntry> ent, int sb, int se,
                                        cptr<float> proj, float* __restrict__ img, const ImageDims& d,
                                        const BpData& data, int lane, int4* __restrict__ slots) {
  constexpr int P = 32 / GL;
  constexpr int U = 4;  // steps whose gathers are in flight together
  const int g = lane / GL;
  const int par = g & 1, part = g >> 1;
  const int zr = 2 * (lane % GL) + par;  // plane relative to min_z
  const int zi = d.min_z + zr;
  float acc = 0.F;
  for (int s = sb; s < se; ++s) {
    const int64_t beg = list_off[static_cast<int64_t>(s) * d.ncols + col];
    const int64_t end = list_off[static_cast<int64_t>(s) * d.ncols + col + 1];
    for (int64_t c0 = beg; c0 < end; c0 += kWave) {
      const int64_t left = end - c0;
      const int n = left < kWave ? static_cast<int>(left) : kWave;  // wave-uniform, >= 1
      int4 mine = make_int4(0, 0, 0, 0);
      if (lane < n) mine = reinterpret_cast<cptr<int4>>(ent)[c0 + lane];
      const bool odd = ((mine.y - d.min_z) & 1) != 0;
      const uint64_t m1 = __ballot(lane < n && odd);
      const uint64_t m0 = __ballot(lane < n && !odd);
      const int ne = __popcll(m0), no = __popcll(m1);  // wave-uniform
      const uint64_t mm = odd ? m1 : m0;
      const int below = __popcll(mm & ((uint64_t{1} << lane) - 1));
      if (lane < n) slots[odd ? ne + below : below] = mine;
      // the wave's own LDS writes precede its reads (in-order LDS queue)
      const int cnt = par ? no : ne;  // this lane's parity
      const int first = par ? ne : 0;
      const int nsteps = ((ne > no ? ne : no) + P - 1) / P;
      for (int t = 0; t < nsteps; t += U) {
        float pv[U], wv[U];
        bool okv[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
          const int r = (t + j) * P + part;  // entry of this parity for this lane's part
          const bool valid = r < cnt;
          const int4 E = slots[valid ? first + r : 0];  // slot 0 always holds an entry of this chunk
          const int a = (zi - E.y) >> 1;  // zi - k is even for the entries of this parity
          const int nax = E.z & 0xffff;
          okv[j] = valid && static_cast<unsigned>(a) < static_cast<unsigned>(nax);
          const int ac = min(max(a, 0), nax - 1);
          pv[j] = load_u32off(proj + E.x, static_cast<unsigned>(ac));
          wv[j] = __int_as_float(E.w);
        }
#pragma unroll
        for (int j = 0; j < U; ++j)
          if (okv[j]) acc = acc + wv[j] * pv[j];
      }
    }
  }
  // parts' sums in part order, on the lanes of part 0
  float total = acc;
#pragma unroll
  for (int q = 1; q < P; ++q) total = total + __shfl(acc, lane + 2 * q * GL, kWave);
  const int z = zr;
  if (part != 0 || z >= d.nzp) return;
  gptr<float> out = as_global(img) + static_cast<int64_t>(col) * d.nzp;
  if constexpr (MODE == BP_PLAIN) {
    out[z] = z < d.nz ? total : 0.F;
  } else {
    if (z < d.nz) {
      const int64_t idx = static_cast<int64_t>(col) * d.nzp + z;
      const float u = osem_factor(total, as_const(data.sens)[idx], data.threshold, data.thr_lo, data.thr_hi);
      as_global(data.estimate)[idx] = as_global(data.estimate)[idx] * u;
      if (data.update) as_global(data.update)[idx] = u;
    }
  }
}

template <int MODE, int PPA, int NSLOT, bool SKIP, bool LO32>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_kernel(const int64_t* __restrict__ list_off,
                                                                   const BpEntry* __restrict__ ent, int sb, int se,
                                                                   const float* __restrict__ proj,
                                                                   float* __restrict__ img, ImageDims d, int nwork,
                                                                   int2 tile, int remap, BpData data, int accumulate) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int b = remap == 1 ? xcd_remap(blockIdx.x, gridDim.x)
                : remap > 1  ? xcd_chunk_remap(blockIdx.x, remap)
                             : static_cast<int>(blockIdx.x);
  const int w = b * kWavesPerBlock + wave;
  if (w >= nwork) return;
  const int col = tiled_column(w, d, tile);
  if (col < 0) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  if constexpr (PPA == 2 && NSLOT == 1) {
    __shared__ int4 slots[kWavesPerBlock * kWave];
    bp_column_parity<MODE, 32>(col, as_const(list_off), as_const(ent), sb, se, as_const(proj), img, d, data, lane,
                               slots + wave * kWave);
  } else
    bp_column<MODE, PPA, NSLOT, SKIP, LO32>(col, as_const(list_off), as_const(ent), sb, se, as_const(proj), img, d,
                                            data, lane, accumulate != 0);
}

// ---- batches: one work item per wave, the problem's descriptor read
// uniformly, the same per-family / per-column bodies as above. The host
// sorts the items by slot count (and back projection by planes per axial
// position), so each launch instantiates one body only and keeps that
// body's register use.

// forward: work item = (problem, task of that problem's task list); every
// task of the launch has at most 64 * NSLOT bins
template <int MODE, int NSLOT>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) fp_batch_kernel(const ProblemDesc* __restrict__ descs,
                                                                         const FpTask* const* __restrict__ task_lists,
                                                                         const WorkItem* __restrict__ work, int nwork) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int wi = batch_block() * kWavesPerBlock + wave;
  if (wi >= nwork) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int p = __builtin_amdgcn_readfirstlane(work[wi].problem);
  const int ti = __builtin_amdgcn_readfirstlane(work[wi].index);
  const ProblemDesc& P = descs[p];
  const FpTask task = as_const(task_lists[p])[ti];
  cptr<Family> fams = as_const(P.fams);
  const Family f1 = fams[task.first];
  const bool pair = task.second >= 0;
  const Family f2 = fams[pair ? task.second : task.first];
  fp_task<MODE, NSLOT>(f1, f2, pair, as_const(P.pool), as_const(P.fp_in), P.fp_out, P.d, P.fd, wi, lane,
                       P.pool_ops ? as_const(P.pool_ops) : nullptr, P.nelem);
}

// back projection: work item = (problem, image column), subset `subset`;
// every image of the launch has at most 64 * NSLOT planes and PPA planes per
// axial position (0: any)
template <int MODE, int PPA, int NSLOT, int NPART>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) bp_batch_kernel(const ProblemDesc* __restrict__ descs,
                                                                         const WorkItem* __restrict__ work, int nwork,
                                                                         int subset, int threshold) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int wi = batch_block() * kWavesPerBlock + wave;
  if (wi >= nwork) return;
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int p = __builtin_amdgcn_readfirstlane(work[wi].problem);
  const int col = __builtin_amdgcn_readfirstlane(work[wi].index);
  if (col < 0) return;
  const ProblemDesc& P = descs[p];
  BpData bd{};
  bd.sens = P.sens ? as_const(P.sens)[subset] : nullptr;
  bd.estimate = P.estimate;
  bd.update = nullptr;
  bd.threshold = threshold;
  bd.thr_lo = P.thr_lo;
  bd.thr_hi = P.thr_hi;
  if constexpr (PPA == 2 && NSLOT == 1) {  // NPART 1, 2, 4: groups of 32, 16, 8 lanes per parity part
    __shared__ int4 slots[kWavesPerBlock * kWave];
    bp_column_parity<MODE, 32 / NPART>(col, as_const(P.list_off), as_const(P.entries), subset, subset + 1,
                                       as_const(P.bp_in), P.bp_out, P.d, bd, lane, slots + wave * kWave);
  } else if constexpr (NPART > 1)
    bp_column_parts<MODE, PPA, NPART>(col, as_const(P.list_off), as_const(P.entries), subset, subset + 1,
                                      as_const(P.bp_in), P.bp_out, P.d, bd, lane);
  else
    bp_column<MODE, PPA, NSLOT, false>(col, as_const(P.list_off), as_const(P.entries), subset, subset + 1,
                                       as_const(P.bp_in), P.bp_out, P.d, bd, lane);
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

// Projection data of the objective's data-term Hessian products, one
// workgroup per viewgram: q = divide_and_truncate(num, den) over the viewgram
// with the small value 1e-6 times the numerator's maximum
// (recon_array_functions.cxx:172-266), where
//   exact (...AndProjData.cxx:1066-1218): num = y * fin, den = (fcur + b)^2;
//   approximate (:957-1064): num = fin, den = y / e / e with e the bin
//   efficiency (the multiplicative factor, at least 1e-20; none: den = y).
// fin = A input, fcur = A current; q may be fin. Workgroup i takes viewgram
// vgs[i] (viewgram i without a list). *negative is set when fin holds a
// negative value (the reference aborts then).
__global__ void __launch_bounds__(256) hessian_ratio_kernel(const int32_t* __restrict__ vgs,
                                                            const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ len, const float* y,
                                                            const float* fin, const float* fcur, const float* add,
                                                            float add_const, const float* mult, int approximate,
                                                            float* q, int* negative) {
  __shared__ float red[256];
  const int vg = vgs ? vgs[blockIdx.x] : static_cast<int>(blockIdx.x);
  const int64_t o = off[vg];
  const int n = len[vg];
  float m = -FLT_MAX;
  bool neg = false;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float t = fin[o + i];
    neg = neg || t < 0.F;
    m = fmaxf(m, approximate ? t : y[o + i] * t);
  }
  if (neg) *negative = 1;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float sv = red[0] * 0.000001F;
  const float small_value = (0.F < sv) ? sv : 0.F;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t j = o + i;
    const float t = fin[j];
    float num, den;
    if (approximate) {
      num = t;
      den = y[j];
      if (mult) {
        const float e = (1.E-20F < mult[j]) ? mult[j] : 1.E-20F;
        den = den / e;
        den = den / e;
      }
    } else {
      num = y[j] * t;
      const float ybar = fcur[j] + (add ? add[j] : add_const);
      den = ybar * ybar;
    }
    const float max_quotient = 10000.F;
    q[j] = num <= small_value ? 0.F : (num > max_quotient * den ? max_quotient : num / den);
  }
}

// The list-mode objective's Hessian product (LM_Hessian,
// ...ListModeDataWithProjMatrixByBin.cxx:790-822) over the events
// histogrammed into y, one workgroup per viewgram vgs[i]: every event of a
// bin (bin value 1) has fwd = fcur + b and adds row^T (-1 / (fwd * fwd) * fin)
// unless 1 > 1e4 fwd or fin == 0, so q = y * (-1 / (fwd * fwd) * fin) or 0.
__global__ void __launch_bounds__(256) lm_hessian_kernel(const int32_t* __restrict__ vgs,
                                                         const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ len, const float* y,
                                                         const float* fin, const float* fcur, const float* add,
                                                         float add_const, float* q) {
  const int vg = vgs[blockIdx.x];
  const int64_t o = off[vg];
  const int n = len[vg];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t j = o + i;
    const float yv = y[j];
    const float fwd = fcur[j] + (add ? add[j] : add_const);
    const float r = fin[j];
    float v = 0.F;
    if (yv != 0.F && !(1.F > 10000.F * fwd) && r != 0.F) {
      const float fwd_sq = fwd * fwd;
      const float per_event = -1.F / fwd_sq;
      v = yv * (per_event * r);
    }
    q[j] = v;
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

// Per in-plane operation op, element e of the basic rows as
// ops[op * n + e] = ((column << 12) | (z + 2048), weight), the column being
// the image column of the element's (x, y) under op (0 where op leaves the
// image: the symmetries then do not use op).
__global__ void expand_rows_kernel(const int2* __restrict__ pool, int64_t n, ImageDims d, int2* __restrict__ ops) {
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int2 el = pool[e];
    int x, y, z;
    decode(el.x, x, y, z);
    for (int op = 0; op < 8; ++op) {
      int xo, yo;
      switch (op) {
        case IP_ID: inplane_op<IP_ID>(x, y, xo, yo); break;
        case IP_XMX: inplane_op<IP_XMX>(x, y, xo, yo); break;
        case IP_YMY: inplane_op<IP_YMY>(x, y, xo, yo); break;
        case IP_XMX_YMY: inplane_op<IP_XMX_YMY>(x, y, xo, yo); break;
        case IP_XY_YX: inplane_op<IP_XY_YX>(x, y, xo, yo); break;
        case IP_XMY_YX: inplane_op<IP_XMY_YX>(x, y, xo, yo); break;
        case IP_XY_YMX: inplane_op<IP_XY_YMX>(x, y, xo, yo); break;
        default: inplane_op<IP_XMY_YMX>(x, y, xo, yo); break;
      }
      const int cx = xo - d.min_x, cy = yo - d.min_y;
      const int col = (cx >= 0 && cx < d.nx && cy >= 0 && cy < d.ny) ? cy * d.nx + cx : 0;
      ops[op * n + e] = make_int2(static_cast<int>((static_cast<unsigned>(col) << 12) | static_cast<unsigned>(z + 2048)),
                                  el.y);
    }
  }
}

void launch_expand_rows(const int2* d_pool, int64_t n, const ImageDims& dims, int2* d_ops, hipStream_t stream) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(expand_rows_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, d_pool, n, dims,
                     d_ops);
}

void launch_forward_project(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams, const int2* d_pool,
                            const float* d_img, float* d_out, const ImageDims& dims, const FpData& data,
                            hipStream_t stream, int xcd_chunk, const int2* d_ops, int64_t nelem) {
  if (ntasks <= 0) return;
  int nblocks = (ntasks + kWavesPerBlock - 1) / kWavesPerBlock;
  if (xcd_chunk > 0) {  // chunked remap: whole rounds of 8 chunks
    const int round = 8 * xcd_chunk;
    nblocks = (nblocks + round - 1) / round * round;
  }
  const dim3 grid(nblocks), block(kWave * kWavesPerBlock);
  // the image inside one 4 GB-aligned window of the address space
  const uint64_t first = reinterpret_cast<uint64_t>(d_img);
  const uint64_t last = first + static_cast<uint64_t>(dims.ncols) * dims.nzp * sizeof(float) - 1;
  const bool lo32 = (first >> 32) == (last >> 32);
  switch (mode) {
#define MI_FP_CASE(M)                                                                                              \
  case M:                                                                                                          \
    if (lo32)                                                                                                      \
      hipLaunchKernelGGL((fp_kernel<M, true>), grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img, d_out, \
                         dims, data, xcd_chunk, d_ops, nelem);                                                     \
    else                                                                                                           \
      hipLaunchKernelGGL((fp_kernel<M, false>), grid, block, 0, stream, d_tasks, ntasks, d_fams, d_pool, d_img,      \
                         d_out, dims, data, xcd_chunk, d_ops, nelem);                                              \
    break;
    MI_FP_CASE(FP_PLAIN)
    MI_FP_CASE(FP_RATIO)
    MI_FP_CASE(FP_GRADIENT)
    MI_FP_CASE(FP_LOGLIK)
    MI_FP_CASE(FP_GRADIENT_LOGLIK)
    MI_FP_CASE(FP_RATIO_LOGLIK)
    MI_FP_CASE(FP_DRATIO)
    MI_FP_CASE(FP_SCALE)
    MI_FP_CASE(FP_LM_RATIO)
    MI_FP_CASE(FP_LM_VALUE)
#undef MI_FP_CASE
  }
}

void launch_back_project(BpMode mode, int ppa, const int64_t* d_list_off, const BpEntry* d_entries, int subset_begin,
                         int subset_end, const float* d_proj, float* d_img, const ImageDims& dims, const BpData& data,
                         const BpTuning& tune, hipStream_t stream, int64_t proj_values) {
  const int nslot = (dims.nzp + kWave - 1) / kWave;
  const int2 tile = make_int2(tune.tile_x > 0 ? tune.tile_x : 1, tune.tile_y > 0 ? tune.tile_y : 1);
  const int nwork = ((dims.nx + tile.x - 1) / tile.x) * ((dims.ny + tile.y - 1) / tile.y) * tile.x * tile.y;
  int nblocks = (nwork + kWavesPerBlock - 1) / kWavesPerBlock;
  if (tune.xcd_remap > 1) {  // chunked remap: whole rounds of 8 chunks
    const int round = 8 * tune.xcd_remap;
    nblocks = (nblocks + round - 1) / round * round;
  }
  const dim3 grid(nblocks), block(kWave * kWavesPerBlock);
  // the projection data inside one 4 GB-aligned window of the address space
  // (the one-plane kernel forms its run buffers with 32-bit adds then)
  const uint64_t first = reinterpret_cast<uint64_t>(d_proj);
  const bool lo32 = proj_values > 0 && (first >> 32) == ((first + static_cast<uint64_t>(proj_values) * 4 - 1) >> 32);
  // a plain projection of several subsets: one launch per subset, each
  // continuing the sums in the image (a working set of one subset's
  // projection data), except for the two-plane parity kernel
  const bool split = tune.split_subsets && mode == BP_PLAIN && subset_end - subset_begin > 1 && !(ppa == 2 && nslot == 1);
  const int nlaunch = split ? subset_end - subset_begin : 1;
  for (int li = 0; li < nlaunch; ++li) {
  const int sb = split ? subset_begin + li : subset_begin, se = split ? sb + 1 : subset_end, acc = split && li > 0;
#define MI_BP_LAUNCH(M, P, S, K)                                                                                    \
  if (lo32)                                                                                                         \
    hipLaunchKernelGGL((bp_kernel<M, P, S, K, true>), grid, block, 0, stream, d_list_off, d_entries, sb, se,       \
                       d_proj, d_img, dims, nwork, tile, tune.xcd_remap, data, acc);                                 \
  else                                                                                                              \
    hipLaunchKernelGGL((bp_kernel<M, P, S, K, false>), grid, block, 0, stream, d_list_off, d_entries, sb, se,      \
                       d_proj, d_img, dims, nwork, tile, tune.xcd_remap, data, acc)
#define MI_BP_SKIP(M, P, S)             \
  if (tune.slot_skip) MI_BP_LAUNCH(M, P, S, true); \
  else MI_BP_LAUNCH(M, P, S, false);
#define MI_BP_SLOTS(M, P)                    \
  switch (nslot) {                           \
    case 1: MI_BP_SKIP(M, P, 1) break;       \
    case 2: MI_BP_SKIP(M, P, 2) break;       \
    case 3: MI_BP_SKIP(M, P, 3) break;       \
    default: MI_BP_SKIP(M, P, 4) break;      \
  }
  if (mode == BP_PLAIN) {
    if (ppa == 1) { MI_BP_SLOTS(BP_PLAIN, 1) }
    else if (ppa == 2) { MI_BP_SLOTS(BP_PLAIN, 2) }
    else { MI_BP_SLOTS(BP_PLAIN, 0) }
  } else {
    if (ppa == 1) { MI_BP_SLOTS(BP_OSEM_UPDATE, 1) }
    else if (ppa == 2) { MI_BP_SLOTS(BP_OSEM_UPDATE, 2) }
    else { MI_BP_SLOTS(BP_OSEM_UPDATE, 0) }
  }
  }
#undef MI_BP_SLOTS
#undef MI_BP_SKIP
#undef MI_BP_LAUNCH
}

void launch_fp_batch(FpMode mode, int nslot, const ProblemDesc* d_descs, const FpTask* const* d_task_lists,
                     const WorkItem* d_work, int nwork, hipStream_t stream) {
  if (nwork <= 0) return;
  const dim3 grid(batch_grid((nwork + kWavesPerBlock - 1) / kWavesPerBlock)), block(kWave * kWavesPerBlock);
#define MI_FPB(M, S) \
  hipLaunchKernelGGL((fp_batch_kernel<M, S>), grid, block, 0, stream, d_descs, d_task_lists, d_work, nwork)
#define MI_FPB_S(M)              \
  switch (nslot) {               \
    case 1: MI_FPB(M, 1); break; \
    case 2: MI_FPB(M, 2); break; \
    case 3: MI_FPB(M, 3); break; \
    default: MI_FPB(M, 4); break; \
  }
  if (mode == FP_RATIO) { MI_FPB_S(FP_RATIO) }
  else { MI_FPB_S(FP_PLAIN) }
#undef MI_FPB_S
#undef MI_FPB
}

void launch_bp_batch(BpMode mode, int ppa, int nslot, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork,
                     int subset, int threshold, hipStream_t stream) {
  if (nwork <= 0) return;
  const dim3 grid(batch_grid((nwork + kWavesPerBlock - 1) / kWavesPerBlock)), block(kWave * kWavesPerBlock);
#define MI_BPB(M, P, S, NP)                                                                                         \
  hipLaunchKernelGGL((bp_batch_kernel<M, P, S, NP>), grid, block, 0, stream, d_descs, d_work, nwork, subset, threshold)
  // nslot < 0: one slot, -nslot parts
#define MI_BPB_S(M, P)                 \
  switch (nslot) {                     \
    case -4: MI_BPB(M, P, 1, 4); break; \
    case -2: MI_BPB(M, P, 1, 2); break; \
    case 1: MI_BPB(M, P, 1, 1); break;  \
    case 2: MI_BPB(M, P, 2, 1); break;  \
    case 3: MI_BPB(M, P, 3, 1); break;  \
    default: MI_BPB(M, P, 4, 1); break; \
  }
#define MI_BPB_P(M)                  \
  if (ppa == 1) { MI_BPB_S(M, 1) }   \
  else if (ppa == 2) { MI_BPB_S(M, 2) } \
  else { MI_BPB_S(M, 0) }
  if (mode == BP_OSEM_UPDATE) { MI_BPB_P(BP_OSEM_UPDATE) }
  else { MI_BPB_P(BP_PLAIN) }
#undef MI_BPB_P
#undef MI_BPB_S
#undef MI_BPB
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

void launch_hessian_ratio(const int32_t* d_vgs, int nvg, const int64_t* d_vg_off, const int32_t* d_vg_len,
                          const float* d_y, const float* d_fin, const float* d_fcur, const float* d_add,
                          float add_const, const float* d_mult, int approximate, float* d_q, int* d_negative,
                          hipStream_t stream) {
  if (nvg <= 0) return;
  hipLaunchKernelGGL(hessian_ratio_kernel, dim3(nvg), dim3(256), 0, stream, d_vgs, d_vg_off, d_vg_len, d_y, d_fin,
                     d_fcur, d_add, add_const, d_mult, approximate, d_q, d_negative);
}

void launch_lm_hessian(const int32_t* d_vgs, int nvg, const int64_t* d_vg_off, const int32_t* d_vg_len,
                       const float* d_y, const float* d_fin, const float* d_fcur, const float* d_add, float add_const,
                       float* d_q, hipStream_t stream) {
  if (nvg <= 0) return;
  hipLaunchKernelGGL(lm_hessian_kernel, dim3(nvg), dim3(256), 0, stream, d_vgs, d_vg_off, d_vg_len, d_y, d_fin, d_fcur,
                     d_add, add_const, d_q);
}

void launch_zero_end_planes(float* d_proj, const int64_t* d_off, int count, int run, int stride, hipStream_t stream) {
  (void)stride;
  if (count <= 0) return;
  hipLaunchKernelGGL(zero_end_planes_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, d_proj, d_off, count, run);
}

}  // namespace mi
