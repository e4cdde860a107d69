// Device helpers shared by the projection kernels (pmrt_project.hip,
// pmrt_tof.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmrt_device.hpp"

namespace mi {
namespace dev {

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

// blockIdx -> work index in chunks of `chunk` consecutive blocks dealt to
// the XCDs in turn (chunk c runs on XCD c % 8): each XCD works through whole
// chunks (neighbouring image columns that share projection-data runs in its
// L2) while every XCD gets chunks from all over the work range (even load).
// The grid is padded to a multiple of 8 * chunk; indices past the work are
// dropped by the caller's bounds check.
__device__ inline int xcd_chunk_remap(int bid, int chunk) {
  const int xcd = bid % kNumXcd;
  const int i = bid / kNumXcd;
  return ((i / chunk) * kNumXcd + xcd) * chunk + i % chunk;
}

// The batch kernels deal their blocks to the XCDs in chunks of kBatchChunk
// (as fp_kernel / bp_kernel do): neighbouring work items (a tile's columns,
// the families of one (view, tangential position)) share one XCD's L2. Their
// grids are padded to whole rounds (batch_grid); blocks past the work find no
// item.
constexpr int kBatchChunk = 32;
__device__ inline int batch_block() { return xcd_chunk_remap(static_cast<int>(blockIdx.x), kBatchChunk); }
inline unsigned batch_grid(int nblocks) {
  const int round = kNumXcd * kBatchChunk;
  return static_cast<unsigned>((nblocks + round - 1) / round * round);
}

// Address spaces. Tables and inputs that a kernel only reads are accessed
// through the constant address space, so that reads at wave-uniform
// addresses become scalar loads; outputs through the global one. Kernel
// arguments get this from their declarations already; pointers read out of
// device descriptors (the batch kernels) carry no address space and would
// otherwise be accessed with per-lane flat loads.
// (The host pass of the compiler only checks these device functions and
// has no such address spaces.)
#if defined(__HIP_DEVICE_COMPILE__)
#define MI_AS_CONST __attribute__((address_space(4)))
#define MI_AS_GLOBAL __attribute__((address_space(1)))
#else
#define MI_AS_CONST
#define MI_AS_GLOBAL
#endif
template <class T>
using cptr = const MI_AS_CONST T*;
template <class T>
using gptr = MI_AS_GLOBAL T*;
template <class T>
__device__ inline cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}
template <class T>
__device__ inline gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

// base[idx] with a wave-uniform base and a 32-bit unsigned lane offset, so
// the address is formed as scalar base + 32-bit vector offset
__device__ inline float load_u32off(const float* base, unsigned idx) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (idx << 2));
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float load_u32off(cptr<float> base, unsigned idx) {
  return *(cptr<float>)((const MI_AS_CONST char*)base + (idx << 2));
}
#endif

// A table entry at a wave-uniform address, its fields made wave-uniform
// (readfirstlane): a later per-lane choice between such entries stays a
// select of registers instead of becoming a per-lane load of the chosen one.
__device__ inline BpEntry load_entry_uniform(cptr<BpEntry> p) {
  const BpEntry E = *p;
  BpEntry r;
  r.sino_base = __builtin_amdgcn_readfirstlane(E.sino_base);
  r.k = __builtin_amdgcn_readfirstlane(E.k);
  r.nax_ppa = __builtin_amdgcn_readfirstlane(E.nax_ppa);
  r.w = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(E.w)));
  return r;
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

// the in-plane operation ip (an IP_* value) at run time
__device__ inline void inplane_any_op(int ip, int x, int y, int& xo, int& yo) {
  switch (ip) {
    case IP_ID: inplane_op<IP_ID>(x, y, xo, yo); break;
    case IP_XMX: inplane_op<IP_XMX>(x, y, xo, yo); break;
    case IP_YMY: inplane_op<IP_YMY>(x, y, xo, yo); break;
    case IP_XMX_YMY: inplane_op<IP_XMX_YMY>(x, y, xo, yo); break;
    case IP_XY_YX: inplane_op<IP_XY_YX>(x, y, xo, yo); break;
    case IP_XMY_YX: inplane_op<IP_XMY_YX>(x, y, xo, yo); break;
    case IP_XY_YMX: inplane_op<IP_XY_YMX>(x, y, xo, yo); break;
    default: inplane_op<IP_XMY_YMX>(x, y, xo, yo); break;
  }
}

__device__ inline void decode(int c, int& x, int& y, int& z) {
  x = (c & 1023) - 512;
  y = ((c >> 10) & 1023) - 512;
  z = static_cast<int>(static_cast<unsigned>(c) >> 20) - 2048;
}

constexpr int kWavesPerBlock = 4;

// Per-lane, per-slot bin of a forward-projection task.
template <int NSLOT>
struct LaneBins {
  int zs[NSLOT];     // z sign of the symmetry operation
  int zb[NSLOT];     // image z of the element z = 0 for this bin (zadd0 + ppa*a - min_z)
  bool valid[NSLOT];
  int64_t out[NSLOT];  // projection-data index
  int vg[NSLOT];     // viewgram index
};

// Axial index of voxel zi for a table entry; `ok` tells whether the entry
// contributes, the returned index is always inside the entry's run. PPA is
// the number of image planes per axial position (0: not known at compile time).
template <int PPA>
__device__ inline unsigned entry_axial(int k, int nax_ppa, int zi, bool& ok) {
  const int nax = nax_ppa & 0xffff;
  const int diff = zi - k;
  int a;
  if constexpr (PPA == 1) {
    a = diff;
    ok = static_cast<unsigned>(a) < static_cast<unsigned>(nax);
  } else if constexpr (PPA == 2) {
    a = diff >> 1;
    ok = (diff & 1) == 0 && static_cast<unsigned>(a) < static_cast<unsigned>(nax);
  } else {
    // families of one table mix 1 and 2 planes per axial position (span > 1
    // with a last segment of ring difference span 1): no division for those
    const int ppa = nax_ppa >> 16;
    if (ppa == 1) {
      a = diff;
      ok = static_cast<unsigned>(a) < static_cast<unsigned>(nax);
    } else if (ppa == 2) {
      a = diff >> 1;
      ok = (diff & 1) == 0 && static_cast<unsigned>(a) < static_cast<unsigned>(nax);
    } else {
      a = diff / ppa;
      ok = diff >= 0 && a * ppa == diff && a < nax;
    }
  }
  return static_cast<unsigned>(min(max(a, 0), nax - 1));
}

// Byte offset, within an entry's run, of the value that reaches plane zi
// (zi4 = 4 * zi), clamped to the run, and whether the entry reaches zi at
// all: the same test and the same value as entry_axial, with the
// one-plane-per-axial-position case reduced to one subtraction, one compare
// and one clamp.
template <int PPA>
__device__ inline unsigned entry_offset(int k, int nax_ppa, int zi, int zi4, bool& ok) {
  const unsigned nax4 = static_cast<unsigned>(nax_ppa & 0xffff) << 2;
  if (PPA == 1 || (PPA == 0 && (nax_ppa >> 16) == 1)) {
    const unsigned off = static_cast<unsigned>(zi4 - 4 * k);  // negative differences wrap above nax4
    ok = off < nax4;
    return min(off, nax4 - 4U);
  }
  return entry_axial<PPA>(k, nax_ppa, zi, ok) << 2;
}

__device__ inline float load_byteoff(const float* base, unsigned off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float load_byteoff(cptr<float> base, unsigned off) {
  return *(cptr<float>)((const MI_AS_CONST char*)base + off);
}
#endif

// n floats at base + first as a raw buffer (stride 0; the third descriptor
// word of the gfx9 family, as in composable_kernel's
// CK_BUFFER_RESOURCE_3RD_DWORD): a load at a byte offset outside [0, 4n),
// negative ones included, returns 0 instead of reading. The range check
// covers the lane offset and the scalar offset together, so the buffer starts
// at the run itself and loads pass no scalar offset.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline __amdgpu_buffer_rsrc_t run_buffer(cptr<float> base, unsigned first, int n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)(base + first), static_cast<short>(0), 4 * n, 0x00020000);
}
// run_buffer for a buffer that lies inside one 4 GB-aligned window of the
// address space (checked by the launcher): the run's start is formed with one
// 32-bit add to the low half of the base address, the high half kept.
__device__ inline __amdgpu_buffer_rsrc_t run_buffer_lo32(cptr<float> base, unsigned first, int n) {
  const uint64_t b = reinterpret_cast<uint64_t>(base);
  const uint32_t lo = static_cast<uint32_t>(b) + (first << 2);
  const uint64_t p = (b & 0xffffffff00000000ull) | lo;
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(p), static_cast<short>(0), 4 * n, 0x00020000);
}
__device__ inline float load_run(__amdgpu_buffer_rsrc_t run, int byte_off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(run, byte_off, 0, 0));
}
// A lane byte offset kept whole for load_run. Offsets of a lane's z slots
// differ by constants (256 bytes a slot), and the compiler would otherwise
// load slot q + 1 as slot q's register offset plus an immediate offset; the
// range check does not wrap that sum, so where slot q's offset is negative
// (a plane below the run) an in-range slot q + 1 would read 0.
__device__ inline int whole_offset(int off) {
  asm("" : "+v"(off));
  return off;
}
#endif

// i clamped to [0, n - 1]
__device__ inline int clamp_index(int i, int n) { return min(max(i, 0), n - 1); }

// work index -> image column, visiting the image in tile.x x tile.y tiles
// (columns of one tile run at the same time and share projection-data runs);
// returns -1 for the padding of partial tiles. A 1 x 1 tile is raster order.
__device__ inline int tiled_column(int w, const ImageDims& d, int2 tile) {
  const int tiles_x = (d.nx + tile.x - 1) / tile.x;
  const int t = w / (tile.x * tile.y), r = w - t * (tile.x * tile.y);
  const int y = (t / tiles_x) * tile.y + r / tile.x;
  const int x = (t % tiles_x) * tile.x + r % tile.x;
  return (y < d.ny && x < d.nx) ? y * d.nx + x : -1;
}

// Multiplicative update of the prior-free OSMAPOSL step from u = BP(y/ybar)
// and the subset sensitivity: divide(u, sens, 0) with 0/0 = 0
// (numerics/divide.inl:26-47), then threshold_upper_lower(lo, hi)
// (thresholding.h:45-54) except in subiteration 1
// (OSMAPOSLReconstruction.cxx:491-507).
__device__ inline float osem_factor(float u, float sv, int threshold, float lo, float hi) {
  if (fabsf(sv) <= 0.F && fabsf(u) <= 0.F) u = 0;
  else u = u / sv;
  if (threshold) {
    if (u > hi) u = hi;
    else if (lo > u) u = lo;
  }
  return u;
}

constexpr bool fp_writes_ratio(int m) {
  return m == FP_RATIO || m == FP_GRADIENT || m == FP_GRADIENT_LOGLIK || m == FP_RATIO_LOGLIK;
}
constexpr bool fp_accumulates_loglik(int m) {
  return m == FP_LOGLIK || m == FP_GRADIENT_LOGLIK || m == FP_RATIO_LOGLIK || m == FP_LM_VALUE;
}

// Forward-projection epilogue of one bin: a = (A x)_i, vg = its viewgram.
template <int MODE>
__device__ inline void fp_epilogue(float a, int64_t i, int vg, const FpData& data, float* __restrict__ out, double& ll) {
  if constexpr (MODE == FP_PLAIN) {
    as_global(out)[i] = a;
  } else if constexpr (MODE == FP_SCALE) {
    as_global(out)[i] = a * as_const(data.scale)[i];
  } else if constexpr (MODE == FP_LM_RATIO || MODE == FP_LM_VALUE) {
    // per event (bin value 1): fwd = row . x + additive; y events of one bin
    // contribute y / fwd to the gradient -- every event, capped or not: the
    // gradient instantiation LM_gradient_and_value<true, false> returns early
    // only when it computes the value (...ListModeDataWithProjMatrixByBin.cxx:
    // 756-787,822-843) -- and, to the value, log(1/1e4) in place of log(fwd)
    // when 1 > 1e4 fwd (<false, true>, :876-911)
    const float fwd = a + (data.additive ? as_const(data.additive)[i] : data.additive_constant);
    const float yv = as_const(data.measured)[i];
    const bool capped = 1.F > 10000.F * fwd;
    if constexpr (MODE == FP_LM_RATIO) {
      as_global(out)[i] = yv == 0.F ? 0.F : yv / fwd;
    } else {
      if (yv != 0.F)
        ll -= static_cast<double>(yv) * (capped ? log(static_cast<double>(1.F / 10000.F)) : log(static_cast<double>(fwd)));
    }
  } else {
    const float ybar = a + (data.additive ? as_const(data.additive)[i] : data.additive_constant);
    const float yv = as_const(data.measured)[i];
    const float small_value = as_const(data.vg_small)[vg];
    const float max_quotient = 10000.F;
    if constexpr (fp_writes_ratio(MODE)) {
      // estimated += additive; divide_and_truncate (recon_array_functions.cxx:172-266)
      float num = yv;
      if (num <= small_value) {
        num = 0;
      } else {
        if (num > max_quotient * ybar) num = max_quotient;
        else num = num / ybar;
      }
      if constexpr (MODE == FP_GRADIENT || MODE == FP_GRADIENT_LOGLIK) num = num - (data.mult ? as_const(data.mult)[i] : 1.F);
      as_global(out)[i] = num;
    }
    if constexpr (MODE == FP_DRATIO) {
      // derivative of the quotient above: -y/ybar^2 where y/ybar is used, 0 where truncated
      float dr = 0.F;
      if (yv > small_value && !(yv > max_quotient * ybar)) {
        const float r = yv / ybar;
        dr = -r / ybar;
      }
      as_global(out)[i] = dr;
    }
    if constexpr (fp_accumulates_loglik(MODE)) {
      // accumulate_loglikelihood (recon_array_functions.cxx:336-377)
      float est = ybar;
      if (data.mult) est = est * as_const(data.mult)[i];
      const float ne = (est < yv / 10000.F) ? yv / 10000.F : est;
      if (yv <= small_value) ll += -static_cast<double>(ne);
      else ll += yv * log(static_cast<double>(ne)) - static_cast<double>(ne);
    }
  }
}

}  // namespace dev
}  // namespace mi
