// force_kernels.h — Lennard-Jones force kernels for gfx950 (MI355X), templated
// over the precision policy; each policy is instantiated in its own
// translation unit (force_f64.hip, force_mixed.hip, force_kahan.hip).
//
// p += F(q) * dt over a Verlet list, the pair term of cpu_ref/force_aos.cpp:176-189
// and cuda/kernel.cuh:22-31. Two kernel families:
//
//  * k_csr: a segment of TPI consecutive lanes of a 64-wide wavefront owns one
//    i-particle and walks its CSR row in batches of U*TPI entries: lane t of
//    the segment takes entries base + u*TPI + t (u < U), so each batch reads
//    U contiguous TPI-wide chunks of the row. All U gathers of a batch are
//    issued before any is used, and the indices of the next batch are loaded
//    while the current batch is computed (a two-stage software pipeline), so
//    a wavefront keeps 2U-3U memory operations in flight instead of one
//    dependent index->gather chain per entry. The per-lane partial forces
//    are summed across the segment with cross-lane shuffles and one lane
//    updates p[i]. TPI = 64 is the wavefront-per-i form of the reference's
//    warp_unroll kernels (cuda/kernel.cuh:821-904, 32-lane warps); smaller
//    TPI packs several rows of ~80-150 entries into one wavefront so fewer
//    lanes idle on the last batch of a row.
//  * k_ell: one lane per i-particle over the reference's transposed (ELL)
//    list, entry k of row i at list[i + k*N] so a wavefront reads 64
//    consecutive indices per entry (force_kernel_memopt2 and relatives,
//    cuda/kernel.cuh:102-236), with the same U-batch pipeline.
//
// Both take full lists (each row updates only its own i, like the
// reference's full-list kernels) or half lists (HALF: every pair is visited
// once and -f is scattered to j with hardware FP64 atomics, like the
// *_with_aar kernels, cuda/kernel.cuh:238-469).
//
// Precision policies (PREC): kF64 — FP64 throughout; kMixedF64 — force scalar
// in FP32, accumulation in FP64; kMixedKahan — force scalar and per-lane sums
// in FP32 with Kahan compensation. Distances and r2 are FP64 in every policy.
//
// Launches are XCD-aware: the grid is padded to a multiple of 8 and block b
// maps to a contiguous logical range per XCD (xcd_logical_block).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "lj_device.h"
#include "ljgpu_internal.h"

namespace ljgpu {
namespace detail {

constexpr int kF64 = 0;
constexpr int kMixedF64 = 1;
constexpr int kMixedKahan = 2;
// Not a force policy: the tiled potential-energy kernel (k_tile_energy)
// runs the tile code with this policy, accumulating per pair
// s^3 (s^3 - 1) with s = 1/r2 and the pair count instead of the force.
constexpr int kEnergy = 3;

// Per-lane accumulator of the i-particle's force for each precision policy.
template <int PREC>
struct Accum {
  double x = 0.0, y = 0.0, z = 0.0;
  __device__ __forceinline__ void add(double df, float /*dff*/, double dx, double dy, double dz) {
    x = fma(df, dx, x);
    y = fma(df, dy, y);
    z = fma(df, dz, z);
  }
  __device__ __forceinline__ double vx() const { return x; }
  __device__ __forceinline__ double vy() const { return y; }
  __device__ __forceinline__ double vz() const { return z; }
};

template <>
struct Accum<kMixedKahan> {
  KahanF32 x, y, z;
  __device__ __forceinline__ void add(double /*df*/, float dff, double dx, double dy, double dz) {
    x.add(dff * static_cast<float>(dx));
    y.add(dff * static_cast<float>(dy));
    z.add(dff * static_cast<float>(dz));
  }
  __device__ __forceinline__ double vx() const { return x.value(); }
  __device__ __forceinline__ double vy() const { return y.value(); }
  __device__ __forceinline__ double vz() const { return z.value(); }
};

template <>
struct Accum<kEnergy> {
  double u = 0.0, c = 0.0;
  // s = 1/r2 (refined reciprocal); nothing unless the pair is inside.
  __device__ __forceinline__ void add_pair(double s, bool inside) {
    const double s3 = s * s * s;
    u += inside ? s3 * (s3 - 1.0) : 0.0;
    c += inside ? 1.0 : 0.0;
  }
  __device__ __forceinline__ double vx() const { return u; }
  __device__ __forceinline__ double vy() const { return c; }
  __device__ __forceinline__ double vz() const { return 0.0; }
};

// One list entry once q_j is in registers: minimum image, cutoff, force
// scalar, accumulation into the i-accumulator and, for half lists, the
// scatter of -f to j. `valid` is false for the padding slots of a batch.
template <int PREC, bool PBC, bool HALF, class M>
__device__ __forceinline__ void pair_compute(const M& mom, int j, bool valid, double dx,
                                             double dy, double dz, const PairParams& pp,
                                             Accum<PREC>& acc) {
  min_image<PBC>(dx, dy, dz, pp);
  const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  double df;
  float dff = 0.0f;
  if constexpr (PREC == kF64) {
    df = lj_df_f64(r2, pp);
    df = valid ? df : 0.0;
  } else {
    dff = lj_df_mixed(r2, pp);
    dff = valid ? dff : 0.0f;
    df = static_cast<double>(dff);
  }
  acc.add(df, dff, dx, dy, dz);
  if constexpr (HALF) {
    if (valid && r2 <= pp.cl2) mom.atomic_add(j, -df * dx, -df * dy, -df * dz);
  }
}

template <int TPI, int U, int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
__global__ void __launch_bounds__(1024)
k_csr(P pos, M mom, int32_t n, const int32_t* __restrict__ nop, const int32_t* __restrict__ ptr,
      const int32_t* __restrict__ list, int64_t n_entries, PairParams pp, int nblocks_padded) {
  const int lb = xcd_logical_block(nblocks_padded);
  const int64_t gt = static_cast<int64_t>(lb) * blockDim.x + threadIdx.x;
  const int64_t i64 = gt / TPI;
  const int sub = static_cast<int>(threadIdx.x) & (TPI - 1);
  const bool active = i64 < n;
  const int i = active ? static_cast<int>(i64) : 0;
  Accum<PREC> acc;
  if (active) {
    double xi, yi, zi;
    pos.ld(i, xi, yi, zi);
    int np = nop[i];
    int64_t kp = ptr[i];
    if constexpr (CLAMP) {
      kp = kp < 0 ? 0 : (kp > n_entries ? n_entries : kp);
      const int64_t room = n_entries - kp;
      np = np < 0 ? 0 : (np > room ? static_cast<int>(room) : np);
    }
    const int32_t* row = list + kp;
    constexpr int kBatch = U * TPI;
    // Indices of the first batch; -1 marks a padding slot past the row end.
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = u * TPI + sub;
      jn[u] = k < np ? row[k] : -1;
    }
    for (int base = 0; base < np; base += kBatch) {
      int j[U];
      double xj[U], yj[U], zj[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        j[u] = jn[u];
        if constexpr (CLAMP) {
          if (j[u] >= 0) j[u] = static_cast<int>(min(static_cast<uint32_t>(j[u]), static_cast<uint32_t>(n - 1)));
        }
        pos.ld(j[u] >= 0 ? j[u] : i, xj[u], yj[u], zj[u]);
      }
      const int next = base + kBatch;
      if (next < np) {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int k = next + u * TPI + sub;
          jn[u] = k < np ? row[k] : -1;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        pair_compute<PREC, PBC, HALF>(mom, j[u], j[u] >= 0, xj[u] - xi, yj[u] - yi, zj[u] - zi,
                                      pp, acc);
      }
    }
  }
  const double fx = segment_sum<TPI>(acc.vx());
  const double fy = segment_sum<TPI>(acc.vy());
  const double fz = segment_sum<TPI>(acc.vz());
  if (active && sub == 0) {
    if constexpr (HALF) mom.atomic_add(i, fx, fy, fz);
    else mom.add(i, fx, fy, fz);
  }
}

template <int U, int PREC, class P, class M, bool PBC, bool HALF>
__global__ void __launch_bounds__(1024)
k_ell(P pos, M mom, int32_t n, const int32_t* __restrict__ nop, const int32_t* __restrict__ ell,
      int32_t width, PairParams pp, int nblocks_padded) {
  const int lb = xcd_logical_block(nblocks_padded);
  const int64_t i64 = static_cast<int64_t>(lb) * blockDim.x + threadIdx.x;
  if (i64 >= n) return;
  const int i = static_cast<int>(i64);
  double xi, yi, zi;
  pos.ld(i, xi, yi, zi);
  int np = nop[i];
  np = np < 0 ? 0 : (np > width ? width : np);
  Accum<PREC> acc;
  const int32_t* col = ell + i;
  int jn[U];
#pragma unroll
  for (int u = 0; u < U; u++) jn[u] = u < np ? col[static_cast<int64_t>(u) * n] : -1;
  for (int base = 0; base < np; base += U) {
    int j[U];
    double xj[U], yj[U], zj[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      j[u] = jn[u];
      pos.ld(j[u] >= 0 ? j[u] : i, xj[u], yj[u], zj[u]);
    }
    if (base + U < np) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = base + U + u;
        jn[u] = k < np ? col[static_cast<int64_t>(k) * n] : -1;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      pair_compute<PREC, PBC, HALF>(mom, j[u], j[u] >= 0, xj[u] - xi, yj[u] - yi, zj[u] - zi, pp,
                                    acc);
  }
  if constexpr (HALF) mom.atomic_add(i, acc.vx(), acc.vy(), acc.vz());
  else mom.add(i, acc.vx(), acc.vy(), acc.vz());
}

// Shift a position by whole box lengths to the periodic image nearest the
// reference point c (no-op in an open box).
template <bool PBC>
__device__ __forceinline__ void unwrap_near(double& x, double& y, double& z, double cx,
                                            double cy, double cz, const PairParams& pp) {
  if constexpr (PBC) {
    x = fma(-pp.box, rint((x - cx) * pp.inv_box), x);
    y = fma(-pp.box, rint((y - cy) * pp.inv_box), y);
    z = fma(-pp.box, rint((z - cz) * pp.inv_box), z);
  }
}

// Loads from the workgroup's LDS at byte address `a`. The tile kernel uses
// dynamic LDS only, which starts at address 0, so an offset into it is its
// LDS address; building the pointer from the integer lets the compiler fold
// constant displacements into the DS instruction's offset field.
template <class T>
__device__ __forceinline__ T lds_at(uint32_t a) {
  return *reinterpret_cast<const __attribute__((address_space(3))) T*>(static_cast<uintptr_t>(a));
}

// One entry of a tile row: in FP64 the unit force scalar (lj_df_unit_f64,
// the row's sum is scaled by 24*dt in tile_body), otherwise pair_compute.
template <int PREC, bool MI>
__device__ __forceinline__ void tile_pair(double dx, double dy, double dz, const PairParams& pp,
                                          Accum<PREC>& acc) {
  if constexpr (PREC == kF64) {
    min_image<MI>(dx, dy, dz, pp);
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    acc.add(lj_df_unit_f64(rcp_f64(r2), !(r2 > pp.cl2)), 0.0f, dx, dy, dz);
  } else {
    pair_compute<PREC, MI, false>(MomAos{nullptr}, 0, true, dx, dy, dz, pp, acc);
  }
}

// One row of an LDS tile: entries [0, np) of the row, lane `sub` of the
// segment taking entries base + u*TPI + sub. The row's entry offsets
// (`lrow`) and the partner positions (`lb`) are both in LDS. Full batches of
// U*TPI entries run unguarded; the remainder is walked one entry per lane at
// a time (exec-masked, so a short tail costs at most one pass per TPI
// entries). MI applies the minimum image per entry (blocks whose unwrapped
// frame is not exact).
template <int TPI, int U, int PREC, bool MI, class M>
__device__ __forceinline__ void tile_row(const M& mom, const uint16_t* lrow, int np, int sub,
                                         double xi, double yi, double zi, const PairParams& pp,
                                         Accum<PREC>& acc) {
  constexpr int kB = U * TPI;
  int base = 0;
  for (; base + kB <= np; base += kB) {
    uint32_t off[U];
#pragma unroll
    for (int u = 0; u < U; u++) off[u] = kTileLdsHeader + lrow[base + u * TPI + sub];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double xj = lds_at<double>(off[u]);
      const double yj = lds_at<double>(off[u] + 8);
      const double zj = lds_at<double>(off[u] + 16);
      tile_pair<PREC, MI>(xj - xi, yj - yi, zj - zi, pp, acc);
    }
  }
  for (int k = base + sub; k < np; k += TPI) {
    const uint32_t o = kTileLdsHeader + lrow[k];
    const double xj = lds_at<double>(o);
    const double yj = lds_at<double>(o + 8);
    const double zj = lds_at<double>(o + 16);
    tile_pair<PREC, MI>(xj - xi, yj - yi, zj - zi, pp, acc);
  }
}

// One batch of a tile row: all U partner positions are read from LDS before
// any is used, then the U entries are computed; the offsets buffer `o` is
// refilled in between with the batch two ahead (loads from global memory
// that then have two batches of arithmetic to land). The offset arrays are
// padded so these loads stay inside them past the end of the last row; what
// they read beyond a row is never used. LIMIT masks entries at or beyond
// `np` (the last, partial batch of a row): their offsets are replaced by 0
// (a valid slot) and their force by 0. `base` is the batch's first entry of
// the row; `refill` the layout index of the batch two ahead, whose slot u
// sits at refill + u * US + LS (LS: the lane's own offset within a slot).
// A last batch in which no active lane holds an entry within the cutoff
// skips the force arithmetic (the wavefront branches as a whole): the
// reference's `if (r2 > CL2) continue;` for a whole batch. Rows hold their
// entries in cutoff classes (cut_partition), so the entries beyond the
// cutoff collect at the rows' ends. In FP64 the force scalar omits the
// factor 24*dt (lj_df_unit_f64), which tile_body applies to the row's sum,
// and the batch's U reciprocals come from one (rcp_f64_batch). PBC is the
// kernel's box (periodic or open); MI whether this row needs the minimum
// image.
template <int TPI, int U, int PREC, bool PBC, bool MI, bool LIMIT, bool REFILL, int US, class M>
__device__ __forceinline__ void tile_batch(const M& mom, const uint16_t* __restrict__ lrow,
                                           uint32_t (&o)[U], int base, int refill, int ls,
                                           int np, int sub, double xi, double yi, double zi,
                                           const PairParams& pp, Accum<PREC>& acc) {
  double dx[U], dy[U], dz[U], r2[U];
  bool valid[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    valid[u] = !LIMIT || base + u * TPI + sub < np;
    const uint32_t a = kTileLdsHeader + (valid[u] ? o[u] : 0u);
    dx[u] = lds_at<double>(a);
    dy[u] = lds_at<double>(a + 8);
    dz[u] = lds_at<double>(a + 16);
  }
  if constexpr (REFILL) {
#pragma unroll
    for (int u = 0; u < U; u++) o[u] = lrow[refill + u * US + ls];
  }
  bool in[U];
  bool inside = false;
#pragma unroll
  for (int u = 0; u < U; u++) {
    dx[u] -= xi;
    dy[u] -= yi;
    dz[u] -= zi;
    min_image<MI>(dx[u], dy[u], dz[u], pp);
    r2[u] = fma(dz[u], dz[u], fma(dy[u], dy[u], dx[u] * dx[u]));
    in[u] = valid[u] && !(r2[u] > pp.cl2);
    inside |= in[u];
  }
  if constexpr (LIMIT) {
    if (!__any(inside)) return;
  }
  double s[U];
  if constexpr (PREC == kF64) {
#pragma unroll
    for (int u = 0; u < U; u++) s[u] = LIMIT && !valid[u] ? 1.0 : r2[u];
    rcp_f64_batch<U>(s);
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    if constexpr (PREC == kF64) {
      acc.add(lj_df_unit_f64(s[u], in[u]), 0.0f, dx[u], dy[u], dz[u]);
    } else if constexpr (PREC == kEnergy) {
      acc.add_pair(rcp_f64(LIMIT && !valid[u] ? 1.0 : r2[u]), in[u]);
    } else {
      float dff = lj_df_mixed(r2[u], pp);
      // Keeps the cutoff select on the FP32 value (one instruction): the
      // compiler would otherwise move it past the conversion to FP64, where
      // a select takes two. Emits nothing. Periodic boxes only: it pushes the
      // open box's kernels into spilling.
      if constexpr (PBC) asm("" : "+v"(dff));
      if constexpr (LIMIT) dff = valid[u] ? dff : 0.0f;
      acc.add(static_cast<double>(dff), dff, dx[u], dy[u], dz[u]);
    }
  }
}

// The first batch's offsets of a row in either layout of tile_row_stream,
// loaded before the tile's staging so that their latency hides behind it.
template <int TPI, int U, bool SL>
__device__ __forceinline__ void row_prefetch(const uint16_t* __restrict__ lrow, int sub,
                                             uint32_t (&oa)[U]) {
  constexpr int US = SL ? 64 : TPI;
  const int ls = SL ? 0 : sub;
#pragma unroll
  for (int u = 0; u < U; u++) oa[u] = lrow[u * US + ls];
}

// As tile_row, with the row's entry offsets read from global memory through
// two alternating offset buffers (oa: the first batch, loaded by the caller
// with row_prefetch; ob: the second), each refilled two batches ahead, and
// the row's last partial batch run as one masked batch. SL selects the
// offset layout: false, the row's own offsets (`lrow` at the row's first
// entry, entry k at k); true, the wavefront slices of dev_slice_tiles
// (`lrow` at the lane's place in its wavefront's slice, batch j slot u at
// j * 64 * U + u * 64), so each offset load reads 64 consecutive values.
template <int TPI, int U, int PREC, bool PBC, bool MI, bool SL, class M>
__device__ __forceinline__ void tile_row_stream(const M& mom, const uint16_t* __restrict__ lrow,
                                                int np, int sub, double xi, double yi, double zi,
                                                const PairParams& pp, Accum<PREC>& acc,
                                                uint32_t (&oa)[U]) {
  constexpr int kB = U * TPI;
  constexpr int JS = SL ? 64 * U : kB;  // layout stride of a batch
  constexpr int US = SL ? 64 : TPI;     // layout stride of a slot
  const int ls = SL ? 0 : sub;
  const int nfull = np / kB;
  uint32_t ob[U];
#pragma unroll
  for (int u = 0; u < U; u++) ob[u] = lrow[JS + u * US + ls];
  int b = 0;
  for (; b + 2 <= nfull; b += 2) {
    tile_batch<TPI, U, PREC, PBC, MI, false, true, US>(mom, lrow, oa, b * kB, (b + 2) * JS, ls,
                                                       np, sub, xi, yi, zi, pp, acc);
    tile_batch<TPI, U, PREC, PBC, MI, false, true, US>(mom, lrow, ob, (b + 1) * kB, (b + 3) * JS,
                                                       ls, np, sub, xi, yi, zi, pp, acc);
  }
  if (b < nfull) {
    tile_batch<TPI, U, PREC, PBC, MI, false, false, US>(mom, lrow, oa, b * kB, 0, ls, np, sub, xi,
                                                        yi, zi, pp, acc);
    b++;
#pragma unroll
    for (int u = 0; u < U; u++) oa[u] = ob[u];
  }
  if (b * kB < np)
    tile_batch<TPI, U, PREC, PBC, MI, true, false, US>(mom, lrow, oa, b * kB, 0, ls, np, sub, xi,
                                                       yi, zi, pp, acc);
}

// LDS-tiled full-list kernel: one workgroup per block of consecutive slots
// of a particle order (slot s holds particle member[s]; the order groups
// nearby particles, or is the identity). The workgroup first stages into
// LDS (a) the positions of every
// particle the block's rows reference (stage[soff .. soff+ns)), as x, y, z
// triples, in a periodic box shifted to the image nearest the block's middle
// particle, and (b) the block's ne per-entry LDS offsets (one contiguous,
// coalesced range of `local`, which holds the rows in slot order: slot s's
// row starts at slot_ptr[s]). A null member means that positions, momenta
// and `stage` are given in slot order (particle index = slot). Rows then run entirely out of LDS: the random
// gathers of the CSR kernel become LDS reads and the list costs 2 bytes per
// entry. In the unwrapped frame a pair difference equals its minimum image
// whenever every particle of the block lies within box/2 - rc of the
// reference point (rc = cutoff); blocks where that fails apply the minimum
// image per entry instead. OFFS_LDS = false skips (b) and streams the
// offsets from global memory instead (less LDS per workgroup, so more
// resident workgroups per CU).
//
// A launch covers tiles [first, first + nblocks). They are assigned
// XCD-aware: the workgroups on XCD x (blockIdx % 8 == x) walk the contiguous
// range [x*per, (x+1)*per) of the launch's tiles with a stride of gridDim/8. With one workgroup per tile each does a single tile; with a
// smaller (persistent) grid each loops over several, so resident workgroups
// drift out of phase and one's staging and memory stalls overlap another's
// arithmetic.
template <int TPI, int U, int PREC, class P, class M, bool PBC, bool OFFS_LDS, bool SL>
__device__ __forceinline__ void tile_body(P pos, M mom, const int32_t* __restrict__ nop,
                                          const int32_t* __restrict__ member,
                                          const int32_t* __restrict__ slot_ptr,
                                          const int4* __restrict__ desc,
                                          const int32_t* __restrict__ stage,
                                          const uint16_t* __restrict__ local,
                                          const uint16_t* __restrict__ sliced,
                                          const int32_t* __restrict__ wave_off, int wpb, int b,
                                          const PairParams& gpp, double greach,
                                          const double2* __restrict__ rep_box, char* lb,
                                          int* flag, double* eacc) {
  const int4 d0 = desc[2 * b];
  const int4 d1 = desc[2 * b + 1];
  const int i0 = d0.x, ni = d0.y, soff = d0.z, ns = d0.w;
  const int eoff = d1.x, ne = d1.y;
  // Batched replicas: the block's own box (box, 1/box) from its replica
  // (descriptor slot 6); an open-box replica has (0, 0), for which every
  // minimum image is the identity.
  PairParams pp = gpp;
  double reach = greach;
  if (PBC && rep_box) {
    const double2 bb = rep_box[d1.z];
    pp.box = bb.x;
    pp.inv_box = bb.y;
    reach = 0.5 * bb.x - pp.rc;
  }
  double* lds = reinterpret_cast<double*>(lb);
  const int off_area = (ns * 24 + 15) & ~15;
  double cx = 0.0, cy = 0.0, cz = 0.0;
  // The block's middle particle (its slot when the arrays are in slot order).
  if constexpr (PBC) pos.ld(member ? d1.w : i0 + ni / 2, cx, cy, cz);
  // The thread's own row, requested before the staging so that its latency
  // overlaps it: first the indices (the first staging pass's particles, own
  // particle, row bounds, offset slice), then what they address (own
  // position, the row's first offset batch).
  constexpr int kStageR = 4;
  const int nt = static_cast<int>(blockDim.x);
  const int seg = static_cast<int>(threadIdx.x) / TPI;
  const int sub = static_cast<int>(threadIdx.x) & (TPI - 1);
  const bool active = seg < ni;
  const int slot = i0 + (active ? seg : 0);
  int jj[kStageR];
#pragma unroll
  for (int r = 0; r < kStageR; r++) {
    const int t = r * nt + static_cast<int>(threadIdx.x);
    jj[r] = t < ns ? stage[soff + t] : i0;
  }
  const int i = member ? member[slot] : slot;
  const int sp = slot_ptr[slot];
  const int sp1 = slot_ptr[slot + 1];  // slot + 1 <= n: slot_ptr has n + 1 entries
  // The wavefront's index, made scalar: its slice offset is one scalar load.
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / 64);
  const int wo = SL && !OFFS_LDS ? wave_off[static_cast<int64_t>(b) * wpb + wave] : 0;
  double xi, yi, zi;
  pos.ld(i, xi, yi, zi);
  const int np = active ? sp1 - sp : 0;
  const int rs = active ? sp - eoff : 0;
  const uint16_t* grow = nullptr;
  uint32_t oa[U];
  if constexpr (!OFFS_LDS) {
    // Inactive lanes read the first batch too: a slice's first batch, and
    // the U * TPI entries from the block's first, lie inside the padded
    // offset arrays.
    grow = SL ? sliced + wo + (threadIdx.x & 63) : local + eoff + rs;
    row_prefetch<TPI, U, SL>(grow, sub, oa);
  }
  // Staging issues all of a thread's loads of a pass before its LDS writes,
  // so each thread has kStageR independent gathers in flight. The first
  // pass (its indices loaded above) runs ahead of the loop, so its gathers
  // wait for those indices only.
#define LJGPU_STAGE_PASS(T0)                                                  \
  {                                                                           \
    double x[kStageR], y[kStageR], z[kStageR];                                \
    _Pragma("unroll") for (int r = 0; r < kStageR; r++)                       \
        pos.ld(jj[r], x[r], y[r], z[r]);                                      \
    _Pragma("unroll") for (int r = 0; r < kStageR; r++) {                     \
      const int t = (T0) + r * nt + static_cast<int>(threadIdx.x);            \
      if (t < ns) {                                                           \
        unwrap_near<PBC>(x[r], y[r], z[r], cx, cy, cz, pp);                   \
        lds[3 * t] = x[r];                                                    \
        lds[3 * t + 1] = y[r];                                                \
        lds[3 * t + 2] = z[r];                                                \
      }                                                                       \
    }                                                                         \
  }
  LJGPU_STAGE_PASS(0)
  for (int t0 = kStageR * nt; t0 < ns; t0 += kStageR * nt) {
#pragma unroll
    for (int r = 0; r < kStageR; r++) {
      const int t = t0 + r * nt + static_cast<int>(threadIdx.x);
      jj[r] = t < ns ? stage[soff + t] : i0;
    }
    LJGPU_STAGE_PASS(t0)
  }
#undef LJGPU_STAGE_PASS
  // The block's entry offsets, copied as 16-byte chunks from the 8-entry
  // boundary at or below eoff (the device array is padded by 8 entries).
  const int e0 = eoff & ~7;
  if constexpr (OFFS_LDS) {
    const int nchunks = (eoff + ne - e0 + 7) >> 3;
    const uint4* src = reinterpret_cast<const uint4*>(local + e0);
    uint4* dst = reinterpret_cast<uint4*>(lb + off_area);
    for (int c0 = 0; c0 < nchunks; c0 += kStageR * nt) {
      uint4 v[kStageR];
#pragma unroll
      for (int r = 0; r < kStageR; r++) {
        const int c = c0 + r * nt + static_cast<int>(threadIdx.x);
        if (c < nchunks) v[r] = src[c];
      }
#pragma unroll
      for (int r = 0; r < kStageR; r++) {
        const int c = c0 + r * nt + static_cast<int>(threadIdx.x);
        if (c < nchunks) dst[c] = v[r];
      }
    }
  }
  const uint16_t* loff = reinterpret_cast<const uint16_t*>(lb + off_area) + (eoff - e0);
  int far = 0;
  if constexpr (PBC) {
    unwrap_near<PBC>(xi, yi, zi, cx, cy, cz, pp);
    far = active && (fabs(xi - cx) >= reach || fabs(yi - cy) >= reach || fabs(zi - cz) >= reach);
  }
  // The frame check ORed over the workgroup through one LDS word: threads
  // that fail it write this tile's index, which no earlier tile of this
  // workgroup used, so the word needs no reset.
  if (far) *flag = b;
  __syncthreads();  // staging complete
  const bool any_far = PBC && *flag == b;
  Accum<PREC> acc;
  if (active) {
    if constexpr (OFFS_LDS) {
      if (PBC && any_far)
        tile_row<TPI, U, PREC, true>(mom, loff + rs, np, sub, xi, yi, zi, pp, acc);
      else
        tile_row<TPI, U, PREC, false>(mom, loff + rs, np, sub, xi, yi, zi, pp, acc);
    } else {
      if (PBC && any_far)
        tile_row_stream<TPI, U, PREC, PBC, true, SL>(mom, grow, np, sub, xi, yi, zi, pp, acc, oa);
      else
        tile_row_stream<TPI, U, PREC, PBC, false, SL>(mom, grow, np, sub, xi, yi, zi, pp, acc, oa);
    }
  }
  if constexpr (PREC == kEnergy) {
    // Row sums of the energy terms and pair counts into the thread's totals.
    const double ru = segment_sum<TPI>(acc.vx());
    const double rc = segment_sum<TPI>(acc.vy());
    if (active && sub == 0) {
      eacc[0] += ru;
      eacc[1] += rc;
    }
  } else {
    double fx = segment_sum<TPI>(acc.vx());
    double fy = segment_sum<TPI>(acc.vy());
    double fz = segment_sum<TPI>(acc.vz());
    if constexpr (PREC == kF64) {
      fx *= pp.c24;
      fy *= pp.c24;
      fz *= pp.c24;
    }
    if (active && sub == 0) mom.add(i, fx, fy, fz);
  }
}

// Register budget (measured on MI355X): five waves per SIMD, at most 96
// VGPRs (the default FP64 kernel takes 90, the mixed one 86, neither
// spills). With the first staging pass peeled off its loop and the batched
// reciprocal, FP64 no longer fits six waves' 80 VGPRs without spilling; the
// mixed kernels run 7 % faster at five waves than at six and 5 % faster
// than at four.
template <int TPI, int U, int PREC, class P, class M, bool PBC, bool OFFS_LDS, bool SL>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(5)))
k_tile(P pos, M mom, const int32_t* __restrict__ nop, const int32_t* __restrict__ member,
       const int32_t* __restrict__ slot_ptr, const int4* __restrict__ desc,
       const int32_t* __restrict__ stage,
       const uint16_t* __restrict__ local, const uint16_t* __restrict__ sliced,
       const int32_t* __restrict__ wave_off, int32_t wpb, int32_t first, int32_t nblocks,
       PairParams pp, double reach, const double2* __restrict__ rep_box) {
  // Dynamic LDS only (no static LDS, so its base is address 0 and constant
  // offsets fold into the LDS instructions): a 16-byte header holding the
  // frame-check word, then the tile's positions (and offsets).
  extern __shared__ double lds[];
  int* flag = reinterpret_cast<int*>(lds);
  char* lb = reinterpret_cast<char*>(lds) + kTileLdsHeader;
  if (threadIdx.x == 0) *flag = -1;
  __syncthreads();
  const int xcd = static_cast<int>(blockIdx.x) % kXcds;
  const int per = (nblocks + kXcds - 1) / kXcds;
  const int stride = static_cast<int>(gridDim.x) / kXcds;
  const int t_end = min(nblocks, (xcd + 1) * per);
  for (int b = xcd * per + static_cast<int>(blockIdx.x) / kXcds; b < t_end; b += stride) {
    tile_body<TPI, U, PREC, P, M, PBC, OFFS_LDS, SL>(pos, mom, nop, member, slot_ptr, desc, stage,
                                                     local, sliced, wave_off, wpb, first + b, pp,
                                                     reach, rep_box, lb, flag, nullptr);
    __syncthreads();  // the next tile overwrites this tile's LDS
  }
}

// Potential energy over the tiles (the force kernel's staging, offsets and
// pair arithmetic, with the energy policy): each workgroup walks its tiles
// as k_tile does and writes weight * 4 * sum s^3 (s^3 - 1) and weight * the
// pair count into part_u / part_c[blockIdx.x] (fixed-order sums, so the
// result does not depend on scheduling). Row offsets from `local`.
template <int TPI, int U, bool PBC>
__global__ void __launch_bounds__(1024)
k_tile_energy(PosAos pos, const int32_t* __restrict__ member,
              const int32_t* __restrict__ slot_ptr, const int4* __restrict__ desc,
              const int32_t* __restrict__ stage, const uint16_t* __restrict__ local,
              int32_t nblocks, PairParams pp, double reach, double weight,
              double* __restrict__ part_u, double* __restrict__ part_c) {
  extern __shared__ double lds[];
  int* flag = reinterpret_cast<int*>(lds);
  char* lb = reinterpret_cast<char*>(lds) + kTileLdsHeader;
  if (threadIdx.x == 0) *flag = -1;
  __syncthreads();
  double eacc[2] = {0.0, 0.0};
  const int xcd = static_cast<int>(blockIdx.x) % kXcds;
  const int per = (nblocks + kXcds - 1) / kXcds;
  const int stride = static_cast<int>(gridDim.x) / kXcds;
  const int t_end = min(nblocks, (xcd + 1) * per);
  for (int b = xcd * per + static_cast<int>(blockIdx.x) / kXcds; b < t_end; b += stride) {
    tile_body<TPI, U, kEnergy, PosAos, MomAos, PBC, false, false>(
        pos, MomAos{nullptr}, nullptr, member, slot_ptr, desc, stage, local, nullptr, nullptr, 0,
        b, pp, reach, nullptr, lb, flag, eacc);
    __syncthreads();
  }
  // Wavefront sums, then the wavefronts' sums in order by thread 0.
  const double wu = segment_sum<64>(eacc[0]);
  const double wc = segment_sum<64>(eacc[1]);
  double* red = lds + 2;  // past the header word, free after the last barrier
  const int w = static_cast<int>(threadIdx.x) / 64;
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = wu;
    red[2 * w + 1] = wc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double u = 0.0, c = 0.0;
    for (int k = 0; k < static_cast<int>(blockDim.x) / 64; k++) {
      u += red[2 * k];
      c += red[2 * k + 1];
    }
    part_u[blockIdx.x] = 4.0 * weight * u;
    part_c[blockIdx.x] = weight * c;
  }
}

inline int grid_blocks(int64_t threads, int block, int* padded) {
  int64_t blocks = (threads + block - 1) / block;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;
  *padded = static_cast<int>(blocks);
  return static_cast<int>(blocks);
}

inline int launch_status(const char* what) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? LJGPU_OK
                         : set_error(LJGPU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Launch k_tile_energy over the installed tiling with `grid` workgroups (a
// multiple of 8, at most the partial arrays' length). Four lanes per
// particle and four entries per batch, whatever the force configuration
// (the `local` offsets serve every configuration).
inline int run_energy_tiles(hipStream_t s, const double* q, const ListView& lv, double cl2,
                            double box, double weight, int grid, double* part_u,
                            double* part_c) {
  constexpr int kLanes = 4;
  const int threads = (lv.tile_particles * kLanes + 63) / 64 * 64;
  if (threads > 1024 || grid % kXcds != 0 || grid <= 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "tiled energy: launch shape");
  PairParams pp{};
  pp.cl2 = cl2;
  pp.rc = sqrt(cl2);
  pp.box = box > 0.0 ? box : 0.0;
  pp.inv_box = box > 0.0 ? 1.0 / box : 0.0;
  const double reach = 0.5 * pp.box - sqrt(cl2);
  // The header, the staged positions, and room for the final reduction.
  const size_t lds_bytes =
      kTileLdsHeader + std::max<size_t>(static_cast<size_t>(lv.tile_max_stage) * 24, 16 * 16);
  PosAos pos{q};
  if (box > 0.0)
    hipLaunchKernelGGL((k_tile_energy<kLanes, 4, true>), dim3(grid), dim3(threads), lds_bytes, s,
                       pos, lv.tile_member, lv.tile_slot_ptr,
                       reinterpret_cast<const int4*>(lv.tile_desc), lv.tile_stage, lv.tile_local,
                       lv.tile_blocks, pp, reach, weight, part_u, part_c);
  else
    hipLaunchKernelGGL((k_tile_energy<kLanes, 4, false>), dim3(grid), dim3(threads), lds_bytes, s,
                       pos, lv.tile_member, lv.tile_slot_ptr,
                       reinterpret_cast<const int4*>(lv.tile_desc), lv.tile_stage, lv.tile_local,
                       lv.tile_blocks, pp, reach, weight, part_u, part_c);
  return launch_status("tiled energy kernel launch");
}

template <int TPI, int U, int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
int run_csr(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp, int block) {
  int padded = 0;
  const int blocks = grid_blocks(static_cast<int64_t>(lv.n) * TPI, block, &padded);
  hipLaunchKernelGGL((k_csr<TPI, U, PREC, P, M, PBC, HALF, CLAMP>), dim3(blocks), dim3(block), 0,
                     s, pos, mom, lv.n, lv.nop, lv.ptr, lv.list, lv.n_entries, pp, padded);
  return launch_status("force kernel launch");
}

template <int U, int PREC, class P, class M, bool PBC, bool HALF>
int run_ell(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp, int block) {
  int padded = 0;
  const int blocks = grid_blocks(lv.n, block, &padded);
  hipLaunchKernelGGL((k_ell<U, PREC, P, M, PBC, HALF>), dim3(blocks), dim3(block), 0, s, pos, mom,
                     lv.n, lv.nop, lv.list, lv.ell_width, pp, padded);
  return launch_status("ELL force kernel launch");
}

template <int TPI, int U, int PREC, class P, class M, bool PBC, bool OFFS_LDS, bool SL>
int run_tile(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
             const LaunchConfig& lc) {
  const int threads = (lv.tile_particles * TPI + 63) / 64 * 64;
  if (threads > 1024)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "tile_particles * lanes_per_particle must not exceed 1024");
  int grid = (lv.tile_blocks + kXcds - 1) / kXcds * kXcds;
  if (lc.tile_resident > 0 && lc.cus > 0) {
    const int persistent = (lc.tile_resident * lc.cus + kXcds - 1) / kXcds * kXcds;
    grid = std::min(grid, persistent);
  }
  const size_t lds_bytes =
      kTileLdsHeader + (OFFS_LDS ? static_cast<size_t>(lv.tile_lds_bytes)
                                 : static_cast<size_t>(lv.tile_max_stage) * 24);
  const double reach = 0.5 * pp.box - sqrt(pp.cl2);
  hipLaunchKernelGGL((k_tile<TPI, U, PREC, P, M, PBC, OFFS_LDS, SL>), dim3(grid), dim3(threads),
                     lds_bytes, s, pos, mom, lv.nop, lv.tile_member, lv.tile_slot_ptr,
                     reinterpret_cast<const int4*>(lv.tile_desc), lv.tile_stage, lv.tile_local,
                     lv.tile_sliced, lv.tile_wave_off, lv.tile_waves_per_block, lv.tile_first,
                     lv.tile_blocks, pp, reach, reinterpret_cast<const double2*>(lv.replica_box));
  return launch_status("tiled force kernel launch");
}

template <int TPI, int U, int PREC, class P, class M, bool PBC>
int run_tile_layout(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
                    const LaunchConfig& lc) {
  if (lc.tile_offsets_lds) return run_tile<TPI, U, PREC, P, M, PBC, true, false>(s, pos, mom, lv, pp, lc);
  // The wavefront slices serve the launch they were cut for.
  const bool sl = lv.tile_sliced && lv.tile_sliced_lanes == TPI && lv.tile_sliced_unroll == U;
  return sl ? run_tile<TPI, U, PREC, P, M, PBC, false, true>(s, pos, mom, lv, pp, lc)
            : run_tile<TPI, U, PREC, P, M, PBC, false, false>(s, pos, mom, lv, pp, lc);
}

template <int TPI, int PREC, class P, class M, bool PBC>
int run_tile_unroll(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
                    const LaunchConfig& lc) {
  switch (lc.unroll) {
    case 1: return run_tile_layout<TPI, 1, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 2: return run_tile_layout<TPI, 2, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 4: return run_tile_layout<TPI, 4, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    default: return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unroll must be 1, 2 or 4");
  }
}

template <int PREC, class P, class M, bool PBC>
int run_tile_lanes(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
                   const LaunchConfig& lc) {
  switch (lc.lanes) {
    case 2: return run_tile_unroll<2, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 4: return run_tile_unroll<4, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 8: return run_tile_unroll<8, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 16: return run_tile_unroll<16, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 32: return run_tile_unroll<32, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    case 64: return run_tile_unroll<64, PREC, P, M, PBC>(s, pos, mom, lv, pp, lc);
    default:
      return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                       "lanes_per_particle must be 2, 4, 8, 16, 32 or 64 for the tiled kernel");
  }
}

template <int TPI, int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
int run_unroll(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
               const LaunchConfig& lc) {
  switch (lc.unroll) {
    case 1: return run_csr<TPI, 1, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 2: return run_csr<TPI, 2, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 4: return run_csr<TPI, 4, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    default: return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unroll must be 1, 2 or 4");
  }
}

template <int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
int run_lanes(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
              const LaunchConfig& lc) {
  if (lv.ell_width > 0) {
    if constexpr (CLAMP) {
      return set_error(LJGPU_ERR_UNSUPPORTED, "ELL list with clamping");
    } else {
      switch (lc.unroll) {
        case 1: return run_ell<1, PREC, P, M, PBC, HALF>(s, pos, mom, lv, pp, lc.block);
        case 2: return run_ell<2, PREC, P, M, PBC, HALF>(s, pos, mom, lv, pp, lc.block);
        case 4: return run_ell<4, PREC, P, M, PBC, HALF>(s, pos, mom, lv, pp, lc.block);
        default: return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unroll must be 1, 2 or 4");
      }
    }
  }
  switch (lc.lanes) {
    case 2: return run_unroll<2, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    case 4: return run_unroll<4, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    case 8: return run_unroll<8, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    case 16: return run_unroll<16, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    case 32: return run_unroll<32, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    case 64: return run_unroll<64, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc);
    default:
      return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                       "lanes_per_particle must be 2, 4, 8, 16, 32 or 64 for a CSR list");
  }
}

template <int PREC, class P, class M>
int run_layout(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
               const LaunchConfig& lc) {
  const bool pbc = pp.box > 0.0;
  if (lc.clamp) {
    // The reference-kernel-ABI entry: AoS, open box, FP64 only.
    if constexpr (PREC == kF64 && std::is_same<P, PosAos>::value) {
      if (pbc) return set_error(LJGPU_ERR_UNSUPPORTED, "clamped launch with periodic box");
      return lc.half_atomic ? run_lanes<PREC, P, M, false, true, true>(s, pos, mom, lv, pp, lc)
                            : run_lanes<PREC, P, M, false, false, true>(s, pos, mom, lv, pp, lc);
    } else {
      return set_error(LJGPU_ERR_UNSUPPORTED, "clamped launch outside AoS FP64");
    }
  }
  if (lv.replica_box) {
    // Batched replicas: tiles only, each block with its replica's box.
    if (lv.tile_blocks <= 0 || lc.half_atomic)
      return set_error(LJGPU_ERR_UNSUPPORTED,
                       "batched replicas need the tiled kernel (lanes_per_particle > 1, "
                       "tiling on, no atomic half list)");
    return run_tile_lanes<PREC, P, M, true>(s, pos, mom, lv, pp, lc);
  }
  if (lv.tile_blocks > 0 && !lc.half_atomic) {
    return pbc ? run_tile_lanes<PREC, P, M, true>(s, pos, mom, lv, pp, lc)
               : run_tile_lanes<PREC, P, M, false>(s, pos, mom, lv, pp, lc);
  }
  if (pbc) {
    return lc.half_atomic ? run_lanes<PREC, P, M, true, true, false>(s, pos, mom, lv, pp, lc)
                          : run_lanes<PREC, P, M, true, false, false>(s, pos, mom, lv, pp, lc);
  }
  return lc.half_atomic ? run_lanes<PREC, P, M, false, true, false>(s, pos, mom, lv, pp, lc)
                        : run_lanes<PREC, P, M, false, false, false>(s, pos, mom, lv, pp, lc);
}

inline int pair_params(const ForceArgs& fa, const LaunchConfig& lc, PairParams* pp) {
  if (lc.block < 64 || lc.block > 1024 || (lc.block % 64) != 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "block size must be a multiple of 64 in [64, 1024]");
  pp->c24 = 24.0 * fa.dt;
  pp->c48 = 48.0 * fa.dt;
  pp->cl2 = fa.cl2;
  pp->box = fa.box > 0.0 ? fa.box : 0.0;
  pp->inv_box = fa.box > 0.0 ? 1.0 / fa.box : 0.0;
  pp->c24f = static_cast<float>(pp->c24);
  pp->c48f = static_cast<float>(pp->c48);
  pp->rc = sqrt(fa.cl2);
  return LJGPU_OK;
}

// AoS (double[N][4]) or SoA arrays.
template <int PREC>
int run_force(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  if (lv.n <= 0) return LJGPU_OK;
  PairParams pp;
  const int rc = pair_params(fa, lc, &pp);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fa.soa) {
    PosSoa pos{fa.q[0], fa.q[1], fa.q[2]};
    MomSoa mom{fa.p[0], fa.p[1], fa.p[2]};
    return run_layout<PREC>(s, pos, mom, lv, pp, lc);
  }
  PosAos pos{fa.q[0]};
  MomAos mom{fa.p[0]};
  return run_layout<PREC>(s, pos, mom, lv, pp, lc);
}

// Packed AoS (double[N][3]) arrays.
template <int PREC>
int run_force_aos3(void* stream, const ListView& lv, const ForceArgs& fa,
                   const LaunchConfig& lc) {
  if (lv.n <= 0) return LJGPU_OK;
  PairParams pp;
  const int rc = pair_params(fa, lc, &pp);
  if (rc) return rc;
  PosAos3 pos{fa.q[0]};
  MomAos3 mom{fa.p[0]};
  return run_layout<PREC>(static_cast<hipStream_t>(stream), pos, mom, lv, pp, lc);
}

}  // namespace detail
}  // namespace ljgpu
