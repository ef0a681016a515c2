// force_kernels.hip — Lennard-Jones force kernels for gfx950 (MI355X).
//
// p += F(q) * dt over a Verlet list, the pair term of cpu_ref/force_aos.cpp:176-189
// and cuda/kernel.cuh:22-31. Two kernel families:
//
//  * k_csr: a segment of TPI consecutive lanes of a 64-wide wavefront owns one
//    i-particle and strides over its CSR row (lane t takes entries t, t+TPI,
//    ...), so each segment reads its row as one contiguous chunk and 64/TPI
//    rows share a wavefront. The per-lane partial forces are summed across
//    the segment with cross-lane shuffles and one lane updates p[i].
//    TPI = 64 is the wavefront-per-i form of the reference's warp_unroll
//    kernels (cuda/kernel.cuh:821-904, 32-lane warps); smaller TPI packs
//    several rows of ~80-150 entries into one wavefront so fewer lanes idle
//    on the last pass over a row.
//  * k_ell: one lane per i-particle over the reference's transposed (ELL)
//    list, entry k of row i at list[i + k*N] so a wavefront reads 64
//    consecutive indices per step (force_kernel_memopt2 and relatives,
//    cuda/kernel.cuh:102-236).
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
#include <hip/hip_runtime.h>

#include "lj_device.h"
#include "ljgpu_internal.h"

namespace ljgpu {
namespace {

constexpr int kF64 = 0;
constexpr int kMixedF64 = 1;
constexpr int kMixedKahan = 2;

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

// One list entry (i, j): distance, cutoff, force scalar, accumulation into
// the i-accumulator and, for half lists, the scatter of -f to j.
template <int PREC, bool PBC, bool HALF, class P, class M>
__device__ __forceinline__ void pair_entry(const P& pos, const M& mom, int j, double xi,
                                           double yi, double zi, const PairParams& pp,
                                           Accum<PREC>& acc) {
  double xj, yj, zj;
  pos.ld(j, xj, yj, zj);
  double dx = xj - xi, dy = yj - yi, dz = zj - zi;
  min_image<PBC>(dx, dy, dz, pp);
  const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  double df;
  float dff = 0.0f;
  if constexpr (PREC == kF64) {
    df = lj_df_f64(r2, pp);
  } else {
    dff = lj_df_mixed(r2, pp);
    df = static_cast<double>(dff);
  }
  acc.add(df, dff, dx, dy, dz);
  if constexpr (HALF) {
    if (r2 <= pp.cl2) mom.atomic_add(j, -df * dx, -df * dy, -df * dz);
  }
}

template <int TPI, int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
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
    for (int k = sub; k < np; k += TPI) {
      int j = row[k];
      if constexpr (CLAMP) j = static_cast<int>(min(static_cast<uint32_t>(j), static_cast<uint32_t>(n - 1)));
      pair_entry<PREC, PBC, HALF>(pos, mom, j, xi, yi, zi, pp, acc);
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

template <int PREC, class P, class M, bool PBC, bool HALF>
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
  for (int k = 0; k < np; k++) {
    const int j = col[static_cast<int64_t>(k) * n];
    pair_entry<PREC, PBC, HALF>(pos, mom, j, xi, yi, zi, pp, acc);
  }
  if constexpr (HALF) mom.atomic_add(i, acc.vx(), acc.vy(), acc.vz());
  else mom.add(i, acc.vx(), acc.vy(), acc.vz());
}

inline int grid_blocks(int64_t threads, int block, int* padded) {
  int64_t blocks = (threads + block - 1) / block;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;
  *padded = static_cast<int>(blocks);
  return static_cast<int>(blocks);
}

template <int TPI, int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
int run_csr(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp, int block) {
  int padded = 0;
  const int blocks = grid_blocks(static_cast<int64_t>(lv.n) * TPI, block, &padded);
  hipLaunchKernelGGL((k_csr<TPI, PREC, P, M, PBC, HALF, CLAMP>), dim3(blocks), dim3(block), 0, s,
                     pos, mom, lv.n, lv.nop, lv.ptr, lv.list, lv.n_entries, pp, padded);
  return hipGetLastError() == hipSuccess ? LJGPU_OK
                                         : set_error(LJGPU_ERR_HIP, "force kernel launch failed");
}

template <int PREC, class P, class M, bool PBC, bool HALF, bool CLAMP>
int run_lanes(hipStream_t s, P pos, M mom, const ListView& lv, const PairParams& pp,
              const LaunchConfig& lc) {
  if (lv.ell_width > 0) {
    if constexpr (CLAMP) {
      return set_error(LJGPU_ERR_UNSUPPORTED, "ELL list with clamping");
    } else {
      int padded = 0;
      const int blocks = grid_blocks(lv.n, lc.block, &padded);
      hipLaunchKernelGGL((k_ell<PREC, P, M, PBC, HALF>), dim3(blocks), dim3(lc.block), 0, s, pos,
                         mom, lv.n, lv.nop, lv.list, lv.ell_width, pp, padded);
      return hipGetLastError() == hipSuccess
                 ? LJGPU_OK
                 : set_error(LJGPU_ERR_HIP, "ELL force kernel launch failed");
    }
  }
  switch (lc.lanes) {
    case 1: return run_csr<1, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 4: return run_csr<4, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 8: return run_csr<8, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 16: return run_csr<16, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 32: return run_csr<32, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    case 64: return run_csr<64, PREC, P, M, PBC, HALF, CLAMP>(s, pos, mom, lv, pp, lc.block);
    default: return set_error(LJGPU_ERR_INVALID_ARGUMENT, "lanes_per_particle must be 1,4,8,16,32,64");
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
  if (pbc) {
    return lc.half_atomic ? run_lanes<PREC, P, M, true, true, false>(s, pos, mom, lv, pp, lc)
                          : run_lanes<PREC, P, M, true, false, false>(s, pos, mom, lv, pp, lc);
  }
  return lc.half_atomic ? run_lanes<PREC, P, M, false, true, false>(s, pos, mom, lv, pp, lc)
                        : run_lanes<PREC, P, M, false, false, false>(s, pos, mom, lv, pp, lc);
}

template <int PREC>
int run_force(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  if (lv.n <= 0) return LJGPU_OK;
  if (lc.block < 64 || lc.block > 1024 || (lc.block % 64) != 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "block size must be a multiple of 64 in [64, 1024]");
  PairParams pp;
  pp.c24 = 24.0 * fa.dt;
  pp.c48 = 48.0 * fa.dt;
  pp.cl2 = fa.cl2;
  pp.box = fa.box > 0.0 ? fa.box : 0.0;
  pp.inv_box = fa.box > 0.0 ? 1.0 / fa.box : 0.0;
  pp.c24f = static_cast<float>(pp.c24);
  pp.c48f = static_cast<float>(pp.c48);
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

}  // namespace

int launch_force_f64(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  return run_force<kF64>(stream, lv, fa, lc);
}

int launch_force_mixed(void* stream, const ListView& lv, const ForceArgs& fa,
                       const LaunchConfig& lc, int accum) {
  if (accum == LJGPU_ACCUM_F64) return run_force<kMixedF64>(stream, lv, fa, lc);
  if (accum == LJGPU_ACCUM_KAHAN_F32) return run_force<kMixedKahan>(stream, lv, fa, lc);
  return set_error(LJGPU_ERR_INVALID_ARGUMENT, "accum must be LJGPU_ACCUM_F64 or LJGPU_ACCUM_KAHAN_F32");
}

}  // namespace ljgpu
