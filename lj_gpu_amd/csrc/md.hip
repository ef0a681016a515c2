// md.hip — molecular dynamics around the force path: the pair potential
// energy over the installed list, and a velocity-Verlet driver (drift, force,
// periodic list rebuild on the device, energy samples) for unit masses.
//
// The reference has no integrator and no energy: it only repeats force
// evaluations on fixed positions (measure()'s loop, cuda/force_cuda.cu:
// 331-335). The driver reuses that same force evaluation (p += F(q) * dt) as
// its kick, so every step runs the library's hot path unchanged:
//
//   p += F(q) dt/2                          (t = 0 -> 1/2)
//   for s = 1 .. steps:
//     q += p dt, wrapped into [0, box)       (drift)
//     rebuild the list every rebuild_every steps
//     p += F(q) dt                           (kick, t = s-1/2 -> s+1/2)
//   with the kick split into two halves around an energy sample at t = s,
//   and a final half kick, so p is at integer time when the run ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "context.h"

namespace ljgpu {

namespace {

constexpr int kBlock = 256;

// x mapped into [0, box): one multiply by 1/box instead of a divide, then
// the last-bit cases (a result just below 0, or equal to box) folded in.
__device__ __forceinline__ double wrap(double x, double box, double inv_box) {
  x = fma(-box, floor(x * inv_box), x);
  if (x < 0.0) x += box;
  return x >= box ? x - box : x;
}

__device__ __forceinline__ double image(double d, double box, double inv_box) {
  return box > 0.0 ? fma(-box, rint(d * inv_box), d) : d;
}

// Drift q += p dt (wrapped into [0, box) for box > 0). With qref, also the
// largest squared displacement |q - qref|^2 (minimum image) since the last
// build, as the bit pattern of a non-negative double (integer order = numeric
// order) maximised into *out.
__global__ void k_drift(int32_t n, double* __restrict__ q, const double* __restrict__ p, double dt,
                        double box, double inv_box, const double* __restrict__ qref,
                        unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double2 xy = reinterpret_cast<double2*>(q + 4 * i)[0];
    double z = q[4 * i + 2];
    const double2 pxy = reinterpret_cast<const double2*>(p + 4 * i)[0];
    xy.x += pxy.x * dt;
    xy.y += pxy.y * dt;
    z += p[4 * i + 2] * dt;
    if (box > 0.0) {
      xy.x = wrap(xy.x, box, inv_box);
      xy.y = wrap(xy.y, box, inv_box);
      z = wrap(z, box, inv_box);
    }
    reinterpret_cast<double2*>(q + 4 * i)[0] = xy;
    q[4 * i + 2] = z;
    if (qref) {
      const double2 rxy = reinterpret_cast<const double2*>(qref + 4 * i)[0];
      const double dx = image(xy.x - rxy.x, box, inv_box);
      const double dy = image(xy.y - rxy.y, box, inv_box);
      const double dz = image(z - qref[4 * i + 2], box, inv_box);
      m = fmax(m, dx * dx + dy * dy + dz * dz);
    }
  }
  if (qref) {
    // One atomic per workgroup: atomics on one address serialise.
    __shared__ double sm[kBlock / 64];
    for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kBlock / 64; w++) m = fmax(m, sm[w]);
      atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
  }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kBlock / 64; w++) t += sh[w];
  __syncthreads();
  return t;
}

// Potential energy: sum over the list of 4 (r^-12 - r^-6) for r2 <= cl2
// (the force kernels' cutoff rule), and the number of those pairs, times
// `weight` (1/2 for a full list, which holds every pair twice). Eight lanes
// share a row (lane k takes entries k, k + 8, ...), so index loads are
// contiguous; each lane loads kEnergyBatch partners before using any. One
// partial sum of each quantity per workgroup.
constexpr int kEnergyLanes = 8;
constexpr int kEnergyBatch = 4;

__global__ __launch_bounds__(kBlock) void k_potential(int32_t n, const int32_t* __restrict__ nop,
                                                      const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ list,
                                                      const double* __restrict__ q, double cl2,
                                                      double box, double inv_box, double weight,
                                                      double* __restrict__ partial,
                                                      double* __restrict__ partial_count) {
  __shared__ double sh[kBlock / 64];
  double u = 0.0, cnt = 0.0;
  const int sub = static_cast<int>(threadIdx.x) % kEnergyLanes;
  const int64_t rows_per_pass = static_cast<int64_t>(gridDim.x) * (kBlock / kEnergyLanes);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock / kEnergyLanes) +
                   threadIdx.x / kEnergyLanes;
       i < n; i += rows_per_pass) {
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    const int64_t kp = ptr[i];
    const int32_t np = nop[i];
    for (int32_t k0 = sub; k0 < np; k0 += kEnergyLanes * kEnergyBatch) {
      int64_t j[kEnergyBatch];
#pragma unroll
      for (int b = 0; b < kEnergyBatch; b++) {
        const int32_t k = k0 + b * kEnergyLanes;
        j[b] = k < np ? list[kp + k] : i;  // past the row: the particle itself (r2 = 0)
      }
      double x[kEnergyBatch], y[kEnergyBatch], z[kEnergyBatch];
#pragma unroll
      for (int b = 0; b < kEnergyBatch; b++) {
        x[b] = q[4 * j[b]];
        y[b] = q[4 * j[b] + 1];
        z[b] = q[4 * j[b] + 2];
      }
#pragma unroll
      for (int b = 0; b < kEnergyBatch; b++) {
        const double dx = image(x[b] - xi, box, inv_box);
        const double dy = image(y[b] - yi, box, inv_box);
        const double dz = image(z[b] - zi, box, inv_box);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (k0 + b * kEnergyLanes >= np || r2 > cl2) continue;
        const double s = 1.0 / r2;
        const double s3 = s * s * s;
        u += 4.0 * s3 * (s3 - 1.0);
        cnt += 1.0;
      }
    }
  }
  const double t = block_sum(u * weight, sh);
  const double c = block_sum(cnt * weight, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = t;
    partial_count[blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(kBlock) void k_kinetic(int32_t n, const double* __restrict__ p,
                                                    double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double k = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double px = p[4 * i], py = p[4 * i + 1], pz = p[4 * i + 2];
    k += 0.5 * (px * px + py * py + pz * pz);
  }
  const double t = block_sum(k, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Working copies in slot order: out[s] = in[order[s]] (gather) or
// out[order[s]] = in[s] (scatter), for 4-double rows and int32 values.
__global__ void k_gather_rows(int32_t n, const int32_t* __restrict__ order,
                              const double4* __restrict__ in, double4* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[t] = in[order[t]];
}

__global__ void k_scatter_rows(int32_t n, const int32_t* __restrict__ order,
                               const double4* __restrict__ in, double4* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[order[t]] = in[t];
}

__global__ void k_gather_i32(int32_t n, const int32_t* __restrict__ order,
                             const int32_t* __restrict__ in, int32_t* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[t] = in[order[t]];
}

__global__ void k_iota_i32(int32_t n, int32_t* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[t] = static_cast<int32_t>(t);
}

int grid_for(int64_t n) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 65535)));
}

// Sums of the energy partials in block order b = 0, 1, ... from 0.0 (the
// order of energy()'s host loop, so both give the same doubles): out =
// (U, K, pairs); grid_q partials of U and pairs, grid_p of K.
__global__ void k_energy_total(int grid_q, int grid_p, int with_q, int with_p,
                               const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double u = 0.0, k = 0.0, c = 0.0;
  if (with_q)
    for (int b = 0; b < grid_q; b++) u += part[b];
  if (with_p)
    for (int b = 0; b < grid_p; b++) k += part[kEnergyBlocks + b];
  if (with_q)
    for (int b = 0; b < grid_q; b++) c += part[2 * kEnergyBlocks + b];
  out[0] = u;
  out[1] = k;
  out[2] = c;
}

// Launch the energy kernels of the current state into `part` (3 *
// kEnergyBlocks doubles: potential, kinetic and pair-count partials). Over
// an installed tiling (the force kernels' layout, used as it stands) the
// potential runs on the tiles (k_tile_energy), otherwise over the CSR rows
// (k_potential). *grid_q / *grid_p: the partials of U and pairs / of K.
int launch_energy(ljgpu_ctx* ctx, const double* q, const double* p, double box, double cl2,
                  double* part, int* grid_q, int* grid_p) {
  const int grid = std::min(kEnergyBlocks, grid_for(ctx->n));
  *grid_q = *grid_p = grid;
  if (q) {
    const double weight = ctx->eval_half ? 1.0 : 0.5;
    const bool tiled = ctx->tiles.blocks > 0 && !ctx->eval_half && ctx->rep_off.empty() &&
                       ctx->tiles.tile_particles * 4 <= 1024;
    if (tiled) {
      ListView lv;
      tile_fields(ctx, &lv);
      *grid_q = std::min(kEnergyBlocks, (ctx->tiles.blocks + 7) / 8 * 8);
      TRY(launch_energy_tiles(ctx->stream, q, lv, cl2, box, weight, *grid_q, part,
                              part + 2 * kEnergyBlocks));
    } else {
      k_potential<<<grid, kBlock, 0, ctx->stream>>>(ctx->n, ctx->eval.nop.as<int32_t>(),
                                                    ctx->eval.ptr.as<int32_t>(),
                                                    ctx->eval.list.as<int32_t>(), q, cl2, box,
                                                    box > 0.0 ? 1.0 / box : 0.0, weight, part,
                                                    part + 2 * kEnergyBlocks);
    }
  }
  if (p) k_kinetic<<<grid, kBlock, 0, ctx->stream>>>(ctx->n, p, part + kEnergyBlocks);
  HIP_TRY(hipGetLastError(), "energy kernels");
  return LJGPU_OK;
}

// U and K of the current state (synchronises; partial sums added in a fixed
// order, so the result does not depend on scheduling).
int energy(ljgpu_ctx* ctx, const double* q, const double* p, double box, double cl2, double* u,
           double* k, double* pairs) {
  *u = 0.0;
  *k = 0.0;
  *pairs = 0.0;
  if (ctx->n == 0) return LJGPU_OK;
  TRY(dev_reserve(ctx->md_red, sizeof(double) * 3 * kEnergyBlocks + 64, ctx->stream));
  double* part = ctx->md_red.as<double>();
  int grid_q = 0, grid_p = 0;
  TRY(launch_energy(ctx, q, p, box, cl2, part, &grid_q, &grid_p));
  std::vector<double> h(3 * kEnergyBlocks);
  HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                         ctx->stream), "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (q)
    for (int b = 0; b < grid_q; b++) *u += h[b];
  if (p)
    for (int b = 0; b < grid_p; b++) *k += h[kEnergyBlocks + b];
  if (q)
    for (int b = 0; b < grid_q; b++) *pairs += h[2 * kEnergyBlocks + b];
  return LJGPU_OK;
}

}  // namespace

}  // namespace ljgpu

using namespace ljgpu;

extern "C" {

int ljgpu_energy(ljgpu_ctx* ctx, const double* q, const double* p, double box, double cl2,
                 double* potential, double* kinetic, int64_t* pairs_in_cutoff) {
  if (!ctx || !potential || !kinetic) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  if (!ctx->rep_off.empty())
    return set_error(LJGPU_ERR_UNSUPPORTED, "energy of batched replicas");
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!(cl2 >= 0.0)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "cl2 must be >= 0");
  TRY(use_device(ctx));
  double pairs = 0.0;
  TRY(energy(ctx, q, p, box, cl2, potential, kinetic, &pairs));
  if (pairs_in_cutoff) *pairs_in_cutoff = static_cast<int64_t>(pairs);
  return LJGPU_OK;
}

int ljgpu_energy_noalloc(ljgpu_ctx* ctx, const double* q, const double* p, double box,
                         double cl2, double* potential, double* kinetic,
                         int64_t* pairs_in_cutoff) {
  static_assert(kNoAllocScratchDoubles >= 3 * kEnergyBlocks + 3, "scratch too small");
  if (!ctx || !potential || !kinetic) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  if (!ctx->rep_off.empty())
    return set_error(LJGPU_ERR_UNSUPPORTED, "energy of batched replicas");
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!(cl2 >= 0.0)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "cl2 must be >= 0");
  if (!ctx->noalloc_scratch) return set_error(LJGPU_ERR_HIP, "no scratch address");
  TRY(use_device(ctx));
  *potential = 0.0;
  *kinetic = 0.0;
  if (pairs_in_cutoff) *pairs_in_cutoff = 0;
  if (ctx->n == 0) return LJGPU_OK;
  // Every partial the totals read is written by this call's kernels (one per
  // workgroup of the grid), so nothing in the scratch needs clearing.
  double* part = ctx->noalloc_scratch;
  double* total = part + 3 * kEnergyBlocks;
  int grid_q = 0, grid_p = 0;
  TRY(launch_energy(ctx, q, p, box, cl2, part, &grid_q, &grid_p));
  k_energy_total<<<1, 64, 0, ctx->stream>>>(grid_q, grid_p, q != nullptr, p != nullptr, part,
                                            total);
  HIP_TRY(hipGetLastError(), "k_energy_total");
  double h[3] = {0.0, 0.0, 0.0};
  HIP_TRY(hipMemcpyAsync(h, total, sizeof(h), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  *potential = h[0];
  *kinetic = h[1];
  if (pairs_in_cutoff) *pairs_in_cutoff = static_cast<int64_t>(h[2]);
  return LJGPU_OK;
}

int ljgpu_md_run(ljgpu_ctx* ctx, double* q, double* p, int32_t n, double box,
                 const ljgpu_md_config* cfg, int32_t steps, ljgpu_md_stats* stats,
                 double* samples, int32_t sample_capacity) {
  if (!ctx || !cfg || !stats) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 0 || steps < 0 || (n > 0 && (!q || !p)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "md_run: bad particle arrays or steps");
  if ((reinterpret_cast<uintptr_t>(q) | reinterpret_cast<uintptr_t>(p)) & 15u)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 16-byte aligned");
  if (!(cfg->dt > 0.0) || !std::isfinite(cfg->dt) || !(cfg->cutoff > 0.0) || !(cfg->skin >= 0.0) ||
      cfg->rebuild_every < 0 || cfg->sample_every < 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "md_run: bad configuration");
  if (cfg->precision != LJGPU_PREC_F64 && cfg->precision != LJGPU_PREC_MIXED)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "md_run: unknown precision");
  if (cfg->accum != LJGPU_ACCUM_F64 && cfg->accum != LJGPU_ACCUM_KAHAN_F32)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "md_run: unknown accumulator");
  if (samples && sample_capacity < 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "md_run: negative sample capacity");
  if (!ctx->rep_off.empty())
    return set_error(LJGPU_ERR_UNSUPPORTED, "molecular dynamics of batched replicas");
  std::memset(stats, 0, sizeof(*stats));
  TRY(use_device(ctx));
  // Tiles in index order over slot-ordered working copies (below): the
  // copies are already in a spatial order, so consecutive slots are compact
  // tiles without the per-entry slot lookups of spatially ordered tiles over
  // the caller's arrays. The caller's choice is restored on return.
  // Tiles rebuilt every few steps skip the bank-class row order: the pass
  // costs more per rebuild than it saves over the kicks between rebuilds.
  struct OrderGuard {
    ljgpu_ctx* c;
    int saved;
    bool saved_bank;
    ~OrderGuard() {
      if (c->tile_order != saved || c->bank_order != saved_bank) {
        c->tile_order = saved;
        c->bank_order = saved_bank;
        c->tiles_for = -1;
        c->tiles_ordered = false;
      }
    }
  } order_guard{ctx, ctx->tile_order, ctx->bank_order};
  ctx->tile_order = LJGPU_TILE_ORDER_INDEX;
  ctx->bank_order = false;
  const double dt = cfg->dt, cl2 = cfg->cutoff * cfg->cutoff;
  const double search = cfg->cutoff + cfg->skin;
  const bool mixed = cfg->precision == LJGPU_PREC_MIXED;
  const int accum = cfg->accum;
  // In a periodic box the loop runs on copies of q and p in a spatial
  // order of the particles, re-sorted at every list rebuild
  // (dev_tile_order): slot s holds the caller's particle perm[s]. The
  // copies go back to the caller's order on return, and the list, which
  // indexes slots, is uninstalled then. An open box runs on q and p.
  const bool slots = box > 0.0 && n > 0;
  double* wq = q;
  double* wp = p;
  int32_t* perm = nullptr;
  int32_t* perm_spare = nullptr;
  const size_t row_bytes = sizeof(double) * 4 * static_cast<size_t>(std::max(n, 1));
  const int grid_n = grid_for(n);
  if (slots) {
    for (DevMem* m : {&ctx->md_q, &ctx->md_p, &ctx->md_tmp})
      TRY(dev_reserve(*m, row_bytes, ctx->stream));
    for (DevMem* m : {&ctx->md_perm, &ctx->md_perm2})
      TRY(dev_reserve(*m, sizeof(int32_t) * n, ctx->stream));
    wq = ctx->md_q.as<double>();
    wp = ctx->md_p.as<double>();
    perm = ctx->md_perm.as<int32_t>();
    perm_spare = ctx->md_perm2.as<int32_t>();
    HIP_TRY(hipMemcpyAsync(wq, q, row_bytes, hipMemcpyDeviceToDevice, ctx->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(wp, p, row_bytes, hipMemcpyDeviceToDevice, ctx->stream),
            "hipMemcpyAsync");
    k_iota_i32<<<grid_n, kBlock, 0, ctx->stream>>>(n, perm);
    HIP_TRY(hipGetLastError(), "slot index kernel");
  }
  bool returned = !slots;
  auto give_back = [&]() -> int {
    if (returned) return LJGPU_OK;
    returned = true;
    k_scatter_rows<<<grid_n, kBlock, 0, ctx->stream>>>(n, perm, reinterpret_cast<double4*>(wq),
                                                       reinterpret_cast<double4*>(q));
    k_scatter_rows<<<grid_n, kBlock, 0, ctx->stream>>>(n, perm, reinterpret_cast<double4*>(wp),
                                                       reinterpret_cast<double4*>(p));
    HIP_TRY(hipGetLastError(), "slot scatter kernels");
    reset_list(ctx);
    return LJGPU_OK;
  };
  struct GiveBack {
    std::function<int()> f;
    ~GiveBack() { (void)f(); }
  } give_back_guard{give_back};
  PosIn in;
  in.x = wq;
  in.y = wq + 1;
  in.z = wq + 2;
  in.stride = 4;
  const int32_t order_tile = effective_tile(ctx) > 0 ? effective_tile(ctx) : 64;
  hipEvent_t e0 = nullptr, e1 = nullptr, b0 = nullptr, b1 = nullptr;
  auto cleanup = [&]() {
    for (hipEvent_t e : {e0, e1, b0, b1})
      if (e) (void)hipEventDestroy(e);
  };
  struct Guard {
    std::function<void()> f;
    ~Guard() { f(); }
  } guard{cleanup};
  HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
  HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
  HIP_TRY(hipEventCreate(&b0), "hipEventCreate");
  HIP_TRY(hipEventCreate(&b1), "hipEventCreate");
  HIP_TRY(hipEventRecord(e0, ctx->stream), "hipEventRecord");

  float rebuild_ms = 0.0f;
  int32_t rebuilds = 0, stale = 0;
  double max_disp = 0.0;
  TRY(dev_reserve(ctx->md_qref, sizeof(double) * 4 * std::max(n, 1), ctx->stream));
  auto rebuild = [&]() -> int {
    HIP_TRY(hipEventRecord(b0, ctx->stream), "hipEventRecord");
    if (slots) {
      // Re-sort the working copies: new slot s takes old slot order[s].
      bool ordered = false;
      TRY(dev_tile_order(ctx->stream, in, n, box, {}, {}, order_tile, ctx->work, &ordered));
      if (ordered) {
        const int32_t* order = ctx->work.order.as<int32_t>();
        double4* tmp = ctx->md_tmp.as<double4>();
        for (double* w : {wq, wp}) {
          k_gather_rows<<<grid_n, kBlock, 0, ctx->stream>>>(n, order,
                                                            reinterpret_cast<double4*>(w), tmp);
          HIP_TRY(hipMemcpyAsync(w, tmp, row_bytes, hipMemcpyDeviceToDevice, ctx->stream),
                  "hipMemcpyAsync");
        }
        k_gather_i32<<<grid_n, kBlock, 0, ctx->stream>>>(n, order, perm, perm_spare);
        HIP_TRY(hipGetLastError(), "slot gather kernels");
        std::swap(perm, perm_spare);
      }
    }
    // Rows in the build's own order: the dynamics need the pair set, not
    // the reference's row order, and that order is fixed by the positions.
    TRY(build_list(ctx, in, n, box, search, LJGPU_LIST_FULL, false));
    if (n > 0)
      HIP_TRY(hipMemcpyAsync(ctx->md_qref.p, wq, sizeof(double) * 4 * n,
                             hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync");
    HIP_TRY(hipEventRecord(b1, ctx->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(b1), "hipEventSynchronize");
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b0, b1), "hipEventElapsedTime");
    rebuild_ms += ms;
    rebuilds++;
    return LJGPU_OK;
  };
  TRY(rebuild());
  ForceArgs fa;
  fa.soa = false;
  fa.q[0] = wq;
  fa.p[0] = wp;
  fa.cl2 = cl2;
  fa.box = box;
  int32_t nsamples = 0;
  double e_first = 0.0, max_drift = 0.0;
  // The conserved energy of truncated forces is K plus the shifted potential
  // (each pair inside the cutoff offset by -U(cutoff)), which is continuous
  // where pairs cross the cutoff.
  const double s6 = 1.0 / (cl2 * cl2 * cl2);
  const double u_cut = 4.0 * s6 * (s6 - 1.0);
  auto sample = [&](int32_t step) -> int {
    double u = 0.0, k = 0.0, pairs = 0.0;
    TRY(energy(ctx, wq, wp, box, cl2, &u, &k, &pairs));
    u -= pairs * u_cut;
    if (nsamples == 0) {
      stats->potential0 = u;
      stats->kinetic0 = k;
      e_first = u + k;
    }
    stats->potential1 = u;
    stats->kinetic1 = k;
    max_drift = std::max(max_drift, std::fabs(u + k - e_first));
    if (samples && nsamples < sample_capacity) {
      samples[4 * nsamples] = step;
      samples[4 * nsamples + 1] = u;
      samples[4 * nsamples + 2] = k;
      samples[4 * nsamples + 3] = u + k;
    }
    nsamples++;
    return LJGPU_OK;
  };
  auto kick = [&](double h) -> int {
    fa.dt = h;
    TRY(enqueue_force(ctx, fa, mixed, accum));
    stats->pair_entries += ctx->eval.e;
    return LJGPU_OK;
  };
  const double inv_box = box > 0.0 ? 1.0 / box : 0.0;
  // One drift; with `disp`, also the largest displacement since the last
  // build (synchronises).
  auto drift = [&](double* disp = nullptr) -> int {
    if (disp) *disp = 0.0;
    if (n == 0) return LJGPU_OK;
    unsigned long long* slot = nullptr;
    if (disp) {
      TRY(dev_reserve(ctx->md_red, sizeof(double) * 3 * kEnergyBlocks + 64, ctx->stream));
      slot = reinterpret_cast<unsigned long long*>(ctx->md_red.as<double>() + 3 * kEnergyBlocks);
      HIP_TRY(hipMemsetAsync(slot, 0, sizeof(unsigned long long), ctx->stream), "hipMemsetAsync");
    }
    k_drift<<<disp ? std::min(1024, grid_for(n)) : grid_for(n), kBlock, 0, ctx->stream>>>(
        n, wq, wp, dt, box, inv_box, disp ? ctx->md_qref.as<double>() : nullptr, slot);
    HIP_TRY(hipGetLastError(), "drift kernel");
    if (disp) {
      unsigned long long bits = 0;
      HIP_TRY(hipMemcpyAsync(&bits, slot, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpyAsync");
      HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      double d2 = 0.0;
      std::memcpy(&d2, &bits, sizeof(double));
      *disp = std::sqrt(d2);
    }
    return LJGPU_OK;
  };

  TRY(sample(0));
  if (steps > 0) TRY(kick(0.5 * dt));
  hipGraphExec_t seg_exec = nullptr;
  int32_t seg_len = 0;
  StepKey seg_key;
  auto drop_seg = [&]() {
    if (seg_exec) (void)hipGraphExecDestroy(seg_exec);
    seg_exec = nullptr;
    seg_len = 0;
  };
  struct SegGuard {
    std::function<void()> f;
    ~SegGuard() { f(); }
  } seg_guard{drop_seg};

  // rebuild_every = 0: rebuild whenever some particle has moved more than
  // skin/2 since the last build (checked after every drift).
  const bool adaptive = cfg->rebuild_every == 0;
  int32_t s = 1;
  while (s <= steps) {
    const bool rebuild_now = !adaptive && s % cfg->rebuild_every == 0;
    const bool sample_now = (cfg->sample_every > 0 && s % cfg->sample_every == 0) || s == steps;
    if (adaptive || rebuild_now || sample_now || !ctx->use_graphs) {
      double d = 0.0;
      TRY(drift(adaptive || rebuild_now ? &d : nullptr));
      if (adaptive) {
        if (2.0 * d > cfg->skin) {
          TRY(rebuild());
          d = 0.0;
        }
        max_disp = std::max(max_disp, d);
      } else if (rebuild_now) {
        max_disp = std::max(max_disp, d);
        if (2.0 * d > cfg->skin) stale++;
        TRY(rebuild());
      }
      if (sample_now) {
        TRY(kick(0.5 * dt));
        TRY(sample(s));
        if (s < steps) TRY(kick(0.5 * dt));
      } else {
        TRY(kick(dt));
      }
      s++;
      continue;
    }
    // Plain steps up to the next rebuild or sample run as one captured
    // graph of (drift, kick) pairs, re-captured when its arguments change.
    int32_t len = 0;
    while (s + len <= steps && len < kGraphChunk) {
      const int32_t t = s + len;
      if (t % cfg->rebuild_every == 0 || (cfg->sample_every > 0 && t % cfg->sample_every == 0) ||
          t == steps)
        break;
      len++;
    }
    if (len < kGraphMinSteps) {
      for (int32_t t = 0; t < len; t++) {
        TRY(drift());
        TRY(kick(dt));
      }
      s += len;
      continue;
    }
    fa.dt = dt;
    ListView klv;
    LaunchConfig klc;
    TRY(force_view(ctx, fa, &klv, &klc));
    const StepKey key = step_key(fa, klv, klc, mixed, accum, len, 1);
    if (!seg_exec || seg_len != len || !(key == seg_key)) {
      drop_seg();
      TRY(capture(ctx, [&]() {
        for (int32_t t = 0; t < len; t++) {
          TRY(drift());
          TRY(enqueue_force(ctx, fa, mixed, accum));
        }
        return LJGPU_OK;
      }, &seg_exec));
      seg_len = len;
      seg_key = key;
    }
    HIP_TRY(hipGraphLaunch(seg_exec, ctx->stream), "hipGraphLaunch");
    stats->pair_entries += static_cast<int64_t>(len) * ctx->eval.e;
    s += len;
  }

  TRY(give_back());
  HIP_TRY(hipEventRecord(e1, ctx->stream), "hipEventRecord");
  HIP_TRY(hipEventSynchronize(e1), "hipEventSynchronize");
  float total_ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&total_ms, e0, e1), "hipEventElapsedTime");
  stats->steps = steps;
  stats->rebuilds = rebuilds;
  stats->samples = nsamples;
  stats->max_abs_drift = max_drift;
  stats->max_displacement = max_disp;
  stats->stale_lists = stale;
  stats->ms_total = total_ms;
  stats->ms_rebuild = rebuild_ms;
  return LJGPU_OK;
}

int ljgpu_set_graphs(ljgpu_ctx* ctx, int enable) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (enable != 0 && enable != 1) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "enable must be 0 or 1");
  ctx->use_graphs = enable != 0;
  if (!ctx->use_graphs) drop_graph(ctx);
  return LJGPU_OK;
}

int ljgpu_set_chains(ljgpu_ctx* ctx, int chains) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (chains < 1 || chains > kMaxChains)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "chains must be 1 to 4");
  ctx->chains = chains;
  return LJGPU_OK;
}

int ljgpu_graph_stats(ljgpu_ctx* ctx, int64_t* captures, int32_t* cached) {
  if (!ctx || !captures || !cached) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  *captures = ctx->graph_captures;
  *cached = static_cast<int32_t>(ctx->graphs.size());
  return LJGPU_OK;
}

}  // extern "C"
