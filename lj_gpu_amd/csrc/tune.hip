// tune.hip — the autotuner: times the force evaluation over the installed
// list for a set of kernel configurations (lanes per particle, entries per
// batch, LDS tile size, persistent grid) on the caller's positions and keeps
// the fastest. No reference counterpart: the reference takes only its block
// size from the command line (cuda/force_cuda.cu:377-383).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "context.h"

using namespace ljgpu;

namespace {

struct Candidate {
  int lanes, unroll, tile, resident, chains;
};

// The configurations tried: the tiled kernel over lanes x unroll x tile
// size x launch chains, one persistent grid, the untiled wavefront-segment
// kernel and the thread-per-particle ELL kernel.
std::vector<Candidate> candidates() {
  std::vector<Candidate> c;
  for (int lanes : {4, 8, 16})
    for (int unroll : {2, 4})
      for (int tile : {16, 32, 64})
        for (int chains : {1, 2}) c.push_back({lanes, unroll, tile, 0, chains});
  c.push_back({8, 4, 32, 2, 1});
  c.push_back({8, 2, 0, 0, 1});
  c.push_back({16, 2, 0, 0, 1});
  c.push_back({1, 2, 0, 0, 1});
  return c;
}

constexpr int kTuneRepeats = 5;  // timed calls per candidate; the median is kept

}  // namespace

extern "C" {

int ljgpu_autotune(ljgpu_ctx* ctx, const double* q, double box, double cl2, int precision,
                   int accum, int32_t reps, ljgpu_tune_entry* best) {
  if (!ctx || !best) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  if (ctx->n > 0 && !q) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q");
  if (reinterpret_cast<uintptr_t>(q) & 15u)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q must be 16-byte aligned");
  if (precision != LJGPU_PREC_F64 && precision != LJGPU_PREC_MIXED)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "autotune: unknown precision");
  if (accum != LJGPU_ACCUM_F64 && accum != LJGPU_ACCUM_KAHAN_F32)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "autotune: unknown accumulator");
  if (reps < 1) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "autotune: reps must be >= 1");
  if (!ctx->rep_off.empty() && box != 0.0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "with replicas, pass box = 0");
  TRY(use_device(ctx));
  ctx->tune.clear();
  std::memset(best, 0, sizeof(*best));
  if (ctx->n == 0) return LJGPU_OK;
  // Scratch momenta: the caller's are not touched.
  DevMem scratch;
  TRY(dev_reserve(scratch, sizeof(double) * 4 * static_cast<size_t>(ctx->n), ctx->stream));
  struct Release {
    DevMem* m;
    ~Release() { dev_release(*m); }
  } release{&scratch};
  HIP_TRY(hipMemsetAsync(scratch.p, 0, sizeof(double) * 4 * static_cast<size_t>(ctx->n),
                         ctx->stream), "hipMemsetAsync");
  const int saved_lanes = ctx->lanes, saved_unroll = ctx->unroll, saved_chains = ctx->chains;
  const int32_t saved_tile = ctx->tile_particles, saved_res = ctx->tile_resident;
  auto restore = [&]() {
    ctx->lanes = saved_lanes;
    ctx->unroll = saved_unroll;
    ctx->chains = saved_chains;
    ctx->tile_particles = saved_tile;
    ctx->tile_resident = saved_res;
  };
  ForceArgs fa;
  fa.soa = false;
  fa.q[0] = q;
  fa.p[0] = scratch.as<double>();
  fa.dt = 1e-12;  // keeps the scratch momenta tiny
  fa.cl2 = cl2;
  fa.box = box;
  const bool mixed = precision == LJGPU_PREC_MIXED;
  float best_us = 0.0f;
  for (const Candidate& c : candidates()) {
    ctx->lanes = c.lanes;
    ctx->unroll = c.unroll;
    ctx->tile_particles = c.tile;
    ctx->tile_resident = c.resident;
    ctx->chains = c.chains;
    ljgpu_tune_entry e;
    std::memset(&e, 0, sizeof(e));
    e.lanes_per_particle = c.lanes;
    e.unroll = c.unroll;
    e.tile_particles = c.tile;
    e.resident_per_cu = c.resident;
    e.chains = c.chains;
    // One untimed call of the same shape (layouts, slot copies, graphs),
    // then kTuneRepeats timed calls of `reps` steps each.
    int rc = run_steps(ctx, fa, reps, mixed, accum);
    float per_call[kTuneRepeats] = {};
    for (int r = 0; r < kTuneRepeats && rc == LJGPU_OK; r++) {
      rc = ljgpu_timer_start(ctx);
      if (rc == LJGPU_OK) rc = run_steps(ctx, fa, reps, mixed, accum);
      if (rc == LJGPU_OK) rc = ljgpu_timer_stop(ctx, &per_call[r]);
    }
    if (rc == LJGPU_ERR_UNSUPPORTED) {
      clear_error();
      continue;  // not available for this list (e.g. replicas without tiles)
    }
    if (rc) {
      restore();
      return rc;
    }
    std::sort(per_call, per_call + kTuneRepeats);
    e.tile_used =
        ctx->tiles.blocks > 0 && c.lanes > 1 && !ctx->eval_half ? ctx->tiles.tile_particles : 0;
    e.us_per_step = 1000.0f * per_call[kTuneRepeats / 2] / static_cast<float>(reps);
    e.us_min = 1000.0f * per_call[0] / static_cast<float>(reps);
    ctx->tune.push_back(e);
    if (ctx->tune.size() == 1 || e.us_per_step < best_us) {
      best_us = e.us_per_step;
      *best = e;
    }
  }
  if (ctx->tune.empty()) {
    restore();
    return set_error(LJGPU_ERR_UNSUPPORTED, "autotune: no configuration runs on this list");
  }
  ctx->lanes = best->lanes_per_particle;
  ctx->unroll = best->unroll;
  ctx->tile_particles = best->tile_particles;
  ctx->tile_resident = best->resident_per_cu;
  ctx->chains = best->chains;
  const PosIn in = positions_of(fa);
  TRY(prepare_layouts(ctx, &in, box, cl2));
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_query_autotune(ljgpu_ctx* ctx, ljgpu_tune_entry* results, int32_t capacity,
                         int32_t* count) {
  if (!ctx || !count) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  *count = static_cast<int32_t>(ctx->tune.size());
  if (!results) return LJGPU_OK;
  const int32_t m = std::min(capacity, *count);
  for (int32_t i = 0; i < m; i++) results[i] = ctx->tune[i];
  return LJGPU_OK;
}

}  // extern "C"
