// api.hip — the C ABI of ljgpu (include/ljgpu.h): contexts, device memory,
// neighbour-list installation and the force entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "context.h"

using namespace ljgpu;

namespace ljgpu {

// Scratch of the allocation-free entry point ljgpu_energy_noalloc: every GPU
// holds its own copy of this array, at its own address, until the process
// ends. In host code the name is only a stand-in for the symbol, never a
// device address: host code uses the address hipGetSymbolAddress returns for
// the current device (noalloc_scratch_of), for kernel arguments and copies.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_addr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, which must be current: looked
// up on the first call for the device, then kept.
int noalloc_scratch_of(int device, double** out) {
  if (device < 0 || device >= kMaxScratchDevices)
    return set_error(LJGPU_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_addr[device]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_addr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_addr[device];
  return LJGPU_OK;
}

}  // namespace

int use_device(ljgpu_ctx* ctx) {
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  return LJGPU_OK;
}

// Forget the installed list; device buffers are kept for the next one.
void reset_list(ljgpu_ctx* ctx) {
  ctx->n = 0;
  ctx->kind = -1;
  ctx->n_entries = 0;
  ctx->max_np = ctx->min_np = 0;
  ctx->eval.n = 0;
  ctx->eval.e = 0;
  ctx->eval_half = false;
  ctx->half_mode = 0;
  ctx->ell_ready = false;
  ctx->ell_width = 0;
  ctx->tiles_for = 0;
  ctx->tiles_ordered = false;
  ctx->tiles.tile_particles = 0;
  ctx->tiles.blocks = 0;
  ctx->tiles.sliced_lanes = ctx->tiles.sliced_unroll = 0;
}

void release_all(ljgpu_ctx* ctx) {
  reset_list(ctx);
  dev_release(ctx->eval);
  dev_release(ctx->staging);
  dev_release(ctx->ell);
  dev_release(ctx->tiles);
  dev_release(ctx->work);
  dev_release(ctx->slot_q);
  dev_release(ctx->slot_p);
  for (DevMem* m : {&ctx->md_q, &ctx->md_p, &ctx->md_tmp, &ctx->md_perm, &ctx->md_perm2})
    dev_release(*m);
}

// Particles per tile for the current kernel configuration: the requested
// tile size, capped so that a tile's workgroup (tile * lanes threads) fits
// 1024 threads; 0 when tiling is off or the ELL kernel is selected.
int32_t effective_tile(const ljgpu_ctx* ctx) {
  if (ctx->tile_particles <= 0 || ctx->lanes <= 1) return 0;
  return std::min(ctx->tile_particles, 1024 / ctx->lanes);
}

// Derive the ELL copy of the evaluated list (thread-per-particle kernel).
int ensure_ell(ljgpu_ctx* ctx) {
  if (ctx->ell_ready) return LJGPU_OK;
  int32_t mn = 0, mx = 0;
  TRY(dev_minmax(ctx->stream, ctx->eval.nop.as<int32_t>(), ctx->eval.n, ctx->work, &mn, &mx));
  ctx->ell_width = std::max(mx, 1);
  TRY(dev_make_ell(ctx->stream, ctx->eval, ctx->ell_width, ctx->ell));
  ctx->ell_ready = true;
  return LJGPU_OK;
}

// Derive the LDS tiling of the evaluated full list for the current
// configuration. Where a block does not fit the LDS budget the tile size is
// halved; a list whose single rows do not fit keeps the untiled kernels.
// With the spatial tile order the blocks take the particles in the order of
// dev_tile_order, which needs positions: a list built on the device gets it
// at once, a list installed from the caller at its first force evaluation
// (index order until then). Systems without a periodic box stay in index
// order. Tiles derived with positions also take the cutoff classes of cl2
// (cut_partition) when the rows are bank-class ordered; a tiling derived
// without positions is derived again, once, when positions first come.
int ensure_tiles(ljgpu_ctx* ctx, const PosIn* q, double box, double cl2) {
  const int32_t want = ctx->eval_half ? 0 : effective_tile(ctx);
  const bool spatial = ctx->tile_order == LJGPU_TILE_ORDER_SPATIAL && want > 0;
  if (want == ctx->tiles_for && (ctx->tiles_ordered || !q || want == 0)) return LJGPU_OK;
  ctx->tiles.tile_particles = 0;
  ctx->tiles.blocks = 0;
  ctx->tiles.sliced_lanes = ctx->tiles.sliced_unroll = 0;
  bool ordered = false;
  if (spatial && q && ctx->eval.n > 0)
    TRY(dev_tile_order(ctx->stream, *q, ctx->eval.n, box, ctx->rep_off, ctx->rep_boxes, want,
                       ctx->work, &ordered));
  if (ordered) TRY(dev_length_sort(ctx->stream, ctx->eval, want, ctx->rep_off, ctx->work));
  const int32_t* member = ordered ? ctx->work.order.as<int32_t>() : nullptr;
  CutHint cut;
  if (q) {
    cut.q = *q;
    cut.box = box;
    cut.rep_box = ctx->rep_box_dev.as<double>();
    cut.cl2 = cl2;
  }
  for (int32_t tp = want; tp >= 1 && ctx->eval.n > 0; tp /= 2) {
    TRY(dev_make_tiles(ctx->stream, ctx->eval, tp, ctx->rep_off, member, ctx->bank_order,
                       q ? &cut : nullptr, ctx->work, ctx->tiles));
    if (ctx->tiles.tile_particles > 0) break;
  }
  // The slot-order copies that calls of kSlotMinSteps or more steps run over
  // are reserved with the tiles, so a force call never allocates.
  if (ctx->tiles.blocks > 0 && ctx->tiles.order == LJGPU_TILE_ORDER_SPATIAL &&
      ctx->tiles.staged > 0) {
    const size_t bytes = sizeof(double) * 4 * static_cast<size_t>(ctx->eval.n);
    TRY(dev_reserve(ctx->slot_q, bytes, ctx->stream));
    TRY(dev_reserve(ctx->slot_p, bytes, ctx->stream));
  }
  ctx->tiles_for = want;
  ctx->tiles_ordered = q != nullptr;
  return LJGPU_OK;
}

// The tiles' offsets sliced for the configured lanes and unroll (built once
// per tiling and configuration).
int ensure_sliced(ljgpu_ctx* ctx) {
  DevTiles& t = ctx->tiles;
  if (t.blocks <= 0 || (t.sliced_lanes == ctx->lanes && t.sliced_unroll == ctx->unroll))
    return LJGPU_OK;
  const int32_t* sptr = t.order == LJGPU_TILE_ORDER_SPATIAL ? t.slot_ptr.as<int32_t>()
                                                            : ctx->eval.ptr.as<int32_t>();
  return dev_slice_tiles(ctx->stream, sptr, ctx->lanes, ctx->unroll, ctx->work, t);
}

int prepare_layouts(ljgpu_ctx* ctx, const PosIn* q, double box, double cl2) {
  if (ctx->lanes == 1) return ensure_ell(ctx);
  TRY(ensure_tiles(ctx, q, box, cl2));
  return ensure_sliced(ctx);
}

PosIn positions_of(const ForceArgs& fa) {
  PosIn in;
  if (fa.soa) {
    in.x = fa.q[0];
    in.y = fa.q[1];
    in.z = fa.q[2];
    in.stride = 1;
  } else {
    in.x = fa.q[0];
    in.y = fa.q[0] + 1;
    in.z = fa.q[0] + 2;
    in.stride = fa.aos3 ? 3 : 4;
  }
  return in;
}

// Turn a validated device CSR (the caller's, or staged from the host) into
// the evaluated list and its layouts.
int derive_list(ljgpu_ctx* ctx, int32_t n, const int32_t* nop, const int32_t* ptr,
                const int32_t* list, int64_t e, int kind) {
  if (!ctx->rep_off.empty() && n != ctx->rep_off.back())
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "particle_number differs from the replicas' total");
  ctx->n = n;
  ctx->kind = kind;
  ctx->n_entries = e;
  if (kind == LJGPU_LIST_HALF)
    ctx->half_mode = ctx->half_mode_req == LJGPU_HALF_ATOMIC ? LJGPU_HALF_ATOMIC : LJGPU_HALF_EXPAND;
  else
    ctx->half_mode = 0;
  if (kind == LJGPU_LIST_HALF && ctx->half_mode == LJGPU_HALF_EXPAND) {
    TRY(dev_expand_half(ctx->stream, n, nop, ptr, list, ctx->work, ctx->eval));
    ctx->eval_half = false;
  } else {
    // Rows stored in order make every tile's entries one contiguous range.
    TRY(dev_copy_row_ordered(ctx->stream, n, nop, ptr, list, e, ctx->work, ctx->eval));
    ctx->eval_half = (kind == LJGPU_LIST_HALF);
  }
  if (!ctx->rep_off.empty()) {
    const int rc = dev_check_replicas(ctx->stream, ctx->eval, ctx->rep_off, ctx->work);
    if (rc) {
      reset_list(ctx);
      return rc;
    }
  }
  ctx->ell_ready = false;
  ctx->tiles_for = -1;
  ctx->tiles_ordered = false;
  TRY(prepare_layouts(ctx, nullptr, 0.0));
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

// Install a list validated on the host.
int install_host(ljgpu_ctx* ctx, int32_t n, const int32_t* nop, const int32_t* ptr,
                 const int32_t* list, int64_t e, int kind) {
  TRY(use_device(ctx));
  reset_list(ctx);
  int32_t mn = n > 0 ? INT32_MAX : 0, mx = 0;
  for (int32_t i = 0; i < n; i++) {
    mn = std::min(mn, nop[i]);
    mx = std::max(mx, nop[i]);
  }
  DevCSR& st = ctx->staging;
  TRY(dev_reserve(st.nop, sizeof(int32_t) * std::max(n, 1), ctx->stream));
  TRY(dev_reserve(st.ptr, sizeof(int32_t) * std::max(n, 1), ctx->stream));
  TRY(dev_reserve(st.list, sizeof(int32_t) * std::max<int64_t>(e, 1), ctx->stream));
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(st.nop.p, nop, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream),
            "hipMemcpyAsync (list)");
    HIP_TRY(hipMemcpyAsync(st.ptr.p, ptr, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream),
            "hipMemcpyAsync (list)");
  }
  if (e > 0)
    HIP_TRY(hipMemcpyAsync(st.list.p, list, sizeof(int32_t) * e, hipMemcpyHostToDevice, ctx->stream),
            "hipMemcpyAsync (list)");
  TRY(derive_list(ctx, n, st.nop.as<int32_t>(), st.ptr.as<int32_t>(), st.list.as<int32_t>(), e,
                  kind));
  ctx->min_np = mn;
  ctx->max_np = mx;
  return LJGPU_OK;
}

// Build the list from device positions (cell list on the device).
int build_list(ljgpu_ctx* ctx, const PosIn& q, int32_t n, double box, double search, int kind,
               bool sort_rows) {
  if (!(search > 0.0) || n < 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "build_list: bad argument");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "build_list: bad kind");
  if (box > 0.0 && box < 2.0 * search)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "build_list: periodic box shorter than twice the search length");
  if (n > 0 && (!q.x || !q.y || !q.z))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "build_list: null positions");
  if (!ctx->rep_off.empty()) {
    if (n != ctx->rep_off.back())
      return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                       "build_list: particle_number differs from the replicas' total");
    if (box != 0.0)
      return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                       "build_list: with replicas, pass box = 0 (each replica has its own)");
    for (double b : ctx->rep_boxes)
      if (b > 0.0 && b < 2.0 * search)
        return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                         "build_list: a replica box is shorter than twice the search length");
  }
  TRY(use_device(ctx));
  reset_list(ctx);
  const bool expand = kind == LJGPU_LIST_FULL || ctx->half_mode_req != LJGPU_HALF_ATOMIC;
  // A full list is the symmetric closure of the half list (the pair
  // predicate is exactly symmetric), so an expanded half list is built
  // directly as the full list.
  const int build_kind = expand ? LJGPU_LIST_FULL : LJGPU_LIST_HALF;
  if (ctx->rep_off.empty())
    TRY(dev_makepair(ctx->stream, q, n, box, search, build_kind, ctx->work, ctx->eval,
                     sort_rows));
  else
    TRY(dev_makepair_batch(ctx->stream, q, ctx->rep_off, ctx->rep_boxes, search, build_kind,
                           ctx->work, ctx->eval));
  ctx->n = n;
  ctx->kind = kind;
  ctx->eval_half = !expand;
  ctx->half_mode = kind == LJGPU_LIST_HALF ? (expand ? LJGPU_HALF_EXPAND : LJGPU_HALF_ATOMIC) : 0;
  ctx->n_entries = (kind == LJGPU_LIST_HALF && expand) ? ctx->eval.e / 2 : ctx->eval.e;
  if (kind == LJGPU_LIST_HALF && expand)
    TRY(dev_half_row_stats(ctx->stream, ctx->eval, ctx->work, &ctx->min_np, &ctx->max_np));
  else
    TRY(dev_minmax(ctx->stream, ctx->eval.nop.as<int32_t>(), n, ctx->work, &ctx->min_np,
                   &ctx->max_np));
  ctx->ell_ready = false;
  ctx->tiles_for = -1;
  ctx->tiles_ordered = false;
  TRY(prepare_layouts(ctx, &q, box));
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
bool aligned8(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 7u) == 0; }

int check_force_args(ljgpu_ctx* ctx, double dt, double cl2, int32_t steps, double box) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  if (!ctx->rep_off.empty() && box != 0.0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "with replicas, pass box = 0 (each replica has its own box)");
  if (steps < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "steps < 0");
  if (!(cl2 >= 0.0)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "cl2 must be >= 0");
  if (dt != dt) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "dt is NaN");
  return LJGPU_OK;
}

// Kernel arguments of one force evaluation over the installed list.
int force_view(ljgpu_ctx* ctx, const ForceArgs& fa, ListView* lv, LaunchConfig* lc) {
  const PosIn in = positions_of(fa);
  TRY(prepare_layouts(ctx, &in, fa.box, fa.cl2));
  *lv = ListView();
  lv->n = ctx->n;
  lv->nop = ctx->eval.nop.as<int32_t>();
  lv->ptr = ctx->eval.ptr.as<int32_t>();
  lv->n_entries = ctx->eval.e;
  *lc = LaunchConfig();
  lc->lanes = ctx->lanes;
  lc->unroll = ctx->unroll;
  lc->block = ctx->block;
  lc->tile_offsets_lds = ctx->tile_offsets_lds;
  lc->tile_resident = ctx->tile_resident;
  lc->cus = ctx->cus;
  lc->half_atomic = ctx->eval_half;
  lc->clamp = false;
  if (!ctx->rep_off.empty()) lv->replica_box = ctx->rep_box_dev.as<double>();
  if (ctx->lanes == 1) {
    lv->list = ctx->ell.as<int32_t>();
    lv->ell_width = ctx->ell_width;
  } else {
    lv->list = ctx->eval.list.as<int32_t>();
    if (ctx->tiles.blocks > 0 && !ctx->eval_half) tile_fields(ctx, lv);
  }
  return LJGPU_OK;
}

// The installed tiling's fields of a ListView (as they stand: nothing is
// derived or rebuilt).
void tile_fields(const ljgpu_ctx* ctx, ListView* lv) {
  lv->tile_blocks = ctx->tiles.blocks;
  lv->tile_particles = ctx->tiles.tile_particles;
  lv->tile_lds_bytes = ctx->tiles.lds_bytes;
  lv->tile_max_stage = ctx->tiles.max_stage;
  lv->tile_desc = ctx->tiles.desc.as<int32_t>();
  lv->tile_stage = ctx->tiles.stage.as<int32_t>();
  lv->tile_local = ctx->tiles.local.as<uint16_t>();
  if (ctx->tiles.sliced_lanes > 0) {
    lv->tile_sliced = ctx->tiles.sliced.as<uint16_t>();
    lv->tile_wave_off = ctx->tiles.wave_off.as<int32_t>();
    lv->tile_waves_per_block = ctx->tiles.waves_per_block;
    lv->tile_sliced_lanes = ctx->tiles.sliced_lanes;
    lv->tile_sliced_unroll = ctx->tiles.sliced_unroll;
  }
  if (ctx->tiles.order == LJGPU_TILE_ORDER_SPATIAL) {
    lv->tile_member = ctx->tiles.member.as<int32_t>();
    lv->tile_slot_ptr = ctx->tiles.slot_ptr.as<int32_t>();
    if (ctx->tiles.staged > 0) lv->tile_stage_slot = ctx->tiles.stage_slot.as<int32_t>();
  } else {
    lv->tile_slot_ptr = ctx->eval.ptr.as<int32_t>();  // slot = particle
  }
}

int launch_one_on(hipStream_t stream, const ListView& lv, const LaunchConfig& lc,
                  const ForceArgs& fa, bool mixed, int accum) {
  if (fa.aos3)
    return !mixed ? launch_force_aos3_f64(stream, lv, fa, lc)
           : accum == LJGPU_ACCUM_KAHAN_F32 ? launch_force_aos3_mixed_kahan(stream, lv, fa, lc)
                                            : launch_force_aos3_mixed_f64acc(stream, lv, fa, lc);
  return !mixed ? launch_force_f64(stream, lv, fa, lc)
         : accum == LJGPU_ACCUM_KAHAN_F32 ? launch_force_mixed_kahan(stream, lv, fa, lc)
                                          : launch_force_mixed_f64acc(stream, lv, fa, lc);
}

int launch_one(ljgpu_ctx* ctx, const ListView& lv, const LaunchConfig& lc, const ForceArgs& fa,
               bool mixed, int accum) {
  return launch_one_on(ctx->stream, lv, lc, fa, mixed, accum);
}

int enqueue_force(ljgpu_ctx* ctx, const ForceArgs& fa, bool mixed, int accum) {
  if (ctx->n == 0) return LJGPU_OK;
  ListView lv;
  LaunchConfig lc;
  TRY(force_view(ctx, fa, &lv, &lc));
  return launch_one(ctx, lv, lc, fa, mixed, accum);
}

void drop_graph(ljgpu_ctx* ctx) {
  // Replays already enqueued finish before their graphs are destroyed.
  if (!ctx->graphs.empty() && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (GraphEntry& g : ctx->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
}

int capture(ljgpu_ctx* ctx, const std::function<int()>& body, hipGraphExec_t* out) {
  *out = nullptr;
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
          "hipStreamBeginCapture");
  const int rc = body();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  const hipError_t e2 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) return hip_fail(e2, "hipGraphInstantiate");
  return LJGPU_OK;
}

StepKey step_key(const ForceArgs& fa, const ListView& lv, const LaunchConfig& lc, bool mixed,
                 int accum, int32_t steps, int32_t chains) {
  StepKey k;
  k.add(fa.soa);
  k.add(fa.aos3);
  for (int c = 0; c < 3; c++) {
    k.add(fa.q[c]);
    k.add(fa.p[c]);
  }
  k.add(fa.dt);
  k.add(fa.cl2);
  k.add(fa.box);
  k.add(lv.n);
  k.add(lv.nop);
  k.add(lv.ptr);
  k.add(lv.list);
  k.add(lv.n_entries);
  k.add(lv.ell_width);
  k.add(lv.tile_blocks);
  k.add(lv.tile_first);
  k.add(lv.tile_particles);
  k.add(lv.tile_lds_bytes);
  k.add(lv.tile_max_stage);
  k.add(lv.tile_desc);
  k.add(lv.tile_stage);
  k.add(lv.tile_local);
  k.add(lv.tile_member);
  k.add(lv.tile_slot_ptr);
  k.add(lv.tile_stage_slot);
  k.add(lv.tile_sliced);
  k.add(lv.tile_wave_off);
  k.add(lv.tile_waves_per_block);
  k.add(lv.tile_sliced_lanes);
  k.add(lv.tile_sliced_unroll);
  k.add(lv.replica_box);
  k.add(lc.lanes);
  k.add(lc.unroll);
  k.add(lc.block);
  k.add(lc.tile_offsets_lds);
  k.add(lc.tile_resident);
  k.add(lc.cus);
  k.add(lc.half_atomic);
  k.add(lc.clamp);
  k.add(mixed);
  k.add(accum);
  k.add(steps);
  k.add(chains);
  return k;
}

int cached_graph(ljgpu_ctx* ctx, const StepKey& key, const std::function<int()>& body,
                 hipGraphExec_t* out) {
  ctx->graph_clock++;
  for (GraphEntry& g : ctx->graphs) {
    if (g.key == key) {
      g.last_use = ctx->graph_clock;
      *out = g.exec;
      return LJGPU_OK;
    }
  }
  hipGraphExec_t exec = nullptr;
  TRY(capture(ctx, body, &exec));
  ctx->graph_captures++;
  if (static_cast<int>(ctx->graphs.size()) >= kGraphCacheSize) {
    auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                [](const GraphEntry& a, const GraphEntry& b) {
                                  return a.last_use < b.last_use;
                                });
    // Work already enqueued from the evicted graph must finish before it
    // is destroyed.
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    (void)hipGraphExecDestroy(lru->exec);
    ctx->graphs.erase(lru);
  }
  GraphEntry g;
  g.key = key;
  g.exec = exec;
  g.last_use = ctx->graph_clock;
  ctx->graphs.push_back(g);
  *out = exec;
  return LJGPU_OK;
}

namespace {

// Copies between the caller's arrays and slot order: out[s] = in[member[s]]
// (gather) or out[member[s]] = in[s] (scatter), 4-double AoS rows or three
// SoA arrays.
template <bool GATHER>
__global__ void k_slot_copy_aos(int32_t n, const int32_t* __restrict__ member,
                                const double4* __restrict__ in, double4* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = member[t];
    if (GATHER) out[t] = in[i];
    else out[i] = in[t];
  }
}

// q and p gathered into slot order in one pass (out[s] = in[member[s]]).
__global__ void k_slot_gather_qp(int32_t n, const int32_t* __restrict__ member,
                                 const double4* __restrict__ q, const double4* __restrict__ p,
                                 double4* __restrict__ oq, double4* __restrict__ op) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = member[t];
    oq[t] = q[i];
    op[t] = p[i];
  }
}

template <bool GATHER>
__global__ void k_slot_copy_soa(int32_t n, const int32_t* __restrict__ member, const double* x,
                                const double* y, const double* z, double* ox, double* oy,
                                double* oz) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = member[t];
    const int64_t a = GATHER ? i : t, b = GATHER ? t : i;
    ox[b] = x[a];
    oy[b] = y[a];
    oz[b] = z[a];
  }
}

// Packed double[N][3] rows to and from 4-double slot rows (pad 0).
template <bool GATHER>
__global__ void k_slot_copy_aos3(int32_t n, const int32_t* __restrict__ member,
                                 const double* __restrict__ in, double* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = member[t];
    if (GATHER) {
      const double* r = in + 3 * i;
      reinterpret_cast<double4*>(out)[t] = make_double4(r[0], r[1], r[2], 0.0);
    } else {
      const double4 v = reinterpret_cast<const double4*>(in)[t];
      double* r = out + 3 * i;
      r[0] = v.x;
      r[1] = v.y;
      r[2] = v.z;
    }
  }
}

int slot_grid(int32_t n) { return static_cast<int>(std::min<int64_t>((n + 255) / 256, 4096)); }

// Positions and momenta of `fa` copied into the context's slot-order
// buffers; *out addresses the copies.
int slot_gather(ljgpu_ctx* ctx, const ForceArgs& fa, ForceArgs* out) {
  const int32_t n = ctx->n;
  const int32_t* mem = ctx->tiles.member.as<int32_t>();
  const size_t bytes = sizeof(double) * 4 * static_cast<size_t>(n);
  TRY(dev_reserve(ctx->slot_q, bytes, ctx->stream));
  TRY(dev_reserve(ctx->slot_p, bytes, ctx->stream));
  *out = fa;
  double* q = ctx->slot_q.as<double>();
  double* p = ctx->slot_p.as<double>();
  if (fa.aos3) {
    // The slot copies are 4-double rows, so the loop runs the AoS kernels.
    k_slot_copy_aos3<true><<<slot_grid(n), 256, 0, ctx->stream>>>(n, mem, fa.q[0], q);
    k_slot_copy_aos3<true><<<slot_grid(n), 256, 0, ctx->stream>>>(n, mem, fa.p[0], p);
    out->aos3 = false;
    out->q[0] = q;
    out->p[0] = p;
  } else if (!fa.soa) {
    k_slot_gather_qp<<<slot_grid(n), 256, 0, ctx->stream>>>(
        n, mem, reinterpret_cast<const double4*>(fa.q[0]), reinterpret_cast<const double4*>(fa.p[0]),
        reinterpret_cast<double4*>(q), reinterpret_cast<double4*>(p));
    out->q[0] = q;
    out->p[0] = p;
  } else {
    k_slot_copy_soa<true><<<slot_grid(n), 256, 0, ctx->stream>>>(n, mem, fa.q[0], fa.q[1], fa.q[2],
                                                                 q, q + n, q + 2 * n);
    k_slot_copy_soa<true><<<slot_grid(n), 256, 0, ctx->stream>>>(n, mem, fa.p[0], fa.p[1], fa.p[2],
                                                                 p, p + n, p + 2 * n);
    for (int c = 0; c < 3; c++) {
      out->q[c] = q + static_cast<int64_t>(c) * n;
      out->p[c] = p + static_cast<int64_t>(c) * n;
    }
  }
  HIP_TRY(hipGetLastError(), "slot gather kernels");
  return LJGPU_OK;
}

// Momenta in slot order (slot) back to the caller's arrays (fa).
int slot_scatter(ljgpu_ctx* ctx, const ForceArgs& slot, const ForceArgs& fa) {
  const int32_t n = ctx->n;
  const int32_t* mem = ctx->tiles.member.as<int32_t>();
  if (fa.aos3)
    k_slot_copy_aos3<false><<<slot_grid(n), 256, 0, ctx->stream>>>(n, mem, slot.p[0], fa.p[0]);
  else if (!fa.soa)
    k_slot_copy_aos<false><<<slot_grid(n), 256, 0, ctx->stream>>>(
        n, mem, reinterpret_cast<const double4*>(slot.p[0]), reinterpret_cast<double4*>(fa.p[0]));
  else
    k_slot_copy_soa<false><<<slot_grid(n), 256, 0, ctx->stream>>>(
        n, mem, slot.p[0], slot.p[1], slot.p[2], fa.p[0], fa.p[1], fa.p[2]);
  HIP_TRY(hipGetLastError(), "slot scatter kernel");
  return LJGPU_OK;
}

}  // namespace

int run_steps(ljgpu_ctx* ctx, const ForceArgs& caller, int32_t steps, bool mixed, int accum) {
  TRY(use_device(ctx));
  if (ctx->n == 0 || steps == 0) return LJGPU_OK;
  ListView lv;
  LaunchConfig lc;
  TRY(force_view(ctx, caller, &lv, &lc));
  // With spatially ordered tiles, a call of many steps runs over copies of
  // q and p in slot order (q does not change between its steps): the rows'
  // and the staging's accesses are then contiguous runs. The arithmetic is
  // the same, so the momenta equal those of per-step evaluations bit for bit.
  ForceArgs fa = caller;
  const bool slots = steps >= kSlotMinSteps && lv.tile_blocks > 0 && lv.tile_stage_slot;
  if (slots) {
    TRY(slot_gather(ctx, caller, &fa));
    lv.tile_member = nullptr;
    lv.tile_stage = lv.tile_stage_slot;
  }
  // Over enough tiles, the tiles in `chains` contiguous parts, each a chain
  // of launches on its own stream, forked from and joined back into the
  // context stream (see kMaxChains).
  const int chains =
      (lv.tile_blocks >= kMinChainTiles * ctx->chains && steps >= kGraphMinSteps) ? ctx->chains : 1;
  if (chains > 1) {
    HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream), "hipEventRecord");
    ListView part[kMaxChains];
    for (int c = 0; c < chains; c++) {
      part[c] = lv;
      const int32_t a = static_cast<int32_t>(int64_t{lv.tile_blocks} * c / chains);
      const int32_t b = static_cast<int32_t>(int64_t{lv.tile_blocks} * (c + 1) / chains);
      part[c].tile_first = a;
      part[c].tile_blocks = b - a;
      if (c > 0)
        HIP_TRY(hipStreamWaitEvent(ctx->chain_stream[c - 1], ctx->ev_fork, 0),
                "hipStreamWaitEvent");
    }
    for (int32_t s = 0; s < steps; s++)
      for (int c = 0; c < chains; c++)
        TRY(launch_one_on(c == 0 ? ctx->stream : ctx->chain_stream[c - 1], part[c], lc, fa, mixed,
                          accum));
    for (int c = 1; c < chains; c++) {
      HIP_TRY(hipEventRecord(ctx->ev_join[c - 1], ctx->chain_stream[c - 1]), "hipEventRecord");
      HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join[c - 1], 0), "hipStreamWaitEvent");
    }
  } else {
    int32_t done = 0;
    if (ctx->use_graphs && steps >= kGraphMinSteps) {
      // Repeated evaluations run as replays of captured graphs, longest
      // chunk first (kGraphChunks); each graph is captured once per argument
      // set.
      for (const int32_t chunk : kGraphChunks) {
        if (steps - done < chunk) continue;
        const StepKey key = step_key(fa, lv, lc, mixed, accum, chunk, 1);
        hipGraphExec_t exec = nullptr;
        TRY(cached_graph(ctx, key, [&]() {
          for (int32_t s = 0; s < chunk; s++) TRY(launch_one(ctx, lv, lc, fa, mixed, accum));
          return LJGPU_OK;
        }, &exec));
        for (; done + chunk <= steps; done += chunk)
          HIP_TRY(hipGraphLaunch(exec, ctx->stream), "hipGraphLaunch");
      }
    }
    for (; done < steps; done++) TRY(launch_one(ctx, lv, lc, fa, mixed, accum));
  }
  if (slots) TRY(slot_scatter(ctx, fa, caller));
  return LJGPU_OK;
}

}  // namespace ljgpu

extern "C" {

int ljgpu_device_count(int* count) {
  if (!count) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return set_error(LJGPU_ERR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = c;
  return LJGPU_OK;
}

int ljgpu_create(int device, ljgpu_ctx** out) {
  if (!out) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return set_error(LJGPU_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "device index out of range");
  ljgpu_ctx* ctx = new (std::nothrow) ljgpu_ctx();
  if (!ctx) return set_error(LJGPU_ERR_OUT_OF_MEMORY, "context allocation");
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev_start);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev_stop);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
  for (int c = 0; c + 1 < kMaxChains && e == hipSuccess; c++) {
    e = hipStreamCreateWithFlags(&ctx->chain_stream[c], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join[c], hipEventDisableTiming);
  }
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    ljgpu_destroy(ctx);
    return hip_fail(e, "context creation");
  }
  const int rc = noalloc_scratch_of(device, &ctx->noalloc_scratch);
  if (rc) {
    ljgpu_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return LJGPU_OK;
}

int ljgpu_destroy(ljgpu_ctx* ctx) {
  if (!ctx) return LJGPU_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  release_all(ctx);
  dev_release(ctx->rep_box_dev);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  for (int c = 0; c + 1 < kMaxChains; c++) {
    if (ctx->ev_join[c]) (void)hipEventDestroy(ctx->ev_join[c]);
    if (ctx->chain_stream[c]) (void)hipStreamSynchronize(ctx->chain_stream[c]);
    if (ctx->chain_stream[c]) (void)hipStreamDestroy(ctx->chain_stream[c]);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LJGPU_OK;
}

int ljgpu_sync(ljgpu_ctx* ctx) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_malloc(ljgpu_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipMalloc(out, std::max<size_t>(bytes, 1)), "hipMalloc");
  return LJGPU_OK;
}

int ljgpu_free(ljgpu_ctx* ctx, void* ptr) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (!ptr) return LJGPU_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  HIP_TRY(hipFree(ptr), "hipFree");
  return LJGPU_OK;
}

static int copy_sync(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (!ctx || (bytes && (!dst || !src))) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (!bytes) return LJGPU_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream), "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_memcpy_htod(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}
int ljgpu_memcpy_dtoh(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}
int ljgpu_memcpy_dtod(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}

int ljgpu_memset(ljgpu_ctx* ctx, void* dst, int value, size_t bytes) {
  if (!ctx || (bytes && !dst)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (!bytes) return LJGPU_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream), "hipMemsetAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_timer_start(ljgpu_ctx* ctx) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream), "hipEventRecord");
  ctx->timer_running = true;
  return LJGPU_OK;
}

int ljgpu_timer_stop(ljgpu_ctx* ctx, float* ms) {
  if (!ctx || !ms) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (!ctx->timer_running) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "timer not started");
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream), "hipEventRecord");
  HIP_TRY(hipEventSynchronize(ctx->ev_stop), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop), "hipEventElapsedTime");
  ctx->timer_running = false;
  return LJGPU_OK;
}

int ljgpu_set_tiling(ljgpu_ctx* ctx, int32_t tile_particles, int stage_offsets,
                     int32_t resident_per_cu) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (tile_particles < 0 || tile_particles > 1024)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "tile_particles must be in [0, 1024]");
  if (stage_offsets != 0 && stage_offsets != 1)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "stage_offsets must be 0 or 1");
  if (resident_per_cu < 0 || resident_per_cu > 64)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "resident_per_cu must be in [0, 64]");
  ctx->tile_particles = tile_particles;
  ctx->tile_offsets_lds = stage_offsets != 0;
  ctx->tile_resident = resident_per_cu;
  return LJGPU_OK;
}

int ljgpu_set_tile_order(ljgpu_ctx* ctx, int order) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (order != LJGPU_TILE_ORDER_INDEX && order != LJGPU_TILE_ORDER_SPATIAL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown tile order");
  if (order != ctx->tile_order) {
    ctx->tile_order = order;
    ctx->tiles_for = -1;  // re-derived at the next use
    ctx->tiles_ordered = false;
  }
  return LJGPU_OK;
}

int ljgpu_set_half_mode(ljgpu_ctx* ctx, int half_mode) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (half_mode != LJGPU_HALF_AUTO && half_mode != LJGPU_HALF_ATOMIC && half_mode != LJGPU_HALF_EXPAND)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown half mode");
  ctx->half_mode_req = half_mode;
  return LJGPU_OK;
}

int ljgpu_set_kernel_config(ljgpu_ctx* ctx, int lanes, int block, int unroll) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 &&
      lanes != 32 && lanes != 64)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "lanes_per_particle must be 1,2,4,8,16,32 or 64");
  if (block != 0 && (block < 64 || block > 1024 || block % 64 != 0))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "block size must be a multiple of 64 in [64, 1024]");
  if (unroll != 0 && unroll != 1 && unroll != 2 && unroll != 4)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unroll must be 1, 2 or 4");
  if (lanes) ctx->lanes = lanes;
  if (block) ctx->block = block;
  if (unroll) ctx->unroll = unroll;
  return LJGPU_OK;
}

int ljgpu_set_list_csr(ljgpu_ctx* ctx, int32_t n, const int32_t* nop, const int32_t* ptr,
                       const int32_t* list, int64_t n_entries, int kind, int mem) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (mem != LJGPU_MEM_HOST && mem != LJGPU_MEM_DEVICE)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "mem must be LJGPU_MEM_HOST or LJGPU_MEM_DEVICE");
  if (n < 0 || n_entries < 0 || n_entries > INT32_MAX)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "particle_number or n_entries out of range");
  if (n > 0 && (!nop || !ptr)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null list array");
  if (n_entries > 0 && !list) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null sorted_list");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "kind must be LJGPU_LIST_HALF or LJGPU_LIST_FULL");
  if (!ctx->rep_off.empty() && n != ctx->rep_off.back())
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "particle_number differs from the replicas' total");
  if (mem == LJGPU_MEM_HOST) {
    int rc = validate_csr(n, nop, ptr, list, n_entries, kind);
    if (rc) return rc;
    return install_host(ctx, n, nop, ptr, list, n_entries, kind);
  }
  // Device input: validated, measured and derived on the device; the
  // caller's buffers are only read.
  TRY(use_device(ctx));
  TRY(dev_validate(ctx->stream, n, nop, ptr, list, n_entries, kind, ctx->work));
  int32_t mn = 0, mx = 0;
  TRY(dev_minmax(ctx->stream, nop, n, ctx->work, &mn, &mx));
  reset_list(ctx);
  TRY(derive_list(ctx, n, nop, ptr, list, n_entries, kind));
  ctx->min_np = mn;
  ctx->max_np = mx;
  return LJGPU_OK;
}

int ljgpu_set_list_transposed(ljgpu_ctx* ctx, int32_t n, const int32_t* nop,
                              const int32_t* transposed, int32_t width, int kind, int mem) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (n < 0 || width < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "negative size");
  if (mem != LJGPU_MEM_HOST && mem != LJGPU_MEM_DEVICE)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "mem must be LJGPU_MEM_HOST or LJGPU_MEM_DEVICE");
  const size_t cells = static_cast<size_t>(n) * width;
  if (n > 0 && (!nop || (cells && !transposed)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null list array");
  std::vector<int32_t> h_nop(n), h_ell(cells);
  if (mem == LJGPU_MEM_DEVICE) {
    int rc = use_device(ctx);
    if (rc) return rc;
    if (n > 0) HIP_TRY(hipMemcpy(h_nop.data(), nop, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    if (cells) HIP_TRY(hipMemcpy(h_ell.data(), transposed, sizeof(int32_t) * cells, hipMemcpyDeviceToHost), "hipMemcpy");
  } else {
    if (n > 0) std::memcpy(h_nop.data(), nop, sizeof(int32_t) * n);
    if (cells) std::memcpy(h_ell.data(), transposed, sizeof(int32_t) * cells);
  }
  std::vector<int32_t> h_ptr(n);
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    if (h_nop[i] < 0 || h_nop[i] > width)
      return set_error(LJGPU_ERR_BAD_LIST, "number_of_partners[" + std::to_string(i) + "] exceeds the ELL width");
    h_ptr[i] = static_cast<int32_t>(pos);
    pos += h_nop[i];
  }
  if (pos > INT32_MAX) return set_error(LJGPU_ERR_CAPACITY, "list too long");
  std::vector<int32_t> h_list(static_cast<size_t>(pos));
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < h_nop[i]; k++) h_list[h_ptr[i] + k] = h_ell[i + static_cast<size_t>(k) * n];
  int rc = validate_csr(n, h_nop.data(), h_ptr.data(), h_list.data(), pos, kind);
  if (rc) return rc;
  return install_host(ctx, n, h_nop.data(), h_ptr.data(), h_list.data(), pos, kind);
}

int ljgpu_get_list_info(ljgpu_ctx* ctx, ljgpu_list_info* info) {
  if (!ctx || !info) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  info->particle_number = ctx->n;
  info->kind = ctx->kind;
  info->n_entries = ctx->n_entries;
  info->n_pairs = ctx->kind == LJGPU_LIST_HALF ? ctx->n_entries : ctx->n_entries / 2;
  info->max_partners = ctx->max_np;
  info->min_partners = ctx->min_np;
  info->half_mode = ctx->half_mode;
  info->lanes_per_particle = ctx->lanes;
  info->unroll = ctx->unroll;
  info->block_size = ctx->block;
  const bool tiled = ctx->lanes > 1 && ctx->tiles.blocks > 0 && !ctx->eval_half &&
                     ctx->tiles_for == effective_tile(ctx);
  info->tile_particles = tiled ? ctx->tiles.tile_particles : 0;
  info->tile_blocks = tiled ? ctx->tiles.blocks : 0;
  info->tile_max_stage = tiled ? ctx->tiles.max_stage : 0;
  return LJGPU_OK;
}

int ljgpu_get_list_csr(ljgpu_ctx* ctx, int32_t* nop, int32_t* ptr, int32_t* list, int64_t capacity) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  const DevCSR& c = ctx->eval;
  if (c.e > capacity) return set_error(LJGPU_ERR_CAPACITY, "capacity too small");
  if (c.n > 0 && (!nop || !ptr)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null output");
  if (c.e > 0 && !list) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null output");
  TRY(use_device(ctx));
  if (c.n > 0) {
    HIP_TRY(hipMemcpyAsync(nop, c.nop.p, sizeof(int32_t) * c.n, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(ptr, c.ptr.p, sizeof(int32_t) * c.n, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpyAsync");
  }
  if (c.e > 0)
    HIP_TRY(hipMemcpyAsync(list, c.list.p, sizeof(int32_t) * c.e, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_get_tile_order(ljgpu_ctx* ctx, int* order, int32_t* member, int64_t capacity) {
  if (!ctx || !order) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  TRY(use_device(ctx));
  TRY(prepare_layouts(ctx));
  const DevTiles& t = ctx->tiles;
  const bool tiled = ctx->lanes > 1 && t.blocks > 0 && !ctx->eval_half;
  *order = tiled ? t.order : -1;
  if (!tiled || !member) return LJGPU_OK;
  if (capacity < ctx->eval.n) return set_error(LJGPU_ERR_CAPACITY, "get_tile_order: output too small");
  if (t.order == LJGPU_TILE_ORDER_INDEX) {
    for (int32_t s = 0; s < ctx->eval.n; s++) member[s] = s;
    return LJGPU_OK;
  }
  HIP_TRY(hipMemcpyAsync(member, t.member.p, sizeof(int32_t) * ctx->eval.n, hipMemcpyDeviceToHost,
                         ctx->stream), "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return LJGPU_OK;
}

int ljgpu_get_tiles(ljgpu_ctx* ctx, int32_t* desc, int64_t desc_capacity, int32_t* stage,
                    int64_t stage_capacity, uint16_t* local, int64_t local_capacity,
                    int32_t* n_blocks, int64_t* n_stage) {
  if (!ctx || !n_blocks || !n_stage) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  TRY(use_device(ctx));
  TRY(prepare_layouts(ctx));
  const DevTiles& t = ctx->tiles;
  const bool tiled = ctx->lanes > 1 && t.blocks > 0 && !ctx->eval_half;
  *n_blocks = tiled ? t.blocks : 0;
  *n_stage = tiled ? t.staged : 0;
  if (!tiled || (!desc && !stage && !local)) return LJGPU_OK;
  if (!desc || !stage || !local || desc_capacity < 8 * static_cast<int64_t>(t.blocks) ||
      stage_capacity < t.staged || local_capacity < ctx->eval.e)
    return set_error(LJGPU_ERR_CAPACITY, "get_tiles: output too small");
  HIP_TRY(hipMemcpyAsync(desc, t.desc.p, sizeof(int32_t) * 8 * t.blocks, hipMemcpyDeviceToHost,
                         ctx->stream), "hipMemcpyAsync");
  if (ctx->eval.e > 0)
    HIP_TRY(hipMemcpyAsync(local, t.local.p, sizeof(uint16_t) * ctx->eval.e, hipMemcpyDeviceToHost,
                           ctx->stream), "hipMemcpyAsync");
  if (t.stage_stride == 0) {
    if (t.staged > 0)
      HIP_TRY(hipMemcpyAsync(stage, t.stage.p, sizeof(int32_t) * t.staged, hipMemcpyDeviceToHost,
                             ctx->stream), "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LJGPU_OK;
  }
  // One-pass build: each block's staging area sits at a fixed stride on the
  // device; pack them back to back, as ljgpu_make_tiles lays them out.
  std::vector<int32_t> sparse(static_cast<size_t>(t.blocks) * t.stage_stride);
  HIP_TRY(hipMemcpyAsync(sparse.data(), t.stage.p, sizeof(int32_t) * sparse.size(),
                         hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  int64_t pos = 0;
  for (int32_t b = 0; b < t.blocks; b++) {
    const int32_t ns = desc[8 * b + 3];
    std::memcpy(stage + pos, sparse.data() + static_cast<size_t>(b) * t.stage_stride,
                sizeof(int32_t) * ns);
    desc[8 * b + 2] = static_cast<int32_t>(pos);
    pos += ns;
  }
  return LJGPU_OK;
}

int ljgpu_set_replicas(ljgpu_ctx* ctx, int32_t n_replicas, const int32_t* offsets,
                       const double* boxes) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (n_replicas < 0 || (n_replicas > 0 && (!offsets || !boxes)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "set_replicas: bad argument");
  if (n_replicas > 0 && offsets[0] != 0)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "set_replicas: offsets[0] must be 0");
  for (int32_t r = 0; r < n_replicas; r++) {
    if (offsets[r + 1] < offsets[r])
      return set_error(LJGPU_ERR_INVALID_ARGUMENT, "set_replicas: offsets must not decrease");
    if (!(boxes[r] == boxes[r]))
      return set_error(LJGPU_ERR_INVALID_ARGUMENT, "set_replicas: box is NaN");
  }
  TRY(use_device(ctx));
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  drop_graph(ctx);
  reset_list(ctx);
  ctx->rep_off.clear();
  ctx->rep_boxes.clear();
  if (n_replicas == 0) return LJGPU_OK;
  ctx->rep_off.assign(offsets, offsets + n_replicas + 1);
  ctx->rep_boxes.assign(boxes, boxes + n_replicas);
  std::vector<double> bb(2 * static_cast<size_t>(n_replicas));
  for (int32_t r = 0; r < n_replicas; r++) {
    bb[2 * r] = boxes[r] > 0.0 ? boxes[r] : 0.0;
    bb[2 * r + 1] = boxes[r] > 0.0 ? 1.0 / boxes[r] : 0.0;
  }
  TRY(dev_reserve(ctx->rep_box_dev, sizeof(double) * bb.size(), ctx->stream));
  HIP_TRY(hipMemcpy(ctx->rep_box_dev.p, bb.data(), sizeof(double) * bb.size(),
                    hipMemcpyHostToDevice), "hipMemcpy");
  return LJGPU_OK;
}

int ljgpu_build_list(ljgpu_ctx* ctx, const double* q, int32_t n, double box,
                     double search_length, int kind) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (n > 0 && (reinterpret_cast<uintptr_t>(q) & 15u))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q must be 16-byte aligned");
  PosIn in;
  in.x = q;
  in.y = q ? q + 1 : nullptr;
  in.z = q ? q + 2 : nullptr;
  in.stride = 4;
  return build_list(ctx, in, n, box, search_length, kind);
}

int ljgpu_build_list_aos3(ljgpu_ctx* ctx, const double* q, int32_t n, double box,
                          double search_length, int kind) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (n > 0 && !aligned8(q)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q must be 8-byte aligned");
  PosIn in;
  in.x = q;
  in.y = q ? q + 1 : nullptr;
  in.z = q ? q + 2 : nullptr;
  in.stride = 3;
  return build_list(ctx, in, n, box, search_length, kind);
}

int ljgpu_build_list_soa(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                         int32_t n, double box, double search_length, int kind) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  PosIn in;
  in.x = qx;
  in.y = qy;
  in.z = qz;
  in.stride = 1;
  return build_list(ctx, in, n, box, search_length, kind);
}

int ljgpu_force_aos_f64(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                        double box, int32_t steps) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!aligned16(q) || !aligned16(p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 16-byte aligned");
  ForceArgs fa;
  fa.soa = false;
  fa.q[0] = q;
  fa.p[0] = p;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, false, 0);
}

int ljgpu_force_aos3_f64(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                         double box, int32_t steps) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!aligned8(q) || !aligned8(p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 8-byte aligned");
  ForceArgs fa;
  fa.aos3 = true;
  fa.q[0] = q;
  fa.p[0] = p;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, false, 0);
}

int ljgpu_force_aos3_mixed(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                           double box, int32_t steps, int accum) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (accum != LJGPU_ACCUM_F64 && accum != LJGPU_ACCUM_KAHAN_F32)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown accumulator");
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!aligned8(q) || !aligned8(p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 8-byte aligned");
  ForceArgs fa;
  fa.aos3 = true;
  fa.q[0] = q;
  fa.p[0] = p;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, true, accum);
}

int ljgpu_force_soa_f64(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                        double* px, double* py, double* pz, double dt, double cl2, double box,
                        int32_t steps) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (ctx->n > 0 && (!qx || !qy || !qz || !px || !py || !pz))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p component");
  ForceArgs fa;
  fa.soa = true;
  fa.q[0] = qx; fa.q[1] = qy; fa.q[2] = qz;
  fa.p[0] = px; fa.p[1] = py; fa.p[2] = pz;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, false, 0);
}

int ljgpu_force_aos_mixed(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                          double box, int32_t steps, int accum) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (accum != LJGPU_ACCUM_F64 && accum != LJGPU_ACCUM_KAHAN_F32)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown accumulator");
  if (ctx->n > 0 && (!q || !p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p");
  if (!aligned16(q) || !aligned16(p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 16-byte aligned");
  ForceArgs fa;
  fa.soa = false;
  fa.q[0] = q;
  fa.p[0] = p;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, true, accum);
}

int ljgpu_force_soa_mixed(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                          double* px, double* py, double* pz, double dt, double cl2, double box,
                          int32_t steps, int accum) {
  int rc = check_force_args(ctx, dt, cl2, steps, box);
  if (rc) return rc;
  if (accum != LJGPU_ACCUM_F64 && accum != LJGPU_ACCUM_KAHAN_F32)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown accumulator");
  if (ctx->n > 0 && (!qx || !qy || !qz || !px || !py || !pz))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null q or p component");
  ForceArgs fa;
  fa.soa = true;
  fa.q[0] = qx; fa.q[1] = qy; fa.q[2] = qz;
  fa.p[0] = px; fa.p[1] = py; fa.p[2] = pz;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = box;
  return run_steps(ctx, fa, steps, true, accum);
}

int ljgpu_force_kernel_csr_f64(ljgpu_ctx* ctx, const double* q, double* p, int32_t n, double dt,
                               double cl2, const int32_t* list, const int32_t* nop,
                               const int32_t* ptr, int64_t n_entries, int kind) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (n < 0 || n_entries < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "negative size");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "bad kind");
  if (n == 0) return LJGPU_OK;
  if (!q || !p || !nop || !ptr || (n_entries > 0 && !list))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null device pointer");
  if (!aligned16(q) || !aligned16(p)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "q and p must be 16-byte aligned");
  int rc = use_device(ctx);
  if (rc) return rc;
  ListView lv;
  lv.n = n;
  lv.nop = nop;
  lv.ptr = ptr;
  lv.list = list;
  lv.n_entries = n_entries;
  if (n_entries == 0) {
    // Every row is empty after clamping; nothing to add.
    return LJGPU_OK;
  }
  ForceArgs fa;
  fa.q[0] = q;
  fa.p[0] = p;
  fa.dt = dt;
  fa.cl2 = cl2;
  fa.box = 0.0;
  LaunchConfig lc;
  lc.lanes = ctx->lanes == 1 ? 16 : ctx->lanes;
  lc.unroll = ctx->unroll;
  lc.block = ctx->block;
  lc.half_atomic = (kind == LJGPU_LIST_HALF);
  lc.clamp = true;
  return launch_force_f64(ctx->stream, lv, fa, lc);
}

}  // extern "C"
