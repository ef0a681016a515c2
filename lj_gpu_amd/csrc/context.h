// context.h — the library context (struct ljgpu_ctx) and the internal
// operations api.hip and md.hip share. HIP-only header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "ljgpu_internal.h"
#include "lists.h"

namespace ljgpu {

// Graph replay of repeated force evaluations. A force call of `steps`
// launches runs as replays of captured graphs of kGraphChunks[c] launches,
// taken greedily from the longest (so 20 steps = 16 + 4), then fewer than
// kGraphMinSteps single launches. A graph is captured once per set of kernel
// arguments and chunk length and kept in a small cache (kGraphCacheSize
// entries, least recently used evicted), so calls that alternate step counts,
// precisions or arrays replay their graphs instead of re-capturing them.
constexpr int32_t kGraphMinSteps = 4;
constexpr int32_t kGraphChunks[3] = {64, 16, 4};
constexpr int kGraphCacheSize = 12;
// Longest run of (drift, kick) pairs the MD loop captures as one graph.
constexpr int32_t kGraphChunk = 100;
// A call of kGraphMinSteps or more steps over LDS tiles splits the tiles into
// `chains` (1 to kMaxChains) contiguous parts and runs the steps of each part
// as its own chain of launches on its own stream. The parts update disjoint
// particles, so the chains need no ordering between them and every particle
// receives the same additions in the same order; one chain's launch ramp and
// tail overlap the others' work. Chained calls launch directly on the
// streams (a captured graph of the same fork and join ran its branches with
// less overlap on MI355X); single-chain calls replay graphs.
constexpr int kMaxChains = 4;
constexpr int32_t kMinChainTiles = 64;  // fewer tiles: one chain
// Calls with at least kSlotMinSteps steps over spatially ordered tiles run
// over copies of q and p in slot order (two gathers and a scatter per call).
constexpr int32_t kSlotMinSteps = 16;
// Energy sums: a fixed grid of kEnergyBlocks workgroups (each striding over
// the rows) writes one partial per workgroup and quantity (potential,
// kinetic, pairs inside the cutoff), added afterwards in block order, so the
// result does not depend on scheduling or on the particle count.
constexpr int kEnergyBlocks = 1024;
// Doubles in the static per-GPU scratch of the allocation-free entry point
// ljgpu_energy_noalloc: the 3 * kEnergyBlocks partials and the 3 totals
// (24.6 KB, whatever the particle count).
constexpr int kNoAllocScratchDoubles = 3 * kEnergyBlocks + 3;

// Every argument of a captured force loop, written field by field into a
// byte string (so padding inside the argument structs never enters it); a
// graph is replayed only for an identical key.
struct StepKey {
  unsigned char bytes[384];
  size_t size = 0;
  template <class T>
  void add(const T& v) {
    static_assert(sizeof(T) <= 8, "scalar fields only");
    std::memcpy(bytes + size, &v, sizeof(T));
    size += sizeof(T);
  }
  bool operator==(const StepKey& o) const {
    return size == o.size && std::memcmp(bytes, o.bytes, size) == 0;
  }
};
StepKey step_key(const ForceArgs& fa, const ListView& lv, const LaunchConfig& lc, bool mixed,
                 int accum, int32_t steps, int32_t chains);

struct GraphEntry {
  StepKey key;
  hipGraphExec_t exec = nullptr;
  uint64_t last_use = 0;
};

}  // namespace ljgpu

struct ljgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool timer_running = false;
  // Further chains of a many-step call over tiles (see kMaxChains): their
  // own streams (default priority), forked from and joined back into
  // `stream`.
  hipStream_t chain_stream[ljgpu::kMaxChains - 1] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[ljgpu::kMaxChains - 1] = {};
  int chains = 2;

  // List as given by the caller (or built by ljgpu_build_list).
  int32_t n = 0;
  int kind = -1;
  int64_t n_entries = 0;
  int32_t max_np = 0, min_np = 0;

  // List as evaluated on the device: row-ordered CSR (the full list, or the
  // half list itself for atomic mode), and the layouts derived from it.
  int half_mode_req = LJGPU_HALF_AUTO;
  int half_mode = 0;
  bool eval_half = false;
  ljgpu::DevCSR eval;
  ljgpu::DevCSR staging;  // a caller's host list on its way to `eval`
  ljgpu::DevMem ell;
  int32_t ell_width = 0;
  bool ell_ready = false;

  // LDS tiling of the evaluated full list (tile_particles > 0 enables it).
  int32_t tile_particles = 64;
  bool tile_offsets_lds = false;  // stage entry offsets in LDS (else stream them)
  int32_t tile_resident = 0;      // persistent tile grid: workgroups per CU (0: one per tile)
  int cus = 0;                    // compute units of the device
  int32_t tiles_for = 0;          // effective tile size the tiling was derived for (0: none)
  int tile_order = LJGPU_TILE_ORDER_SPATIAL;  // requested order of the particles in tiles
  bool tiles_ordered = false;     // the tiling was derived with positions (order, cutoff classes)
  bool bank_order = true;         // tile rows in the LDS bank-class order (bank_class_order)
  ljgpu::DevTiles tiles;
  ljgpu::DevMem slot_q, slot_p;   // q and p in slot order during a many-step call

  ljgpu::ListWork work;

  // Captured force loops (see StepKey), at most kGraphCacheSize of them.
  bool use_graphs = true;
  std::vector<ljgpu::GraphEntry> graphs;
  uint64_t graph_clock = 0;
  int64_t graph_captures = 0;  // graphs captured over the context's life

  // Batched replicas: particle offsets (n_replicas + 1) and boxes on the
  // host, (box, 1/box) per replica on the device; empty for one system.
  std::vector<int32_t> rep_off;
  std::vector<double> rep_boxes;
  ljgpu::DevMem rep_box_dev;

  // Autotuner results of the last ljgpu_autotune call.
  std::vector<ljgpu_tune_entry> tune;

  // Molecular-dynamics state: positions at the last list build, reduction
  // scratch.
  ljgpu::DevMem md_qref;
  ljgpu::DevMem md_red;
  // The loop's working copies of q and p in a spatial order, a staging
  // buffer, and the caller's index of each slot (two buffers).
  ljgpu::DevMem md_q, md_p, md_tmp, md_perm, md_perm2;
  // Device address of this GPU's static scratch array (looked up once per
  // device at context creation; shared by every context on the device).
  double* noalloc_scratch = nullptr;

  int lanes = 4;
  int unroll = 4;
  int block = 256;
};

namespace ljgpu {

inline int hip_fail(hipError_t e, const char* what) {
  return set_error(e == hipErrorOutOfMemory ? LJGPU_ERR_OUT_OF_MEMORY : LJGPU_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                              \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return ::ljgpu::hip_fail(e_, what); \
  } while (0)

#define TRY(call)         \
  do {                    \
    int rc_ = (call);     \
    if (rc_) return rc_;  \
  } while (0)

int use_device(ljgpu_ctx* ctx);
// Forget the installed list (and its layouts).
void reset_list(ljgpu_ctx* ctx);
// Particles per tile for the current kernel configuration (0: untiled).
int32_t effective_tile(const ljgpu_ctx* ctx);
// Derive the layouts the configured kernel needs. Positions (q, with the
// periodic box, 0 for none) let the tiles take the spatial order and the
// cutoff classes of cl2 (the reference's squared cutoff, 9.0, by default).
constexpr double kDefaultCutHint = 9.0;
int prepare_layouts(ljgpu_ctx* ctx, const PosIn* q = nullptr, double box = 0.0,
                    double cl2 = kDefaultCutHint);
// Device positions of a force evaluation's arguments.
PosIn positions_of(const ForceArgs& fa);
// Build the list from device positions and install it (ljgpu_build_list).
int build_list(ljgpu_ctx* ctx, const PosIn& q, int32_t n, double box, double search, int kind,
               bool sort_rows = true);
// Enqueue one force evaluation p += F(q) * dt over the installed list.
int enqueue_force(ljgpu_ctx* ctx, const ForceArgs& fa, bool mixed, int accum);
// `steps` force evaluations (graph replays where possible).
int run_steps(ljgpu_ctx* ctx, const ForceArgs& fa, int32_t steps, bool mixed, int accum);
void drop_graph(ljgpu_ctx* ctx);
// Kernel arguments of one force evaluation over the installed list (derives
// missing layouts first, from the positions of `fa`).
int force_view(ljgpu_ctx* ctx, const ForceArgs& fa, ListView* lv, LaunchConfig* lc);
// The installed tiling's fields of a ListView, as they stand.
void tile_fields(const ljgpu_ctx* ctx, ListView* lv);
// Record the launches `body` enqueues on the context stream into an
// executable graph.
int capture(ljgpu_ctx* ctx, const std::function<int()>& body, hipGraphExec_t* out);
// The context's cached graph for `key`, captured from `body` on a miss.
int cached_graph(ljgpu_ctx* ctx, const StepKey& key, const std::function<int()>& body,
                 hipGraphExec_t* out);

}  // namespace ljgpu
