// api.hip — the C ABI of ljgpu (include/ljgpu.h): contexts, device memory,
// neighbour-list installation and the force entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ljgpu_internal.h"

using namespace ljgpu;

struct ljgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool timer_running = false;

  // List as given by the caller.
  int32_t n = 0;
  int kind = -1;
  int64_t n_entries = 0;
  int32_t max_np = 0, min_np = 0;

  // List as evaluated on the device: CSR (full, or half for atomic mode).
  int half_mode_req = LJGPU_HALF_AUTO;
  int half_mode = 0;
  bool eval_half = false;
  HostCSR eval;  // host copy (used for derived layouts and get_list)
  int32_t* d_nop = nullptr;
  int32_t* d_ptr = nullptr;
  int32_t* d_list = nullptr;
  int64_t d_list_len = 0;
  int32_t* d_ell = nullptr;
  int32_t ell_width = 0;

  // LDS tiling of the evaluated full list (tile_particles > 0 enables it).
  int32_t tile_particles = 32;
  bool tile_offsets_lds = false;  // stage entry offsets in LDS (else stream them)
  int32_t tile_resident = 0;      // persistent tile grid: workgroups per CU (0: one per tile)
  int cus = 0;                    // compute units of the device
  int32_t tiles_built_for = 0;  // tile_particles the device tiles were built with
  HostTiles tiles;
  int32_t* d_tile_desc = nullptr;
  int32_t* d_tile_stage = nullptr;
  uint16_t* d_tile_local = nullptr;

  int lanes = 8;
  int unroll = 4;
  int block = 256;
};

namespace {

int hip_fail(hipError_t e, const char* what) {
  return set_error(e == hipErrorOutOfMemory ? LJGPU_ERR_OUT_OF_MEMORY : LJGPU_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                 \
  do {                                      \
    hipError_t e_ = (expr);                 \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int use_device(ljgpu_ctx* ctx) {
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  return LJGPU_OK;
}

void free_tiles(ljgpu_ctx* ctx) {
  if (ctx->d_tile_desc) (void)hipFree(ctx->d_tile_desc);
  if (ctx->d_tile_stage) (void)hipFree(ctx->d_tile_stage);
  if (ctx->d_tile_local) (void)hipFree(ctx->d_tile_local);
  ctx->d_tile_desc = ctx->d_tile_stage = nullptr;
  ctx->d_tile_local = nullptr;
  ctx->tiles = HostTiles();
  ctx->tiles_built_for = 0;
}

void free_list(ljgpu_ctx* ctx) {
  free_tiles(ctx);
  if (ctx->d_nop) (void)hipFree(ctx->d_nop);
  if (ctx->d_ptr) (void)hipFree(ctx->d_ptr);
  if (ctx->d_list) (void)hipFree(ctx->d_list);
  if (ctx->d_ell) (void)hipFree(ctx->d_ell);
  ctx->d_nop = ctx->d_ptr = ctx->d_list = ctx->d_ell = nullptr;
  ctx->d_list_len = 0;
  ctx->ell_width = 0;
  ctx->n = 0;
  ctx->kind = -1;
  ctx->n_entries = 0;
  ctx->eval = HostCSR();
}

template <class T>
int upload(ljgpu_ctx* ctx, T** dst, const std::vector<T>& src) {
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), bytes), "hipMalloc (list)");
  if (!src.empty())
    HIP_TRY(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                           ctx->stream),
            "hipMemcpyAsync (list)");
  return LJGPU_OK;
}

int build_ell(ljgpu_ctx* ctx) {
  const HostCSR& c = ctx->eval;
  int32_t width = 0;
  for (int32_t i = 0; i < c.n; i++) width = std::max(width, c.nop[i]);
  std::vector<int32_t> ell(static_cast<size_t>(c.n) * std::max(width, 1), 0);
  for (int32_t i = 0; i < c.n; i++)
    for (int32_t k = 0; k < c.nop[i]; k++)
      ell[i + static_cast<size_t>(k) * c.n] = c.list[c.ptr[i] + k];
  if (ctx->d_ell) (void)hipFree(ctx->d_ell);
  ctx->d_ell = nullptr;
  int rc = upload(ctx, &ctx->d_ell, ell);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  ctx->ell_width = std::max(width, 1);
  return LJGPU_OK;
}

// Derive and upload the LDS tiling of the evaluated (full) list for the
// context's tile_particles. A list whose single rows exceed the LDS staging
// capacity keeps the untiled kernels.
// Particles per tile for the current kernel configuration: the requested
// tile size, capped so that a tile's workgroup (tile * lanes threads) fits
// 1024 threads; 0 when tiling is off or the ELL kernel is selected.
int32_t effective_tile(const ljgpu_ctx* ctx) {
  if (ctx->tile_particles <= 0 || ctx->lanes <= 1) return 0;
  return std::min(ctx->tile_particles, 1024 / ctx->lanes);
}

int build_device_tiles(ljgpu_ctx* ctx) {
  free_tiles(ctx);
  const int32_t tp = effective_tile(ctx);
  if (tp <= 0 || ctx->eval_half || ctx->eval.n == 0) return LJGPU_OK;
  HostTiles t;
  if (!build_tiles(ctx->eval, tp, &t)) return LJGPU_OK;
  int rc;
  if ((rc = upload(ctx, &ctx->d_tile_desc, t.desc))) return rc;
  if ((rc = upload(ctx, &ctx->d_tile_stage, t.stage))) return rc;
  t.local.resize(t.local.size() + kTileLocalPad, 0);  // the kernel copies whole 16-byte chunks
  if ((rc = upload(ctx, &ctx->d_tile_local, t.local))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  ctx->tiles = std::move(t);
  ctx->tiles.stage.clear();
  ctx->tiles.stage.shrink_to_fit();
  ctx->tiles.local.clear();
  ctx->tiles.local.shrink_to_fit();
  ctx->tiles_built_for = tp;
  return LJGPU_OK;
}

// Install a validated CSR list held on the host.
int install_csr(ljgpu_ctx* ctx, int32_t n, std::vector<int32_t>&& nop, std::vector<int32_t>&& ptr,
                std::vector<int32_t>&& list, int kind) {
  int rc = use_device(ctx);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  free_list(ctx);
  ctx->n = n;
  ctx->kind = kind;
  ctx->n_entries = static_cast<int64_t>(list.size());
  ctx->max_np = 0;
  ctx->min_np = n > 0 ? INT32_MAX : 0;
  for (int32_t i = 0; i < n; i++) {
    ctx->max_np = std::max(ctx->max_np, nop[i]);
    ctx->min_np = std::min(ctx->min_np, nop[i]);
  }
  HostCSR given;
  given.n = n;
  given.nop = std::move(nop);
  given.ptr = std::move(ptr);
  given.list = std::move(list);
  if (kind == LJGPU_LIST_HALF) {
    ctx->half_mode = ctx->half_mode_req == LJGPU_HALF_ATOMIC ? LJGPU_HALF_ATOMIC : LJGPU_HALF_EXPAND;
  } else {
    ctx->half_mode = 0;
  }
  if (kind == LJGPU_LIST_HALF && ctx->half_mode == LJGPU_HALF_EXPAND) {
    expand_half_to_full(given, &ctx->eval);
    ctx->eval_half = false;
  } else {
    ctx->eval = std::move(given);
    ctx->eval_half = (kind == LJGPU_LIST_HALF);
  }
  // Rows stored in order make every tile's entries one contiguous range.
  csr_make_row_ordered(&ctx->eval);
  if ((rc = upload(ctx, &ctx->d_nop, ctx->eval.nop))) return rc;
  if ((rc = upload(ctx, &ctx->d_ptr, ctx->eval.ptr))) return rc;
  if ((rc = upload(ctx, &ctx->d_list, ctx->eval.list))) return rc;
  ctx->d_list_len = static_cast<int64_t>(ctx->eval.list.size());
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (ctx->lanes == 1) {
    if ((rc = build_ell(ctx))) return rc;
  } else if (effective_tile(ctx) > 0) {
    if ((rc = build_device_tiles(ctx))) return rc;
  }
  return LJGPU_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int check_force_args(ljgpu_ctx* ctx, double dt, double cl2, int32_t steps) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  if (steps < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "steps < 0");
  if (!(cl2 >= 0.0)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "cl2 must be >= 0");
  if (dt != dt) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "dt is NaN");
  return LJGPU_OK;
}

int run_steps(ljgpu_ctx* ctx, const ForceArgs& fa, int32_t steps, bool mixed, int accum) {
  int rc = use_device(ctx);
  if (rc) return rc;
  if (ctx->n == 0 || steps == 0) return LJGPU_OK;
  ListView lv;
  lv.n = ctx->n;
  lv.nop = ctx->d_nop;
  lv.ptr = ctx->d_ptr;
  lv.n_entries = ctx->d_list_len;
  LaunchConfig lc;
  lc.lanes = ctx->lanes;
  lc.unroll = ctx->unroll;
  lc.block = ctx->block;
  lc.tile_offsets_lds = ctx->tile_offsets_lds;
  lc.tile_resident = ctx->tile_resident;
  lc.cus = ctx->cus;
  lc.half_atomic = ctx->eval_half;
  lc.clamp = false;
  if (ctx->lanes == 1) {
    if (!ctx->d_ell && (rc = build_ell(ctx))) return rc;
    lv.list = ctx->d_ell;
    lv.ell_width = ctx->ell_width;
  } else {
    lv.list = ctx->d_list;
    const int32_t tp = effective_tile(ctx);
    if (tp > 0 && !ctx->eval_half) {
      if (ctx->tiles_built_for != tp && (rc = build_device_tiles(ctx))) return rc;
      if (ctx->tiles_built_for > 0) {
        lv.tile_blocks = static_cast<int32_t>(ctx->tiles.desc.size() / 8);
        lv.tile_particles = ctx->tiles.tile_particles;
        lv.tile_lds_bytes = ctx->tiles.lds_bytes;
        lv.tile_max_stage = ctx->tiles.max_stage;
        lv.tile_desc = ctx->d_tile_desc;
        lv.tile_stage = ctx->d_tile_stage;
        lv.tile_local = ctx->d_tile_local;
      }
    }
  }
  for (int32_t s = 0; s < steps; s++) {
    rc = !mixed ? launch_force_f64(ctx->stream, lv, fa, lc)
         : accum == LJGPU_ACCUM_KAHAN_F32 ? launch_force_mixed_kahan(ctx->stream, lv, fa, lc)
                                          : launch_force_mixed_f64acc(ctx->stream, lv, fa, lc);
    if (rc) return rc;
  }
  return LJGPU_OK;
}

}  // namespace

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
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    ljgpu_destroy(ctx);
    return hip_fail(e, "context creation");
  }
  *out = ctx;
  return LJGPU_OK;
}

int ljgpu_destroy(ljgpu_ctx* ctx) {
  if (!ctx) return LJGPU_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_list(ctx);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
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

int ljgpu_set_half_mode(ljgpu_ctx* ctx, int half_mode) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (half_mode != LJGPU_HALF_AUTO && half_mode != LJGPU_HALF_ATOMIC && half_mode != LJGPU_HALF_EXPAND)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "unknown half mode");
  ctx->half_mode_req = half_mode;
  return LJGPU_OK;
}

int ljgpu_set_kernel_config(ljgpu_ctx* ctx, int lanes, int block, int unroll) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (lanes != 0 && lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "lanes_per_particle must be 1,4,8,16,32 or 64");
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
  std::vector<int32_t> h_nop(nop ? n : 0), h_ptr(ptr ? n : 0), h_list(static_cast<size_t>(n_entries));
  if (mem == LJGPU_MEM_DEVICE) {
    int rc = use_device(ctx);
    if (rc) return rc;
    if (n > 0) {
      HIP_TRY(hipMemcpy(h_nop.data(), nop, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "hipMemcpy");
      HIP_TRY(hipMemcpy(h_ptr.data(), ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    if (n_entries > 0)
      HIP_TRY(hipMemcpy(h_list.data(), list, sizeof(int32_t) * n_entries, hipMemcpyDeviceToHost), "hipMemcpy");
  } else {
    if (n > 0) {
      std::memcpy(h_nop.data(), nop, sizeof(int32_t) * n);
      std::memcpy(h_ptr.data(), ptr, sizeof(int32_t) * n);
    }
    if (n_entries > 0) std::memcpy(h_list.data(), list, sizeof(int32_t) * n_entries);
  }
  int rc = validate_csr(n, h_nop.data(), h_ptr.data(), h_list.data(), n_entries, kind);
  if (rc) return rc;
  return install_csr(ctx, n, std::move(h_nop), std::move(h_ptr), std::move(h_list), kind);
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
  return install_csr(ctx, n, std::move(h_nop), std::move(h_ptr), std::move(h_list), kind);
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
  info->tile_particles = ctx->tiles_built_for;
  info->tile_blocks = static_cast<int32_t>(ctx->tiles.desc.size() / 8);
  info->tile_max_stage = ctx->tiles.max_stage;
  return LJGPU_OK;
}

int ljgpu_get_list_csr(ljgpu_ctx* ctx, int32_t* nop, int32_t* ptr, int32_t* list, int64_t capacity) {
  if (!ctx) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->kind < 0) return set_error(LJGPU_ERR_NO_LIST, "no neighbour list installed");
  const HostCSR& c = ctx->eval;
  if (static_cast<int64_t>(c.list.size()) > capacity) return set_error(LJGPU_ERR_CAPACITY, "capacity too small");
  if (c.n > 0 && (!nop || !ptr)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null output");
  if (!c.list.empty() && !list) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null output");
  if (c.n > 0) {
    std::memcpy(nop, c.nop.data(), sizeof(int32_t) * c.n);
    std::memcpy(ptr, c.ptr.data(), sizeof(int32_t) * c.n);
  }
  if (!c.list.empty()) std::memcpy(list, c.list.data(), sizeof(int32_t) * c.list.size());
  return LJGPU_OK;
}

int ljgpu_force_aos_f64(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                        double box, int32_t steps) {
  int rc = check_force_args(ctx, dt, cl2, steps);
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

int ljgpu_force_soa_f64(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                        double* px, double* py, double* pz, double dt, double cl2, double box,
                        int32_t steps) {
  int rc = check_force_args(ctx, dt, cl2, steps);
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
  int rc = check_force_args(ctx, dt, cl2, steps);
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
  int rc = check_force_args(ctx, dt, cl2, steps);
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
