// lists.h — neighbour lists on the device: the Verlet-list build from device
// positions (cell list, the reference's exact pair set and CSR order) and
// every layout the force kernels evaluate, derived on the device from a CSR
// list (half-to-full expansion, row ordering, ELL, LDS tiles). HIP-only
// header, used by api.hip and lists.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "ljgpu_internal.h"

namespace ljgpu {

// A device allocation that only grows: reserve() reallocates (after
// synchronising `stream`) when more bytes are needed, so repeated list
// builds of similar size reuse their buffers.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};
int dev_reserve(DevMem& m, size_t bytes, hipStream_t stream);
void dev_release(DevMem& m);

// A CSR list on the device: number_of_partners[n], pointer[n + 1] (pointer[n]
// = e when rows are stored in order), sorted_list[e].
struct DevCSR {
  int32_t n = 0;
  int64_t e = 0;
  DevMem nop, ptr, list;
};
void dev_release(DevCSR& c);

// LDS tiling of a row-ordered full list (the layout ListView describes).
// Blocks cover consecutive slots of a particle order: slot s holds particle
// member[s], whose row occupies local[slot_ptr[s] .. slot_ptr[s+1]). In
// index order (member[s] = s) neither array is kept: the list's own
// pointer serves as slot_ptr.
struct DevTiles {
  int32_t tile_particles = 0;  // particles per block (0: no tiling)
  int32_t blocks = 0;
  int64_t staged = 0;
  int32_t max_stage = 0;
  int32_t lds_bytes = 0;
  int order = 0;               // LJGPU_TILE_ORDER_INDEX (member[s] = s) or _SPATIAL
  int32_t stage_stride = 0;    // >0: block b stages at b * stage_stride (one-pass build);
                               // 0: staging areas packed back to back
  DevMem desc, stage, local;
  DevMem member;      // int32 per slot
  DevMem slot_ptr;    // int32 per slot (+1)
  DevMem stage_slot;  // stage as slots (spatial order only)
  // Each block's staged particles are sorted by slot (by index in index
  // order), and local holds 24 * the rank of each entry's partner among them.

  // The same offsets sliced per wavefront for the streamed tile kernel of
  // sliced_lanes lanes per particle and sliced_unroll entries per lane per
  // batch (see dev_slice_tiles); sliced_lanes = 0: not built.
  DevMem sliced;      // uint16
  DevMem wave_off;    // int32 per (block, wavefront slot): start of its slice
  int32_t sliced_lanes = 0, sliced_unroll = 0, waves_per_block = 0;
};
void dev_release(DevTiles& t);

// Scratch buffers reused across builds.
struct ListWork {
  DevMem scan_sums;   // int64 per scan tile
  DevMem small;       // a few int64 results (totals, flags, extrema)
  DevMem cell_count;  // int32 per cell (+1)
  DevMem cell_start;  // int32 per cell (+1)
  DevMem cell_slot;   // int32 per particle: rank within its cell / fill counters
  DevMem cell_of;     // int32 per particle: cell index
  DevMem cell_sorted; // int32 per particle, cell order: cell index
  DevMem cpos;        // double4 per particle, cell order: x, y, z, index
  DevMem counts;      // int32 per row (+1)
  DevMem tmp;         // int32 per list entry
  DevMem tile_ns;     // int32 per tile (+1)
  DevMem tile_soff;   // int32 per tile (+1)
  DevMem tile_geo;    // int32 x 4 per tile: first particle, count, replica
  DevMem rep_of;      // int32 per particle: replica index
  DevMem order;       // int32 per particle: spatial tile order
  DevMem order_grid;  // OrderGrid per system
  DevMem slot_of;     // int32 per particle: its slot in the tile order
  DevCSR part;        // one replica's list during a batched build
  int32_t last_max_row = 0;  // longest row of the last device build (slab size hint)
};
void dev_release(ListWork& w);

// Positions as three strided arrays (AoS: q, q+1, q+2 with stride 4; SoA:
// three arrays with stride 1).
struct PosIn {
  const double* x = nullptr;
  const double* y = nullptr;
  const double* z = nullptr;
  int stride = 4;
};

// Exclusive scan of in[0..n) into out[0..n], out[n] = total; *total receives
// the sum (synchronises the stream).
int dev_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, ListWork& w,
             int64_t* total);

// Strict check of a CSR list on the device (the rules of validate_csr);
// returns LJGPU_OK or LJGPU_ERR_BAD_LIST naming the first bad row.
int dev_validate(hipStream_t s, int32_t n, const int32_t* nop, const int32_t* ptr,
                 const int32_t* list, int64_t e, int kind, ListWork& w);

// Copy a caller's device CSR into `out` with rows stored in order.
int dev_copy_row_ordered(hipStream_t s, int32_t n, const int32_t* nop, const int32_t* ptr,
                         const int32_t* list, int64_t e, ListWork& w, DevCSR& out);

// Full list from a half list (both on the device): row i holds its lower
// partners ascending, then its own half row as stored.
int dev_expand_half(hipStream_t s, int32_t n, const int32_t* nop, const int32_t* ptr,
                    const int32_t* list, ListWork& w, DevCSR& full);

// Verlet list of `kind` from device positions: pair predicate
// r2 = (dx*dx + dy*dy) + dz*dz < search^2, unfused, minimum image for
// box > 0; rows ascending in i. With sort_rows, partners ascending (the
// reference's order); without, each row's partners come in an order fixed
// by the positions (stencil runs, cells, particle index within a cell), the
// same on every build from the same positions.
int dev_makepair(hipStream_t s, const PosIn& q, int32_t n, double box, double search_length,
                 int kind, ListWork& w, DevCSR& out, bool sort_rows = true);

// Smallest and largest of a[0..n) (synchronises); 0, 0 for n = 0.
int dev_minmax(hipStream_t s, const int32_t* a, int32_t n, ListWork& w, int32_t* mn, int32_t* mx);
// Row lengths of the half list contained in a full list with ascending rows
// (partners j > i), as min/max (synchronises).
int dev_half_row_stats(hipStream_t s, const DevCSR& full, ListWork& w, int32_t* min_np,
                       int32_t* max_np);

// ELL ("transposed") copy, entry k of row i at ell[i + k*n], width = longest
// row (at least 1), zero padded.
int dev_make_ell(hipStream_t s, const DevCSR& c, int32_t width, DevMem& ell);

// LDS tiling of a row-ordered full list with blocks of `tile_particles`
// consecutive particles, restarted at each replica boundary of `replicas`
// (host offsets; empty for a single system), descriptor slot 6 holding the
// block's replica. Returns LJGPU_OK with out.tile_particles = 0 when a block
// does not fit the LDS budget (the caller then halves the tile size or
// evaluates untiled).
// `member` (device, one particle per slot, a permutation that keeps every
// replica's particles within its own slot range) orders the particles the
// blocks cover; null takes them in index order. bank_order puts each row's
// entries in the LDS bank-class order (bank_class_order), as the host model
// does; without it rows keep the list order.
// With bank_order and positions in `cut` (cut->q.x non-null), the rows
// first take the cutoff classes (cut_partition) of those positions.
struct CutHint {
  PosIn q;                        // positions in the caller's particle order
  double box = 0.0;               // single system: periodic box (<= 0 open)
  const double* rep_box = nullptr;  // replicas: (box, 1/box) per replica on the device
  double cl2 = 9.0;
};
int dev_make_tiles(hipStream_t s, const DevCSR& c, int32_t tile_particles,
                   const std::vector<int32_t>& replicas, const int32_t* member, bool bank_order,
                   const CutHint* cut, ListWork& w, DevTiles& out);

// The tiles' entry offsets sliced per wavefront, for a tile kernel in
// which `lanes` consecutive lanes own one row and take entries in batches
// of lanes * unroll (entry k of a row: batch k / (lanes * unroll), slot
// (k / lanes) % unroll, lane k % lanes). Wavefront w of block b holds the
// block's rows [w * 64/lanes, (w+1) * 64/lanes) and owns the slice that
// starts at wave_off[b * waves_per_block + w]: for each of its batches
// (as many as its longest row needs) 64 * unroll offsets, slot by slot,
// each slot the 64 lanes' entries in lane order, so one offset load of the
// kernel reads 64 consecutive values. Slots past a row's end hold 0.
// Needs local, desc and the slot row starts of a built tiling; a slicing
// whose total does not fit int32 is not built (sliced_lanes stays 0).
int dev_slice_tiles(hipStream_t s, const int32_t* slot_ptr, int32_t lanes, int32_t unroll,
                    ListWork& w, DevTiles& t);

// Spatial particle order for LDS tiles, into w.order: each periodic system
// (the whole set, or every replica of `replicas` with its box) is cut into
// columns along x about 0.75 * cbrt(tile_particles / density) wide in y and
// z, each column into x bins of about one particle; particles are ordered by
// (system, z column, y column, x bin, index). Consecutive slots then hold
// compact groups of particles, so a block of them stages fewer partners than
// a block of consecutive indices. *ok = false (and no order) for an open box,
// non-finite positions or a bin of more than 256 particles.
int dev_tile_order(hipStream_t s, const PosIn& q, int32_t n, double box,
                   const std::vector<int32_t>& replicas, const std::vector<double>& boxes,
                   int32_t tile_particles, ListWork& w, bool* ok);

// Within each block of tile_particles consecutive slots of w.order (restarted
// at every replica boundary, as dev_make_tiles forms its blocks), the slots
// reordered by the row length of their particles in c, longest first, ties
// in the previous order.
int dev_length_sort(hipStream_t s, const DevCSR& c, int32_t tile_particles,
                    const std::vector<int32_t>& replicas, ListWork& w);

// Batched replicas: lists of every replica (particles [off[r], off[r+1]),
// box boxes[r]) built on the device into one CSR with global indices.
int dev_makepair_batch(hipStream_t s, const PosIn& q, const std::vector<int32_t>& off,
                       const std::vector<double>& boxes, double search_length, int kind,
                       ListWork& w, DevCSR& out);

// LJGPU_ERR_BAD_LIST unless every entry of row i lies in i's replica.
int dev_check_replicas(hipStream_t s, const DevCSR& c, const std::vector<int32_t>& off,
                       ListWork& w);

}  // namespace ljgpu
