// ljgpu_internal.h — shared declarations of the ljgpu library (not installed).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ljgpu.h"

#if defined(__HIPCC__)
#define LJ_HD __host__ __device__
#else
#define LJ_HD
#endif

namespace ljgpu {

// Thread-local message of the last failing call; returns `code`.
int set_error(int code, const std::string& msg);
void clear_error();

// ---------------------------------------------------------------------------
// Host-side list helpers (host.cpp)
// ---------------------------------------------------------------------------
struct HostCSR {
  int32_t n = 0;
  std::vector<int32_t> nop, ptr, list;
};

// Expand a half list (i < j) into the full list: row i holds the lower
// partners (ascending, from rows r < i) followed by its own half row.
void expand_half_to_full(const HostCSR& half, HostCSR* full);

// Return LJGPU_OK or LJGPU_ERR_BAD_LIST with a message naming the defect.
int validate_csr(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                 int64_t n_entries, int kind);

// ---------------------------------------------------------------------------
// Device launchers (force_f64.hip, force_mixed.hip)
// ---------------------------------------------------------------------------
struct ListView {
  int32_t n = 0;
  const int32_t* nop = nullptr;   // device
  const int32_t* ptr = nullptr;   // device (CSR); may be null for ELL
  const int32_t* list = nullptr;  // device CSR list, or ELL list (stride n)
  int64_t n_entries = 0;          // CSR length (for clamping)
  int32_t ell_width = 0;          // >0: `list` is ELL

  // LDS-tiled form of a CSR list (tile_blocks > 0 selects the tile kernel):
  // block b covers slots [i0, i0+ni) of a particle order (slot s holds
  // particle tile_member[s]), whose entries are [eoff, eoff+ne) of the rows
  // stored in slot order (slot s's row at tile_slot_ptr[s]), and stages in
  // LDS the positions of the ns particles stage[soff .. soff+ns) and the ne
  // entry offsets. `local` holds, per entry of the slot-ordered rows, the LDS
  // byte offset (24 * rank) of its j in the block's position area.
  int32_t tile_blocks = 0;
  int32_t tile_first = 0;           // the launch covers blocks [tile_first, tile_first + tile_blocks)
  int32_t tile_particles = 0;       // max particles per block
  int32_t tile_lds_bytes = 0;       // max LDS bytes per block (positions + offsets)
  int32_t tile_max_stage = 0;       // max staged positions per block
  const int32_t* tile_desc = nullptr;   // 8 ints per block: i0, ni, soff, ns, eoff, ne, replica,
                                        // middle particle (member[i0 + ni/2])
  const int32_t* tile_stage = nullptr;  // staged particle indices
  const uint16_t* tile_local = nullptr; // per-entry LDS byte offsets
  const int32_t* tile_member = nullptr;   // particle of each slot; null: positions and
                                          // momenta are given in slot order
  const int32_t* tile_slot_ptr = nullptr; // row start of each slot (+1)
  const int32_t* tile_stage_slot = nullptr;  // stage as slots (spatial order; else null)
  // The offsets sliced per wavefront (dev_slice_tiles) for tile_sliced_lanes
  // lanes per particle and tile_sliced_unroll entries per batch; null: none.
  const uint16_t* tile_sliced = nullptr;
  const int32_t* tile_wave_off = nullptr;
  int32_t tile_waves_per_block = 0, tile_sliced_lanes = 0, tile_sliced_unroll = 0;

  // Batched replicas (tiled kernel only): per replica (box, 1/box), indexed
  // by descriptor slot 6; null for a single system.
  const double* replica_box = nullptr;
};

// Host-built LDS tiling of a full CSR list (see ListView).
struct HostTiles {
  int32_t tile_particles = 0;
  int32_t max_stage = 0;
  int32_t lds_bytes = 0;       // max over blocks of tile_lds_bytes(ns, ne)
  std::vector<int32_t> desc;   // 8 per block
  std::vector<int32_t> stage;
  std::vector<uint16_t> local;
};

// LDS budget of one tile: staged positions (24 bytes each, addressed by
// 16-bit byte offsets), then the tile's 16-bit entry offsets, copied in
// 16-byte chunks from the 8-entry boundary below the first entry (up to 14
// extra entries). The device copy of the offsets is padded by kTileLocalPad
// zero entries: the last chunk copy and the kernel's offset prefetch, which
// reads up to two batches (2 * 4 * 64 entries) past the end of a row, stay
// inside the allocation.
constexpr int32_t kTileLdsHeader = 16;  // kernel's own word(s) ahead of the tile data
constexpr int32_t kTileLdsBudget = 65536 - kTileLdsHeader;
constexpr int32_t kTileLocalPad = 512;
LJ_HD inline int64_t tile_lds_bytes(int64_t ns, int64_t ne) {
  return ((ns * 24 + 15) / 16) * 16 + ((ne + 14 + 7) / 8) * 16;
}

// Entry order within a tile row, chosen so that the partner reads of one
// LDS instruction spread over the banks. The staged position of rank t
// starts at dword 4 + 6t, so with 64 four-byte banks the bank pair of each of
// its 8-byte words depends on t mod 32 alone: the entry's class. With four
// lanes per particle, lane `sub` of the particle at row position r of its
// tile reads entry k = 16 * batch + 4 * u + sub together with the other 31
// lanes of its half-wavefront (8 rows). Entry k of the row at position r
// therefore takes a partner of class (4 * (r & 7) + k) & 31, or, when the row
// has none of that class left, of the next class up (cyclically) that it
// has: while the rows' classes last, the 32 lanes read 32 distinct classes.
// Entries of one class keep their list order. Rows longer than kBankRowMax
// keep the list order. Kernels with other lane counts read the same rows
// with a weaker spread; the result is the same set of pairs either way.
// The row may be the part of a tile row that starts at entry k0 of it (the
// classes then continue from there). buf holds kBankRowMax entries, cnt and
// nxt kBankClasses counts each.
constexpr int kBankClasses = 32;
constexpr int kBankRowMax = 256;
LJ_HD inline void bank_class_order(uint16_t* row, int m, int r, int k0, uint16_t* buf,
                                   uint16_t* cnt, uint16_t* nxt) {
  if (m < 2 || m > kBankRowMax) return;
  for (int c = 0; c < kBankClasses; c++) cnt[c] = 0;
  for (int k = 0; k < m; k++) cnt[(row[k] / 24) & (kBankClasses - 1)]++;
  int run = 0;
  for (int c = 0; c < kBankClasses; c++) {
    nxt[c] = static_cast<uint16_t>(run);
    run += cnt[c];
  }
  for (int k = 0; k < m; k++) {
    const uint16_t o = row[k];
    buf[nxt[(o / 24) & (kBankClasses - 1)]++] = o;
  }
  // nxt[c] is now the end of class c in buf; its next entry is nxt[c] - cnt[c].
  for (int k = 0; k < m; k++) {
    int c = (4 * (r & 7) + k0 + k) & (kBankClasses - 1);
    while (cnt[c] == 0) c = (c + 1) & (kBankClasses - 1);
    row[k] = buf[nxt[c] - cnt[c]];
    cnt[c]--;
  }
}

// Cutoff classes of a tile row. When the tiles are built from positions,
// each row holds first the entries whose pair lay within the cutoff
// (r2 <= cl2) at that time, then the others, each part in list order and
// then in the bank-class order (its classes continuing from the part's
// start). A force batch whose entries are all beyond the cutoff in every
// lane of a wavefront then skips the force arithmetic, which is the
// reference's `if (r2 > CL2) continue;` taken by whole batches; every entry's
// r2 is still evaluated at every step, so pairs that have moved inside the
// cutoff since are computed wherever they sit. r2 is the reference's
// unfused (dx*dx + dy*dy) + dz*dz over the threshold minimum image (box > 0)
// of d = q_j - q_i, the same double on the host and the device. Rows longer
// than kBankRowMax keep the list order.
LJ_HD inline bool cut_inner(double xi, double yi, double zi, double xj, double yj, double zj,
                            double box, double cl2) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double d[3] = {xj - xi, yj - yi, zj - zi};
  if (box > 0.0) {
    const double h = 0.5 * box;
    for (int c = 0; c < 3; c++) {
      if (d[c] > h) d[c] -= box;
      else if (d[c] < -h) d[c] += box;
    }
  }
  const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
  return r2 <= cl2;
}

// Reorder a row as inner entries (inner[k] true), then the others, both in
// their order; returns the inner count. buf holds kBankRowMax entries.
LJ_HD inline int cut_partition(uint16_t* row, int m, const bool* inner, uint16_t* buf) {
  int ni = 0, no = 0;
  for (int k = 0; k < m; k++) {
    const uint16_t o = row[k];
    if (inner[k]) row[ni++] = o;
    else buf[no++] = o;
  }
  for (int k = 0; k < no; k++) row[ni + k] = buf[k];
  return ni;
}

// Cell grid of the Verlet-list builders: cells at least the search length
// wide (so a pair within the search length lies in the same or adjacent
// cells), at most about 8 per particle. A periodic box with fewer than three
// cells along a side falls back to one cell holding every particle (brute).
struct CellGrid {
  int nc[3] = {1, 1, 1};
  double origin[3] = {0, 0, 0};
  double extent[3] = {0, 0, 0};
  double width[3] = {1, 1, 1};
  bool brute = false;
  int64_t ncell = 1;
};
void make_cell_grid(int32_t n, const double lo[3], const double hi[3], double box,
                    double search_length, CellGrid* g);

// True if rows are stored in order: pointer[i+1] == pointer[i] + nop[i].
bool csr_is_row_ordered(const HostCSR& c);
// Rewrite `c` so that its rows are stored in order (pointer = exclusive scan).
void csr_make_row_ordered(HostCSR* c);

// Split the particles of a row-ordered list, taken in the order `member`
// (one particle per slot; null: index order), into blocks of at most
// `tile_particles` consecutive slots that fit the LDS budget (blocks shrink
// where needed). local is indexed by the rows stored in slot order. With
// positions q (AoS double[N][4]; null: none) the rows take the cutoff
// classes (cut_inner with box, cl2) before the bank-class order. Returns
// false if a single row does not fit.
bool build_tiles(const HostCSR& full, int32_t tile_particles, const int32_t* member,
                 const double* q, double box, double cl2, HostTiles* out);

struct ForceArgs {
  bool soa = false;
  bool aos3 = false;  // AoS rows of 3 doubles (double[N][3]) instead of 4
  const double* q[3] = {nullptr, nullptr, nullptr};  // AoS: q[0] is double[N][4] (or [3])
  double* p[3] = {nullptr, nullptr, nullptr};
  double dt = 0, cl2 = 0, box = 0;
};

struct LaunchConfig {
  int lanes = 16;       // lanes per i-particle (CSR kernel); ELL lists use one
  int unroll = 2;       // list entries per lane per pipelined batch: 1, 2 or 4
  int block = 256;      // threads per workgroup
  bool tile_offsets_lds = false;  // tile kernel: stage entry offsets in LDS too
  int tile_resident = 0;     // tile kernel: persistent grid of this many workgroups per CU (0: one per tile)
  int cus = 0;               // compute units of the device
  bool half_atomic = false;  // list is a half list: scatter to j with atomics
  bool clamp = false;   // clamp every index read into range (unvalidated lists)
};

// Enqueue one force evaluation on `stream`. Returns LJGPU_OK or an error.
int launch_force_f64(void* stream, const ListView& lv, const ForceArgs& fa,
                     const LaunchConfig& lc);
int launch_force_mixed_f64acc(void* stream, const ListView& lv, const ForceArgs& fa,
                              const LaunchConfig& lc);
int launch_force_mixed_kahan(void* stream, const ListView& lv, const ForceArgs& fa,
                             const LaunchConfig& lc);
// Potential energy over an installed tiling (AoS double4 q): weight * 4 *
// sum over pairs inside cl2 of s^3 (s^3 - 1), s = 1/r2, and weight * the
// pair count, one partial per workgroup of `grid` (a multiple of 8) in
// part_u / part_c (force_f64.hip).
int launch_energy_tiles(void* stream, const double* q, const ListView& lv, double cl2,
                        double box, double weight, int grid, double* part_u, double* part_c);

// The same three precisions over packed double[N][3] AoS arrays
// (force_aos3_*.hip).
int launch_force_aos3_f64(void* stream, const ListView& lv, const ForceArgs& fa,
                          const LaunchConfig& lc);
int launch_force_aos3_mixed_f64acc(void* stream, const ListView& lv, const ForceArgs& fa,
                                   const LaunchConfig& lc);
int launch_force_aos3_mixed_kahan(void* stream, const ListView& lv, const ForceArgs& fa,
                                  const LaunchConfig& lc);

}  // namespace ljgpu
