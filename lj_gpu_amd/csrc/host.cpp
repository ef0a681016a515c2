// host.cpp — host-side setup of the ljgpu library: lattices, the Verlet-list
// build (cell list, exact reference pair set), list validation, row shuffle
// and the ELL transpose. Plain C++17, compiled with -ffp-contract=off so the
// pair predicate r2 = (dx*dx + dy*dy) + dz*dz is evaluated unfused, exactly
// as the reference's makepair writes it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "ljgpu_internal.h"

namespace ljgpu {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
void clear_error() { g_last_error.clear(); }

// Minimum image for a periodic box of side `box` (box <= 0: open box).
static inline double min_image(double d, double box) {
  if (box > 0.0) {
    const double h = 0.5 * box;
    if (d > h) d -= box;
    else if (d < -h) d += box;
  }
  return d;
}

int validate_csr(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                 int64_t n_entries, int kind) {
  if (n < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "particle_number < 0");
  if (n_entries < 0) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "n_entries < 0");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "kind must be LJGPU_LIST_HALF or LJGPU_LIST_FULL");
  if (n > 0 && (!nop || !ptr)) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null list array");
  if (n_entries > 0 && !list) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "null sorted_list");
  for (int32_t i = 0; i < n; i++) {
    const int64_t np = nop[i], kp = ptr[i];
    if (np < 0 || np >= std::max<int64_t>(n, 1))
      return set_error(LJGPU_ERR_BAD_LIST,
                       "number_of_partners[" + std::to_string(i) + "] = " + std::to_string(np));
    if (kp < 0 || kp + np > n_entries)
      return set_error(LJGPU_ERR_BAD_LIST, "pointer[" + std::to_string(i) + "] = " +
                                               std::to_string(kp) + " (row leaves the list)");
    for (int64_t k = 0; k < np; k++) {
      const int32_t j = list[kp + k];
      if (j < 0 || j >= n)
        return set_error(LJGPU_ERR_BAD_LIST, "sorted_list[" + std::to_string(kp + k) +
                                                 "] = " + std::to_string(j) + " out of range");
      if (j == i)
        return set_error(LJGPU_ERR_BAD_LIST, "row " + std::to_string(i) + " lists itself");
      if (kind == LJGPU_LIST_HALF && j < i)
        return set_error(LJGPU_ERR_BAD_LIST, "half list row " + std::to_string(i) +
                                                 " holds j = " + std::to_string(j) + " < i");
    }
  }
  return LJGPU_OK;
}

void expand_half_to_full(const HostCSR& half, HostCSR* full) {
  const int32_t n = half.n;
  full->n = n;
  std::vector<int32_t> lower(n, 0);
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < half.nop[i]; k++) lower[half.list[half.ptr[i] + k]]++;
  full->nop.assign(n, 0);
  full->ptr.assign(n, 0);
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    full->nop[i] = lower[i] + half.nop[i];
    full->ptr[i] = static_cast<int32_t>(pos);
    pos += full->nop[i];
  }
  full->list.assign(static_cast<size_t>(pos), 0);
  std::vector<int32_t> cursor(n, 0);
  for (int32_t i = 0; i < n; i++) {
    const int32_t kp = half.ptr[i];
    for (int32_t k = 0; k < half.nop[i]; k++) {
      const int32_t j = half.list[kp + k];
      full->list[full->ptr[j] + cursor[j]++] = i;  // i < j, visited in ascending i
    }
  }
  for (int32_t i = 0; i < n; i++) {
    std::memcpy(full->list.data() + full->ptr[i] + lower[i],
                half.list.data() + half.ptr[i], sizeof(int32_t) * half.nop[i]);
  }
}

bool csr_is_row_ordered(const HostCSR& c) {
  int64_t pos = 0;
  for (int32_t i = 0; i < c.n; i++) {
    if (c.ptr[i] != pos) return false;
    pos += c.nop[i];
  }
  return true;
}

void csr_make_row_ordered(HostCSR* c) {
  if (csr_is_row_ordered(*c)) return;
  std::vector<int32_t> ptr(c->n), list;
  list.reserve(c->list.size());
  for (int32_t i = 0; i < c->n; i++) {
    ptr[i] = static_cast<int32_t>(list.size());
    list.insert(list.end(), c->list.begin() + c->ptr[i], c->list.begin() + c->ptr[i] + c->nop[i]);
  }
  c->ptr.swap(ptr);
  c->list.swap(list);
}

bool build_tiles(const HostCSR& full, int32_t tile_particles, const int32_t* member,
                 const double* q, double box, double cl2, HostTiles* out) {
  const int32_t n = full.n;
  out->tile_particles = tile_particles;
  out->max_stage = 0;
  out->lds_bytes = 0;
  out->desc.clear();
  out->stage.clear();
  out->local.assign(full.list.size(), 0);
  if (tile_particles <= 0 || !csr_is_row_ordered(full)) return false;
  auto mem = [&](int32_t s) { return member ? member[s] : s; };
  // Blocks key their particles by slot.
  std::vector<int32_t> slot_of(static_cast<size_t>(n));
  for (int32_t s = 0; s < n; s++) slot_of[mem(s)] = s;
  // Row start of each slot in the rows stored in slot order.
  std::vector<int64_t> sptr(static_cast<size_t>(n) + 1, 0);
  for (int32_t s = 0; s < n; s++) sptr[s + 1] = sptr[s] + full.nop[mem(s)];
  auto entries = [&](int32_t i0, int32_t ni) -> int64_t { return sptr[i0 + ni] - sptr[i0]; };
  std::vector<int32_t> set;
  auto collect = [&](int32_t i0, int32_t ni) {  // the block's slots, ascending
    set.clear();
    for (int32_t s = i0; s < i0 + ni; s++) {
      const int32_t i = mem(s);
      set.push_back(s);
      const int32_t* r = full.list.data() + full.ptr[i];
      for (int32_t k = 0; k < full.nop[i]; k++) set.push_back(slot_of[r[k]]);
    }
    std::sort(set.begin(), set.end());
    set.erase(std::unique(set.begin(), set.end()), set.end());
  };
  int32_t i0 = 0;
  while (i0 < n) {
    int32_t ni = std::min(tile_particles, n - i0);
    collect(i0, ni);
    while (tile_lds_bytes(static_cast<int64_t>(set.size()), entries(i0, ni)) > kTileLdsBudget &&
           ni > 1) {
      ni = std::max(1, ni / 2);
      collect(i0, ni);
    }
    const int64_t ne = entries(i0, ni);
    if (tile_lds_bytes(static_cast<int64_t>(set.size()), ne) > kTileLdsBudget) return false;
    const int32_t soff = static_cast<int32_t>(out->stage.size());
    const int32_t ns = static_cast<int32_t>(set.size());
    out->desc.insert(out->desc.end(), {i0, ni, soff, ns, static_cast<int32_t>(sptr[i0]),
                                       static_cast<int32_t>(ne), 0, mem(i0 + ni / 2)});
    for (int32_t v : set) out->stage.push_back(mem(v));
    out->max_stage = std::max(out->max_stage, ns);
    out->lds_bytes = std::max(out->lds_bytes, static_cast<int32_t>(tile_lds_bytes(ns, ne)));
    for (int32_t s = i0; s < i0 + ni; s++) {
      const int32_t i = mem(s);
      const int32_t kp = full.ptr[i];
      for (int32_t k = 0; k < full.nop[i]; k++) {
        const int32_t j = slot_of[full.list[kp + k]];
        const int32_t rank = static_cast<int32_t>(
            std::lower_bound(set.begin(), set.end(), j) - set.begin());
        out->local[sptr[s] + k] = static_cast<uint16_t>(rank * 24);
      }
      uint16_t buf[kBankRowMax], cnt[kBankClasses], nxt[kBankClasses];
      uint16_t* row = out->local.data() + sptr[s];
      const int m = full.nop[i];
      int ni = 0;
      if (q && m <= kBankRowMax) {
        bool inner[kBankRowMax];
        for (int k = 0; k < m; k++) {
          const int32_t j = full.list[kp + k];
          inner[k] = cut_inner(q[4 * static_cast<int64_t>(i)], q[4 * static_cast<int64_t>(i) + 1],
                               q[4 * static_cast<int64_t>(i) + 2], q[4 * static_cast<int64_t>(j)],
                               q[4 * static_cast<int64_t>(j) + 1],
                               q[4 * static_cast<int64_t>(j) + 2], box, cl2);
        }
        ni = cut_partition(row, m, inner, buf);
      }
      bank_class_order(row, ni, s - i0, 0, buf, cnt, nxt);
      bank_class_order(row + ni, m - ni, s - i0, ni, buf, cnt, nxt);
    }
    i0 += ni;
  }
  return true;
}

void make_cell_grid(int32_t n, const double lo[3], const double hi[3], double box,
                    double search_length, CellGrid* g) {
  g->brute = false;
  for (int c = 0; c < 3; c++) {
    const double extent = box > 0.0 ? box : (hi[c] - lo[c]);
    g->origin[c] = box > 0.0 ? 0.0 : lo[c];
    g->extent[c] = extent;
    // Cells are kept a hair wider than the search length so that rounding
    // in x / width can never put two particles closer than the search
    // length two cells apart.
    int m = static_cast<int>(extent / (search_length * (1.0 + 1e-9)));
    g->nc[c] = std::min(std::max(m, 1), 1024);
  }
  // At most a few cells per particle: coarser cells stay correct (wider than
  // the search length) and keep the cell arrays proportional to N.
  const int64_t limit = 8 * static_cast<int64_t>(std::max(n, 1)) + 64;
  while (static_cast<int64_t>(g->nc[0]) * g->nc[1] * g->nc[2] > limit) {
    int* m = std::max_element(g->nc, g->nc + 3);
    *m = (*m + 1) / 2;
  }
  for (int c = 0; c < 3; c++) {
    g->width[c] = g->extent[c] > 0.0 ? g->extent[c] / g->nc[c] : 1.0;
    if (box > 0.0 && g->nc[c] < 3) g->brute = true;
  }
  if (g->brute) g->nc[0] = g->nc[1] = g->nc[2] = 1;
  g->ncell = static_cast<int64_t>(g->nc[0]) * g->nc[1] * g->nc[2];
}

}  // namespace ljgpu

using namespace ljgpu;

extern "C" {

const char* ljgpu_last_error(void) { return g_last_error.c_str(); }
const char* ljgpu_version(void) { return "ljgpu 0.1 (gfx950)"; }

// cpu_ref/force_aos.cpp:43-51,142-169: FCC lattice, open box, U(0,0.1) jitter.
int ljgpu_lattice_reference(double density, double box_length, double* q, int32_t capacity,
                            int32_t* particle_number) {
  if (!(density > 0.0) || !(box_length > 0.0) || !q || !particle_number)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "lattice_reference: bad argument");
  const double s = 1.0 / std::pow(density * 0.25, 1.0 / 3.0);
  const double hs = s * 0.5;
  const int sx = static_cast<int>(box_length / s);
  const int64_t total = 4LL * sx * sx * sx;
  if (total > capacity) {
    *particle_number = static_cast<int32_t>(std::min<int64_t>(total, INT32_MAX));
    return set_error(LJGPU_ERR_CAPACITY, "lattice_reference: capacity too small");
  }
  std::mt19937 mt(2);
  std::uniform_real_distribution<double> ud(0.0, 0.1);
  int32_t n = 0;
  auto add = [&](double x, double y, double z) {
    const double jx = ud(mt);
    const double jy = ud(mt);
    const double jz = ud(mt);
    double* r = q + 4 * static_cast<int64_t>(n);
    r[0] = x + jx;
    r[1] = y + jy;
    r[2] = z + jz;
    r[3] = 0.0;
    n++;
  };
  for (int iz = 0; iz < sx; iz++)
    for (int iy = 0; iy < sx; iy++)
      for (int ix = 0; ix < sx; ix++) {
        const double x = ix * s, y = iy * s, z = iz * s;
        add(x, y, z);
        add(x, y + hs, z + hs);
        add(x + hs, y, z + hs);
        add(x + hs, y + hs, z);
      }
  *particle_number = n;
  return LJGPU_OK;
}

int ljgpu_lattice_periodic(int32_t cells, double density, uint32_t seed, double jitter_lo,
                           double jitter_hi, double* q, int32_t capacity, double* box) {
  if (cells <= 0 || !(density > 0.0) || !q || !box || !(jitter_hi >= jitter_lo))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "lattice_periodic: bad argument");
  const int64_t total = 4LL * cells * cells * cells;
  if (total > capacity) return set_error(LJGPU_ERR_CAPACITY, "lattice_periodic: capacity too small");
  const double s = 1.0 / std::pow(density * 0.25, 1.0 / 3.0);
  const double hs = s * 0.5;
  const double L = cells * s;
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> ud(jitter_lo, jitter_hi);
  auto wrap = [L](double v) {
    if (v < 0.0) v += L;
    if (v >= L) v -= L;
    return v;
  };
  int64_t n = 0;
  auto add = [&](double x, double y, double z) {
    const double jx = ud(mt);
    const double jy = ud(mt);
    const double jz = ud(mt);
    double* r = q + 4 * n;
    r[0] = wrap(x + jx);
    r[1] = wrap(y + jy);
    r[2] = wrap(z + jz);
    r[3] = 0.0;
    n++;
  };
  for (int iz = 0; iz < cells; iz++)
    for (int iy = 0; iy < cells; iy++)
      for (int ix = 0; ix < cells; ix++) {
        const double x = ix * s, y = iy * s, z = iz * s;
        add(x, y, z);
        add(x, y + hs, z + hs);
        add(x + hs, y, z + hs);
        add(x + hs, y + hs, z);
      }
  *box = L;
  return LJGPU_OK;
}

// Cell-list Verlet build with the reference's exact pair set and CSR order
// (cpu_ref/force_aos.cpp:100-140; cuda/force_cuda.cu:122-163).
int ljgpu_makepair_host(int32_t n, const double* q, double box, double search_length, int kind,
                        int32_t* nop, int32_t* pointer, int32_t* sorted_list, int64_t capacity,
                        int64_t* n_entries) {
  if (n < 0 || (n > 0 && (!q || !nop || !pointer)) || !n_entries || !(search_length > 0.0))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "makepair_host: bad argument");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "makepair_host: bad kind");
  if (box > 0.0 && box < 2.0 * search_length)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT,
                     "makepair_host: periodic box shorter than twice the search length");
  const double sl2 = search_length * search_length;
  *n_entries = 0;
  if (n == 0) return LJGPU_OK;

  // Cell grid: cells of side >= search_length; a periodic box needs at least
  // three cells per side for the 27-cell stencil to visit each cell once,
  // otherwise every particle is a candidate (brute force).
  double lo[3], hi[3];
  for (int c = 0; c < 3; c++) {
    lo[c] = hi[c] = q[c];
  }
  for (int32_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min(lo[c], q[4 * (int64_t)i + c]);
      hi[c] = std::max(hi[c], q[4 * (int64_t)i + c]);
    }
  CellGrid g;
  make_cell_grid(n, lo, hi, box, search_length, &g);
  const int* nc = g.nc;
  const double* origin = g.origin;
  const double* width = g.width;
  const bool brute = g.brute;
  const int64_t ncell = static_cast<int64_t>(nc[0]) * nc[1] * nc[2];
  std::vector<int32_t> cell_of(n), cell_start(ncell + 1, 0), cell_items(n);
  for (int32_t i = 0; i < n; i++) {
    int ci[3];
    for (int c = 0; c < 3; c++) {
      double x = q[4 * (int64_t)i + c] - origin[c];
      if (box > 0.0) x -= box * std::floor(x / box);  // cell of the wrapped image
      int v = static_cast<int>(std::floor(x / width[c]));
      ci[c] = std::min(std::max(v, 0), nc[c] - 1);
    }
    cell_of[i] = static_cast<int32_t>((static_cast<int64_t>(ci[2]) * nc[1] + ci[1]) * nc[0] + ci[0]);
    cell_start[cell_of[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cell_start[c + 1] += cell_start[c];
  {
    std::vector<int32_t> fill(cell_start.begin(), cell_start.end() - 1);
    for (int32_t i = 0; i < n; i++) cell_items[fill[cell_of[i]]++] = i;  // ascending i per cell
  }

  // Rows are built in parallel into per-row vectors, then concatenated in i
  // order, so the CSR does not depend on the thread count.
  std::vector<std::vector<int32_t>> rows(n);
  auto build_rows = [&](int32_t i0, int32_t i1) {
    std::vector<int32_t> cand;
    for (int32_t i = i0; i < i1; i++) {
      const double xi = q[4 * (int64_t)i], yi = q[4 * (int64_t)i + 1], zi = q[4 * (int64_t)i + 2];
      std::vector<int32_t>& row = rows[i];
      cand.clear();
      if (brute) {
        for (int32_t j = 0; j < n; j++) cand.push_back(j);
      } else {
        const int32_t c = cell_of[i];
        const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
        for (int dz = -1; dz <= 1; dz++) {
          int z = cz + dz;
          if (box > 0.0) z = (z + nc[2]) % nc[2];
          else if (z < 0 || z >= nc[2]) continue;
          for (int dy = -1; dy <= 1; dy++) {
            int y = cy + dy;
            if (box > 0.0) y = (y + nc[1]) % nc[1];
            else if (y < 0 || y >= nc[1]) continue;
            for (int dx = -1; dx <= 1; dx++) {
              int x = cx + dx;
              if (box > 0.0) x = (x + nc[0]) % nc[0];
              else if (x < 0 || x >= nc[0]) continue;
              const int64_t cc = (static_cast<int64_t>(z) * nc[1] + y) * nc[0] + x;
              for (int32_t t = cell_start[cc]; t < cell_start[cc + 1]; t++)
                cand.push_back(cell_items[t]);
            }
          }
        }
      }
      for (int32_t j : cand) {
        if (kind == LJGPU_LIST_HALF ? (j <= i) : (j == i)) continue;
        const double dx = min_image(xi - q[4 * (int64_t)j], box);
        const double dy = min_image(yi - q[4 * (int64_t)j + 1], box);
        const double dz = min_image(zi - q[4 * (int64_t)j + 2], box);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < sl2) row.push_back(j);
      }
      std::sort(row.begin(), row.end());
    }
  };
  unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nthreads = 1;
  std::vector<std::thread> pool;
  const int32_t chunk = (n + nthreads - 1) / nthreads;
  for (unsigned t = 0; t < nthreads; t++) {
    const int32_t a = static_cast<int32_t>(std::min<int64_t>((int64_t)t * chunk, n));
    const int32_t b = static_cast<int32_t>(std::min<int64_t>((int64_t)a + chunk, n));
    if (a < b) pool.emplace_back(build_rows, a, b);
  }
  for (auto& th : pool) th.join();

  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    nop[i] = static_cast<int32_t>(rows[i].size());
    pointer[i] = static_cast<int32_t>(std::min<int64_t>(pos, INT32_MAX));
    if (sorted_list && pos + static_cast<int64_t>(rows[i].size()) <= capacity)
      std::memcpy(sorted_list + pos, rows[i].data(), rows[i].size() * sizeof(int32_t));
    pos += static_cast<int64_t>(rows[i].size());
  }
  *n_entries = pos;
  if (pos > capacity || pos > INT32_MAX)
    return set_error(LJGPU_ERR_CAPACITY, "makepair_host: list exceeds capacity (" +
                                             std::to_string(pos) + " entries)");
  return LJGPU_OK;
}

int ljgpu_maxwell_boltzmann(int32_t n, double temperature, uint32_t seed, double* p) {
  if (n < 0 || (n > 0 && !p) || !(temperature >= 0.0))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "maxwell_boltzmann: bad argument");
  std::mt19937 mt(seed);
  std::normal_distribution<double> nd(0.0, std::sqrt(temperature));
  double mean[3] = {0.0, 0.0, 0.0};
  for (int32_t i = 0; i < n; i++) {
    for (int c = 0; c < 3; c++) {
      p[4 * (int64_t)i + c] = nd(mt);
      mean[c] += p[4 * (int64_t)i + c];
    }
    p[4 * (int64_t)i + 3] = 0.0;
  }
  for (int c = 0; c < 3; c++) mean[c] = n > 0 ? mean[c] / n : 0.0;
  for (int32_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) p[4 * (int64_t)i + c] -= mean[c];
  return LJGPU_OK;
}

int ljgpu_validate_csr(int32_t n, const int32_t* nop, const int32_t* ptr,
                       const int32_t* list, int64_t n_entries, int kind) {
  return validate_csr(n, nop, ptr, list, n_entries, kind);
}

// cuda/force_cuda.cu:255-263.
int ljgpu_random_shfl(int32_t n, const int32_t* nop, const int32_t* ptr, int32_t* list,
                      uint32_t seed) {
  if (n < 0 || (n > 0 && (!nop || !ptr || !list)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "random_shfl: bad argument");
  std::mt19937 mt(seed);
  for (int32_t i = 0; i < n; i++) std::shuffle(list + ptr[i], list + ptr[i] + nop[i], mt);
  return LJGPU_OK;
}

int ljgpu_make_tiles_cut(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                         int64_t n_entries, int32_t tile_particles, const int32_t* member,
                         const double* q, double box, double cl2, int32_t* desc,
                         int64_t desc_capacity, int32_t* stage, int64_t stage_capacity,
                         uint16_t* local, int32_t* n_blocks, int64_t* n_stage,
                         int32_t* max_stage) {
  if (!n_blocks || !n_stage || !max_stage || tile_particles <= 0 || tile_particles > 1024)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "make_tiles: bad argument");
  if (q && !(cl2 >= 0.0))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "make_tiles: cl2 must be >= 0");
  int rc = validate_csr(n, nop, ptr, list, n_entries, LJGPU_LIST_FULL);
  if (rc) return rc;
  if (member) {
    std::vector<char> seen(static_cast<size_t>(n), 0);
    for (int32_t s = 0; s < n; s++) {
      if (member[s] < 0 || member[s] >= n || seen[member[s]])
        return set_error(LJGPU_ERR_INVALID_ARGUMENT, "make_tiles: member is not a permutation");
      seen[member[s]] = 1;
    }
  }
  HostCSR c;
  c.n = n;
  c.nop.assign(nop, nop + n);
  c.ptr.assign(ptr, ptr + n);
  c.list.assign(list, list + n_entries);
  if (!csr_is_row_ordered(c))
    return set_error(LJGPU_ERR_BAD_LIST, "make_tiles: rows must be stored in order");
  HostTiles t;
  if (!build_tiles(c, tile_particles, member, q, box, cl2, &t))
    return set_error(LJGPU_ERR_CAPACITY, "make_tiles: a row does not fit one tile");
  *n_blocks = static_cast<int32_t>(t.desc.size() / 8);
  *n_stage = static_cast<int64_t>(t.stage.size());
  *max_stage = t.max_stage;
  if (!desc && !stage && !local) return LJGPU_OK;
  if (!desc || !stage || !local || desc_capacity < static_cast<int64_t>(t.desc.size()) ||
      stage_capacity < static_cast<int64_t>(t.stage.size()))
    return set_error(LJGPU_ERR_CAPACITY, "make_tiles: output too small");
  std::memcpy(desc, t.desc.data(), t.desc.size() * sizeof(int32_t));
  std::memcpy(stage, t.stage.data(), t.stage.size() * sizeof(int32_t));
  std::memcpy(local, t.local.data(), t.local.size() * sizeof(uint16_t));
  return LJGPU_OK;
}

int ljgpu_make_tiles_ordered(int32_t n, const int32_t* nop, const int32_t* ptr,
                             const int32_t* list, int64_t n_entries, int32_t tile_particles,
                             const int32_t* member, int32_t* desc, int64_t desc_capacity,
                             int32_t* stage, int64_t stage_capacity, uint16_t* local,
                             int32_t* n_blocks, int64_t* n_stage, int32_t* max_stage) {
  return ljgpu_make_tiles_cut(n, nop, ptr, list, n_entries, tile_particles, member, nullptr, 0.0,
                              0.0, desc, desc_capacity, stage, stage_capacity, local, n_blocks,
                              n_stage, max_stage);
}

int ljgpu_make_tiles(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                     int64_t n_entries, int32_t tile_particles, int32_t* desc,
                     int64_t desc_capacity, int32_t* stage, int64_t stage_capacity,
                     uint16_t* local, int32_t* n_blocks, int64_t* n_stage, int32_t* max_stage) {
  return ljgpu_make_tiles_ordered(n, nop, ptr, list, n_entries, tile_particles, nullptr, desc,
                                  desc_capacity, stage, stage_capacity, local, n_blocks, n_stage,
                                  max_stage);
}

// cuda/force_cuda.cu:229-240.
int ljgpu_make_transposed(int32_t n, const int32_t* nop, const int32_t* ptr,
                          const int32_t* list, int32_t width, int32_t* out) {
  if (n < 0 || width < 0 || (n > 0 && (!nop || !ptr || !out)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "make_transposed: bad argument");
  for (int32_t i = 0; i < n; i++)
    if (nop[i] > width) return set_error(LJGPU_ERR_CAPACITY, "make_transposed: row longer than width");
  std::memset(out, 0, sizeof(int32_t) * static_cast<size_t>(n) * width);
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < nop[i]; k++) out[i + static_cast<int64_t>(k) * n] = list[ptr[i] + k];
  return LJGPU_OK;
}

}  // extern "C"
