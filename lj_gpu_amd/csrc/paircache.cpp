// paircache.cpp — the reference's two on-disk pair-list caches, read and
// written bit-compatibly so lists can be exchanged with the reference's
// programs:
//
//   pair.dat           binary, cpu_ref/force_aos.cpp:603-620 (loadpair /
//                      savepair; same in cpu_ref/force_soa.cpp:360-376):
//                        int32 number_of_pairs
//                        int32 number_of_partners[N_CAP]
//                        int32 i_particles[PAIR_CAP]
//                        int32 j_particles[PAIR_CAP]
//                      where N_CAP and PAIR_CAP are the reference's
//                      compile-time array sizes (400000 and 30 * 400000);
//                      unused tails are zero. Pairs are (i < j) in the order
//                      makepair() registered them; sortpair() rebuilds the CSR.
//   .cache_pair_*.dat  text, cuda/force_cuda.cu:165-227 (makepaircache /
//                      loadpair): "N E\n", N lines "number_of_partners
//                      pointer\n", then E lines "sorted_list\n".
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ljgpu_internal.h"

using ljgpu::set_error;

namespace {

constexpr int32_t kRefParticleCapacity = 400000;  // cpu_ref/force_aos.cpp:11
constexpr int64_t kRefPairsPerParticle = 30;      // cpu_ref/force_aos.cpp:12

struct File {
  FILE* f = nullptr;
  explicit File(FILE* fp) : f(fp) {}
  ~File() {
    if (f) std::fclose(f);
  }
  File(const File&) = delete;
  File& operator=(const File&) = delete;
};

int io_error(const char* what, const char* path) {
  return set_error(LJGPU_ERR_IO, std::string(what) + " '" + (path ? path : "(null)") +
                                     "': " + std::strerror(errno));
}

void resolve_caps(int32_t* n_cap, int64_t* pair_cap) {
  if (*n_cap <= 0) *n_cap = kRefParticleCapacity;
  if (*pair_cap <= 0) *pair_cap = kRefPairsPerParticle * static_cast<int64_t>(*n_cap);
}

// Write `count` int32 values of which the first `valid` come from `src` and
// the rest are zero (the reference writes its whole zero-initialised array).
bool write_padded(FILE* f, const int32_t* src, int64_t valid, int64_t count) {
  if (valid > 0 && std::fwrite(src, sizeof(int32_t), static_cast<size_t>(valid), f) !=
                       static_cast<size_t>(valid))
    return false;
  static const int32_t zeros[4096] = {};
  for (int64_t left = count - valid; left > 0;) {
    const size_t m = static_cast<size_t>(std::min<int64_t>(left, 4096));
    if (std::fwrite(zeros, sizeof(int32_t), m, f) != m) return false;
    left -= static_cast<int64_t>(m);
  }
  return true;
}

// A minimal reader of whitespace-separated decimal integers (the text cache
// is written with "%d"); fails on anything else.
struct IntReader {
  std::vector<char> buf;
  size_t pos = 0;
  bool next(int64_t* v) {
    while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\r' ||
                                buf[pos] == '\t'))
      pos++;
    if (pos >= buf.size()) return false;
    bool neg = false;
    if (buf[pos] == '-' || buf[pos] == '+') {
      neg = buf[pos] == '-';
      pos++;
    }
    if (pos >= buf.size() || buf[pos] < '0' || buf[pos] > '9') return false;
    int64_t x = 0;
    int digits = 0;
    while (pos < buf.size() && buf[pos] >= '0' && buf[pos] <= '9') {
      x = x * 10 + (buf[pos] - '0');
      if (++digits > 12) return false;
      pos++;
    }
    *v = neg ? -x : x;
    return true;
  }
  bool next32(int32_t* v) {
    int64_t x;
    if (!next(&x) || x < INT32_MIN || x > INT32_MAX) return false;
    *v = static_cast<int32_t>(x);
    return true;
  }
};

}  // namespace

extern "C" {

int ljgpu_pair_dat_write(const char* path, int32_t n, const int32_t* nop, const int32_t* ptr,
                         const int32_t* list, int64_t n_entries, int32_t n_capacity,
                         int64_t pair_capacity) {
  if (!path) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_dat_write: null path");
  resolve_caps(&n_capacity, &pair_capacity);
  int rc = ljgpu::validate_csr(n, nop, ptr, list, n_entries, LJGPU_LIST_HALF);
  if (rc) return rc;
  if (n > n_capacity || n_entries > pair_capacity || n_entries > INT32_MAX)
    return set_error(LJGPU_ERR_CAPACITY,
                     "pair_dat_write: list larger than the file's fixed arrays (N " +
                         std::to_string(n) + " of " + std::to_string(n_capacity) + ", pairs " +
                         std::to_string(n_entries) + " of " + std::to_string(pair_capacity) + ")");
  // Pairs in CSR storage order: row by row, in each row's stored order (for
  // a list built by makepair+sortpair this is makepair's registration order).
  std::vector<int32_t> ip(static_cast<size_t>(n_entries)), jp(static_cast<size_t>(n_entries));
  int64_t k = 0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t t = 0; t < nop[i]; t++, k++) {
      ip[k] = i;
      jp[k] = list[static_cast<int64_t>(ptr[i]) + t];
    }
  if (k != n_entries)
    return set_error(LJGPU_ERR_BAD_LIST,
                     "pair_dat_write: rows hold " + std::to_string(k) + " entries, list has " +
                         std::to_string(n_entries));
  File f(std::fopen(path, "wb"));
  if (!f.f) return io_error("pair_dat_write: cannot open", path);
  const int32_t np32 = static_cast<int32_t>(n_entries);
  if (std::fwrite(&np32, sizeof(int32_t), 1, f.f) != 1 ||
      !write_padded(f.f, nop, n, n_capacity) || !write_padded(f.f, ip.data(), n_entries, pair_capacity) ||
      !write_padded(f.f, jp.data(), n_entries, pair_capacity))
    return io_error("pair_dat_write: write failed", path);
  if (std::fclose(f.f) != 0) {
    f.f = nullptr;
    return io_error("pair_dat_write: close failed", path);
  }
  f.f = nullptr;
  return LJGPU_OK;
}

int ljgpu_pair_dat_read(const char* path, int32_t n, int32_t n_capacity, int64_t pair_capacity,
                        int32_t* nop, int32_t* ptr, int32_t* list, int64_t capacity,
                        int64_t* n_entries) {
  if (!path || !n_entries || n < 0 || (n > 0 && (!nop || !ptr)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_dat_read: bad argument");
  resolve_caps(&n_capacity, &pair_capacity);
  if (n > n_capacity)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_dat_read: particle_number exceeds the file's capacity");
  File f(std::fopen(path, "rb"));
  if (!f.f) return io_error("pair_dat_read: cannot open", path);
  const int64_t expect = 4 * (1 + static_cast<int64_t>(n_capacity) + 2 * pair_capacity);
  if (std::fseek(f.f, 0, SEEK_END) != 0) return io_error("pair_dat_read: seek failed", path);
  const long size = std::ftell(f.f);
  if (size != expect)
    return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: file holds " + std::to_string(size) +
                                             " bytes, expected " + std::to_string(expect) +
                                             " for the given capacities");
  std::rewind(f.f);
  int32_t npairs = 0;
  std::vector<int32_t> counts(static_cast<size_t>(n_capacity));
  if (std::fread(&npairs, sizeof(int32_t), 1, f.f) != 1 ||
      std::fread(counts.data(), sizeof(int32_t), counts.size(), f.f) != counts.size())
    return io_error("pair_dat_read: read failed", path);
  if (npairs < 0 || npairs > pair_capacity)
    return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: number_of_pairs = " + std::to_string(npairs));
  int64_t total = 0;
  for (int32_t i = 0; i < n_capacity; i++) {
    if (counts[i] < 0 || (i >= n && counts[i] != 0))
      return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: number_of_partners[" +
                                               std::to_string(i) + "] = " +
                                               std::to_string(counts[i]) +
                                               " (file is for another system size?)");
    total += counts[i];
  }
  if (total != npairs)
    return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: partner counts sum to " +
                                             std::to_string(total) + ", number_of_pairs is " +
                                             std::to_string(npairs));
  *n_entries = npairs;
  for (int32_t i = 0; i < n; i++) nop[i] = counts[i];
  // sortpair(): pointer = exclusive scan of the counts (cpu_ref/force_aos.cpp:100-120).
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    ptr[i] = static_cast<int32_t>(pos);
    pos += counts[i];
  }
  if (!list) return LJGPU_OK;  // size query
  if (capacity < npairs)
    return set_error(LJGPU_ERR_CAPACITY, "pair_dat_read: sorted_list capacity " +
                                             std::to_string(capacity) + " < " + std::to_string(npairs));
  std::vector<int32_t> ip(static_cast<size_t>(npairs)), jp(static_cast<size_t>(npairs));
  if (std::fread(ip.data(), sizeof(int32_t), ip.size(), f.f) != ip.size() ||
      std::fseek(f.f, static_cast<long>(4 * (1 + static_cast<int64_t>(n_capacity) + pair_capacity)),
                 SEEK_SET) != 0 ||
      std::fread(jp.data(), sizeof(int32_t), jp.size(), f.f) != jp.size())
    return io_error("pair_dat_read: read failed", path);
  // sortpair(): place each pair at pointer[i] + (pairs of row i seen so far).
  std::vector<int32_t> fill(static_cast<size_t>(n), 0);
  for (int64_t k = 0; k < npairs; k++) {
    const int32_t i = ip[k], j = jp[k];
    if (i < 0 || i >= n || j <= i || j >= n)
      return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: pair " + std::to_string(k) + " = (" +
                                               std::to_string(i) + ", " + std::to_string(j) +
                                               ") is not a half-list pair of this system");
    if (fill[i] >= counts[i])
      return set_error(LJGPU_ERR_BAD_LIST, "pair_dat_read: row " + std::to_string(i) +
                                               " holds more pairs than number_of_partners");
    list[ptr[i] + fill[i]++] = j;
  }
  return LJGPU_OK;
}

int ljgpu_pair_cache_write(const char* path, int32_t n, const int32_t* nop, const int32_t* ptr,
                           const int32_t* list, int64_t n_entries) {
  if (!path) return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_cache_write: null path");
  int rc = ljgpu::validate_csr(n, nop, ptr, list, n_entries, LJGPU_LIST_FULL);
  if (rc) return rc;
  if (n_entries > INT32_MAX)
    return set_error(LJGPU_ERR_CAPACITY, "pair_cache_write: more than 2^31-1 entries");
  File f(std::fopen(path, "w"));
  if (!f.f) return io_error("pair_cache_write: cannot open", path);
  bool ok = std::fprintf(f.f, "%d %d\n", n, static_cast<int32_t>(n_entries)) > 0;
  for (int32_t i = 0; ok && i < n; i++) ok = std::fprintf(f.f, "%d %d\n", nop[i], ptr[i]) > 0;
  for (int64_t k = 0; ok && k < n_entries; k++) ok = std::fprintf(f.f, "%d\n", list[k]) > 0;
  if (!ok) return io_error("pair_cache_write: write failed", path);
  if (std::fclose(f.f) != 0) {
    f.f = nullptr;
    return io_error("pair_cache_write: close failed", path);
  }
  f.f = nullptr;
  return LJGPU_OK;
}

int ljgpu_pair_cache_read(const char* path, int32_t n, int kind, int32_t* nop, int32_t* ptr,
                          int32_t* list, int64_t capacity, int64_t* n_entries) {
  if (!path || !n_entries || n < 0 || (n > 0 && (!nop || !ptr)))
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_cache_read: bad argument");
  if (kind != LJGPU_LIST_HALF && kind != LJGPU_LIST_FULL)
    return set_error(LJGPU_ERR_INVALID_ARGUMENT, "pair_cache_read: bad kind");
  IntReader r;
  {
    File f(std::fopen(path, "rb"));
    if (!f.f) return io_error("pair_cache_read: cannot open", path);
    if (std::fseek(f.f, 0, SEEK_END) != 0) return io_error("pair_cache_read: seek failed", path);
    const long size = std::ftell(f.f);
    if (size < 0) return io_error("pair_cache_read: tell failed", path);
    std::rewind(f.f);
    r.buf.resize(static_cast<size_t>(size));
    if (size > 0 && std::fread(r.buf.data(), 1, r.buf.size(), f.f) != r.buf.size())
      return io_error("pair_cache_read: read failed", path);
  }
  int32_t file_n = 0, npairs = 0;
  if (!r.next32(&file_n) || !r.next32(&npairs))
    return set_error(LJGPU_ERR_BAD_LIST, "pair_cache_read: missing header");
  // cuda/force_cuda.cu:211-214: a cache for another particle number is refused.
  if (file_n != n)
    return set_error(LJGPU_ERR_BAD_LIST, "pair_cache_read: cache holds " + std::to_string(file_n) +
                                             " particles, expected " + std::to_string(n) +
                                             " (pairlist cache data may be broken)");
  if (npairs < 0) return set_error(LJGPU_ERR_BAD_LIST, "pair_cache_read: negative pair count");
  for (int32_t i = 0; i < n; i++)
    if (!r.next32(&nop[i]) || !r.next32(&ptr[i]))
      return set_error(LJGPU_ERR_BAD_LIST,
                       "pair_cache_read: truncated or malformed row line " + std::to_string(i));
  *n_entries = npairs;
  if (!list) return LJGPU_OK;  // size query
  if (capacity < npairs)
    return set_error(LJGPU_ERR_CAPACITY, "pair_cache_read: sorted_list capacity " +
                                             std::to_string(capacity) + " < " + std::to_string(npairs));
  for (int64_t k = 0; k < npairs; k++)
    if (!r.next32(&list[k]))
      return set_error(LJGPU_ERR_BAD_LIST,
                       "pair_cache_read: truncated or malformed list line " + std::to_string(k));
  int64_t trailing;
  if (r.next(&trailing))
    return set_error(LJGPU_ERR_BAD_LIST, "pair_cache_read: data after the last list entry");
  // check_loadedpair() (cuda/force_cuda.cu:183-201), strict form.
  return ljgpu::validate_csr(n, nop, ptr, list, npairs, kind);
}

}  // extern "C"
