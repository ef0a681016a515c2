// lists.hip — neighbour lists on the device (see lists.h).
//
// The Verlet-list build follows the reference's makepair()/sortpair()
// (cpu_ref/force_aos.cpp:100-140; full list cuda/force_cuda.cu:122-163): the
// same pair predicate, evaluated unfused so every pair decision is
// bit-identical to the reference's, and the same CSR (rows ascending in i,
// partners ascending in j). The O(N^2) pair scan is replaced by a cell list:
//   1. bin particles into cells at least half the search length wide (a
//      radius-2 stencil; the search length and radius 1 for small boxes);
//   2. search each cell's stencil with one wavefront per cell, writing rows
//      into a per-row slab (an exact count-then-fill pass when a row
//      overflows it), and scan the row lengths into row pointers;
//   3. sort every row into place (or, for the MD loop, copy rows in the
//      build's own, position-determined order).
// Derived layouts (full list from a half list, ELL, LDS tiles, and the
// spatial particle order the tiles may follow) are built here from the
// device CSR, so a list never round-trips through the host.
#include "lists.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

namespace ljgpu {

namespace {

int hip_err(hipError_t e, const char* what) {
  return set_error(e == hipErrorOutOfMemory ? LJGPU_ERR_OUT_OF_MEMORY : LJGPU_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define LTRY(expr, what)                          \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return hip_err(e_, what); \
  } while (0)

#define LCHECK(call)        \
  do {                      \
    int rc_ = (call);       \
    if (rc_) return rc_;    \
  } while (0)

constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

int grid_for(int64_t items, int block, int cap = 65535 * 8) {
  const int64_t g = (items + block - 1) / block;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g, cap)));
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_fill(T* __restrict__ a, int64_t n, T v) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[t] = v;
}

template <class T>
int fill(hipStream_t s, T* a, int64_t n, T v) {
  if (n <= 0) return LJGPU_OK;
  k_fill<T><<<grid_for(n, kBlock), kBlock, 0, s>>>(a, n, v);
  LTRY(hipGetLastError(), "fill kernel");
  return LJGPU_OK;
}

// Inclusive scan of one value per lane over a wavefront.
__device__ __forceinline__ int64_t wave_inclusive(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread over a kBlock-thread workgroup;
// *total receives the workgroup sum. `sh` holds kBlock/64 values.
__device__ __forceinline__ int64_t block_exclusive(int64_t v, int64_t* total, int64_t* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t inc = wave_inclusive(v);
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  int64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; w++) {
    const int64_t x = sh[w];
    off += (w < wid) ? x : 0;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Scan pass 1: per-tile sums.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int64_t* __restrict__ sums) {
  __shared__ int64_t sh[kBlock / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t v = 0;
  for (int r = 0; r < kScanItems; r++) {
    const int64_t t = base + r * kBlock + threadIdx.x;
    if (t < n) v += in[t];
  }
  int64_t total;
  (void)block_exclusive(v, &total, sh);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Scan pass 2: exclusive scan of the tile sums in place (one workgroup).
__global__ __launch_bounds__(kBlock) void k_scan_sums(int64_t* __restrict__ sums, int64_t nb,
                                                      int64_t* __restrict__ total_out) {
  __shared__ int64_t sh[kBlock / 64];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kBlock) {
    const int64_t t = b0 + threadIdx.x;
    const int64_t v = t < nb ? sums[t] : 0;
    int64_t tot;
    const int64_t ex = block_exclusive(v, &tot, sh);
    if (t < nb) sums[t] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

// Scan pass 3: exclusive scan within each tile plus the tile offset; the
// last tile also writes out[n] = total (clamped to int32; the host rejects
// totals above INT32_MAX).
__global__ __launch_bounds__(kBlock) void k_scan_apply(const int32_t* __restrict__ in, int64_t n,
                                                       const int64_t* __restrict__ sums,
                                                       const int64_t* __restrict__ total,
                                                       int32_t* __restrict__ out) {
  __shared__ int64_t sh[kBlock / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t carry = sums[blockIdx.x];
  for (int r = 0; r < kScanItems; r++) {
    const int64_t t = base + r * kBlock + threadIdx.x;
    const int64_t v = t < n ? in[t] : 0;
    int64_t tot;
    const int64_t ex = block_exclusive(v, &tot, sh);
    if (t < n) out[t] = static_cast<int32_t>(std::min<int64_t>(carry + ex, INT32_MAX));
    carry += tot;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    out[n] = static_cast<int32_t>(std::min<int64_t>(*total, INT32_MAX));
}

// ---------------------------------------------------------------------------
// Validation (validate_csr's rules; reason codes name the defect)
// ---------------------------------------------------------------------------
enum : int { kBadCount = 1, kBadPointer = 2, kBadIndex = 3, kSelf = 4, kHalfOrder = 5 };

__global__ void k_validate(int32_t n, const int32_t* __restrict__ nop,
                           const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                           int64_t e, int half, unsigned long long* __restrict__ bad) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = static_cast<int32_t>(t);
    const int64_t np = nop[i], kp = ptr[i];
    int reason = 0;
    if (np < 0 || np >= (n > 1 ? n : 1)) reason = kBadCount;
    else if (kp < 0 || kp + np > e) reason = kBadPointer;
    else
      for (int64_t k = 0; k < np; k++) {
        const int32_t j = list[kp + k];
        if (j < 0 || j >= n) { reason = kBadIndex; break; }
        if (j == i) { reason = kSelf; break; }
        if (half && j < i) { reason = kHalfOrder; break; }
      }
    if (reason) atomicMin(bad, static_cast<unsigned long long>(i) * 8ull + reason);
  }
}

// ---------------------------------------------------------------------------
// Row operations
// ---------------------------------------------------------------------------
// Copy row i of (ptr, list) to dst[dptr[i] ..): one wavefront per row.
__global__ __launch_bounds__(kBlock) void k_copy_rows(int32_t n, const int32_t* __restrict__ nop,
                                                      const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ dptr,
                                                      int32_t* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
  for (int64_t r = blockIdx.x * static_cast<int64_t>(kBlock / 64) + (threadIdx.x >> 6); r < n;
       r += waves) {
    const int32_t m = nop[r];
    const int64_t s = ptr[r], d = dptr[r];
    for (int32_t k = lane; k < m; k += 64) dst[d + k] = list[s + k];
  }
}

// Sort of each segment [start[r], start[r] + len[r]) of src into the same
// positions of dst, ascending, one wavefront per segment. Segments of up to
// 64*R elements are sorted in registers by a bitonic network (element
// e = r*64 + lane in register r of that lane; partners closer than 64 are
// exchanged across lanes, farther ones between a lane's own registers), padded
// with INT32_MAX. Longer segments fall back to a rank sort from memory: each
// element goes to the number of smaller elements plus the equal ones before
// it.
template <int R>
__device__ __forceinline__ void wave_bitonic(int32_t (&v)[R], int lane) {
#pragma unroll
  for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int pr = r ^ (j / 64);
          if (pr > r) {
            const bool up = ((r * 64 + lane) & k) == 0;
            const int32_t a = v[r], b = v[pr];
            v[r] = up ? min(a, b) : max(a, b);
            v[pr] = up ? max(a, b) : min(a, b);
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int32_t other = __shfl_xor(v[r], j, 64);
          const bool up = ((r * 64 + lane) & k) == 0;
          const bool keep_min = ((lane & j) == 0) == up;
          v[r] = keep_min ? min(v[r], other) : max(v[r], other);
        }
      }
    }
  }
}

template <int R>
__device__ __forceinline__ void sort_segment(const int32_t* __restrict__ src,
                                             int32_t* __restrict__ dst, int32_t m, int lane) {
  int32_t v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int e = r * 64 + lane;
    v[r] = e < m ? src[e] : INT32_MAX;
  }
  wave_bitonic<R>(v, lane);
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int e = r * 64 + lane;
    if (e < m) dst[e] = v[r];
  }
}

// Row r is read from src + r * src_stride (src_stride > 0, a slab) or from
// src + start[r], and written to dst + start[r].
__global__ __launch_bounds__(kBlock) void k_row_sort(int32_t n, const int32_t* __restrict__ len,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ src,
                                                     int64_t src_stride,
                                                     int32_t* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
  for (int64_t r = blockIdx.x * static_cast<int64_t>(kBlock / 64) + (threadIdx.x >> 6); r < n;
       r += waves) {
    const int32_t m = len[r];
    const int64_t d = start[r];
    const int32_t* in = src + (src_stride > 0 ? r * src_stride : d);
    int32_t* o = dst + d;
    if (m <= 1) {
      if (m == 1 && lane == 0) o[0] = in[0];
    } else if (m <= 64) {
      sort_segment<1>(in, o, m, lane);
    } else if (m <= 128) {
      sort_segment<2>(in, o, m, lane);
    } else if (m <= 256) {
      sort_segment<4>(in, o, m, lane);
    } else if (m <= 512) {
      sort_segment<8>(in, o, m, lane);
    } else {
      for (int32_t a = lane; a < m; a += 64) {
        const int32_t x = in[a];
        int32_t rank = 0;
        for (int32_t b = 0; b < m; b++) {
          const int32_t y = in[b];
          rank += (y < x) | ((y == x) & (b < a));
        }
        o[rank] = x;
      }
    }
  }
}

int row_sort(hipStream_t s, int32_t n, const int32_t* len, const int32_t* start,
             const int32_t* src, int32_t* dst, int64_t src_stride = 0) {
  if (n <= 0) return LJGPU_OK;
  k_row_sort<<<grid_for(n, kBlock / 64), kBlock, 0, s>>>(n, len, start, src, src_stride, dst);
  LTRY(hipGetLastError(), "row sort kernel");
  return LJGPU_OK;
}

// Half-to-full expansion.
__global__ void k_count_lower(int32_t n, const int32_t* __restrict__ nop,
                              const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                              int32_t* __restrict__ lower) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t kp = ptr[t];
    for (int32_t k = 0; k < nop[t]; k++) atomicAdd(&lower[list[kp + k]], 1);
  }
}

__global__ void k_add(int32_t n, const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                      int32_t* __restrict__ c) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    c[t] = a[t] + b[t];
}

__global__ void k_fill_expanded(int32_t n, const int32_t* __restrict__ nop,
                                const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                                const int32_t* __restrict__ lower,
                                const int32_t* __restrict__ fptr, int32_t* __restrict__ fill,
                                int32_t* __restrict__ lower_tmp, int32_t* __restrict__ flist) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t r = static_cast<int32_t>(t);
    const int64_t kp = ptr[r];
    const int64_t own = static_cast<int64_t>(fptr[r]) + lower[r];
    for (int32_t k = 0; k < nop[r]; k++) {
      const int32_t j = list[kp + k];
      const int32_t pos = atomicAdd(&fill[j], 1);
      lower_tmp[static_cast<int64_t>(fptr[j]) + pos] = r;
      flist[own + k] = j;
    }
  }
}

// Row-length extrema (int32 atomics on small[0] = min, small[1] = max).
__global__ void k_minmax(int32_t n, const int32_t* __restrict__ a, int32_t* __restrict__ mm) {
  int32_t lo = INT32_MAX, hi = INT32_MIN;
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    lo = min(lo, a[t]);
    hi = max(hi, a[t]);
  }
  for (int d = 32; d > 0; d >>= 1) {
    lo = min(lo, __shfl_xor(lo, d, 64));
    hi = max(hi, __shfl_xor(hi, d, 64));
  }
  // One atomic pair per workgroup: atomics on one address serialise.
  __shared__ int32_t slo[kBlock / 64], shi[kBlock / 64];
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; w++) {
      lo = min(lo, slo[w]);
      hi = max(hi, shi[w]);
    }
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
  }
}

// Partners j > i in each ascending full row.
__global__ void k_upper_counts(int32_t n, const int32_t* __restrict__ nop,
                               const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                               int32_t* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = static_cast<int32_t>(t);
    const int32_t* row = list + ptr[i];
    int32_t lo = 0, hi = nop[i];
    while (lo < hi) {  // first position with row[pos] > i
      const int32_t mid = (lo + hi) >> 1;
      if (row[mid] <= i) lo = mid + 1;
      else hi = mid;
    }
    out[i] = nop[i] - lo;
  }
}

__global__ void k_ell(int32_t n, int32_t width, const int32_t* __restrict__ nop,
                      const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                      int32_t* __restrict__ ell) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t kp = ptr[t];
    const int32_t m = nop[t];
    for (int32_t k = 0; k < width; k++) ell[t + static_cast<int64_t>(k) * n] = k < m ? list[kp + k] : 0;
  }
}

// Minimum and maximum of a strided coordinate stream (one workgroup).
__global__ __launch_bounds__(kBlock) void k_extrema(int32_t n, const double* __restrict__ a,
                                                    int stride, double* __restrict__ out) {
  __shared__ double smin[kBlock], smax[kBlock];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t t = threadIdx.x; t < n; t += kBlock) {
    const double v = a[t * stride];
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + d]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + d]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smin[0];
    out[1] = smax[0];
  }
}

// ---------------------------------------------------------------------------
// Cell list and pair passes
// ---------------------------------------------------------------------------
struct GridArgs {
  int nc0, nc1, nc2;
  double o0, o1, o2;
  double w0, w1, w2;
  double box;
};

__device__ __forceinline__ int cell_coord(double x, double origin, double width, int m, double box) {
  x -= origin;
  if (box > 0.0) x -= box * floor(x / box);  // the wrapped image's cell
  const int v = static_cast<int>(floor(x / width));
  return min(max(v, 0), m - 1);
}

// Also sets *outside when a periodic box has a coordinate outside [0, box)
// (the pair search then applies the minimum image in every cell).
__global__ void k_cell_assign(int32_t n, PosIn q, GridArgs g, int32_t* __restrict__ count,
                              int32_t* __restrict__ slot, int32_t* __restrict__ cell_of,
                              int32_t* __restrict__ outside) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t o = t * q.stride;
    if (g.box > 0.0) {
      const double x = q.x[o], y = q.y[o], z = q.z[o];
      // Written as negations so that NaN counts as outside.
      if (!(x >= 0.0 && x < g.box && y >= 0.0 && y < g.box && z >= 0.0 && z < g.box))
        atomicOr(outside, 1);
    }
    const int cx = cell_coord(q.x[o], g.o0, g.w0, g.nc0, g.box);
    const int cy = cell_coord(q.y[o], g.o1, g.w1, g.nc1, g.box);
    const int cz = cell_coord(q.z[o], g.o2, g.w2, g.nc2, g.box);
    const int32_t c = (cz * g.nc1 + cy) * g.nc0 + cx;
    cell_of[t] = c;
    slot[t] = atomicAdd(&count[c], 1);
  }
}

__global__ void k_cell_scatter(int32_t n, PosIn q, const int32_t* __restrict__ cell_of,
                               const int32_t* __restrict__ slot,
                               const int32_t* __restrict__ cell_start,
                               double4* __restrict__ cpos, int32_t* __restrict__ cell_sorted) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t o = t * q.stride;
    const int32_t c = cell_of[t];
    const int64_t d = static_cast<int64_t>(cell_start[c]) + slot[t];
    cpos[d] = make_double4(q.x[o], q.y[o], q.z[o], static_cast<double>(t));
    cell_sorted[d] = c;
  }
}

// Each cell's particles by ascending index (insertion sort, one thread per
// cell): the scatter's atomic slots leave them in arrival order. Used only
// when no cell holds more than kCellOrderMax particles.
constexpr int32_t kCellOrderMax = 256;
__global__ void k_cell_order(int64_t ncell, const int32_t* __restrict__ cell_start,
                             double4* __restrict__ cpos) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < ncell;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t a = cell_start[c], b = cell_start[c + 1];
    for (int32_t k = a + 1; k < b; k++) {
      const double4 v = cpos[k];
      int32_t m = k;
      while (m > a && cpos[m - 1].w > v.w) {
        cpos[m] = cpos[m - 1];
        m--;
      }
      cpos[m] = v;
    }
  }
}

// Row r of a slab (src + r * stride, len[r] entries) to dst + start[r], in
// order. One wavefront copies kSlabRows rows at a time, 128 entries of each
// per pass, every load of a pass issued before its stores (rows are about
// as long as the search sphere holds partners, so one or two passes).
constexpr int kSlabRows = 4;
__global__ __launch_bounds__(kBlock) void k_copy_slab(int32_t n, const int32_t* __restrict__ len,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ src,
                                                      int64_t stride, int32_t* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
  for (int64_t r0 = (blockIdx.x * static_cast<int64_t>(kBlock / 64) + (threadIdx.x >> 6)) *
                    kSlabRows;
       r0 < n; r0 += waves * kSlabRows) {
    int32_t m[kSlabRows], mx = 0;
#pragma unroll
    for (int k = 0; k < kSlabRows; k++) {
      m[k] = r0 + k < n ? len[r0 + k] : 0;
      mx = max(mx, m[k]);
    }
    for (int32_t e0 = 0; e0 < mx; e0 += 2 * 64) {
      int32_t v[kSlabRows][2];
#pragma unroll
      for (int k = 0; k < kSlabRows; k++) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const int32_t e = e0 + u * 64 + lane;
          v[k][u] = e < m[k] ? src[(r0 + k) * stride + e] : 0;
        }
      }
#pragma unroll
      for (int k = 0; k < kSlabRows; k++) {
        if (e0 >= m[k]) continue;
        int32_t* o = dst + start[r0 + k];
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const int32_t e = e0 + u * 64 + lane;
          if (e < m[k]) o[e] = v[k][u];
        }
      }
    }
  }
}

// The reference's pair predicate (cpu_ref/force_aos.cpp:131-135) with the
// minimum image of the periodic box: every operation rounded separately, in
// the reference's order, (dx*dx + dy*dy) + dz*dz < SL2.
template <bool IMAGE = true>
__device__ __forceinline__ bool within_search(const double4& a, const double4& b, double box,
                                              double half_box, double sl2) {
#pragma clang fp contract(off)
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  if (IMAGE && box > 0.0) {
    if (dx > half_box) dx -= box;
    else if (dx < -half_box) dx += box;
    if (dy > half_box) dy -= box;
    else if (dy < -half_box) dy += box;
    if (dz > half_box) dz -= box;
    else if (dz < -half_box) dz += box;
  }
  const double r2 = dx * dx + dy * dy + dz * dz;
  return r2 < sl2;
}

// One thread per particle in cell order (lanes of a wavefront share their
// home cell, so candidate loads are mostly broadcasts), over the cells within
// R of its own (R = 1 for cells at least the search length wide, 2 for cells
// at least half of it; R = 0: one cell holding every particle). MODE 0 counts
// each row's partners into nop[i]; MODE 1 writes them to out[ptr[i] ..) in
// candidate order; MODE 2 does both, writing at most `cap` of them to the
// row's slab out[i * cap ..).
template <int MODE, bool HALF, int R>
__global__ __launch_bounds__(kBlock) void k_pairs(int32_t n, const double4* __restrict__ cpos,
                                                  const int32_t* __restrict__ cell_sorted,
                                                  const int32_t* __restrict__ cell_start,
                                                  GridArgs g, double sl2, int32_t cap,
                                                  int32_t* __restrict__ nop,
                                                  const int32_t* __restrict__ ptr,
                                                  int32_t* __restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const double4 pi = cpos[t];
  const int32_t i = static_cast<int32_t>(pi.w);
  const int32_t c = cell_sorted[t];
  const int cx = c % g.nc0, cy = (c / g.nc0) % g.nc1, cz = c / (g.nc0 * g.nc1);
  const bool periodic = g.box > 0.0;
  const double half_box = 0.5 * g.box;
  int64_t base = 0;
  if (MODE == 1) base = ptr[i];
  if (MODE == 2) base = static_cast<int64_t>(i) * cap;
  int32_t cnt = 0;
  for (int dz = -R; dz <= R; dz++) {
    int z = cz + dz;
    if (periodic) z = (z + g.nc2) % g.nc2;
    else if (z < 0 || z >= g.nc2) continue;
    for (int dy = -R; dy <= R; dy++) {
      int y = cy + dy;
      if (periodic) y = (y + g.nc1) % g.nc1;
      else if (y < 0 || y >= g.nc1) continue;
      for (int dx = -R; dx <= R; dx++) {
        int x = cx + dx;
        if (periodic) x = (x + g.nc0) % g.nc0;
        else if (x < 0 || x >= g.nc0) continue;
        const int32_t cc = (z * g.nc1 + y) * g.nc0 + x;
        const int32_t u1 = cell_start[cc + 1];
        auto visit = [&](const double4& pj) {
          const int32_t j = static_cast<int32_t>(pj.w);
          if (HALF ? (j <= i) : (j == i)) return;
          if (within_search(pi, pj, g.box, half_box, sl2)) {
            if (MODE == 1 || (MODE == 2 && cnt < cap)) out[base + cnt] = j;
            cnt++;
          }
        };
        int32_t u = cell_start[cc];
        // Four candidates loaded before any is tested, so their loads overlap.
        for (; u + 4 <= u1; u += 4) {
          const double4 a = cpos[u], b = cpos[u + 1], c4 = cpos[u + 2], d = cpos[u + 3];
          visit(a);
          visit(b);
          visit(c4);
          visit(d);
        }
        for (; u < u1; u++) visit(cpos[u]);
      }
    }
  }
  if (MODE != 1) nop[i] = cnt;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// One wavefront per cell, for cell grids with a stencil (R >= 1). The cell's
// particles sit one per lane (in groups of 64); the candidates of the
// (2R+1)^3 cells around it are walked as one flattened sequence, 64 per
// round, so each candidate is loaded once for all of the cell's particles
// instead of once per particle. A round's hits for one particle go to
// consecutive addresses of its row (ballot, then each hit lane's rank among
// them), so the stores coalesce. The runs of the flattened sequence are the
// x-rows of the stencil, contiguous in cell order: (2R+1)^2 rows, each split
// at most once where it wraps around a periodic box (the grid has at least
// 2R+1 cells per side). Same predicate (within_search, in the same argument
// order), exclusions and MODEs as k_pairs; a row's entries come out in
// candidate order and are sorted afterwards.
// The candidates of one cell (see k_pairs_wave), with the minimum image
// (IMAGE) or without it: a cell whose stencil does not wrap around the box,
// with every position inside [0, box), has no pair that the image would
// shift, so the predicate without it gives the same result.
template <int MODE, bool HALF, int R, bool IMAGE>
__device__ __forceinline__ void pairs_cell(int32_t p0, int32_t p1, int32_t total, int32_t pre,
                                           int32_t rs, int32_t excl,
                                           const double4* __restrict__ cpos, const GridArgs& g,
                                           double half_box, double sl2, int32_t cap,
                                           int32_t* __restrict__ nop,
                                           const int32_t* __restrict__ ptr,
                                           int32_t* __restrict__ out) {
  constexpr int kSpan = 2 * R + 1;
  constexpr int kRuns = 2 * kSpan * kSpan;
  const int lane = threadIdx.x & 63;
  for (int32_t g0 = p0; g0 < p1; g0 += 64) {
    const int32_t mg = min(64, p1 - g0);
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    if (lane < mg) pi = cpos[g0 + lane];
    const int32_t iv = static_cast<int32_t>(pi.w);
    int64_t basev = 0;
    if (MODE == 1) basev = lane < mg ? ptr[iv] : 0;
    if (MODE == 2) basev = static_cast<int64_t>(iv) * cap;
    int32_t cntv = 0;
    for (int32_t f0 = 0; f0 < total; f0 += 64) {
      const int32_t f = f0 + lane;
      const bool live = f < total;
      // The run holding f: the number of runs ending at or before f.
      int r = 0;
#pragma unroll
      for (int bit = 32; bit > 0; bit >>= 1) {
        const int t = r + bit;
        const int32_t v = __shfl(pre, min(t, 64) - 1, 64);
        if (t <= kRuns && v <= f) r = t;
      }
      const int rr = min(r, kRuns - 1);
      const int32_t u = __shfl(rs, rr, 64) + (f - __shfl(excl, rr, 64));
      // Lanes past the end read candidate 0 (a valid address) and test nothing.
      const double4 pj = cpos[live ? u : 0];
      const int32_t j = static_cast<int32_t>(pj.w);
      for (int k = 0; k < mg; k++) {
        double4 a;
        a.x = readlane_d(pi.x, k);
        a.y = readlane_d(pi.y, k);
        a.z = readlane_d(pi.z, k);
        a.w = 0.0;
        const int32_t ik = __builtin_amdgcn_readlane(iv, k);
        const bool hit = live && (HALF ? j > ik : j != ik) &&
                         within_search<IMAGE>(a, pj, g.box, half_box, sl2);
        const uint64_t mask = __ballot(hit);
        if (mask == 0) continue;
        const int32_t ck = __builtin_amdgcn_readlane(cntv, k);
        if (MODE != 0 && hit) {
          const int32_t rank = static_cast<int32_t>(__builtin_amdgcn_mbcnt_hi(
              static_cast<uint32_t>(mask >> 32),
              __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
          const int64_t bk =
              (static_cast<int64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(
                   static_cast<int32_t>(basev >> 32), k)))
               << 32) |
              static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(basev), k));
          if (MODE == 1 || ck + rank < cap) out[bk + ck + rank] = j;
        }
        if (lane == k) cntv = ck + __popcll(mask);
      }
    }
    if (MODE != 1 && lane < mg) nop[iv] = cntv;
  }
}

// One wavefront per cell, for cell grids with a stencil (R >= 1). The cell's
// particles sit one per lane (in groups of 64); the candidates of the
// (2R+1)^3 cells around it are walked as one flattened sequence, 64 per
// round, so each candidate is loaded once for all of the cell's particles
// instead of once per particle. A round's hits for one particle go to
// consecutive addresses of its row (ballot, then each hit lane's rank among
// them), so the stores coalesce. The runs of the flattened sequence are the
// x-rows of the stencil, contiguous in cell order: (2R+1)^2 rows, each split
// at most once where it wraps around a periodic box (the grid has at least
// 2R+1 cells per side). Same predicate (within_search, in the same argument
// order), exclusions and MODEs as k_pairs; a row's entries come out in
// candidate order and are sorted afterwards. *outside (from k_cell_assign)
// keeps the minimum image on in every cell.
template <int MODE, bool HALF, int R>
__global__ __launch_bounds__(kBlock) void k_pairs_wave(int32_t ncell,
                                                       const double4* __restrict__ cpos,
                                                       const int32_t* __restrict__ cell_start,
                                                       GridArgs g, double sl2, int32_t cap,
                                                       int32_t* __restrict__ nop,
                                                       const int32_t* __restrict__ ptr,
                                                       int32_t* __restrict__ out,
                                                       const int32_t* __restrict__ outside) {
  constexpr int kSpan = 2 * R + 1;
  constexpr int kRuns = 2 * kSpan * kSpan;
  static_assert(kRuns <= 64, "one run per lane");
  const int lane = threadIdx.x & 63;
  const bool periodic = g.box > 0.0;
  const bool all_inside = *outside == 0;
  const double half_box = 0.5 * g.box;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
  for (int64_t c64 = (blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x) >> 6; c64 < ncell;
       c64 += waves) {
    const int32_t c = static_cast<int32_t>(c64);
    const int32_t p0 = cell_start[c], p1 = cell_start[c + 1];
    if (p0 == p1) continue;  // the whole wavefront
    const int cx = c % g.nc0, cy = (c / g.nc0) % g.nc1, cz = c / (g.nc0 * g.nc1);
    // Lane r < kRuns: run r of the stencil as [rs, rs + rl) in cpos.
    int32_t rs = 0, rl = 0;
    if (lane < kRuns) {
      const int yz = lane >> 1;
      int z = cz + yz / kSpan - R, y = cy + yz % kSpan - R;
      bool ok = true;
      if (periodic) {
        z = (z + g.nc2) % g.nc2;
        y = (y + g.nc1) % g.nc1;
      } else {
        ok = z >= 0 && z < g.nc2 && y >= 0 && y < g.nc1;
      }
      const int x0 = cx - R, x1 = cx + R;
      int a = 1, b = 0;  // inclusive x range of cells; empty when a > b
      if ((lane & 1) == 0) {
        a = max(x0, 0);
        b = min(x1, g.nc0 - 1);
      } else if (periodic && x0 < 0) {
        a = x0 + g.nc0;
        b = g.nc0 - 1;
      } else if (periodic && x1 >= g.nc0) {
        a = 0;
        b = x1 - g.nc0;
      }
      if (ok && a <= b) {
        const int32_t rowc = (z * g.nc1 + y) * g.nc0;
        rs = cell_start[rowc + a];
        rl = cell_start[rowc + b + 1] - rs;
      }
    }
    int32_t pre = rl;  // inclusive prefix of the run lengths
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t v = __shfl_up(pre, o, 64);
      if (lane >= o) pre += v;
    }
    const int32_t excl = pre - rl;
    const int32_t total = __shfl(pre, 63, 64);
    const bool wraps = periodic && (!all_inside || cx < R || cx + R >= g.nc0 || cy < R ||
                                    cy + R >= g.nc1 || cz < R || cz + R >= g.nc2);
    if (wraps)
      pairs_cell<MODE, HALF, R, true>(p0, p1, total, pre, rs, excl, cpos, g, half_box, sl2, cap,
                                      nop, ptr, out);
    else
      pairs_cell<MODE, HALF, R, false>(p0, p1, total, pre, rs, excl, cpos, g, half_box, sl2, cap,
                                       nop, ptr, out);
  }
}

// ---------------------------------------------------------------------------
// LDS tiles
// ---------------------------------------------------------------------------
constexpr int kTileHashBits = 12;
constexpr int kTileHash = 1 << kTileHashBits;          // open-addressing table slots
constexpr int kTileHashLimit = 3072;                   // distinct keys before giving up
constexpr int kTileMaxStage = kTileLdsBudget / 24 + 1; // staged positions that can fit
constexpr int kTileUniq = 4096;                        // sort buffer (power of two)
constexpr int kTileBatch = 8;                          // list entries loaded ahead per thread
constexpr int64_t kTileOnePassBytes = int64_t{4} << 30;  // staging slots of the one-pass build
static_assert(kTileMaxStage <= kTileHashLimit, "the hash limit must cover every tile that fits");
static_assert(kTileHashLimit + kBlock < kTileHash, "probing must always find a free slot");
static_assert(kTileHashLimit <= kTileUniq, "the sort buffer must hold every distinct key");

__device__ __forceinline__ uint32_t tile_hash(int32_t key) {
  return (static_cast<uint32_t>(key) * 2654435761u) >> (32 - kTileHashBits);
}

// Wavefronts per tile of the sliced offsets: at most 1024 threads.
constexpr int kSliceWavesMax = 16;
// A block's rows, staged in LDS: slot i0 + s holds particle mem[s], whose
// row is list[src[s] ..) and covers entries [beg[s], beg[s + 1]) of the
// block (beg[ni] = ne).
constexpr int kTileRowsMax = 1024;
struct TileRows {
  int32_t mem[kTileRowsMax];
  int32_t src[kTileRowsMax];
  int32_t beg[kTileRowsMax + 1];
};

struct TileGeom {
  int32_t i0, ni, ne, rid;
  int64_t e0;
};

// Block b covers slots [geo[4b], geo[4b] + geo[4b+1]) of replica geo[4b+2];
// its entries are [slot_ptr[i0], slot_ptr[i0 + ni]) of the slot-ordered rows.
__device__ __forceinline__ TileGeom tile_geom(int32_t b, const int32_t* __restrict__ geo,
                                              const int32_t* __restrict__ slot_ptr) {
  TileGeom g;
  g.i0 = geo[4 * b];
  g.ni = geo[4 * b + 1];
  g.rid = geo[4 * b + 2];
  g.e0 = slot_ptr[g.i0];
  g.ne = slot_ptr[g.i0 + g.ni] - static_cast<int32_t>(g.e0);
  return g;
}

// Fill the row table of block g. Returns the list position of the block's
// first entry when its rows lie back to back in the list (index order), so
// entry u is list[base + u]; -1 when they are scattered (entry u is then
// found by a binary search over beg).
__device__ int32_t tile_rows(const TileGeom& g, const int32_t* __restrict__ member,
                             const int32_t* __restrict__ ptr, const int32_t* __restrict__ slot_ptr,
                             TileRows& r) {
  for (int32_t t = threadIdx.x; t < g.ni; t += kBlock) {
    const int32_t i = member ? member[g.i0 + t] : g.i0 + t;
    r.mem[t] = i;
    r.src[t] = ptr[i];
    r.beg[t] = slot_ptr[g.i0 + t] - static_cast<int32_t>(g.e0);
  }
  if (threadIdx.x == 0) r.beg[g.ni] = g.ne;
  __syncthreads();
  const int32_t base = r.src[0];
  int ok = 1;
  for (int32_t t = threadIdx.x; t < g.ni; t += kBlock) ok &= (r.src[t] - r.beg[t] == base);
  return __syncthreads_and(ok) ? base : -1;
}

// Partner of the block's entry u (0 <= u < ne).
__device__ __forceinline__ int32_t tile_entry(const TileRows& r, int32_t ni, int32_t base,
                                              const int32_t* __restrict__ list, int32_t u) {
  if (base >= 0) return list[base + u];
  int32_t lo = 0, hi = ni;  // beg[lo] <= u < beg[hi]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (r.beg[mid] <= u) lo = mid;
    else hi = mid;
  }
  return list[r.src[lo] + (u - r.beg[lo])];
}

// Insert the block's keys (the slots of its own particles and of every
// entry's partner; slot_of null: index order, slot = particle) into an LDS
// hash set. Returns the number of distinct keys, or -1 when there are more
// than kTileHashLimit (such a block cannot fit the LDS budget).
__device__ int32_t tile_insert(const TileGeom& g, const TileRows& rows, int32_t base,
                               const int32_t* __restrict__ list,
                               const int32_t* __restrict__ slot_of, int32_t* table,
                               int32_t* counter) {
  for (int32_t t = threadIdx.x; t < kTileHash; t += kBlock) table[t] = -1;
  if (threadIdx.x == 0) *counter = 0;
  __syncthreads();
  const int64_t m = g.ni + static_cast<int64_t>(g.ne);
  // Keys are loaded kTileBatch at a time before any is inserted, so each
  // thread waits for memory once per batch rather than once per key.
  for (int64_t t0 = 0; t0 < m; t0 += kTileBatch * kBlock) {
    int32_t keys[kTileBatch];
#pragma unroll
    for (int r = 0; r < kTileBatch; r++) {
      const int64_t t = t0 + r * kBlock + threadIdx.x;
      keys[r] = t >= m     ? -1
                : t < g.ni ? -2 - static_cast<int32_t>(t)  // own slot g.i0 + t, below
                           : tile_entry(rows, g.ni, base, list, static_cast<int32_t>(t - g.ni));
    }
#pragma unroll
    for (int r = 0; r < kTileBatch; r++) {
      if (keys[r] <= -2) keys[r] = g.i0 - 2 - keys[r];
      else if (keys[r] >= 0 && slot_of) keys[r] = slot_of[keys[r]];
    }
#pragma unroll
    for (int r = 0; r < kTileBatch; r++) {
      const int32_t key = keys[r];
      if (key < 0) continue;
      uint32_t h = tile_hash(key);
      while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) <=
             kTileHashLimit) {
        const int32_t prev = atomicCAS(&table[h], -1, key);
        if (prev == -1) {
          atomicAdd(counter, 1);
          break;
        }
        if (prev == key) break;
        h = (h + 1) & (kTileHash - 1);
      }
    }
  }
  __syncthreads();
  const int32_t ns = *counter;
  __syncthreads();
  return ns > kTileHashLimit ? -1 : ns;
}

// Slot of a key known to be in the table.
__device__ __forceinline__ int32_t tile_find(const int32_t* table, int32_t key) {
  uint32_t h = tile_hash(key);
  while (table[h] != key) h = (h + 1) & (kTileHash - 1);
  return static_cast<int32_t>(h);
}

// Ascending sort of uniq[0 .. kBlock * R) by the workgroup's wavefronts:
// element e = w*64R + r*64 + lane sits in register r of lane `lane` of
// wavefront w. Bitonic network: exchanges closer than 64 go through lane
// shuffles, closer than 64R between a lane's own registers, and only the
// farther ones (log2(waves) stages per merge size beyond 64R) through LDS.
template <int R>
__device__ void block_bitonic(int32_t* uniq) {
  constexpr int kSpan = 64 * R;
  constexpr int P = kBlock * R;
  const int lane = threadIdx.x & 63;
  const int wbase = static_cast<int>(threadIdx.x >> 6) * kSpan;
  int32_t v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = uniq[wbase + r * 64 + lane];
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= kSpan) {
#pragma unroll
        for (int r = 0; r < R; r++) uniq[wbase + r * 64 + lane] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int e = wbase + r * 64 + lane;
          const int32_t other = uniq[e ^ j];
          const bool keep_min = ((e & j) == 0) == ((e & k) == 0);
          v[r] = keep_min ? min(v[r], other) : max(v[r], other);
        }
        __syncthreads();
      } else if (j >= 64) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int pr = r ^ (j / 64);
          if (pr > r) {
            const bool up = ((wbase + r * 64 + lane) & k) == 0;
            const int32_t a = v[r], c = v[pr];
            v[r] = up ? min(a, c) : max(a, c);
            v[pr] = up ? max(a, c) : min(a, c);
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int32_t other = __shfl_xor(v[r], j, 64);
          const bool up = ((wbase + r * 64 + lane) & k) == 0;
          const bool keep_min = ((lane & j) == 0) == up;
          v[r] = keep_min ? min(v[r], other) : max(v[r], other);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) uniq[wbase + r * 64 + lane] = v[r];
}

// Pass 1: distinct staged particles per block; flags blocks that exceed the
// LDS budget. stats: [0] overflow, [1] max staged, [2] max LDS bytes.
__global__ __launch_bounds__(kBlock) void k_tile_count(const int32_t* __restrict__ geo, int32_t nblocks,
                                                       const int32_t* __restrict__ member,
                                                       const int32_t* __restrict__ slot_of,
                                                       const int32_t* __restrict__ slot_ptr,
                                                       const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ list,
                                                       int32_t* __restrict__ ns_out,
                                                       int32_t* __restrict__ stats) {
  __shared__ int32_t table[kTileHash];
  __shared__ TileRows rows;
  __shared__ int32_t counter;
  for (int32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const TileGeom g = tile_geom(b, geo, slot_ptr);
    const int32_t base = tile_rows(g, member, ptr, slot_ptr, rows);
    const int32_t ns = tile_insert(g, rows, base, list, slot_of, table, &counter);
    if (threadIdx.x == 0) {
      ns_out[b] = ns < 0 ? 0 : ns;
      const int64_t bytes = ns < 0 ? INT32_MAX : tile_lds_bytes(ns, g.ne);
      if (bytes > kTileLdsBudget) atomicOr(&stats[0], 1);
      atomicMax(&stats[1], ns);
      atomicMax(&stats[2], static_cast<int32_t>(std::min<int64_t>(bytes, INT32_MAX)));
    }
  }
}

// Pass 2 (every block fits): the distinct keys sorted ascending become the
// block's staged particles; each entry's LDS offset is 24 * the rank of its
// partner, looked up through the hash table.
__global__ __launch_bounds__(kBlock) void k_tile_fill(const int32_t* __restrict__ geo, int32_t nblocks,
                                                      const int32_t* __restrict__ member,
                                                      const int32_t* __restrict__ slot_of,
                                                      const int32_t* __restrict__ slot_ptr,
                                                      const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ soff,
                                                      int32_t* __restrict__ desc,
                                                      int32_t* __restrict__ stage,
                                                      int32_t* __restrict__ stage_slot,
                                                      uint16_t* __restrict__ local,
                                                      int32_t* __restrict__ stats,
                                                      unsigned long long* __restrict__ staged) {
  __shared__ int32_t table[kTileHash];
  __shared__ uint16_t rank_of[kTileHash];
  __shared__ int32_t uniq[kTileUniq];
  __shared__ TileRows rows;
  __shared__ int32_t counter;
  __shared__ int64_t sh[kBlock / 64];
  constexpr int32_t C = kTileHash / kBlock;
  for (int32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const TileGeom g = tile_geom(b, geo, slot_ptr);
    const int32_t base = tile_rows(g, member, ptr, slot_ptr, rows);
    const int32_t ns = tile_insert(g, rows, base, list, slot_of, table, &counter);
    if (!soff) {
      // One pass: block b stages at b * kTileMaxStage, and the budget check
      // of k_tile_count happens here (stats as there, plus the total staged).
      const int64_t bytes = ns < 0 ? INT32_MAX : tile_lds_bytes(ns, g.ne);
      const bool over = bytes > kTileLdsBudget;
      if (threadIdx.x == 0) {
        if (over) atomicOr(&stats[0], 1);
        atomicMax(&stats[1], ns);
        atomicMax(&stats[2], static_cast<int32_t>(std::min<int64_t>(bytes, INT32_MAX)));
        if (!over) atomicAdd(staged, static_cast<unsigned long long>(ns));
      }
      if (over) continue;  // the whole block (ns is uniform); the caller retries smaller
    }
    // Compact the occupied slots (C per thread) into uniq[0..ns), pad, sort.
    int32_t cnt = 0;
    for (int32_t t = threadIdx.x * C; t < (threadIdx.x + 1) * C; t++) cnt += table[t] != -1;
    int64_t total;
    int64_t off = block_exclusive(cnt, &total, sh);
    for (int32_t t = threadIdx.x * C; t < (threadIdx.x + 1) * C; t++)
      if (table[t] != -1) uniq[off++] = table[t];
    int32_t P = kBlock;
    while (P < ns) P <<= 1;
    for (int32_t t = ns + threadIdx.x; t < P; t += kBlock) uniq[t] = INT32_MAX;
    __syncthreads();
    switch (P / kBlock) {
      case 1: block_bitonic<1>(uniq); break;
      case 2: block_bitonic<2>(uniq); break;
      case 4: block_bitonic<4>(uniq); break;
      case 8: block_bitonic<8>(uniq); break;
      default: block_bitonic<16>(uniq); break;
    }
    __syncthreads();
    const int64_t s0 = soff ? soff[b] : static_cast<int64_t>(b) * kTileMaxStage;
    // Staged particles in slot order: stage holds the particles, stage_slot
    // (spatial order only) their slots.
    for (int32_t t = threadIdx.x; t < ns; t += kBlock) {
      const int32_t key = uniq[t];
      stage[s0 + t] = member ? member[key] : key;
      if (stage_slot) stage_slot[s0 + t] = key;
      rank_of[tile_find(table, key)] = static_cast<uint16_t>(t);
    }
    if (threadIdx.x == 0) {
      int32_t* d = desc + static_cast<int64_t>(b) * 8;
      d[0] = g.i0;
      d[1] = g.ni;
      d[2] = static_cast<int32_t>(s0);
      d[3] = ns;
      d[4] = static_cast<int32_t>(g.e0);
      d[5] = g.ne;
      d[6] = g.rid;
      d[7] = rows.mem[g.ni / 2];  // middle particle: the periodic unwrap's reference
    }
    __syncthreads();
    for (int32_t e0 = 0; e0 < g.ne; e0 += kTileBatch * kBlock) {
      int32_t js[kTileBatch];
#pragma unroll
      for (int r = 0; r < kTileBatch; r++) {
        const int32_t e = e0 + r * kBlock + threadIdx.x;
        js[r] = e < g.ne ? tile_entry(rows, g.ni, base, list, e) : -1;
      }
      if (slot_of) {
#pragma unroll
        for (int r = 0; r < kTileBatch; r++)
          if (js[r] >= 0) js[r] = slot_of[js[r]];
      }
#pragma unroll
      for (int r = 0; r < kTileBatch; r++) {
        const int32_t e = e0 + r * kBlock + threadIdx.x;
        if (js[r] >= 0) local[g.e0 + e] = static_cast<uint16_t>(rank_of[tile_find(table, js[r])] * 24);
      }
    }
    __syncthreads();
  }
}

// The row order of every tile row: with positions (cut.q.x non-null) the
// cutoff classes (cut_partition over cut_inner, partners found through the
// block's staged particles), then the bank-class order of each class
// (bank_class_order). One wavefront per tile, one row per lane, each lane's
// buffers in LDS.
struct CutDev {
  PosIn q;
  const int32_t* desc;
  const int32_t* stage;
  const int32_t* member;
  double box;
  const double* rep_box;
  double cl2;
};
constexpr int kBankBlock = 64;
__global__ __launch_bounds__(kBankBlock) void k_bank_order(const int32_t* __restrict__ geo,
                                                           int32_t nblocks,
                                                           const int32_t* __restrict__ slot_ptr,
                                                           uint16_t* __restrict__ local,
                                                           CutDev cut) {
  __shared__ uint16_t buf[kBankBlock][kBankRowMax];
  __shared__ uint16_t cnt[kBankBlock][kBankClasses];
  __shared__ uint16_t nxt[kBankBlock][kBankClasses];
  const int t = static_cast<int>(threadIdx.x);
  const int64_t st = cut.q.stride;
  for (int32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const int32_t i0 = geo[4 * b], ni = geo[4 * b + 1];
    for (int32_t r = t; r < ni; r += kBankBlock) {
      const int32_t a = slot_ptr[i0 + r];
      const int m = slot_ptr[i0 + r + 1] - a;
      uint16_t* row = local + a;
      int nin = 0;
      if (cut.q.x && m <= kBankRowMax) {
        const int64_t i = cut.member ? cut.member[i0 + r] : i0 + r;
        const double xi = cut.q.x[i * st], yi = cut.q.y[i * st], zi = cut.q.z[i * st];
        const int32_t soff = cut.desc[8 * static_cast<int64_t>(b) + 2];
        const double box = cut.rep_box ? cut.rep_box[2 * geo[4 * b + 2]] : cut.box;
        int nout = 0;
        for (int k = 0; k < m; k++) {
          const uint16_t o = row[k];
          const int64_t j = cut.stage[soff + o / 24];
          if (cut_inner(xi, yi, zi, cut.q.x[j * st], cut.q.y[j * st], cut.q.z[j * st], box,
                        cut.cl2))
            row[nin++] = o;
          else
            buf[t][nout++] = o;
        }
        for (int k = 0; k < nout; k++) row[nin + k] = buf[t][k];
      }
      bank_class_order(row, nin, r, 0, buf[t], cnt[t], nxt[t]);
      bank_class_order(row + nin, m - nin, r, nin, buf[t], cnt[t], nxt[t]);
    }
  }
}

// Slice lengths: per (block, wavefront slot), 64 * unroll offsets for each
// batch its longest row needs (0 for slots past the block's rows).
__global__ void k_wave_slices(const int32_t* __restrict__ desc, int32_t nblocks, int32_t wpb,
                              const int32_t* __restrict__ slot_ptr, int32_t lanes, int32_t unroll,
                              int32_t* __restrict__ len) {
  const int64_t id = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (id >= static_cast<int64_t>(nblocks) * wpb) return;
  const int32_t b = static_cast<int32_t>(id / wpb), w = static_cast<int32_t>(id % wpb);
  const int32_t i0 = desc[8 * static_cast<int64_t>(b)], ni = desc[8 * static_cast<int64_t>(b) + 1];
  const int32_t rpw = 64 / lanes, kb = lanes * unroll;
  int32_t batches = 0;
  for (int32_t r = w * rpw; r < min(ni, (w + 1) * rpw); r++) {
    const int32_t np = slot_ptr[i0 + r + 1] - slot_ptr[i0 + r];
    batches = max(batches, (np + kb - 1) / kb);
  }
  len[id] = batches * 64 * unroll;
}

// The slices of one block per workgroup, written in order: slice position
// p of wavefront w (batch j = p / (64 unroll), slot u, lane l) holds entry
// k = j * lanes * unroll + u * lanes + l % lanes of the block's row
// w * (64 / lanes) + l / lanes, or 0 past the row's end (padding). The
// writes are contiguous; the reads gather short runs of `local` (the
// block's rows, a few KB, served by the cache). Row starts and the block's
// slice offsets come from LDS; lanes and unroll are powers of two, so the
// position arithmetic is shifts and masks.
__global__ __launch_bounds__(kBlock) void k_slice_gather(
    const int32_t* __restrict__ desc, int32_t nblocks, int32_t wpb,
    const int32_t* __restrict__ slot_ptr, const uint16_t* __restrict__ local, int32_t lanes,
    int32_t unroll, const int32_t* __restrict__ wave_off, uint16_t* __restrict__ sliced) {
  __shared__ int32_t rs[kTileRowsMax + 1];
  __shared__ int32_t wo[kSliceWavesMax + 1];
  const int32_t lsh = 31 - __clz(lanes), ush = 31 - __clz(unroll);
  const int32_t bsh = 6 + ush;  // log2 of a batch's 64 * unroll positions
  for (int32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const int32_t i0 = desc[8 * static_cast<int64_t>(b)], ni = desc[8 * static_cast<int64_t>(b) + 1];
    for (int32_t r = threadIdx.x; r <= ni; r += kBlock) rs[r] = slot_ptr[i0 + r];
    for (int32_t w = threadIdx.x; w <= wpb; w += kBlock)
      wo[w] = wave_off[static_cast<int64_t>(b) * wpb + w];
    __syncthreads();
    for (int32_t w = 0; w < wpb; w++) {
      const int32_t p0 = wo[w], len = wo[w + 1] - p0;
      const int32_t row0 = w << (6 - lsh);
      for (int32_t off = threadIdx.x; off < len; off += kBlock) {
        const int32_t j = off >> bsh, u = (off >> 6) & (unroll - 1), l = off & 63;
        const int32_t row = row0 + (l >> lsh);
        const int32_t k = (j << (lsh + ush)) + (u << lsh) + (l & (lanes - 1));
        uint16_t v = 0;
        if (row < ni && k < rs[row + 1] - rs[row]) v = local[rs[row] + k];
        sliced[p0 + off] = v;
      }
    }
    __syncthreads();
  }
}

// Batched replicas.
__global__ void k_copy_add(int64_t n, const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                           int32_t add) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[t] = src[t] + add;
}

__global__ void k_replica_of(int32_t n, const int32_t* __restrict__ off, int32_t nrep,
                             int32_t* __restrict__ rep) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int32_t lo = 0, hi = nrep;  // last r with off[r] <= t
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (off[mid] <= t) lo = mid;
      else hi = mid;
    }
    rep[t] = lo;
  }
}

__global__ void k_check_replicas(int32_t n, const int32_t* __restrict__ nop,
                                 const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                                 const int32_t* __restrict__ rep,
                                 unsigned long long* __restrict__ bad) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t kp = ptr[t];
    for (int32_t k = 0; k < nop[t]; k++)
      if (rep[list[kp + k]] != rep[t]) {
        atomicMin(bad, static_cast<unsigned long long>(t));
        break;
      }
  }
}

// Spatial tile order: bins of one system (dev_tile_order).
struct OrderGrid {
  double w[3];     // bin widths: x bins, y and z columns
  double box;      // period
  int32_t nc[3];   // x bins, y columns, z columns
  int32_t key0;    // first key of the system
};

__global__ void k_order_key(int32_t n, PosIn q, const OrderGrid* __restrict__ grids,
                            const int32_t* __restrict__ rep_of, int32_t* __restrict__ count,
                            int32_t* __restrict__ slot, int32_t* __restrict__ key_of) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const OrderGrid g = grids[rep_of ? rep_of[t] : 0];
    const int64_t o = t * q.stride;
    const int cx = cell_coord(q.x[o], 0.0, g.w[0], g.nc[0], g.box);
    const int cy = cell_coord(q.y[o], 0.0, g.w[1], g.nc[1], g.box);
    const int cz = cell_coord(q.z[o], 0.0, g.w[2], g.nc[2], g.box);
    const int32_t k = g.key0 + (cz * g.nc[1] + cy) * g.nc[0] + cx;
    key_of[t] = k;
    slot[t] = atomicAdd(&count[k], 1);
  }
}

__global__ void k_order_scatter(int32_t n, const int32_t* __restrict__ key_of,
                                const int32_t* __restrict__ slot,
                                const int32_t* __restrict__ start, int32_t* __restrict__ order) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    order[start[key_of[t]] + slot[t]] = static_cast<int32_t>(t);
}

// Each bin's particles by ascending index (insertion sort, one thread per
// bin; bins hold at most kCellOrderMax particles).
__global__ void k_order_sort(int64_t nkey, const int32_t* __restrict__ start,
                             int32_t* __restrict__ order) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < nkey;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t a = start[c], b = start[c + 1];
    for (int32_t k = a + 1; k < b; k++) {
      const int32_t v = order[k];
      int32_t m = k;
      while (m > a && order[m - 1] > v) {
        order[m] = order[m - 1];
        m--;
      }
      order[m] = v;
    }
  }
}

// slot_of[member[s]] = s.
__global__ void k_inverse(int32_t n, const int32_t* __restrict__ member,
                          int32_t* __restrict__ slot_of) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    slot_of[member[t]] = static_cast<int32_t>(t);
}

// Each group of slots (geo: first slot, count) reordered by the row length
// of its particles, longest first, ties in the previous slot order: the
// wavefronts of a tile then hold rows of similar length, which need similar
// numbers of batches.
__global__ __launch_bounds__(kBlock) void k_group_length_sort(const int32_t* __restrict__ geo,
                                                              int32_t ngroups,
                                                              const int32_t* __restrict__ nop,
                                                              int32_t* __restrict__ order) {
  __shared__ int32_t mem[kTileRowsMax];
  __shared__ int32_t len[kTileRowsMax];
  for (int32_t b = blockIdx.x; b < ngroups; b += gridDim.x) {
    const int32_t i0 = geo[4 * b], ni = geo[4 * b + 1];
    for (int32_t t = threadIdx.x; t < ni; t += blockDim.x) {
      const int32_t i = order[i0 + t];
      mem[t] = i;
      len[t] = nop[i];
    }
    __syncthreads();
    for (int32_t t = threadIdx.x; t < ni; t += blockDim.x) {
      const int32_t l = len[t];
      int32_t r = 0;
      for (int32_t u = 0; u < ni; u++) r += len[u] > l || (len[u] == l && u < t);
      order[i0 + r] = mem[t];
    }
    __syncthreads();
  }
}

// Row length of the particle in each slot.
__global__ void k_slot_nop(int32_t n, const int32_t* __restrict__ member,
                           const int32_t* __restrict__ nop, int32_t* __restrict__ out) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[t] = nop[member[t]];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int dev_reserve(DevMem& m, size_t bytes, hipStream_t stream) {
  bytes = std::max<size_t>(bytes, 16);
  if (m.p && m.bytes >= bytes) return LJGPU_OK;
  if (m.p) {
    LTRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    (void)hipFree(m.p);
    m.p = nullptr;
    m.bytes = 0;
  }
  const size_t grow = bytes + bytes / 8;  // headroom for the next, slightly larger build
  LTRY(hipMalloc(&m.p, grow), "hipMalloc (list buffer)");
  m.bytes = grow;
  return LJGPU_OK;
}

void dev_release(DevMem& m) {
  if (m.p) (void)hipFree(m.p);
  m.p = nullptr;
  m.bytes = 0;
}

void dev_release(DevCSR& c) {
  dev_release(c.nop);
  dev_release(c.ptr);
  dev_release(c.list);
  c.n = 0;
  c.e = 0;
}

void dev_release(DevTiles& t) {
  dev_release(t.desc);
  dev_release(t.stage);
  dev_release(t.local);
  dev_release(t.sliced);
  dev_release(t.wave_off);
  t.sliced_lanes = t.sliced_unroll = t.waves_per_block = 0;
  dev_release(t.member);
  dev_release(t.slot_ptr);
  dev_release(t.stage_slot);
  t.tile_particles = 0;
  t.blocks = 0;
  t.staged = 0;
  t.max_stage = 0;
  t.lds_bytes = 0;
}

void dev_release(ListWork& w) {
  for (DevMem* m : {&w.scan_sums, &w.small, &w.cell_count, &w.cell_start, &w.cell_slot, &w.cell_of,
                    &w.cell_sorted, &w.cpos, &w.counts, &w.tmp, &w.tile_ns, &w.tile_soff,
                    &w.tile_geo, &w.rep_of, &w.order, &w.order_grid, &w.slot_of})
    dev_release(*m);
  dev_release(w.part);
}

int dev_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, ListWork& w,
             int64_t* total) {
  *total = 0;
  if (n <= 0) {
    LCHECK(fill<int32_t>(s, out, 1, 0));
    return LJGPU_OK;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  LCHECK(dev_reserve(w.scan_sums, sizeof(int64_t) * nb, s));
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  int64_t* sums = w.scan_sums.as<int64_t>();
  int64_t* tot = w.small.as<int64_t>();
  k_scan_reduce<<<static_cast<int>(nb), kBlock, 0, s>>>(in, n, sums);
  k_scan_sums<<<1, kBlock, 0, s>>>(sums, nb, tot);
  k_scan_apply<<<static_cast<int>(nb), kBlock, 0, s>>>(in, n, sums, tot, out);
  LTRY(hipGetLastError(), "scan kernels");
  LTRY(hipMemcpyAsync(total, tot, sizeof(int64_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  return LJGPU_OK;
}

int dev_validate(hipStream_t s, int32_t n, const int32_t* nop, const int32_t* ptr,
                 const int32_t* list, int64_t e, int kind, ListWork& w) {
  if (n <= 0) return LJGPU_OK;
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  unsigned long long* bad = w.small.as<unsigned long long>();
  LCHECK(fill<unsigned long long>(s, bad, 1, ~0ull));
  k_validate<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, nop, ptr, list, e,
                                                   kind == LJGPU_LIST_HALF ? 1 : 0, bad);
  LTRY(hipGetLastError(), "validate kernel");
  unsigned long long h = 0;
  LTRY(hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (h == ~0ull) return LJGPU_OK;
  const int64_t row = static_cast<int64_t>(h / 8);
  static const char* what[] = {"", "number_of_partners out of range", "row leaves the list",
                               "partner index out of range", "row lists itself",
                               "half list holds j < i"};
  return set_error(LJGPU_ERR_BAD_LIST,
                   "row " + std::to_string(row) + ": " + what[std::min<unsigned long long>(h % 8, 5)]);
}

int dev_copy_row_ordered(hipStream_t s, int32_t n, const int32_t* nop, const int32_t* ptr,
                         const int32_t* list, int64_t e, ListWork& w, DevCSR& out) {
  (void)e;
  out.n = n;
  LCHECK(dev_reserve(out.nop, sizeof(int32_t) * std::max(n, 1), s));
  LCHECK(dev_reserve(out.ptr, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  if (n > 0)
    LTRY(hipMemcpyAsync(out.nop.p, nop, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s),
         "hipMemcpyAsync");
  int64_t total = 0;
  LCHECK(dev_scan(s, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), n, w, &total));
  if (total > INT32_MAX) return set_error(LJGPU_ERR_CAPACITY, "list longer than 2^31-1 entries");
  out.e = total;
  LCHECK(dev_reserve(out.list, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  if (n > 0) {
    k_copy_rows<<<grid_for(n, kBlock / 64), kBlock, 0, s>>>(n, nop, ptr, list, out.ptr.as<int32_t>(),
                                                           out.list.as<int32_t>());
    LTRY(hipGetLastError(), "copy rows kernel");
  }
  return LJGPU_OK;
}

int dev_expand_half(hipStream_t s, int32_t n, const int32_t* hnop, const int32_t* hptr,
                    const int32_t* hlist, ListWork& w, DevCSR& full) {
  full.n = n;
  LCHECK(dev_reserve(w.counts, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  LCHECK(dev_reserve(w.cell_slot, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  LCHECK(dev_reserve(full.nop, sizeof(int32_t) * std::max(n, 1), s));
  LCHECK(dev_reserve(full.ptr, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  int32_t* lower = w.counts.as<int32_t>();
  int32_t* fillc = w.cell_slot.as<int32_t>();
  LCHECK(fill<int32_t>(s, lower, n, 0));
  LCHECK(fill<int32_t>(s, fillc, n, 0));
  if (n > 0) {
    k_count_lower<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, hnop, hptr, hlist, lower);
    k_add<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, lower, hnop, full.nop.as<int32_t>());
    LTRY(hipGetLastError(), "expand kernels");
  }
  int64_t total = 0;
  LCHECK(dev_scan(s, full.nop.as<int32_t>(), full.ptr.as<int32_t>(), n, w, &total));
  if (total > INT32_MAX) return set_error(LJGPU_ERR_CAPACITY, "full list longer than 2^31-1 entries");
  full.e = total;
  LCHECK(dev_reserve(full.list, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  LCHECK(dev_reserve(w.tmp, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  if (n > 0) {
    k_fill_expanded<<<grid_for(n, kBlock), kBlock, 0, s>>>(
        n, hnop, hptr, hlist, lower, full.ptr.as<int32_t>(), fillc, w.tmp.as<int32_t>(),
        full.list.as<int32_t>());
    LTRY(hipGetLastError(), "expand fill kernel");
    // Lower partners ascending (the order in which the reference's rows
    // r < i list i).
    LCHECK(row_sort(s, n, lower, full.ptr.as<int32_t>(), w.tmp.as<int32_t>(),
                    full.list.as<int32_t>()));
  }
  return LJGPU_OK;
}

int dev_makepair(hipStream_t s, const PosIn& q, int32_t n, double box, double search_length,
                 int kind, ListWork& w, DevCSR& out, bool sort_rows) {
  const double sl2 = search_length * search_length;
  out.n = n;
  LCHECK(dev_reserve(out.nop, sizeof(int32_t) * std::max(n, 1), s));
  LCHECK(dev_reserve(out.ptr, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  if (n == 0) {
    out.e = 0;
    LCHECK(dev_reserve(out.list, sizeof(int32_t), s));
    return fill<int32_t>(s, out.ptr.as<int32_t>(), 1, 0);
  }
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  if (box <= 0.0) {
    // Bounding box of the open system, reduced on the device.
    std::vector<double> ext(6);
    DevMem& m = w.scan_sums;
    LCHECK(dev_reserve(m, sizeof(double) * 6, s));
    const double* comp[3] = {q.x, q.y, q.z};
    for (int c = 0; c < 3; c++)
      k_extrema<<<1, kBlock, 0, s>>>(n, comp[c], q.stride, m.as<double>() + 2 * c);
    LTRY(hipGetLastError(), "bounding box kernel");
    LTRY(hipMemcpyAsync(ext.data(), m.p, sizeof(double) * 6, hipMemcpyDeviceToHost, s),
         "hipMemcpyAsync");
    LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (int c = 0; c < 3; c++) {
      lo[c] = ext[2 * c];
      hi[c] = ext[2 * c + 1];
      if (!std::isfinite(lo[c]) || !std::isfinite(hi[c]))
        return set_error(LJGPU_ERR_INVALID_ARGUMENT, "positions are not finite");
    }
  }
  // Cells at least half the search length wide with a radius-2 stencil
  // visit about 45 % fewer candidates than search-length cells with a
  // radius-1 stencil; a periodic box needs 5 such cells per side (each cell
  // visited once), else the coarser grid is used.
  CellGrid cg;
  make_cell_grid(n, lo, hi, box, 0.5 * search_length, &cg);
  bool fine = !cg.brute;
  for (int c = 0; c < 3; c++)
    if (box > 0.0 && cg.nc[c] < 5) fine = false;
  if (!fine) make_cell_grid(n, lo, hi, box, search_length, &cg);
  GridArgs g{cg.nc[0], cg.nc[1], cg.nc[2], cg.origin[0], cg.origin[1], cg.origin[2],
             cg.width[0], cg.width[1], cg.width[2], box};
  const int64_t ncell = cg.ncell;
  LCHECK(dev_reserve(w.cell_count, sizeof(int32_t) * (ncell + 1), s));
  LCHECK(dev_reserve(w.cell_start, sizeof(int32_t) * (ncell + 1), s));
  LCHECK(dev_reserve(w.cell_slot, sizeof(int32_t) * n, s));
  LCHECK(dev_reserve(w.cell_of, sizeof(int32_t) * n, s));
  LCHECK(dev_reserve(w.cell_sorted, sizeof(int32_t) * n, s));
  LCHECK(dev_reserve(w.cpos, sizeof(double4) * n, s));
  int32_t* count = w.cell_count.as<int32_t>();
  int32_t* cstart = w.cell_start.as<int32_t>();
  LCHECK(fill<int32_t>(s, count, ncell, 0));
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  int32_t* outside = w.small.as<int32_t>() + 12;
  LCHECK(fill<int32_t>(s, outside, 1, 0));
  k_cell_assign<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, q, g, count, w.cell_slot.as<int32_t>(),
                                                       w.cell_of.as<int32_t>(), outside);
  LTRY(hipGetLastError(), "cell assign kernel");
  int64_t total = 0;
  LCHECK(dev_scan(s, count, cstart, ncell, w, &total));
  k_cell_scatter<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, q, w.cell_of.as<int32_t>(),
                                                        w.cell_slot.as<int32_t>(), cstart,
                                                        w.cpos.as<double4>(),
                                                        w.cell_sorted.as<int32_t>());
  LTRY(hipGetLastError(), "cell scatter kernel");
  const bool half = kind == LJGPU_LIST_HALF;
  const int R = cg.brute ? 0 : (fine ? 2 : 1);
  // Unsorted rows need a candidate order fixed by the positions alone: the
  // stencil kernel's run order with each cell's particles by index. The
  // single-cell kernel has no such order, so it always sorts.
  if (R == 0) sort_rows = true;
  if (!sort_rows) {
    // The per-cell insertion sort is quadratic in the cell's population:
    // crowded cells (a collapsed or clustered system) sort rows instead.
    int32_t cmin = 0, cmax = 0;
    LCHECK(dev_minmax(s, count, static_cast<int32_t>(ncell), w, &cmin, &cmax));
    if (cmax > kCellOrderMax) sort_rows = true;
  }
  if (!sort_rows) {
    k_cell_order<<<grid_for(ncell, kBlock), kBlock, 0, s>>>(ncell, cstart, w.cpos.as<double4>());
    LTRY(hipGetLastError(), "cell order kernel");
  }
  const int grid = grid_for(n, kBlock, INT32_MAX);
  const int wave_grid = grid_for(ncell * 64, kBlock);  // one wavefront per cell, grid-stride
  auto pairs = [&](auto mode_tag, int32_t cap, int32_t* nop, const int32_t* ptr, int32_t* outl) {
    constexpr int M = decltype(mode_tag)::value;
    const double4* cp = w.cpos.as<double4>();
    const int32_t* cs = w.cell_sorted.as<int32_t>();
    // A single all-particle cell (R = 0) runs one thread per particle; a
    // stencil grid one wavefront per cell.
#define LJ_PAIRS(H) \
  k_pairs<M, H, 0><<<grid, kBlock, 0, s>>>(n, cp, cs, cstart, g, sl2, cap, nop, ptr, outl)
#define LJ_PAIRS_WAVE(H, RR)                                                                     \
  k_pairs_wave<M, H, RR><<<wave_grid, kBlock, 0, s>>>(static_cast<int32_t>(ncell), cp, cstart, g, \
                                                      sl2, cap, nop, ptr, outl, outside)
    if (half) {
      if (R == 0) LJ_PAIRS(true);
      else if (R == 1) LJ_PAIRS_WAVE(true, 1);
      else LJ_PAIRS_WAVE(true, 2);
    } else {
      if (R == 0) LJ_PAIRS(false);
      else if (R == 1) LJ_PAIRS_WAVE(false, 1);
      else LJ_PAIRS_WAVE(false, 2);
    }
#undef LJ_PAIRS
#undef LJ_PAIRS_WAVE
  };
  // Single pass into a slab of `cap` entries per row, sized from the last
  // build's longest row (or from the mean density), then one sort from the
  // slab into place; a row longer than the slab falls back to a second,
  // exact-size pass. Without room for a slab: count pass, then fill pass.
  int32_t cap = 0;
  {
    double vol = 1.0;
    for (int c = 0; c < 3; c++) vol *= std::max(cg.extent[c], search_length);
    const double expect = 4.18879020478639 * search_length * search_length * search_length *
                          (static_cast<double>(n) / vol) * (half ? 0.5 : 1.0);
    const double want = w.last_max_row > 0 ? 1.2 * w.last_max_row + 16 : 1.5 * expect + 32;
    cap = static_cast<int32_t>(std::min(4096.0, std::max(64.0, want)));
    cap = (cap + 31) / 32 * 32;
    if (static_cast<double>(n) * cap * 4.0 > 16.0 * (1u << 30)) cap = 0;
  }
  int32_t mx = 0;
  if (cap > 0) {
    LCHECK(dev_reserve(w.tmp, sizeof(int32_t) * static_cast<size_t>(n) * cap, s));
    pairs(std::integral_constant<int, 2>{}, cap, out.nop.as<int32_t>(), nullptr,
          w.tmp.as<int32_t>());
    LTRY(hipGetLastError(), "pair kernel");
    int32_t mn = 0;
    LCHECK(dev_minmax(s, out.nop.as<int32_t>(), n, w, &mn, &mx));
  } else {
    pairs(std::integral_constant<int, 0>{}, 0, out.nop.as<int32_t>(), nullptr, nullptr);
    LTRY(hipGetLastError(), "pair count kernel");
  }
  LCHECK(dev_scan(s, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), n, w, &total));
  if (total > INT32_MAX) return set_error(LJGPU_ERR_CAPACITY, "list longer than 2^31-1 entries");
  out.e = total;
  LCHECK(dev_reserve(out.list, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  if (cap > 0 && mx <= cap) {
    w.last_max_row = mx;
    if (!sort_rows) {
      k_copy_slab<<<grid_for((n + kSlabRows - 1) / kSlabRows, kBlock / 64), kBlock, 0, s>>>(
          n, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), w.tmp.as<int32_t>(), cap,
          out.list.as<int32_t>());
      LTRY(hipGetLastError(), "slab copy kernel");
      return LJGPU_OK;
    }
    return row_sort(s, n, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), w.tmp.as<int32_t>(),
                    out.list.as<int32_t>(), cap);
  }
  if (cap > 0) w.last_max_row = mx;
  if (!sort_rows) {
    pairs(std::integral_constant<int, 1>{}, 0, out.nop.as<int32_t>(), out.ptr.as<int32_t>(),
          out.list.as<int32_t>());
    LTRY(hipGetLastError(), "pair fill kernel");
    return LJGPU_OK;
  }
  LCHECK(dev_reserve(w.tmp, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  pairs(std::integral_constant<int, 1>{}, 0, out.nop.as<int32_t>(), out.ptr.as<int32_t>(),
        w.tmp.as<int32_t>());
  LTRY(hipGetLastError(), "pair fill kernel");
  return row_sort(s, n, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), w.tmp.as<int32_t>(),
                  out.list.as<int32_t>());
}

int dev_minmax(hipStream_t s, const int32_t* a, int32_t n, ListWork& w, int32_t* mn,
               int32_t* mx) {
  *mn = 0;
  *mx = 0;
  if (n <= 0) return LJGPU_OK;
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  int32_t* mm = w.small.as<int32_t>();
  LCHECK(fill<int32_t>(s, mm, 1, INT32_MAX));
  LCHECK(fill<int32_t>(s, mm + 1, 1, INT32_MIN));
  k_minmax<<<grid_for(n, kBlock, 256), kBlock, 0, s>>>(n, a, mm);
  LTRY(hipGetLastError(), "minmax kernel");
  int32_t h[2];
  LTRY(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  *mn = h[0];
  *mx = h[1];
  return LJGPU_OK;
}

int dev_half_row_stats(hipStream_t s, const DevCSR& full, ListWork& w, int32_t* min_np,
                       int32_t* max_np) {
  if (full.n <= 0) return dev_minmax(s, nullptr, 0, w, min_np, max_np);
  LCHECK(dev_reserve(w.counts, sizeof(int32_t) * (static_cast<size_t>(full.n) + 1), s));
  k_upper_counts<<<grid_for(full.n, kBlock), kBlock, 0, s>>>(
      full.n, full.nop.as<int32_t>(), full.ptr.as<int32_t>(), full.list.as<int32_t>(),
      w.counts.as<int32_t>());
  LTRY(hipGetLastError(), "upper count kernel");
  return dev_minmax(s, w.counts.as<int32_t>(), full.n, w, min_np, max_np);
}

int dev_make_ell(hipStream_t s, const DevCSR& c, int32_t width, DevMem& ell) {
  width = std::max(width, 1);
  const int64_t cells = static_cast<int64_t>(c.n) * width;
  LCHECK(dev_reserve(ell, sizeof(int32_t) * std::max<int64_t>(cells, 1), s));
  if (c.n > 0) {
    k_ell<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(c.n, width, c.nop.as<int32_t>(),
                                                   c.ptr.as<int32_t>(), c.list.as<int32_t>(),
                                                   ell.as<int32_t>());
    LTRY(hipGetLastError(), "ELL kernel");
  }
  return LJGPU_OK;
}

int dev_tile_order(hipStream_t s, const PosIn& q, int32_t n, double box,
                   const std::vector<int32_t>& replicas, const std::vector<double>& boxes,
                   int32_t T, ListWork& w, bool* ok) {
  *ok = false;
  if (n <= 0 || T <= 0) return LJGPU_OK;
  const std::vector<int32_t> single = {0, n};
  const std::vector<int32_t>& off = replicas.empty() ? single : replicas;
  const int32_t nsys = static_cast<int32_t>(off.size()) - 1;
  std::vector<OrderGrid> grids(nsys);
  int64_t keys = 0;
  for (int32_t r = 0; r < nsys; r++) {
    const double L = replicas.empty() ? box : boxes[r];
    if (!(L > 0.0) || !std::isfinite(L)) return LJGPU_OK;  // open systems keep index order
    const int32_t m = std::max(off[r + 1] - off[r], 1);
    OrderGrid& g = grids[r];
    g.box = L;
    const double rho = m / (L * L * L);
    const double side = 0.75 * std::cbrt(T / rho);
    // Rounded to the nearest count: for lattices of 4k^3 particles L / side
    // often lands exactly on an integer, where truncation would follow
    // last-bit rounding.
    for (int c = 1; c < 3; c++) {
      const double v = std::floor(L / side + 0.5);
      g.nc[c] = v >= 1024.0 ? 1024 : std::max(1, static_cast<int>(v));
    }
    // A few particles in a large box: at most about 2 columns per particle,
    // so the key arrays stay proportional to N.
    while (static_cast<int64_t>(g.nc[1]) * g.nc[2] > 2 * static_cast<int64_t>(m) + 16) {
      int& big = g.nc[1] >= g.nc[2] ? g.nc[1] : g.nc[2];
      big = (big + 1) / 2;
    }
    g.nc[0] = std::max<int32_t>(1, m / (g.nc[1] * g.nc[2]));
    for (int c = 0; c < 3; c++) g.w[c] = L / g.nc[c];
    g.key0 = static_cast<int32_t>(keys);
    keys += static_cast<int64_t>(g.nc[0]) * g.nc[1] * g.nc[2];
    if (keys > INT32_MAX - 1) return LJGPU_OK;
  }
  const size_t gbytes = (sizeof(OrderGrid) * grids.size() + 255) / 256 * 256;
  LCHECK(dev_reserve(w.order_grid, gbytes + sizeof(int32_t) * off.size(), s));
  LTRY(hipMemcpyAsync(w.order_grid.p, grids.data(), sizeof(OrderGrid) * grids.size(),
                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  const int32_t* rep_of = nullptr;
  if (!replicas.empty()) {
    int32_t* doff = reinterpret_cast<int32_t*>(static_cast<char*>(w.order_grid.p) + gbytes);
    LTRY(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, s),
         "hipMemcpyAsync");
    LCHECK(dev_reserve(w.rep_of, sizeof(int32_t) * n, s));
    k_replica_of<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, doff, nsys, w.rep_of.as<int32_t>());
    LTRY(hipGetLastError(), "replica index kernel");
    rep_of = w.rep_of.as<int32_t>();
  }
  LCHECK(dev_reserve(w.cell_count, sizeof(int32_t) * (keys + 1), s));
  LCHECK(dev_reserve(w.cell_start, sizeof(int32_t) * (keys + 1), s));
  LCHECK(dev_reserve(w.cell_slot, sizeof(int32_t) * n, s));
  LCHECK(dev_reserve(w.cell_of, sizeof(int32_t) * n, s));
  LCHECK(dev_reserve(w.order, sizeof(int32_t) * n, s));
  int32_t* count = w.cell_count.as<int32_t>();
  LCHECK(fill<int32_t>(s, count, keys, 0));
  k_order_key<<<grid_for(n, kBlock), kBlock, 0, s>>>(
      n, q, static_cast<const OrderGrid*>(w.order_grid.p), rep_of, count,
      w.cell_slot.as<int32_t>(), w.cell_of.as<int32_t>());
  LTRY(hipGetLastError(), "tile order key kernel");
  int32_t kmin = 0, kmax = 0;
  LCHECK(dev_minmax(s, count, static_cast<int32_t>(keys), w, &kmin, &kmax));
  if (kmax > kCellOrderMax) return LJGPU_OK;  // a crowded bin: keep index order
  int64_t total = 0;
  LCHECK(dev_scan(s, count, w.cell_start.as<int32_t>(), keys, w, &total));
  k_order_scatter<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, w.cell_of.as<int32_t>(),
                                                         w.cell_slot.as<int32_t>(),
                                                         w.cell_start.as<int32_t>(),
                                                         w.order.as<int32_t>());
  k_order_sort<<<grid_for(keys, kBlock), kBlock, 0, s>>>(keys, w.cell_start.as<int32_t>(),
                                                         w.order.as<int32_t>());
  LTRY(hipGetLastError(), "tile order kernels");
  *ok = true;
  return LJGPU_OK;
}

int dev_length_sort(hipStream_t s, const DevCSR& c, int32_t T, const std::vector<int32_t>& reps,
                    ListWork& w) {
  if (T <= 0 || T > kTileRowsMax || c.n <= 0) return LJGPU_OK;
  // The groups are dev_make_tiles's blocks of T consecutive slots.
  std::vector<int32_t> geo;
  const std::vector<int32_t> single = {0, c.n};
  const std::vector<int32_t>& off = reps.empty() ? single : reps;
  for (size_t r = 0; r + 1 < off.size(); r++)
    for (int32_t i0 = off[r]; i0 < off[r + 1]; i0 += T)
      geo.insert(geo.end(), {i0, std::min(T, off[r + 1] - i0), 0, 0});
  const int32_t ng = static_cast<int32_t>(geo.size() / 4);
  if (ng == 0) return LJGPU_OK;
  LCHECK(dev_reserve(w.tile_geo, sizeof(int32_t) * geo.size(), s));
  LTRY(hipMemcpyAsync(w.tile_geo.p, geo.data(), sizeof(int32_t) * geo.size(),
                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  k_group_length_sort<<<std::min(ng, 65535), kBlock, 0, s>>>(
      w.tile_geo.as<int32_t>(), ng, c.nop.as<int32_t>(), w.order.as<int32_t>());
  LTRY(hipGetLastError(), "tile length sort kernel");
  // The upload's host vector must outlive the copy.
  LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  return LJGPU_OK;
}

int dev_make_tiles(hipStream_t s, const DevCSR& c, int32_t T, const std::vector<int32_t>& reps,
                   const int32_t* member, bool bank_order, const CutHint* cut, ListWork& w,
                   DevTiles& out) {
  out.tile_particles = 0;
  out.blocks = 0;
  if (T <= 0 || T > kTileRowsMax || c.n <= 0) return LJGPU_OK;
  // The slot order and each slot's row start in the slot-ordered rows.
  // Index order: slot = particle and the list's own row starts. Otherwise
  // the slot order, the slot of each particle (blocks key their particles
  // by slot) and each slot's row start in the rows stored in slot order.
  out.order = member ? LJGPU_TILE_ORDER_SPATIAL : LJGPU_TILE_ORDER_INDEX;
  const int32_t* mem = nullptr;
  const int32_t* slot_of = nullptr;
  const int32_t* sptr = c.ptr.as<int32_t>();
  if (member) {
    LCHECK(dev_reserve(out.member, sizeof(int32_t) * c.n, s));
    LCHECK(dev_reserve(out.slot_ptr, sizeof(int32_t) * (static_cast<size_t>(c.n) + 1), s));
    LCHECK(dev_reserve(w.slot_of, sizeof(int32_t) * c.n, s));
    LCHECK(dev_reserve(w.counts, sizeof(int32_t) * (static_cast<size_t>(c.n) + 1), s));
    int32_t* m = out.member.as<int32_t>();
    LTRY(hipMemcpyAsync(m, member, sizeof(int32_t) * c.n, hipMemcpyDeviceToDevice, s),
         "hipMemcpyAsync");
    k_inverse<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(c.n, m, w.slot_of.as<int32_t>());
    k_slot_nop<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(c.n, m, c.nop.as<int32_t>(),
                                                         w.counts.as<int32_t>());
    LTRY(hipGetLastError(), "tile order kernels");
    int64_t e = 0;
    LCHECK(dev_scan(s, w.counts.as<int32_t>(), out.slot_ptr.as<int32_t>(), c.n, w, &e));
    if (e != c.e) return set_error(LJGPU_ERR_HIP, "tile order is not a permutation of the rows");
    mem = m;
    slot_of = w.slot_of.as<int32_t>();
    sptr = out.slot_ptr.as<int32_t>();
  }
  // Blocks of T consecutive slots, restarted at every replica boundary.
  std::vector<int32_t> geo;
  const std::vector<int32_t> single = {0, c.n};
  const std::vector<int32_t>& off = reps.empty() ? single : reps;
  for (size_t r = 0; r + 1 < off.size(); r++)
    for (int32_t i0 = off[r]; i0 < off[r + 1]; i0 += T)
      geo.insert(geo.end(), {i0, std::min(T, off[r + 1] - i0), static_cast<int32_t>(r), 0});
  const int32_t nb = static_cast<int32_t>(geo.size() / 4);
  if (nb == 0) return LJGPU_OK;
  LCHECK(dev_reserve(w.tile_geo, sizeof(int32_t) * geo.size(), s));
  LTRY(hipMemcpyAsync(w.tile_geo.p, geo.data(), sizeof(int32_t) * geo.size(),
                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  const int32_t* dgeo = w.tile_geo.as<int32_t>();
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  int32_t* stats = w.small.as<int32_t>() + 4;
  unsigned long long* staged_dev = w.small.as<unsigned long long>() + 4;
  LCHECK(fill<int32_t>(s, stats, 3, 0));
  LCHECK(fill<unsigned long long>(s, staged_dev, 1, 0ull));
  const int grid = std::min(nb, 65535);
  // The force kernel copies entry offsets in whole 16-byte chunks: pad.
  const int64_t nlocal = c.e + kTileLocalPad;
  LCHECK(dev_reserve(out.local, sizeof(uint16_t) * nlocal, s));
  LCHECK(fill<uint16_t>(s, out.local.as<uint16_t>() + c.e, kTileLocalPad, 0));
  LCHECK(dev_reserve(out.desc, sizeof(int32_t) * 8 * static_cast<size_t>(nb), s));
  // One pass when every block can own kTileMaxStage staging slots within
  // kTileOnePassBytes (and the offsets fit int32); otherwise count, scan, fill.
  const int64_t slots = static_cast<int64_t>(nb) * kTileMaxStage;
  const bool one_pass =
      slots * static_cast<int64_t>(sizeof(int32_t)) * (member ? 2 : 1) <= kTileOnePassBytes &&
      slots <= INT32_MAX;
  int64_t staged = 0;
  int32_t h[3];
  if (one_pass) {
    LCHECK(dev_reserve(out.stage, sizeof(int32_t) * slots, s));
    if (member) LCHECK(dev_reserve(out.stage_slot, sizeof(int32_t) * slots, s));
    k_tile_fill<<<grid, kBlock, 0, s>>>(dgeo, nb, mem, slot_of, sptr, c.ptr.as<int32_t>(),
                                        c.list.as<int32_t>(), nullptr, out.desc.as<int32_t>(),
                                        out.stage.as<int32_t>(),
                                        member ? out.stage_slot.as<int32_t>() : nullptr,
                                        out.local.as<uint16_t>(), stats, staged_dev);
    LTRY(hipGetLastError(), "tile fill kernel");
    unsigned long long st = 0;
    LTRY(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    LTRY(hipMemcpyAsync(&st, staged_dev, sizeof(st), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (h[0] != 0) return LJGPU_OK;  // does not fit: caller retries smaller or runs untiled
    staged = static_cast<int64_t>(st);
    out.stage_stride = kTileMaxStage;
  } else {
    LCHECK(dev_reserve(w.tile_ns, sizeof(int32_t) * (static_cast<size_t>(nb) + 1), s));
    LCHECK(dev_reserve(w.tile_soff, sizeof(int32_t) * (static_cast<size_t>(nb) + 1), s));
    k_tile_count<<<grid, kBlock, 0, s>>>(dgeo, nb, mem, slot_of, sptr, c.ptr.as<int32_t>(),
                                         c.list.as<int32_t>(), w.tile_ns.as<int32_t>(), stats);
    LTRY(hipGetLastError(), "tile count kernel");
    LTRY(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (h[0] != 0) return LJGPU_OK;  // does not fit: caller retries smaller or runs untiled
    LCHECK(dev_scan(s, w.tile_ns.as<int32_t>(), w.tile_soff.as<int32_t>(), nb, w, &staged));
    if (staged > INT32_MAX) return LJGPU_OK;
    LCHECK(dev_reserve(out.stage, sizeof(int32_t) * std::max<int64_t>(staged, 1), s));
    // The staged particles as slots too, for evaluations over copies of the
    // positions and momenta in slot order.
    if (member)
      LCHECK(dev_reserve(out.stage_slot, sizeof(int32_t) * std::max<int64_t>(staged, 1), s));
    k_tile_fill<<<grid, kBlock, 0, s>>>(dgeo, nb, mem, slot_of, sptr, c.ptr.as<int32_t>(),
                                        c.list.as<int32_t>(), w.tile_soff.as<int32_t>(),
                                        out.desc.as<int32_t>(), out.stage.as<int32_t>(),
                                        member ? out.stage_slot.as<int32_t>() : nullptr,
                                        out.local.as<uint16_t>(), stats, staged_dev);
    LTRY(hipGetLastError(), "tile fill kernel");
    out.stage_stride = 0;
  }
  if (bank_order) {
    CutDev cd{};
    if (cut && cut->q.x) {
      cd.q = cut->q;
      cd.desc = out.desc.as<int32_t>();
      cd.stage = out.stage.as<int32_t>();
      cd.member = mem;
      cd.box = cut->box;
      cd.rep_box = reps.empty() ? nullptr : cut->rep_box;
      cd.cl2 = cut->cl2;
    }
    k_bank_order<<<grid, kBankBlock, 0, s>>>(dgeo, nb, sptr, out.local.as<uint16_t>(), cd);
    LTRY(hipGetLastError(), "tile bank order kernel");
  }
  out.tile_particles = T;
  out.blocks = nb;
  out.staged = staged;
  out.max_stage = h[1];
  out.lds_bytes = h[2];
  return LJGPU_OK;
}

int dev_slice_tiles(hipStream_t s, const int32_t* slot_ptr, int32_t lanes, int32_t unroll,
                    ListWork& w, DevTiles& t) {
  t.sliced_lanes = t.sliced_unroll = t.waves_per_block = 0;
  if (t.blocks <= 0 || lanes < 1 || lanes > 64 || (64 % lanes) != 0 || unroll < 1 ||
      (unroll & (unroll - 1)) != 0)
    return LJGPU_OK;
  const int32_t wpb = (t.tile_particles * lanes + 63) / 64;
  if (wpb > kSliceWavesMax) return LJGPU_OK;
  const int64_t nw = static_cast<int64_t>(t.blocks) * wpb;
  LCHECK(dev_reserve(w.counts, sizeof(int32_t) * (static_cast<size_t>(nw) + 1), s));
  LCHECK(dev_reserve(t.wave_off, sizeof(int32_t) * (static_cast<size_t>(nw) + 1), s));
  k_wave_slices<<<grid_for(nw, kBlock), kBlock, 0, s>>>(t.desc.as<int32_t>(), t.blocks, wpb,
                                                        slot_ptr, lanes, unroll,
                                                        w.counts.as<int32_t>());
  LTRY(hipGetLastError(), "slice length kernel");
  int64_t total = 0;
  LCHECK(dev_scan(s, w.counts.as<int32_t>(), t.wave_off.as<int32_t>(), nw, w, &total));
  // The kernel's offset prefetch reads up to two batches past a slice.
  const int64_t pad = 2 * 64 * static_cast<int64_t>(unroll) + 64;
  if (total + pad > INT32_MAX) return LJGPU_OK;
  LCHECK(dev_reserve(t.sliced, sizeof(uint16_t) * (total + pad), s));
  LCHECK(fill<uint16_t>(s, t.sliced.as<uint16_t>() + total, pad, 0));
  k_slice_gather<<<std::min(t.blocks, 65535), kBlock, 0, s>>>(
      t.desc.as<int32_t>(), t.blocks, wpb, slot_ptr, t.local.as<uint16_t>(), lanes, unroll,
      t.wave_off.as<int32_t>(), t.sliced.as<uint16_t>());
  LTRY(hipGetLastError(), "slice gather kernel");
  t.sliced_lanes = lanes;
  t.sliced_unroll = unroll;
  t.waves_per_block = wpb;
  return LJGPU_OK;
}

int dev_makepair_batch(hipStream_t s, const PosIn& q, const std::vector<int32_t>& off,
                       const std::vector<double>& boxes, double search_length, int kind,
                       ListWork& w, DevCSR& out) {
  const int32_t nrep = static_cast<int32_t>(off.size()) - 1;
  const int32_t n = off.back();
  out.n = n;
  LCHECK(dev_reserve(out.nop, sizeof(int32_t) * std::max(n, 1), s));
  LCHECK(dev_reserve(out.ptr, sizeof(int32_t) * (static_cast<size_t>(n) + 1), s));
  auto part_of = [&](int32_t r) {
    PosIn qr = q;
    const int64_t shift = static_cast<int64_t>(off[r]) * q.stride;
    qr.x += shift;
    qr.y += shift;
    qr.z += shift;
    return qr;
  };
  // Pass 1: row counts of every replica; pass 2 (after the scan) rebuilds
  // each replica's list and stores it at its place with global indices.
  std::vector<int64_t> e(nrep, 0);
  for (int32_t r = 0; r < nrep; r++) {
    const int32_t nr = off[r + 1] - off[r];
    if (nr == 0) continue;
    LCHECK(dev_makepair(s, part_of(r), nr, boxes[r], search_length, kind, w, w.part));
    e[r] = w.part.e;
    LTRY(hipMemcpyAsync(out.nop.as<int32_t>() + off[r], w.part.nop.p, sizeof(int32_t) * nr,
                        hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  }
  int64_t total = 0;
  LCHECK(dev_scan(s, out.nop.as<int32_t>(), out.ptr.as<int32_t>(), n, w, &total));
  if (total > INT32_MAX) return set_error(LJGPU_ERR_CAPACITY, "list longer than 2^31-1 entries");
  out.e = total;
  LCHECK(dev_reserve(out.list, sizeof(int32_t) * std::max<int64_t>(total, 1), s));
  int64_t pos = 0;
  for (int32_t r = 0; r < nrep; r++) {
    const int32_t nr = off[r + 1] - off[r];
    if (nr == 0) continue;
    LCHECK(dev_makepair(s, part_of(r), nr, boxes[r], search_length, kind, w, w.part));
    if (w.part.e != e[r]) return set_error(LJGPU_ERR_HIP, "batched list build: replica changed size");
    if (e[r] > 0) {
      k_copy_add<<<grid_for(e[r], kBlock), kBlock, 0, s>>>(e[r], w.part.list.as<int32_t>(),
                                                           out.list.as<int32_t>() + pos, off[r]);
      LTRY(hipGetLastError(), "copy kernel");
    }
    pos += e[r];
  }
  return LJGPU_OK;
}

int dev_check_replicas(hipStream_t s, const DevCSR& c, const std::vector<int32_t>& off,
                       ListWork& w) {
  if (c.n <= 0) return LJGPU_OK;
  const int32_t nrep = static_cast<int32_t>(off.size()) - 1;
  LCHECK(dev_reserve(w.rep_of, sizeof(int32_t) * c.n, s));
  LCHECK(dev_reserve(w.tile_geo, sizeof(int32_t) * off.size(), s));
  LTRY(hipMemcpyAsync(w.tile_geo.p, off.data(), sizeof(int32_t) * off.size(),
                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  k_replica_of<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(c.n, w.tile_geo.as<int32_t>(), nrep,
                                                        w.rep_of.as<int32_t>());
  LCHECK(dev_reserve(w.small, sizeof(int64_t) * 8, s));
  unsigned long long* bad = w.small.as<unsigned long long>();
  LCHECK(fill<unsigned long long>(s, bad, 1, ~0ull));
  k_check_replicas<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(
      c.n, c.nop.as<int32_t>(), c.ptr.as<int32_t>(), c.list.as<int32_t>(), w.rep_of.as<int32_t>(),
      bad);
  LTRY(hipGetLastError(), "replica check kernels");
  unsigned long long h = 0;
  LTRY(hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  LTRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (h == ~0ull) return LJGPU_OK;
  return set_error(LJGPU_ERR_BAD_LIST,
                   "row " + std::to_string(h) + " lists a particle of another replica");
}

}  // namespace ljgpu
