// cpu_baseline.cpp — the CPU baselines bench.py times beside the GPU kernels.
//
// TEST INFRASTRUCTURE ONLY, like oracle.cpp: only bench.py's cpu_baseline leg
// and tests/ load the library built from this file (oracle/liblj_cpubase.so).
// The checker is oracle.cpp (-O2, no contraction); this file restates the
// reference's CPU force routines for speed, the way the reference builds
// them: -O3 with FMA contraction allowed (cpu_ref/Makefile). The library is
// built in one container and run on the GPU host, so the common AVX2 + FMA
// level (-march=x86-64-v3) stands in for the reference's -march=native.
//
// Baselines (every one accumulates p += F(q) * dt, `steps` times):
//   kind 0  force_sorted: half CSR list, i in registers, j scattered; one core.
//           cpu_ref/force_aos.cpp:192-224 (the reference's "Reference" row).
//   kind 1  force_intrin_mat_transpose: the reference's fastest CPU routine,
//           four partners per AVX2 step with a 4x4 transpose into dx/dy/dz
//           vectors, scalar tail; one core. cpu_ref/force_aos.cpp:413-500.
//   kind 2  full CSR list, one row per i, rows split over all cores with
//           OpenMP (no scatter, so no races); the all-core figure, labelled
//           separately because the reference has no threaded CPU path.
// A periodic box (box > 0) applies the oracle's threshold minimum image
// (d > box/2: d - box; d < -box/2: d + box) to each difference; box <= 0 is
// the reference's open box.
#include <immintrin.h>
#include <omp.h>

#include <chrono>
#include <cstdint>

#ifndef CB_FLAGS
#define CB_FLAGS "unknown"
#endif

namespace {

inline double min_image(double d, double box) {
  if (box > 0.0) {
    const double h = 0.5 * box;
    if (d > h) d -= box;
    else if (d < -h) d += box;
  }
  return d;
}

// Scalar pair term of every reference routine: false beyond the cutoff.
inline bool pair_df(double dx, double dy, double dz, double dt, double cl2, double* df) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 > cl2) return false;
  const double r6 = r2 * r2 * r2;
  *df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2)) * dt;
  return true;
}

void sorted_aos(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                const double* q, double* p, double dt, double cl2, double box) {
  for (int32_t i = 0; i < n; i++) {
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t* row = list + ptr[i];
    for (int32_t k = 0; k < nop[i]; k++) {
      const int32_t j = row[k];
      const double dx = min_image(q[4 * j] - xi, box);
      const double dy = min_image(q[4 * j + 1] - yi, box);
      const double dz = min_image(q[4 * j + 2] - zi, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      fx += df * dx;
      fy += df * dy;
      fz += df * dz;
      p[4 * j] -= df * dx;
      p[4 * j + 1] -= df * dy;
      p[4 * j + 2] -= df * dz;
    }
    p[4 * i] += fx;
    p[4 * i + 1] += fy;
    p[4 * i + 2] += fz;
  }
}

// Minimum image of the four lanes of d (threshold form, see the header).
inline __m256d min_image4(__m256d d, __m256d box, __m256d half, __m256d nhalf) {
  const __m256d up = _mm256_and_pd(_mm256_cmp_pd(d, nhalf, _CMP_LT_OQ), box);
  const __m256d dn = _mm256_and_pd(_mm256_cmp_pd(d, half, _CMP_GT_OQ), box);
  return _mm256_add_pd(_mm256_sub_pd(d, dn), up);
}

void mat_transpose_aos(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                       const double* q, double* p, double dt, double cl2, double box) {
  const bool pbc = box > 0.0;
  const __m256d vbox = _mm256_set1_pd(pbc ? box : 0.0);
  const __m256d vh = _mm256_set1_pd(0.5 * box);
  const __m256d vnh = _mm256_set1_pd(-0.5 * box);
  const __m256d vcl2 = _mm256_set1_pd(cl2);
  const __m256d v24 = _mm256_set1_pd(24.0 * dt);
  const __m256d v48 = _mm256_set1_pd(48.0 * dt);
  for (int32_t i = 0; i < n; i++) {
    const __m256d qi = _mm256_load_pd(q + 4 * i);
    __m256d pi = _mm256_load_pd(p + 4 * i);
    const int32_t np = nop[i];
    const int32_t* row = list + ptr[i];
    const int32_t n4 = np & ~3;
    for (int32_t k = 0; k < n4; k += 4) {
      const int32_t j[4] = {row[k], row[k + 1], row[k + 2], row[k + 3]};
      __m256d d[4];
      for (int u = 0; u < 4; u++) d[u] = _mm256_sub_pd(_mm256_load_pd(q + 4 * j[u]), qi);
      // Rows (x, y, z, w) of the four partners into columns x[4], y[4], z[4].
      const __m256d ab_lo = _mm256_unpacklo_pd(d[0], d[1]);  // ax bx az bz
      const __m256d ab_hi = _mm256_unpackhi_pd(d[0], d[1]);  // ay by aw bw
      const __m256d cd_lo = _mm256_unpacklo_pd(d[2], d[3]);
      const __m256d cd_hi = _mm256_unpackhi_pd(d[2], d[3]);
      __m256d dx = _mm256_permute2f128_pd(ab_lo, cd_lo, 0x20);
      __m256d dy = _mm256_permute2f128_pd(ab_hi, cd_hi, 0x20);
      __m256d dz = _mm256_permute2f128_pd(ab_lo, cd_lo, 0x31);
      if (pbc) {
        dx = min_image4(dx, vbox, vh, vnh);
        dy = min_image4(dy, vbox, vh, vnh);
        dz = min_image4(dz, vbox, vh, vnh);
      }
      const __m256d r2 = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dx, dx), _mm256_mul_pd(dy, dy)),
                                       _mm256_mul_pd(dz, dz));
      const __m256d r6 = _mm256_mul_pd(_mm256_mul_pd(r2, r2), r2);
      __m256d df = _mm256_div_pd(_mm256_sub_pd(_mm256_mul_pd(v24, r6), v48),
                                 _mm256_mul_pd(_mm256_mul_pd(r6, r6), r2));
      df = _mm256_andnot_pd(_mm256_cmp_pd(r2, vcl2, _CMP_GT_OQ), df);  // r2 > cl2: 0
      // Back to rows: force on partner u is df[u] * (dx[u], dy[u], dz[u], 0).
      const __m256d fx = _mm256_mul_pd(df, dx), fy = _mm256_mul_pd(df, dy),
                    fz = _mm256_mul_pd(df, dz), fw = _mm256_setzero_pd();
      const __m256d xy_lo = _mm256_unpacklo_pd(fx, fy);  // ax ay cx cy
      const __m256d xy_hi = _mm256_unpackhi_pd(fx, fy);  // bx by dx dy
      const __m256d zw_lo = _mm256_unpacklo_pd(fz, fw);
      const __m256d zw_hi = _mm256_unpackhi_pd(fz, fw);
      const __m256d f[4] = {_mm256_permute2f128_pd(xy_lo, zw_lo, 0x20),
                            _mm256_permute2f128_pd(xy_hi, zw_hi, 0x20),
                            _mm256_permute2f128_pd(xy_lo, zw_lo, 0x31),
                            _mm256_permute2f128_pd(xy_hi, zw_hi, 0x31)};
      for (int u = 0; u < 4; u++) {
        pi = _mm256_add_pd(pi, f[u]);
        double* pj = p + 4 * j[u];
        _mm256_store_pd(pj, _mm256_sub_pd(_mm256_load_pd(pj), f[u]));
      }
    }
    _mm256_store_pd(p + 4 * i, pi);
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    for (int32_t k = n4; k < np; k++) {
      const int32_t j = row[k];
      const double dx = min_image(q[4 * j] - xi, box);
      const double dy = min_image(q[4 * j + 1] - yi, box);
      const double dz = min_image(q[4 * j + 2] - zi, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      p[4 * i] += df * dx;
      p[4 * i + 1] += df * dy;
      p[4 * i + 2] += df * dz;
      p[4 * j] -= df * dx;
      p[4 * j + 1] -= df * dy;
      p[4 * j + 2] -= df * dz;
    }
  }
}

void full_aos_omp(int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
                  const double* q, double* p, double dt, double cl2, double box) {
#pragma omp parallel for schedule(static, 256)
  for (int32_t i = 0; i < n; i++) {
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t* row = list + ptr[i];
    for (int32_t k = 0; k < nop[i]; k++) {
      const int32_t j = row[k];
      const double dx = min_image(q[4 * j] - xi, box);
      const double dy = min_image(q[4 * j + 1] - yi, box);
      const double dz = min_image(q[4 * j + 2] - zi, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      fx += df * dx;
      fy += df * dy;
      fz += df * dz;
    }
    p[4 * i] += fx;
    p[4 * i + 1] += fy;
    p[4 * i + 2] += fz;
  }
}

}  // namespace

extern "C" {

// Run `steps` evaluations of baseline `kind` (see the header) and return the
// wall time in seconds (steady_clock around the loop, the reference's
// measure(), cpu_ref/force_aos.cpp:592-601); -1 for an unknown kind. q and
// p are AoS double[N][4], 32-byte aligned for kind 1.
double cb_run(int kind, int32_t n, const int32_t* nop, const int32_t* ptr, const int32_t* list,
              const double* q, double* p, double dt, double cl2, double box, int32_t steps) {
  using Fn = void (*)(int32_t, const int32_t*, const int32_t*, const int32_t*, const double*,
                      double*, double, double, double);
  Fn fn = kind == 0 ? sorted_aos : kind == 1 ? mat_transpose_aos : kind == 2 ? full_aos_omp
                                                                             : nullptr;
  if (!fn) return -1.0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t s = 0; s < steps; s++) fn(n, nop, ptr, list, q, p, dt, cl2, box);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Threads kind 2 uses (OpenMP's maximum for this process).
int cb_threads() { return omp_get_max_threads(); }

// Compiler and flags the library was built with.
const char* cb_build_info() { return "g++ " __VERSION__ ", " CB_FLAGS; }

}  // extern "C"
