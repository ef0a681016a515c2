// oracle.cpp — CPU restatement of kohnakagawa/lj_gpu's Lennard-Jones force path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
// shared library built from it (oracle/liblj_oracle.so). The product library
// (lj_gpu_amd/libljgpu.so) neither links nor calls anything here.
//
// Compiled with -O2 -ffp-contract=off so every expression is evaluated exactly
// as written (no FMA contraction): this is the canonical arithmetic the HIP
// kernels are compared against.
//
// Every function cites the reference lines it restates (paths relative to the
// kohnakagawa/lj_gpu checkout). Parity pins, by function:
//   orc_lattice_reference   ref_data/density{0.5,1}.dat via the full pipeline,
//                           and positions dumped from the compiled cpu_ref
//   orc_lattice_pbc         no reference counterpart (the reference has no
//                           periodic box); pinned by construction tests:
//                           zero force on the unjittered lattice, particle
//                           count, box length, jitter bounds and shared
//                           std::mt19937 stream with orc_lattice_reference
//   orc_makepair_*          pair counts and CSR checksums of the compiled
//                           cpu_ref makepair/sortpair; brute-force reflexivity
//   orc_force_*             ref_data (10 decimals) and 17-digit momenta from
//                           the compiled cpu_ref force_pair/force_sorted;
//                           two-body closed forms, Newton's third law,
//                           half-vs-full-list agreement
//   orc_random_shfl         an independent restatement of the permutation
//                           (tests/test_random_shfl.py): numpy's legacy
//                           MT19937 seeded with 10 is std::mt19937(10), and
//                           libstdc++'s std::shuffle over its
//                           uniform_int_distribution is restated in Python;
//                           every row, one stream across rows, must match
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <algorithm>
#include <vector>
#include <chrono>

extern "C" {

// ---------------------------------------------------------------------------
// Lattices
// ---------------------------------------------------------------------------

// The reference's init()/add_particle(): FCC lattice in an open box of side L,
// lattice constant s = 1/((density/4)^(1/3)), sx = int(L/s) cells per side,
// particle order (iz, iy, ix, basis), each coordinate jittered by U(0, 0.1)
// drawn from one std::mt19937(2) stream in x, y, z order.
// cpu_ref/force_aos.cpp:43-51,142-169; cuda/force_cuda.cu:47-94.
// q is AoS with stride 4 (x, y, z, 0). Returns the particle count, or -1 if
// cap is too small.
int32_t orc_lattice_reference(double density, double L, double* q, int32_t cap) {
  std::mt19937 mt(2);
  std::uniform_real_distribution<double> ud(0.0, 0.1);
  const double s = 1.0 / pow(density * 0.25, 1.0 / 3.0);
  const double hs = s * 0.5;
  const int sx = static_cast<int>(L / s);
  const int sy = static_cast<int>(L / s);
  const int sz = static_cast<int>(L / s);
  const int64_t total = 4LL * sx * sy * sz;
  if (total > cap) return -1;
  int32_t n = 0;
  auto add = [&](double x, double y, double z) {
    const double jx = ud(mt);
    const double jy = ud(mt);
    const double jz = ud(mt);
    q[4 * n + 0] = x + jx;
    q[4 * n + 1] = y + jy;
    q[4 * n + 2] = z + jz;
    q[4 * n + 3] = 0.0;
    n++;
  };
  for (int iz = 0; iz < sz; iz++)
    for (int iy = 0; iy < sy; iy++)
      for (int ix = 0; ix < sx; ix++) {
        const double x = ix * s, y = iy * s, z = iz * s;
        add(x, y, z);
        add(x, y + hs, z + hs);
        add(x + hs, y, z + hs);
        add(x + hs, y + hs, z);
      }
  return n;
}

// Periodic FCC lattice of 4*k^3 particles: same lattice constant formula and
// particle order as the reference's init(), box side L = k*s, jitter
// U(jlo, jhi) per coordinate from std::mt19937(seed) in x, y, z order, then
// each coordinate wrapped into [0, L). The reference has no periodic mode;
// this is the generator of the periodic workloads (jitter U(-0.05, 0.05)).
// Returns N and writes L to *box_out.
int32_t orc_lattice_pbc(int32_t k, double density, uint32_t seed, double jlo, double jhi,
                        double* q, int32_t cap, double* box_out) {
  const double s = 1.0 / pow(density * 0.25, 1.0 / 3.0);
  const double hs = s * 0.5;
  const double L = k * s;
  const int64_t total = 4LL * k * k * k;
  if (total > cap) return -1;
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> ud(jlo, jhi);
  int32_t n = 0;
  auto wrap = [L](double v) {
    if (v < 0.0) v += L;
    if (v >= L) v -= L;
    return v;
  };
  auto add = [&](double x, double y, double z) {
    const double jx = ud(mt);
    const double jy = ud(mt);
    const double jz = ud(mt);
    q[4 * n + 0] = wrap(x + jx);
    q[4 * n + 1] = wrap(y + jy);
    q[4 * n + 2] = wrap(z + jz);
    q[4 * n + 3] = 0.0;
    n++;
  };
  for (int iz = 0; iz < k; iz++)
    for (int iy = 0; iy < k; iy++)
      for (int ix = 0; ix < k; ix++) {
        const double x = ix * s, y = iy * s, z = iz * s;
        add(x, y, z);
        add(x, y + hs, z + hs);
        add(x + hs, y, z + hs);
        add(x + hs, y + hs, z);
      }
  *box_out = L;
  return n;
}

// ---------------------------------------------------------------------------
// Minimum image (periodic box only; box <= 0 means the reference's open box)
// ---------------------------------------------------------------------------
static inline double min_image(double d, double box) {
  if (box > 0.0) {
    const double h = 0.5 * box;
    if (d > h) d -= box;
    else if (d < -h) d += box;
  }
  return d;
}

// ---------------------------------------------------------------------------
// Neighbour lists
// ---------------------------------------------------------------------------

// Brute-force O(N^2) half list (i < j) in the reference's order: rows in
// ascending i, partners in ascending j, predicate r2 < SL2 with
// r2 = (dx*dx + dy*dy) + dz*dz and dx = q_i - q_j (minimum image if box > 0).
// cpu_ref/force_aos.cpp:60-74,122-140 (makepair + register_pair) followed by
// :100-120 (sortpair), which for this insertion order leaves each row sorted.
// Output is CSR: nop[N], pointer[N] (exclusive scan), sorted_list[E].
// Returns E, or -1 if cap is exceeded. Rows are found in parallel (OpenMP) and
// concatenated in i order, so the result is independent of the thread count.
int64_t orc_makepair_half(int32_t n, const double* q, double box, double sl2,
                          int32_t* nop, int32_t* pointer, int32_t* sorted_list, int64_t cap) {
  std::vector<std::vector<int32_t>> rows(n > 0 ? n : 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t i = 0; i < n - 1; i++) {
    std::vector<int32_t>& r = rows[i];
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    for (int32_t j = i + 1; j < n; j++) {
      const double dx = min_image(xi - q[4 * j], box);
      const double dy = min_image(yi - q[4 * j + 1], box);
      const double dz = min_image(zi - q[4 * j + 2], box);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < sl2) r.push_back(j);
    }
  }
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    nop[i] = static_cast<int32_t>(rows[i].size());
    pointer[i] = static_cast<int32_t>(pos);
    if (pos + static_cast<int64_t>(rows[i].size()) > cap) return -1;
    std::memcpy(sorted_list + pos, rows[i].data(), rows[i].size() * sizeof(int32_t));
    pos += static_cast<int64_t>(rows[i].size());
  }
  return pos;
}

// Brute-force full list (every ordered pair i != j), the reference's non
// action-reaction makepair: for i, for j != i ascending, r2 < SL2.
// cuda/force_cuda.cu:102-163 (EN_ACTION_REACTION undefined).
int64_t orc_makepair_full(int32_t n, const double* q, double box, double sl2,
                          int32_t* nop, int32_t* pointer, int32_t* sorted_list, int64_t cap) {
  std::vector<std::vector<int32_t>> rows(n > 0 ? n : 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t i = 0; i < n; i++) {
    std::vector<int32_t>& r = rows[i];
    const double xi = q[4 * i], yi = q[4 * i + 1], zi = q[4 * i + 2];
    for (int32_t j = 0; j < n; j++) {
      if (i == j) continue;
      const double dx = min_image(xi - q[4 * j], box);
      const double dy = min_image(yi - q[4 * j + 1], box);
      const double dz = min_image(zi - q[4 * j + 2], box);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < sl2) r.push_back(j);
    }
  }
  int64_t pos = 0;
  for (int32_t i = 0; i < n; i++) {
    nop[i] = static_cast<int32_t>(rows[i].size());
    pointer[i] = static_cast<int32_t>(pos);
    if (pos + static_cast<int64_t>(rows[i].size()) > cap) return -1;
    std::memcpy(sorted_list + pos, rows[i].data(), rows[i].size() * sizeof(int32_t));
    pos += static_cast<int64_t>(rows[i].size());
  }
  return pos;
}

// The reference's sortpair(): counting sort of a flat (i, j) pair array into
// CSR, keeping insertion order inside each row. nop must hold the per-i counts.
// cpu_ref/force_aos.cpp:100-120.
void orc_sortpair(int32_t n, const int32_t* nop, const int32_t* i_particles,
                  const int32_t* j_particles, int64_t npairs, int32_t* pointer,
                  int32_t* sorted_list) {
  if (n <= 0) return;
  std::vector<int32_t> pointer2(n, 0);
  int32_t pos = 0;
  pointer[0] = 0;
  for (int32_t i = 0; i < n - 1; i++) {
    pos += nop[i];
    pointer[i + 1] = pos;
  }
  for (int64_t k = 0; k < npairs; k++) {
    const int32_t i = i_particles[k];
    sorted_list[pointer[i] + pointer2[i]] = j_particles[k];
    pointer2[i]++;
  }
}

// The CUDA driver's random_shfl(): shuffle every CSR row with one
// std::mt19937(seed) stream (seed 10 in the reference).
// cuda/force_cuda.cu:255-263.
void orc_random_shfl(int32_t n, const int32_t* nop, const int32_t* pointer,
                     int32_t* sorted_list, uint32_t seed) {
  std::mt19937 mt(seed);
  for (int32_t i = 0; i < n; i++) {
    std::shuffle(sorted_list + pointer[i], sorted_list + pointer[i] + nop[i], mt);
  }
}

// The CUDA driver's make_transposed_pairlist(): ELL layout with stride N,
// entry k of row i at transposed[i + k*N], zero padded to `width` entries.
// cuda/force_cuda.cu:229-240.
void orc_make_transposed(int32_t n, const int32_t* nop, const int32_t* pointer,
                         const int32_t* sorted_list, int32_t width, int32_t* transposed) {
  std::memset(transposed, 0, sizeof(int32_t) * static_cast<size_t>(n) * width);
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < nop[i]; k++)
      transposed[i + static_cast<int64_t>(k) * n] = sorted_list[pointer[i] + k];
}

// ---------------------------------------------------------------------------
// Forces: p += F(q) * dt.  The pair term of every reference force routine:
//   d = q_j - q_i; r2 = (dx*dx + dy*dy) + dz*dz; skip if r2 > CL2;
//   r6 = r2*r2*r2; df = ((24*r6 - 48) / (r6*r6*r2)) * dt;
//   p_i += df*d  (and p_j -= df*d for half lists).
// cpu_ref/force_aos.cpp:176-182; cuda/kernel.cuh:22-28.
// ---------------------------------------------------------------------------
static inline bool pair_df(double dx, double dy, double dz, double dt, double cl2, double* df) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 > cl2) return false;
  const double r6 = r2 * r2 * r2;
  *df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2)) * dt;
  return true;
}

// force_pair over a flat half-pair array, AoS stride 4.
// cpu_ref/force_aos.cpp:171-190.
void orc_force_pair_aos(int64_t npairs, const int32_t* ip, const int32_t* jp, const double* q,
                        double* p, double dt, double cl2, double box) {
  for (int64_t k = 0; k < npairs; k++) {
    const int32_t i = ip[k], j = jp[k];
    const double dx = min_image(q[4 * j] - q[4 * i], box);
    const double dy = min_image(q[4 * j + 1] - q[4 * i + 1], box);
    const double dz = min_image(q[4 * j + 2] - q[4 * i + 2], box);
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

// force_pair, SoA (q[3][N] as three arrays). cpu_ref/force_soa.cpp:142-161.
void orc_force_pair_soa(int64_t npairs, const int32_t* ip, const int32_t* jp, const double* qx,
                        const double* qy, const double* qz, double* px, double* py, double* pz,
                        double dt, double cl2, double box) {
  for (int64_t k = 0; k < npairs; k++) {
    const int32_t i = ip[k], j = jp[k];
    const double dx = min_image(qx[j] - qx[i], box);
    const double dy = min_image(qy[j] - qy[i], box);
    const double dz = min_image(qz[j] - qz[i], box);
    double df;
    if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
    px[i] += df * dx;
    py[i] += df * dy;
    pz[i] += df * dz;
    px[j] -= df * dx;
    py[j] -= df * dy;
    pz[j] -= df * dz;
  }
}

// force_sorted: CSR half list, i accumulated in registers, j scattered.
// cpu_ref/force_aos.cpp:192-224.
void orc_force_sorted_aos(int32_t n, const int32_t* nop, const int32_t* pointer,
                          const int32_t* sorted_list, const double* q, double* p, double dt,
                          double cl2, double box) {
  for (int32_t i = 0; i < n; i++) {
    const double qx = q[4 * i], qy = q[4 * i + 1], qz = q[4 * i + 2];
    const int32_t np = nop[i], kp = pointer[i];
    double pfx = 0, pfy = 0, pfz = 0;
    for (int32_t k = 0; k < np; k++) {
      const int32_t j = sorted_list[kp + k];
      const double dx = min_image(q[4 * j] - qx, box);
      const double dy = min_image(q[4 * j + 1] - qy, box);
      const double dz = min_image(q[4 * j + 2] - qz, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      pfx += df * dx;
      pfy += df * dy;
      pfz += df * dz;
      p[4 * j] -= df * dx;
      p[4 * j + 1] -= df * dy;
      p[4 * j + 2] -= df * dz;
    }
    p[4 * i] += pfx;
    p[4 * i + 1] += pfy;
    p[4 * i + 2] += pfz;
  }
}

// force_sorted, SoA. cpu_ref/force_soa.cpp:163-195.
void orc_force_sorted_soa(int32_t n, const int32_t* nop, const int32_t* pointer,
                          const int32_t* sorted_list, const double* qx, const double* qy,
                          const double* qz, double* px, double* py, double* pz, double dt,
                          double cl2, double box) {
  for (int32_t i = 0; i < n; i++) {
    const double xi = qx[i], yi = qy[i], zi = qz[i];
    const int32_t np = nop[i], kp = pointer[i];
    double pfx = 0, pfy = 0, pfz = 0;
    for (int32_t k = 0; k < np; k++) {
      const int32_t j = sorted_list[kp + k];
      const double dx = min_image(qx[j] - xi, box);
      const double dy = min_image(qy[j] - yi, box);
      const double dz = min_image(qz[j] - zi, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      pfx += df * dx;
      pfy += df * dy;
      pfz += df * dz;
      px[j] -= df * dx;
      py[j] -= df * dy;
      pz[j] -= df * dz;
    }
    px[i] += pfx;
    py[i] += pfy;
    pz[i] += pfz;
  }
}

// Full-list force (every ordered pair present): each row only updates its own
// i, the semantics of the CUDA full-list kernels (force_kernel_memopt and its
// ELL/warp relatives accumulate p[i] only). cuda/kernel.cuh:67-100,821-860.
void orc_force_full_aos(int32_t n, const int32_t* nop, const int32_t* pointer,
                        const int32_t* sorted_list, const double* q, double* p, double dt,
                        double cl2, double box) {
  for (int32_t i = 0; i < n; i++) {
    const double qx = q[4 * i], qy = q[4 * i + 1], qz = q[4 * i + 2];
    const int32_t np = nop[i], kp = pointer[i];
    double pfx = 0, pfy = 0, pfz = 0;
    for (int32_t k = 0; k < np; k++) {
      const int32_t j = sorted_list[kp + k];
      const double dx = min_image(q[4 * j] - qx, box);
      const double dy = min_image(q[4 * j + 1] - qy, box);
      const double dz = min_image(q[4 * j + 2] - qz, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      pfx += df * dx;
      pfy += df * dy;
      pfz += df * dz;
    }
    p[4 * i] += pfx;
    p[4 * i + 1] += pfy;
    p[4 * i + 2] += pfz;
  }
}

// Full-list force, SoA.
void orc_force_full_soa(int32_t n, const int32_t* nop, const int32_t* pointer,
                        const int32_t* sorted_list, const double* qx, const double* qy,
                        const double* qz, double* px, double* py, double* pz, double dt,
                        double cl2, double box) {
  for (int32_t i = 0; i < n; i++) {
    const double xi = qx[i], yi = qy[i], zi = qz[i];
    const int32_t np = nop[i], kp = pointer[i];
    double pfx = 0, pfy = 0, pfz = 0;
    for (int32_t k = 0; k < np; k++) {
      const int32_t j = sorted_list[kp + k];
      const double dx = min_image(qx[j] - xi, box);
      const double dy = min_image(qy[j] - yi, box);
      const double dz = min_image(qz[j] - zi, box);
      double df;
      if (!pair_df(dx, dy, dz, dt, cl2, &df)) continue;
      pfx += df * dx;
      pfy += df * dy;
      pfz += df * dz;
    }
    px[i] += pfx;
    py[i] += pfy;
    pz[i] += pfz;
  }
}

// Run `steps` repetitions of force_sorted (AoS) and return the wall time in
// seconds: the reference's measure() loop, cpu_ref/force_aos.cpp:592-601,
// timed with steady_clock. Used as the CPU baseline.
double orc_time_force_sorted_aos(int32_t n, const int32_t* nop, const int32_t* pointer,
                                 const int32_t* sorted_list, const double* q, double* p,
                                 double dt, double cl2, double box, int32_t steps) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t s = 0; s < steps; s++)
    orc_force_sorted_aos(n, nop, pointer, sorted_list, q, p, dt, cl2, box);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
