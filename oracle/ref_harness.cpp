// ref_harness.cpp — a C driver around the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles the reference's
// cpu_ref/force_aos.cpp (or force_soa.cpp) unmodified, where it lies, with its
// main() renamed, and links it with this file into oracle/_ref/libljref_aos.so
// (libljref_soa.so). The reference's globals and functions are reached by
// declaration; nothing of the reference is copied here. Used to write golden
// fixtures (tests/golden/make_golden.py) and, optionally, as a CPU timing
// baseline. Built only where the reference checkout exists.
//
// The reference keeps its particle data in fixed global arrays sized
// N = 400000 and MAX_PAIRS = 30*N (cpu_ref/force_aos.cpp:11-12,17-29).
#include <cstdint>
#include <cstring>
#include <chrono>

#ifndef LJREF_SOA
// cpu_ref/force_aos.cpp:17-29 (AoS: q[N][4], p[N][4])
extern double q[400000][4];
extern double p[400000][4];
#else
// cpu_ref/force_soa.cpp:17-26 (SoA: q[4][N], p[4][N])
extern double q[4][400000];
extern double p[4][400000];
#endif
extern int particle_number;
extern int number_of_pairs;
extern int number_of_partners[400000];
extern int i_particles[12000000];
extern int j_particles[12000000];
extern int pointer[400000];
extern int sorted_list[12000000];
extern double L;

void init(void);
void makepair(void);
void savepair(void);
void loadpair(void);
void sortpair(void);
void force_pair(void);
void force_sorted(void);
void force_intrin(void);
#ifndef LJREF_SOA
void force_next(void);
void force_intrin_mat_transpose(void);
void copy_to_z(void);
void copy_from_z(void);
void force_sorted_z(void);
void force_sorted_z_intrin(void);
#endif

namespace {
constexpr int kMaxN = 400000;
constexpr int kMaxPairs = 12000000;
bool g_init_done = false;
}  // namespace

extern "C" {

int ljref_capacity_particles(void) { return kMaxN; }
int ljref_capacity_pairs(void) { return kMaxPairs; }

// Run the reference's init() once per process (its add_particle keeps a static
// std::mt19937(2), so a second call would continue that stream). Returns N.
int ljref_init(double box_length) {
  if (g_init_done) return -1;
  L = box_length;
  init();
  g_init_done = true;
  return particle_number;
}

// Replace the positions with caller data (AoS stride-4 input for both
// layouts) and zero the momenta.
int ljref_set_positions(int n, const double* q4) {
  if (n < 0 || n > kMaxN) return -1;
  particle_number = n;
  for (int i = 0; i < n; i++) {
#ifndef LJREF_SOA
    q[i][0] = q4[4 * i]; q[i][1] = q4[4 * i + 1]; q[i][2] = q4[4 * i + 2]; q[i][3] = 0.0;
    p[i][0] = p[i][1] = p[i][2] = p[i][3] = 0.0;
#else
    q[0][i] = q4[4 * i]; q[1][i] = q4[4 * i + 1]; q[2][i] = q4[4 * i + 2];
    p[0][i] = p[1][i] = p[2][i] = 0.0;
#endif
  }
  return n;
}

void ljref_get_positions(double* q4) {
  for (int i = 0; i < particle_number; i++) {
#ifndef LJREF_SOA
    q4[4 * i] = q[i][0]; q4[4 * i + 1] = q[i][1]; q4[4 * i + 2] = q[i][2]; q4[4 * i + 3] = 0.0;
#else
    q4[4 * i] = q[0][i]; q4[4 * i + 1] = q[1][i]; q4[4 * i + 2] = q[2][i]; q4[4 * i + 3] = 0.0;
#endif
  }
}

void ljref_zero_momenta(void) {
  for (int i = 0; i < particle_number; i++) {
#ifndef LJREF_SOA
    p[i][0] = p[i][1] = p[i][2] = p[i][3] = 0.0;
#else
    p[0][i] = p[1][i] = p[2][i] = 0.0;
#endif
  }
}

void ljref_get_momenta(double* p4) {
  for (int i = 0; i < particle_number; i++) {
#ifndef LJREF_SOA
    p4[4 * i] = p[i][0]; p4[4 * i + 1] = p[i][1]; p4[4 * i + 2] = p[i][2]; p4[4 * i + 3] = 0.0;
#else
    p4[4 * i] = p[0][i]; p4[4 * i + 1] = p[1][i]; p4[4 * i + 2] = p[2][i]; p4[4 * i + 3] = 0.0;
#endif
  }
}

// The reference's makepair() then sortpair(): O(N^2) half list and its CSR.
// Returns the number of pairs.
int ljref_makepair(void) {
  number_of_pairs = 0;
  makepair();
  sortpair();
  return number_of_pairs;
}

// The reference's savepair(): makepair() then the binary pair.dat in the
// current directory (cpu_ref/force_aos.cpp:612-620), from an empty pair
// count as at program start. Returns the number of pairs.
int ljref_savepair(void) {
  number_of_pairs = 0;
  savepair();
  return number_of_pairs;
}

// The reference's loadpair() of ./pair.dat then sortpair()
// (cpu_ref/force_aos.cpp:603-610,644-645). Returns the number of pairs.
int ljref_loadpair(void) {
  loadpair();
  sortpair();
  return number_of_pairs;
}

int ljref_particle_number(void) { return particle_number; }
int ljref_number_of_pairs(void) { return number_of_pairs; }

void ljref_get_csr(int32_t* nop, int32_t* ptr, int32_t* list) {
  std::memcpy(nop, number_of_partners, sizeof(int) * particle_number);
  std::memcpy(ptr, pointer, sizeof(int) * particle_number);
  std::memcpy(list, sorted_list, sizeof(int) * number_of_pairs);
}

void ljref_get_pairs(int32_t* ip, int32_t* jp) {
  std::memcpy(ip, i_particles, sizeof(int) * number_of_pairs);
  std::memcpy(jp, j_particles, sizeof(int) * number_of_pairs);
}

// Run one reference force routine `loops` times (the reference's measure()
// loop, cpu_ref/force_aos.cpp:592-601) and return the elapsed seconds.
// kind: 0 force_pair, 1 force_sorted, 2 force_intrin, 3 force_next,
//       4 force_intrin_mat_transpose, 5 force_sorted_z, 6 force_sorted_z_intrin
double ljref_run(int kind, int loops) {
  void (*fn)(void) = nullptr;
  bool z = false;
  switch (kind) {
    case 0: fn = force_pair; break;
    case 1: fn = force_sorted; break;
    case 2: fn = force_intrin; break;
#ifndef LJREF_SOA
    case 3: fn = force_next; break;
    case 4: fn = force_intrin_mat_transpose; break;
    case 5: fn = force_sorted_z; z = true; break;
    case 6: fn = force_sorted_z_intrin; z = true; break;
#endif
    default: return -1.0;
  }
#ifndef LJREF_SOA
  if (z) copy_to_z();
#endif
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < loops; s++) fn();
  const auto t1 = std::chrono::steady_clock::now();
#ifndef LJREF_SOA
  if (z) copy_from_z();
#endif
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
