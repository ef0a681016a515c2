// force_ljgpu — the reference's GPU benchmark program (cuda/force_cuda.cu)
// rebuilt on libljgpu: same problem, same command line, same output.
//
//   force_ljgpu [THREAD_BLOCK] [--density D] [--box-length L] [--loop K]
//               [--dt DT] [--half] [--no-shuffle] [--cache FILE]
//               [--layout aos|soa] [--precision f64|mixed|kahan]
//
// Flow of the reference's main() (cuda/force_cuda.cu:377-420):
//   init()                      -> ljgpu_lattice_reference   (:47-94)
//   loadpair() / makepair()     -> ljgpu_pair_cache_read / ljgpu_makepair_host
//   random_shfl(), makepaircache() -> ljgpu_random_shfl / ljgpu_pair_cache_write
//   measure()                   -> copy in, LOOP force evaluations, copy out (:319-342)
//   print_results()             -> first and last five momenta, "%.10f" (:344-352)
// The full list is the default (the reference's force_cuda.cu without
// EN_ACTION_REACTION); --half builds the half list of its action-reaction
// build. --cache names the pair cache file; without it no cache is read or
// written (the reference always uses .cache_pair_{all,half}.dat in the
// working directory).
#include <sys/time.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ljgpu.h"

namespace {

constexpr double kCutoffLength = 3.0;  // cuda/force_cuda.cu:37-39
constexpr double kSearchLength = 3.3;

struct Options {
  int thread_block = 128;
  double density = 0.5;
  double box_length = 50.0;
  int loop = 100;
  double dt = 0.001;
  int kind = LJGPU_LIST_FULL;
  bool shuffle = true;
  std::string cache;
  bool soa = false;
  int precision = LJGPU_PREC_F64;
  int accum = LJGPU_ACCUM_F64;
};

void check(int rc, const char* what) {
  if (rc != LJGPU_OK) {
    std::fprintf(stderr, "%s: %s (%d)\n", what, ljgpu_last_error(), rc);
    std::exit(1);
  }
}

double myclock() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec * 1e-6;
}

[[noreturn]] void usage(const char* argv0) {
  std::fprintf(stderr,
               "usage: %s [THREAD_BLOCK] [--density D] [--box-length L] [--loop K] [--dt DT]\n"
               "          [--half] [--no-shuffle] [--cache FILE] [--layout aos|soa]\n"
               "          [--precision f64|mixed|kahan]\n",
               argv0);
  std::exit(1);
}

Options parse(int argc, char** argv) {
  Options o;
  for (int a = 1; a < argc; a++) {
    const std::string s = argv[a];
    auto value = [&]() -> const char* {
      if (a + 1 >= argc) usage(argv[0]);
      return argv[++a];
    };
    if (s == "--density") o.density = std::atof(value());
    else if (s == "--box-length") o.box_length = std::atof(value());
    else if (s == "--loop") o.loop = std::atoi(value());
    else if (s == "--dt") o.dt = std::atof(value());
    else if (s == "--half") o.kind = LJGPU_LIST_HALF;
    else if (s == "--no-shuffle") o.shuffle = false;
    else if (s == "--cache") o.cache = value();
    else if (s == "--layout") {
      const std::string v = value();
      if (v != "aos" && v != "soa") usage(argv[0]);
      o.soa = v == "soa";
    } else if (s == "--precision") {
      const std::string v = value();
      if (v == "f64") o.precision = LJGPU_PREC_F64;
      else if (v == "mixed") o.precision = LJGPU_PREC_MIXED;
      else if (v == "kahan") o.precision = LJGPU_PREC_MIXED, o.accum = LJGPU_ACCUM_KAHAN_F32;
      else usage(argv[0]);
    } else if (!s.empty() && s[0] != '-' && a == 1) {
      o.thread_block = std::atoi(s.c_str());
    } else {
      usage(argv[0]);
    }
  }
  // The reference's bounds on argv[1] (cuda/force_cuda.cu:379-383); the
  // library also needs a multiple of 64.
  if (o.thread_block < 64 || o.thread_block > 1024 || o.thread_block % 64) {
    std::fprintf(stderr, "THREAD_BLOCK must be a multiple of 64 in [64, 1024].\n");
    std::exit(1);
  }
  if (o.loop < 0) usage(argv[0]);
  return o;
}

struct List {
  std::vector<int32_t> nop, ptr, lst;
};

// loadpair(): the cache is used when it exists and holds a list for this
// particle number (cuda/force_cuda.cu:203-227).
bool load_cache(const Options& o, int32_t n, List& l) {
  if (o.cache.empty()) return false;
  FILE* f = std::fopen(o.cache.c_str(), "r");
  if (!f) return false;
  std::fclose(f);
  l.nop.assign(n, 0);
  l.ptr.assign(n, 0);
  int64_t e = 0;
  if (ljgpu_pair_cache_read(o.cache.c_str(), n, o.kind, l.nop.data(), l.ptr.data(), nullptr, 0,
                            &e) != LJGPU_OK) {
    std::fprintf(stderr, "%s: %s; rebuilding it.\n", o.cache.c_str(), ljgpu_last_error());
    return false;
  }
  l.lst.assign(static_cast<size_t>(e), 0);
  check(ljgpu_pair_cache_read(o.cache.c_str(), n, o.kind, l.nop.data(), l.ptr.data(),
                              l.lst.data(), e, &e),
        "ljgpu_pair_cache_read");
  return true;
}

// makepair() + random_shfl() + makepaircache() (cuda/force_cuda.cu:122-163,
// 165-181, 255-263).
void make_list(const Options& o, int32_t n, const std::vector<double>& q, List& l) {
  std::fprintf(stderr, "Now make pairlist %s.\n", o.cache.empty() ? "(no cache)" : o.cache.c_str());
  l.nop.assign(n, 0);
  l.ptr.assign(n, 0);
  int64_t e = 0;
  const int rc = ljgpu_makepair_host(n, q.data(), 0.0, kSearchLength, o.kind, l.nop.data(),
                                     l.ptr.data(), nullptr, 0, &e);
  if (rc != LJGPU_OK && rc != LJGPU_ERR_CAPACITY) check(rc, "ljgpu_makepair_host");
  l.lst.assign(static_cast<size_t>(e), 0);
  check(ljgpu_makepair_host(n, q.data(), 0.0, kSearchLength, o.kind, l.nop.data(), l.ptr.data(),
                            l.lst.data(), e, &e),
        "ljgpu_makepair_host");
  if (o.shuffle) check(ljgpu_random_shfl(n, l.nop.data(), l.ptr.data(), l.lst.data(), 10),
                       "ljgpu_random_shfl");
  if (!o.cache.empty())
    check(ljgpu_pair_cache_write(o.cache.c_str(), n, l.nop.data(), l.ptr.data(), l.lst.data(),
                                 static_cast<int64_t>(l.lst.size())),
          "ljgpu_pair_cache_write");
}

const char* kernel_name(const Options& o) {
  if (o.precision == LJGPU_PREC_F64) return o.soa ? "ljgpu_force_soa_f64" : "ljgpu_force_aos_f64";
  if (o.accum == LJGPU_ACCUM_KAHAN_F32)
    return o.soa ? "ljgpu_force_soa_mixed_kahan" : "ljgpu_force_aos_mixed_kahan";
  return o.soa ? "ljgpu_force_soa_mixed" : "ljgpu_force_aos_mixed";
}

// measure() (cuda/force_cuda.cu:319-342): host-to-device copies, LOOP force
// evaluations, device-to-host copy; the two timings it prints. p_host is
// AoS double[N][4] in and out.
void measure(ljgpu_ctx* ctx, const Options& o, int32_t n, const std::vector<double>& q_host,
             std::vector<double>& p_host) {
  const size_t nn = static_cast<size_t>(n);
  const char* name = kernel_name(o);
  if (!o.soa) {
    const size_t bytes = nn * 4 * sizeof(double);
    double *q = nullptr, *p = nullptr;
    check(ljgpu_malloc(ctx, bytes, reinterpret_cast<void**>(&q)), "ljgpu_malloc");
    check(ljgpu_malloc(ctx, bytes, reinterpret_cast<void**>(&p)), "ljgpu_malloc");
    const double st_all = myclock();
    check(ljgpu_memcpy_htod(ctx, q, q_host.data(), bytes), "ljgpu_memcpy_htod");
    check(ljgpu_memcpy_htod(ctx, p, p_host.data(), bytes), "ljgpu_memcpy_htod");
    const double st_calc = myclock();
    if (o.precision == LJGPU_PREC_F64)
      check(ljgpu_force_aos_f64(ctx, q, p, o.dt, kCutoffLength * kCutoffLength, 0.0, o.loop),
            name);
    else
      check(ljgpu_force_aos_mixed(ctx, q, p, o.dt, kCutoffLength * kCutoffLength, 0.0, o.loop,
                                  o.accum),
            name);
    check(ljgpu_sync(ctx), "ljgpu_sync");
    const double diff_calc = myclock() - st_calc;
    check(ljgpu_memcpy_dtoh(ctx, p_host.data(), p, bytes), "ljgpu_memcpy_dtoh");
    const double diff_all = myclock() - st_all;
    std::fprintf(stderr, "N=%d, %s %f [sec]\n", n, name, diff_all);
    std::fprintf(stderr, "N=%d, %s %f [sec] (without Host<->Device)\n", n, name, diff_calc);
    check(ljgpu_free(ctx, q), "ljgpu_free");
    check(ljgpu_free(ctx, p), "ljgpu_free");
    return;
  }
  // SoA: the reference CPU programs' q[X][N] layout (cpu_ref/force_soa.cpp).
  const size_t bytes = nn * sizeof(double);
  std::vector<double> hq(3 * nn), hp(3 * nn);
  for (size_t i = 0; i < nn; i++)
    for (int d = 0; d < 3; d++) {
      hq[d * nn + i] = q_host[4 * i + d];
      hp[d * nn + i] = p_host[4 * i + d];
    }
  double* buf[6] = {};
  for (double*& b : buf) check(ljgpu_malloc(ctx, bytes, reinterpret_cast<void**>(&b)), "ljgpu_malloc");
  double **q = buf, **p = buf + 3;
  const double st_all = myclock();
  for (int d = 0; d < 3; d++) {
    check(ljgpu_memcpy_htod(ctx, q[d], hq.data() + d * nn, bytes), "ljgpu_memcpy_htod");
    check(ljgpu_memcpy_htod(ctx, p[d], hp.data() + d * nn, bytes), "ljgpu_memcpy_htod");
  }
  const double st_calc = myclock();
  const double cl2 = kCutoffLength * kCutoffLength;
  if (o.precision == LJGPU_PREC_F64)
    check(ljgpu_force_soa_f64(ctx, q[0], q[1], q[2], p[0], p[1], p[2], o.dt, cl2, 0.0, o.loop),
          name);
  else
    check(ljgpu_force_soa_mixed(ctx, q[0], q[1], q[2], p[0], p[1], p[2], o.dt, cl2, 0.0, o.loop,
                                o.accum),
          name);
  check(ljgpu_sync(ctx), "ljgpu_sync");
  const double diff_calc = myclock() - st_calc;
  for (int d = 0; d < 3; d++)
    check(ljgpu_memcpy_dtoh(ctx, hp.data() + d * nn, p[d], bytes), "ljgpu_memcpy_dtoh");
  const double diff_all = myclock() - st_all;
  std::fprintf(stderr, "N=%d, %s %f [sec]\n", n, name, diff_all);
  std::fprintf(stderr, "N=%d, %s %f [sec] (without Host<->Device)\n", n, name, diff_calc);
  for (size_t i = 0; i < nn; i++)
    for (int d = 0; d < 3; d++) p_host[4 * i + d] = hp[d * nn + i];
  for (double* b : buf) check(ljgpu_free(ctx, b), "ljgpu_free");
}

// print_results() (cuda/force_cuda.cu:344-352).
void print_results(int32_t n, const std::vector<double>& p) {
  auto row = [&](int32_t i) {
    std::fprintf(stdout, "%.10f %.10f %.10f\n", p[4 * i], p[4 * i + 1], p[4 * i + 2]);
  };
  for (int32_t i = 0; i < 5 && i < n; i++) row(i);
  for (int32_t i = n > 5 ? n - 5 : 0; i < n; i++) row(i);
}

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse(argc, argv);

  // init(): query the particle count, then fill.
  int32_t n = 0;
  std::vector<double> q(4);
  int rc = ljgpu_lattice_reference(o.density, o.box_length, q.data(), 0, &n);
  if (rc != LJGPU_OK && rc != LJGPU_ERR_CAPACITY) check(rc, "ljgpu_lattice_reference");
  q.assign(4 * static_cast<size_t>(n), 0.0);
  check(ljgpu_lattice_reference(o.density, o.box_length, q.data(), n, &n),
        "ljgpu_lattice_reference");
  std::vector<double> p(q.size(), 0.0);

  List l;
  if (!load_cache(o, n, l)) make_list(o, n, q, l);

  ljgpu_ctx* ctx = nullptr;
  check(ljgpu_create(0, &ctx), "ljgpu_create");
  check(ljgpu_set_kernel_config(ctx, 0, o.thread_block, 0), "ljgpu_set_kernel_config");
  check(ljgpu_set_list_csr(ctx, n, l.nop.data(), l.ptr.data(), l.lst.data(),
                           static_cast<int64_t>(l.lst.size()), o.kind, LJGPU_MEM_HOST),
        "ljgpu_set_list_csr");
  measure(ctx, o, n, q, p);
  print_results(n, p);
  check(ljgpu_destroy(ctx), "ljgpu_destroy");
  return 0;
}
