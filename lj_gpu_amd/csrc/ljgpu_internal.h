// ljgpu_internal.h — shared declarations of the ljgpu library (not installed).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ljgpu.h"

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
};

struct ForceArgs {
  bool soa = false;
  const double* q[3] = {nullptr, nullptr, nullptr};  // AoS: q[0] is double[N][4]
  double* p[3] = {nullptr, nullptr, nullptr};
  double dt = 0, cl2 = 0, box = 0;
};

struct LaunchConfig {
  int lanes = 16;       // lanes per i-particle; 1 selects the ELL kernel
  int block = 256;      // threads per workgroup
  bool half_atomic = false;  // list is a half list: scatter to j with atomics
  bool clamp = false;   // clamp every index read into range (unvalidated lists)
};

// Enqueue one force evaluation on `stream`. Returns LJGPU_OK or an error.
int launch_force_f64(void* stream, const ListView& lv, const ForceArgs& fa,
                     const LaunchConfig& lc);
int launch_force_mixed(void* stream, const ListView& lv, const ForceArgs& fa,
                       const LaunchConfig& lc, int accum);

}  // namespace ljgpu
