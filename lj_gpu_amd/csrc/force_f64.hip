// force_f64.hip — instantiates the force kernels of force_kernels.h for the
// kF64 precision policy (one translation unit per policy, compiled in
// parallel), and the tiled potential-energy kernel.
#include "force_kernels.h"

namespace ljgpu {

int launch_force_f64(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  return detail::run_force<detail::kF64>(stream, lv, fa, lc);
}

int launch_energy_tiles(void* stream, const double* q, const ListView& lv, double cl2,
                        double box, double weight, int grid, double* part_u, double* part_c) {
  return detail::run_energy_tiles(static_cast<hipStream_t>(stream), q, lv, cl2, box, weight, grid,
                                  part_u, part_c);
}

}  // namespace ljgpu
