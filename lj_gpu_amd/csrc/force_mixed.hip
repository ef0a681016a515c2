// force_mixed.hip — instantiates the force kernels of force_kernels.h for the
// kMixedF64 precision policy (one translation unit per policy, compiled in
// parallel).
#include "force_kernels.h"

namespace ljgpu {

int launch_force_mixed_f64acc(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  return detail::run_force<detail::kMixedF64>(stream, lv, fa, lc);
}

}  // namespace ljgpu
