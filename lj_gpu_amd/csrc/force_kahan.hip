// force_kahan.hip — instantiates the force kernels of force_kernels.h for the
// kMixedKahan precision policy (one translation unit per policy, compiled in
// parallel).
#include "force_kernels.h"

namespace ljgpu {

int launch_force_mixed_kahan(void* stream, const ListView& lv, const ForceArgs& fa, const LaunchConfig& lc) {
  return detail::run_force<detail::kMixedKahan>(stream, lv, fa, lc);
}

}  // namespace ljgpu
