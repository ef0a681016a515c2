// force_aos3_mixed.hip — instantiates the force kernels of force_kernels.h for
// the kMixedF64 precision policy over packed double[N][3] AoS arrays (one
// translation unit per policy and layout, compiled in parallel).
#include "force_kernels.h"

namespace ljgpu {

int launch_force_aos3_mixed_f64acc(void* stream, const ListView& lv, const ForceArgs& fa,
                                   const LaunchConfig& lc) {
  return detail::run_force_aos3<detail::kMixedF64>(stream, lv, fa, lc);
}

}  // namespace ljgpu
