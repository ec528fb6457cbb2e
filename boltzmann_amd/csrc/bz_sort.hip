// Energy/cell binning, stable radix sort and ES-WPNC resampling (gfx950).
#include "bz_launch.h"

namespace bz {

size_t sort_workspace_bytes(int64_t capacity, int32_t max_keys) {
  (void)max_keys;
  return (size_t)(capacity > 0 ? capacity : 1) * 4;
}

int launch_bin_sort(const LaunchEnv&, const bz_state_t*, const bz_params_t*, const double*, int,
                    int32_t*, int32_t*, uint32_t*) {
  return fail("bz_bin_sort: not built yet");
}

int launch_resample(const LaunchEnv&, const bz_state_t*, const bz_state_t*, const bz_params_t*,
                    const int32_t*, const int32_t*, const uint32_t*, const int32_t*, int32_t,
                    const bz_resample_policy_t*) {
  return fail("bz_resample: not built yet");
}

}  // namespace bz
