// Shadow of the reference's include/constants.hpp for the config-1 harness build
// (make -C oracle ref, target _ref/libswe_ref_g981.so): the same namespace and names, with
// the gravity config 1 states (9.81 instead of the reference's 9.80665).  Placed ahead of
// the reference root on the include path, it is the header the reference's kernels pick up
// for "include/constants.hpp"; the kernel headers themselves compile unmodified.
// TEST INFRASTRUCTURE ONLY.
#pragma once

namespace constants {
inline constexpr double g = 9.81;
inline constexpr double eps = 1e-8;
}  // namespace constants
