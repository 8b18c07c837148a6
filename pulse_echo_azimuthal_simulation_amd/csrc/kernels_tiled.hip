// Tiled z-marching kernels (PE_KERNEL_TILED): not yet built.
#include "common.hpp"

namespace pe {

template <typename R>
hipError_t launch_stress_tiled(const Shot<R> &, hipStream_t) { return hipErrorNotSupported; }
template <typename R>
hipError_t launch_velocity_tiled(const Shot<R> &, hipStream_t) { return hipErrorNotSupported; }

template hipError_t launch_stress_tiled<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_stress_tiled<double>(const Shot<double> &, hipStream_t);
template hipError_t launch_velocity_tiled<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_velocity_tiled<double>(const Shot<double> &, hipStream_t);

}  // namespace pe
