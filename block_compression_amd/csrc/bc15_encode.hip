// BC1-BC5 encode kernels (placeholder until the kernels land).
#include "bcn_internal.hpp"

namespace bcn {
hipError_t launch_bc15(const Task&, hipStream_t) { return hipErrorNotSupported; }
}  // namespace bcn
