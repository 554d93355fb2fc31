// BC6H encode kernel (placeholder until the kernel lands).
#include "bcn_internal.hpp"

namespace bcn {
hipError_t launch_bc6h(const Task&, hipStream_t) { return hipErrorNotSupported; }
}  // namespace bcn
