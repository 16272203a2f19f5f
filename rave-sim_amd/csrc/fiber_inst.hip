// Instantiates the fiber passes of one field precision and x-space variant;
// the build compiles this file once per (RS_FIBER_S, RS_FIBER_XV) so that the
// kernel variants compile in parallel.
#include "fiber.hpp"

#ifndef RS_FIBER_S
#error "RS_FIBER_S (float or double) must be defined"
#endif
#ifndef RS_FIBER_XV
#error "RS_FIBER_XV (0, 1 or 2) must be defined"
#endif

namespace rs {
template hipError_t launch_fiber_xv<RS_FIBER_S, RS_FIBER_XV>(int, bool, const FieldArgs *, int, const Geom &,
                                                             hipStream_t, hipEvent_t, hipEvent_t);
}  // namespace rs
