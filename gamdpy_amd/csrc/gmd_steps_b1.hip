// gmd_steps_b1.hip -- the step kernels of pair-function base 1 (LJ_12_6),
// compiled apart from gmd_core.hip so the instances build in parallel.
#include <hip/hip_runtime.h>

#include "gamdpy_amd.h"
#include "gmd_steps.h"

GMD_STEP_KERNELS_BASE(, 1)
