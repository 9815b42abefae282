// gmd_steps_b2.hip -- the step kernels of pair-function base 2 (the IPL_n / LJ_m_n / harmonic / hertzian family),
// compiled apart from gmd_core.hip so the instances build in parallel.
#include <hip/hip_runtime.h>

#include "gamdpy_amd.h"
#include "gmd_steps.h"

GMD_STEP_KERNELS_BASE(, 2)
