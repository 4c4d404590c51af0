/*
 * D3Q19 mesoscale oracle -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's
 * cpu_baseline). libaero_lbm never links or calls it.
 *
 * Restates native/src/aero_lbm_mesoscale.cpp:110-738 (mesoscale_cpu_step and
 * helpers) in FP32 (prefix m32_, the reference's operation order) and FP64
 * (prefix m64_). The reference TU includes <jni.h>, absent from this image, so
 * it is not compiled; tests/test_oracle_meso.py pins this restatement through
 * invariants (quiescent isothermal state with flat terrain stays fixed, TRT with
 * equal odd/even rates reduces to BGK, mass conservation in the interior) and
 * the transport formulas against the reference's hydro/thermal core.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_API __attribute__((visibility("default")))

/* aero_lbm_mesoscale.cpp:146-176 */
static const int kMCX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int kMCY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int kMCZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int kMOPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

#define R float
#define MS(n) m32_##n
#define K(x) ((float)(x##f))
#include "meso_oracle_impl.h"
#undef R
#undef MS
#undef K

#define R double
#define MS(n) m64_##n
#define K(x) ((double)(x))
#include "meso_oracle_impl.h"
