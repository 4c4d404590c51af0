/*
 * Wind-tunnel solver C ABI over the aero_lbm library: one solver object per
 * handle, a solid mask, an optional starting flow field, and steps with a
 * wind-tunnel, pressure-outlet, closed or periodic box. Same names, argument
 * meaning, return convention (1 = ok, 0 = failure) and struct layouts as the
 * reference's native/include/aero_solver_capi.h, so the reference's ctypes tools
 * bind unchanged. Flow buffers hold 4 floats per cell (vx, vy, vz in m/s, then
 * the pressure proxy), cell = (x * ny + y) * nz + z.
 */
#ifndef AERO_SOLVER_CAPI_H
#define AERO_SOLVER_CAPI_H

#include <stdint.h>

#include "aero_lbm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ref: native/include/aero_solver_capi.h:12-17 */
enum {
    AERO_SOLVER_ABI_VERSION = 1,
    AERO_SOLVER_FLOW_CHANNELS = 4,
    AERO_SOLVER_STATUS_OK = 1,
    AERO_SOLVER_STATUS_ERROR = 0
};

/* ref: native/include/aero_solver_capi.h:19-24 */
typedef enum AeroSolverBoundaryMode {
    AERO_SOLVER_BOUNDARY_WIND_TUNNEL = 1,
    AERO_SOLVER_BOUNDARY_WIND_TUNNEL_PRESSURE_OUTLET = 2,
    AERO_SOLVER_BOUNDARY_CLOSED = 3,
    AERO_SOLVER_BOUNDARY_PERIODIC = 4
} AeroSolverBoundaryMode;

/* ref: native/include/aero_solver_capi.h:26-32 (dx in m, dt in s) */
typedef struct AeroGridDesc {
    int nx;
    int ny;
    int nz;
    float dx;
    float dt;
} AeroGridDesc;

/* ref: native/include/aero_solver_capi.h:34-42 (velocities in m/s, viscosity in m^2/s) */
typedef struct AeroBoundaryDesc {
    int mode;
    float inlet_vx;
    float inlet_vy;
    float inlet_vz;
    float outlet_pressure;
    float density;
    float viscosity;
} AeroBoundaryDesc;

/* ref: native/include/aero_solver_capi.h:44-50 */
typedef struct AeroStepInput {
    const uint8_t* solid_mask;
    const float* prev_flow;
    AeroGridDesc grid;
    AeroBoundaryDesc boundary;
    int steps;
} AeroStepInput;

/* ref: native/include/aero_solver_capi.h:52-56 */
typedef struct AeroStepOutput {
    float* flow_out;
    float max_velocity;
    int status;
} AeroStepOutput;

/* ref: native/include/aero_solver_capi.h:58-59: 64^3, dx 1 m, dt 0.05 s; wind tunnel at 5 m/s */
AERO_LBM_CAPI_EXPORT void aero_solver_default_grid(AeroGridDesc* out_grid);
AERO_LBM_CAPI_EXPORT void aero_solver_default_boundary(AeroBoundaryDesc* out_boundary);

/* ref: native/include/aero_solver_capi.h:61-73. Creating a solver replaces every
 * earlier one (the library holds one grid shape at a time). */
AERO_LBM_CAPI_EXPORT int aero_solver_create(int nx, int ny, int nz, float dx, float dt, long long* out_handle);
AERO_LBM_CAPI_EXPORT int aero_solver_create_with_grid(const AeroGridDesc* grid, long long* out_handle);

/* ref: native/include/aero_solver_capi.h:75-85; cell_count / value_count must be
 * nx*ny*nz / 4*nx*ny*nz. Either call restarts the flow from the new state. */
AERO_LBM_CAPI_EXPORT int aero_solver_set_solid_mask(long long handle, const uint8_t* solid_mask, int cell_count);
AERO_LBM_CAPI_EXPORT int aero_solver_set_flow_state(long long handle, const float* flow, int value_count);

/* ref: native/include/aero_solver_capi.h:87-99; boundary NULL = default
 * boundary; out_flow (optional) receives the flow after the last step. */
AERO_LBM_CAPI_EXPORT int aero_solver_step_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary, int steps,
                                                      float* out_flow, int out_value_count);
AERO_LBM_CAPI_EXPORT int aero_solver_advance_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary,
                                                         int steps);

/* ref: native/include/aero_solver_capi.h:101-104: create, set mask / flow, step,
 * read back, destroy in one call; output->max_velocity in m/s */
AERO_LBM_CAPI_EXPORT int aero_solver_run_wind_tunnel(const AeroStepInput* input, AeroStepOutput* output);

/* ref: native/include/aero_solver_capi.h:106-109 */
AERO_LBM_CAPI_EXPORT void aero_solver_destroy(long long handle);
AERO_LBM_CAPI_EXPORT const char* aero_solver_last_error(void);
AERO_LBM_CAPI_EXPORT const char* aero_solver_runtime_info(void);
AERO_LBM_CAPI_EXPORT int aero_solver_finish(void);

#ifdef __cplusplus
}
#endif

#endif
