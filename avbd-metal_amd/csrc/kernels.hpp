// kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once
#include "state.hpp"

namespace avbd {

// util.hip
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t n, uint32_t value);
// zero up to 6 buffers in one launch
void launch_zero_multi(hipStream_t st, uint32_t* const* ptrs, const uint32_t* counts, int num);
void launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch);
uint32_t scan_scratch_words(uint32_t n);

// collide.hip
void launch_broadphase(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_narrowphase(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_map_rebuild(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_map_find(hipStream_t st, const DeviceState& S, const StepParams& P, const U2* q, int32_t* out,
                     uint32_t n);

// solve.hip
void launch_warmstart(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_contact_adjacency(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
// static palette: one Jacobi pass of c_v = mex{c_u : u < v, u ~ v}; sets CTR_COLOR_CHANGED when any color moved
void launch_static_color_pass(hipStream_t st, const DeviceState& S, const StepParams& P, const uint32_t* in,
                              uint32_t* out);
void launch_dynamic_coloring(hipStream_t st, const DeviceState& S, const StepParams& P);
// softOnly: every body of the launch is a particle whose static entries are tets
void launch_primal_color(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color,
                         uint32_t colorCount, int lanes, bool softOnly);
// colors >= first of a device-sized palette, one workgroup, in order
void launch_primal_tail(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t first, int lanes);
constexpr uint32_t kDynamicColorLaunches = 8;  // dynamic palette: colors launched one kernel each
void launch_dual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_finalize(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_residual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t blocks);
void launch_diag_error(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t* out, uint32_t gridCap);
// persistent cooperative sweep: all iterations x colors x dual in one launch
hipError_t launch_persistent_solve(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t numColors,
                                   int lanes, bool softOnly, uint32_t gridBlocks, uint32_t barrierBase);
int persistent_max_blocks(int lanes, bool softOnly);

}  // namespace avbd
