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
// exclusive scan of min(*nDev, cap) + 1 elements (length read on the device; nMax bounds it)
void launch_exclusive_scan_dev(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t nMax, const uint32_t* nDev,
                               uint32_t cap, uint32_t* scratch);

// collide.hip
void launch_broadphase(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_narrowphase(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_map_rebuild(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_map_find(hipStream_t st, const DeviceState& S, const StepParams& P, const U2* q, int32_t* out,
                     uint32_t n);

// upload.hip: boundary faces of the tet meshes, wound outward, ascending;
// false when the scene is outside the device path (2^21 or more bodies) or
// a HIP call fails, and the caller takes the host path
bool gpu_tet_boundary_faces(const uint32_t* tetIds, uint32_t nt, const float* bodyPos, uint32_t nb, hipStream_t st,
                            std::vector<std::array<uint32_t, 3>>& out);

// upload.hip: per-tet rest shape into tetsDev and the slot of every tet entry
// of the (already uploaded) static adjacency into slotsDev
hipError_t gpu_tet_precompute(const uint32_t* tetIds, const float* mu, const float* lam, uint32_t nt,
                              const float* bodyPos, uint32_t nb, const uint32_t* sStartDev, const uint32_t* sCountDev,
                              const uint32_t* sListDev, TetRec* tetsDev, TetSlot* slotsDev, hipStream_t st);

// cloth.hip (no-ops without cloth triangles)
// rigid feature vs triangle contacts at the step-start pose: the prefix of this frame's soft records
void launch_rt_emit(hipStream_t st, const DeviceState& S, const StepParams& P);
// element grid + V-T / E-E pair stream at the current (predictor) pose
void launch_dat_pairs(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
// legacy servo rates: each listed joint's motor target advances by rate * dt, wrapped to [-pi, pi]
void launch_spinners(hipStream_t st, const DeviceState& S, const uint32_t* bodies, const uint32_t* start,
                     const F4* dq, uint32_t n);
void launch_rate_motors(hipStream_t st, const DeviceState& S, const uint32_t* joints, const float* rates, uint32_t n,
                        float dt);
// OGC records from the pair stream, soft counts, soft persistence map rebuild
void launch_dat_contacts(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
// isotropic limiter: tracked sets, clearance bounds and tracked contacts at
// the step-start pose (after the rigid-feature records); the conditional
// bound refresh after each iteration's dual
void launch_iso_track(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_ogc_refresh(hipStream_t st, const DeviceState& S, const StepParams& P);
// Planar-DAT truncation: the full pass after the predictor (anchor = step
// start), the anchor at the accepted predictor pose, the pass after each
// static color, and the failed-surface restore before finalize
void launch_dat_predictor_pass(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_dat_reanchor(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_dat_color_pass(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color, uint32_t start,
                           uint32_t count, uint32_t gridCap);
void launch_dat_restore(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_viscosity(hipStream_t st, const DeviceState& S, const StepParams& P);
// debug lookup in the element-contact persistence map (keys -> previous record index or -1)
void launch_soft_map_find(hipStream_t st, const DeviceState& S, const StepParams& P, const U2* q, int32_t* out,
                          uint32_t n);

// solve.hip
void launch_warmstart(hipStream_t st, const DeviceState& S, const StepParams& P);
void launch_contact_adjacency(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
// static palette: one Jacobi pass of c_v = mex{c_u : u < v, u ~ v}; sets CTR_COLOR_CHANGED when any color moved
void launch_static_color_pass(hipStream_t st, const DeviceState& S, const StepParams& P, const uint32_t* in,
                              uint32_t* out);
void launch_dynamic_coloring(hipStream_t st, const DeviceState& S, const StepParams& P);
// primal specialisations of a color's bodies
constexpr int kModeGeneral = 0;  // rigid bodies and particles, any entries
constexpr int kModeTet = 1;      // particles whose static entries are all tets
constexpr int kModeParticle = 2; // particles only (cloth slots, springs, contacts)
// colorCount 0xFFFFFFFF: device-sized color of the dynamic palette; otherwise
// the color's `colorCount` work records start at colorWork[base]
void launch_primal_color(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color,
                         uint32_t colorCount, uint32_t base, int lanes, int mode);
// colors >= first of a device-sized palette, one workgroup, in order
void launch_primal_tail(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t first, int lanes);
constexpr uint32_t kDynamicColorLaunches = 8;  // dynamic palette: colors launched one kernel each
void launch_dual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap);
void launch_finalize(hipStream_t st, const DeviceState& S, const StepParams& P);
// guard: skip once a residual-terminated frame has ended its iterations
void launch_residual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t blocks, bool guard = false);
void launch_residual_check(hipStream_t st, const DeviceState& S, uint32_t blocks, float tol, uint32_t maxIt,
                           uint32_t n, ConvHost* host);
void launch_diag_error(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t* out, uint32_t gridCap);
// persistent cooperative sweep: all iterations x colors x dual in one launch
hipError_t launch_persistent_solve(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t numColors,
                                   int lanes, int mode, uint32_t gridBlocks, uint32_t barrierBase);
int persistent_max_blocks(int lanes, int mode);

}  // namespace avbd
