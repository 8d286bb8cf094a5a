// solver.hip -- host runtime of the gfx950 AVBD backend: scene upload and
// precompute, device buffers, per-step orchestration (captured in hipGraphs),
// the lane-split / sweep-mode autotuner, failure latching and readback.
//
// Follows GPUSolver (Sources/PhysicsAVBD/GPU/GPUSolver.swift): init :338,
// upload :901-2081, encodeAndSubmitStep :3138-4405, retire :2829,
// synchronize :2924, setBodyStates :2409.
#include <math.h>
#include <string.h>
#include <stdio.h>
#include <algorithm>
#include <array>
#include <functional>
#include <map>
#include <chrono>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/avbd.h"
#include "kernels.hpp"
#include "lifetime.hpp"
#include "mlp.hpp"

// upload-time geometry must round exactly as written (radii feed the
// hashed/global partition and the cell size)
#pragma clang fp contract(off)

using namespace avbd;

namespace {

uint32_t next_pow2(uint32_t v) {
  uint32_t n = 1;
  while (n < v) n <<= 1;
  return n;
}

struct HostBody {
  float mass;
  V3 moment;
  float radius;
};

}  // namespace

struct avbd_solver {
  avbd_settings settings{};
  avbd_options opts{};
  uint32_t nb = 0, nc = 0, nj = 0, ns = 0, nt = 0;
  uint32_t nHashed = 0, nGlobal = 0, nHashedRigid = 0, nExcl = 0;
  uint32_t numBpLane = 0, numBpWave = 0;  // pair producers on the lane / wave paths
  uint32_t maxPairs = 0, mapCapacity = 0, gridHashSize = 0;
  float cellSize = 0.1f;
  float fineCellSize = 0.0f;  // candidate search grid (0 = off)
  bool dynamicColoring = false;
  uint32_t staticUsedColors = 1;
  std::vector<uint32_t> staticColorCounts;  // per color
  std::vector<uint32_t> staticColorStarts;  // per color: first position in the color order
  std::vector<uint8_t> colorMode;           // per color: primal specialisation (kMode*)
  std::vector<ClothSlot> clothSlotsHost;    // static cloth slots without same-color bits
  bool hasClothSlots = false;               // membrane / bend / particle-spring entries exist
  std::vector<uint32_t> sAdjStartHost;
  int allMode = kModeGeneral;               // common mode of every used color, else general
  // cloth surfaces
  uint32_t nTris = 0, nEdges = 0, nParticles = 0, nMembranes = 0, nBends = 0;
  uint32_t numContactEdges = 0;  // authored-triangle edges: the E-E contact prefix of the edge list
  bool pairStream = false;       // V-T / E-E pair stream and its contacts (some pair can be accepted)
  bool datOn = false;            // Planar-DAT limiter passes (a self-colliding surface exists)
  bool isoOn = false;            // isotropic OGC limiter and its tracked V-T / E-E contacts
  uint32_t lastOgcTrunc = 0, lastOgcRefresh = 0;  // the last retired frame's truncations / refreshes
  // wall time of the create phases (host clock, device work synchronised)
  std::vector<std::pair<const char*, double>> createPhases;
  std::chrono::steady_clock::time_point phaseT0;
  std::vector<uint32_t> rateJoints;  // implicit-PD motors with a legacy servo rate
  std::vector<float> rateValues;
  // kinematic spinners grouped by body (bodies in first-appearance order,
  // each body's spinners in scene order): CSR over spinAxisOmega / spinDq
  std::vector<uint32_t> spinBodies, spinStart;
  std::vector<F4> spinAxisOmega;
  uint32_t* spinBodyDev = nullptr;
  uint32_t* spinStartDev = nullptr;
  F4* spinDqDev = nullptr;
  float spinDt = 0.0f;  // the step dt spinDqDev was computed for
  F4* jointPenInit = nullptr;      // upload-time joint penalties (lin [nj], ang [nj]), restored on reset
  int device = 0;                  // the HIP device current at create
  uint32_t* fixedScratch = nullptr;  // this device's address of the no-allocation staging array
  uint32_t* rateJointDev = nullptr;
  float* rateValueDev = nullptr;
  uint32_t elemHashSize = 0, maxSoft = 0, softMapCapacity = 0, maxDatPairs = 0;
  float elemCellSize = 0.0f, elemMargin = 0.0f, datQueryRadius = 0.0f, surfaceContactCellSize = 0.0f;
  float datQueryRadiusDerived = 0.0f;  // rq derived at create (a test hook may replace datQueryRadius)
  int lanes = 8, sweepMode = 1;
  int persistentBlocks = 0;
  uint32_t gridCap = 1024;
  hipStream_t stream = nullptr;
  DeviceState S{};
  std::vector<void*> allocs;
  // host mirrors
  std::vector<float> mass;
  std::vector<float> inertia;  // n*3
  std::vector<uint8_t> isParticle;
  std::vector<uint32_t> sAdjCount;
  std::vector<uint32_t> sAdjListHost;
  std::vector<U4> jointHeaders;
  // frames in flight
  uint32_t* pinnedCounters = nullptr;  // [2][CTR_TOTAL]
  hipEvent_t frameDone[2] = {nullptr, nullptr};
  int64_t slotFrame[2] = {-1, -1};
  bool slotBusy[2] = {false, false};
  int64_t frameIndex = 0;
  // graphs (even / odd frame: manifold buffers alternate)
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  bool graphsValid = false;
  bool graphsDisabled = false;
  int64_t graphBaseFrame = 1;  // frame that replays graphs[0]
  // failure latch
  avbd_failure failure{};
  bool failed = false;
  // stats
  avbd_stats stats{};
  // test hook
  std::string deniedStage;
  int completionFailKind = 0;        // completion failure injected at retire (0 = none)
  int64_t completionFailFrame = -1;
  int64_t deniedFrame = -1;
  // profiling
  bool profiling = false;
  std::vector<std::string> profNames;
  std::vector<float> profMs;
  // scratch for queries
  uint32_t* devScratchU32 = nullptr;
  uint32_t devScratchWords = 0;
  double* residualPartialHost = nullptr;
  uint32_t residualBlocks = 0;
  // residual-terminated frames: the device's progress record (mapped host
  // memory) and the captured pre-solve / chunk / post-solve graphs of the
  // even and odd frame roles, for one (chunk, tolerance, cap) key
  ConvHost* convHost = nullptr;
  ConvHost* convDev = nullptr;
  hipGraphExec_t resPre[2] = {nullptr, nullptr}, resChunk[2] = {nullptr, nullptr}, resPost[2] = {nullptr, nullptr};
  bool resValid = false;
  int32_t resCheck = 0, resMax = 0;
  float resTol = 0.0f;
  int64_t resBase = 1;  // frame whose roles resPre/resChunk/resPost[0] were captured with
  uint32_t cus = 256;
  // learned warm start
  avbd_mlp* mlp = nullptr;       // 3 outputs: dynamic particles
  avbd_mlp* mlpRigid = nullptr;  // 6 outputs: dynamic rigid bodies
  int wsMode = 0;
  bool wsBuilt = false;
  WarmState W{};
  std::vector<U4> tetIdsHost;
  std::vector<U2> springEndsHost;
};

namespace {

avbd_status latch(avbd_solver* s, int kind, int64_t frame, int64_t required, int64_t capacity, int hipErr,
                  const char* msg) {
  if (!s->failed) {
    s->failed = true;
    s->failure.kind = kind;
    s->failure.frame = frame;
    s->failure.required = required;
    s->failure.capacity = capacity;
    s->failure.hip_error = hipErr;
    snprintf(s->failure.message, sizeof(s->failure.message), "%s", msg);
  }
  return AVBD_ERR_RUNTIME_FAILURE;
}

#define HIP_OR_LATCH(s, expr, frame)                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) {                                                                       \
      char _m[160];                                                                               \
      snprintf(_m, sizeof(_m), "%s: %s", #expr, hipGetErrorString(_e));                           \
      return latch(s, AVBD_FAIL_COMMAND_EXECUTION, frame, 0, 0, (int)_e, _m);                     \
    }                                                                                             \
  } while (0)

template <typename T>
hipError_t dev_alloc(avbd_solver* s, T** p, size_t count) {
  size_t bytes = count * sizeof(T);
  if (bytes < 16) bytes = 16;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return e;
  e = hipMemset(q, 0, bytes);
  s->allocs.push_back(q);
  *p = (T*)q;
  return e;
}

template <typename T>
hipError_t upload(T* dst, const std::vector<T>& src) {
  if (src.empty()) return hipSuccess;
  return hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

StepParams make_params(const avbd_solver* s) {
  StepParams P{};
  P.dt = s->settings.dt;
  P.gravity = s->settings.gravity;
  P.alpha = s->settings.alpha;
  P.betaLin = s->settings.beta_lin;
  P.betaAng = s->settings.beta_ang;
  P.gamma = s->settings.gamma;
  P.lambdaMax = s->settings.lambda_max;
  P.cellSize = s->cellSize;
  P.fineCellSize = s->fineCellSize;
  P.maxSpeed = fmaxf(30.0f, 1.5f * s->cellSize / s->settings.dt * 0.5f);
  P.collisionMargin = fmaxf(s->settings.collision_margin, 0.0f);
  P.particleDamping = s->settings.particle_damping;
  P.rigidLinearDamping = fmaxf(s->settings.rigid_linear_damping, 0.0f);
  P.rigidAngularDamping = fmaxf(s->settings.rigid_angular_damping, 0.0f);
  P.numBodies = s->nb;
  P.numColliders = s->nc;
  P.numJoints = s->nj;
  P.numSprings = s->ns;
  P.numTets = s->nt;
  P.mapCapacity = s->mapCapacity;
  P.maxPairs = s->maxPairs;
  P.gridHashSize = s->gridHashSize;
  P.numHashed = s->nHashed;
  P.numGlobals = s->nGlobal;
  P.numHashedRigid = s->nHashedRigid;
  P.numBpLane = s->numBpLane;
  P.numBpWave = s->numBpWave;
  P.frictionCombineMode = s->settings.friction_combine_mode;
  P.numExclusions = s->nExcl;
  P.iterations = (uint32_t)std::max(0, s->settings.iterations);
  P.staticPalette = s->dynamicColoring ? 0u : 1u;
  P.numTris = s->nTris;
  P.numEdges = s->nEdges;
  P.numContactEdges = s->numContactEdges;
  P.datOn = s->datOn ? 1u : 0u;
  P.isoOn = s->isoOn ? 1u : 0u;
  P.pairStream = s->pairStream ? 1u : 0u;
  P.numParticles = s->nParticles;
  P.elemHashSize = s->elemHashSize;
  P.maxSoft = s->maxSoft;
  P.softMapCapacity = s->softMapCapacity;
  P.maxDatPairs = s->maxDatPairs;
  P.elemCellSize = s->elemCellSize;
  P.elemMargin = s->elemMargin;
  P.datQueryRadius = s->datQueryRadius;
  P.surfaceContactCellSize = s->surfaceContactCellSize;
  P.clothViscosity = s->settings.cloth_viscosity;
  return P;
}

// the learned stage's per-frame parameters (L follows the current settings)
WarmState ws_params(const avbd_solver* s) {
  WarmState W = s->W;
  const float dt = s->settings.dt;
  W.L = fmaxf(fabsf(s->settings.gravity) * dt * dt, 1.0e-6f);
  W.force = (s->wsMode & AVBD_WS_FORCE) ? 1 : 0;
  W.log = (s->wsMode & AVBD_WS_LOG) ? 1 : 0;
  return W;
}

bool stage_denied(const avbd_solver* s, const char* stage, int64_t frame) {
  return !s->deniedStage.empty() && s->deniedStage == stage && (s->deniedFrame < 0 || s->deniedFrame == frame);
}

// One frame is pre-solve (collision, warm start, adjacency, coloring), the
// iterations, and post-solve (finalize, viscosity); `stage` (profiling) is
// called after each stage with its name.
using StageFn = std::function<void(const char*)>;

void enqueue_pre_solve(avbd_solver* s, const DeviceState& S, const StepParams& P, const StageFn& stage) {
  hipStream_t st = s->stream;
  uint32_t* zp[5] = {S.counters, S.cellCount, S.cellRigid, S.dAdjCount, S.fineCount};
  uint32_t zn[5] = {CTR_TOTAL, s->gridHashSize, s->gridHashSize, s->nb, s->fineCellSize > 0.0f ? s->gridHashSize : 0u};
  launch_zero_multi(st, zp, zn, 5);
  if (s->datOn) launch_fill_u32(st, S.datT, s->nb, 0x3F800000u);  // 1.0f: no pass has lowered a fraction
  if (!s->rateJoints.empty())
    launch_rate_motors(st, S, s->rateJointDev, s->rateValueDev, (uint32_t)s->rateJoints.size(), P.dt);
  if (!s->spinBodies.empty())
    launch_spinners(st, S, s->spinBodyDev, s->spinStartDev, s->spinDqDev, (uint32_t)s->spinBodies.size());
  launch_broadphase(st, S, P);
  stage("broadphase");
  launch_narrowphase(st, S, P, s->gridCap);
  launch_map_rebuild(st, S, P, s->gridCap);
  stage("narrowphase");
  // cloth (GPUSolver.swift:3685-3870): the V-T / E-E query and the
  // rigid-feature contacts at the step-start pose; the predictor and its
  // full Planar-DAT pass; the query again at the accepted predictor pose,
  // which feeds the OGC contacts and every per-color pass, anchored there
  if (s->nTris) {
    if (s->datOn) launch_dat_pairs(st, S, P, s->gridCap);
    stage("dat-pairs");
    launch_rt_emit(st, S, P);
    stage("rt-emit");
    launch_iso_track(st, S, P, s->gridCap);
    stage("vt-ee-track");
  }
  launch_warmstart(st, S, P);
  stage("warmstart");
  if (s->nTris) {
    if (s->datOn) launch_dat_predictor_pass(st, S, P, s->gridCap);
    stage("dat-predict");
    launch_dat_pairs(st, S, P, s->gridCap);
    stage("dat-accepted-pairs");
    launch_dat_contacts(st, S, P, s->gridCap);
    stage("vt-ee-emit");
    if (s->datOn) launch_dat_reanchor(st, S, P);
    stage("dat-reanchor");
  }
  launch_contact_adjacency(st, S, P, s->gridCap);
  stage("adjacency");
  if (s->dynamicColoring) launch_dynamic_coloring(st, S, P);
  stage("coloring");
  // learned warm start: features, then (with a network) the learned start
  // and the residual comparison against the plain guess
  if (s->wsMode && (s->W.n || s->W.nR)) {
    const WarmState W = ws_params(s);
    launch_ws_features(st, S, P, W);
    launch_ws_features_rigid(st, S, P, W);
    const bool lp = s->mlp && W.n, lr = s->mlpRigid && W.nR;
    if ((s->wsMode & AVBD_WS_LEARNED) && (lp || lr)) {
      launch_residual(st, S, P, s->residualBlocks);
      launch_ws_reduce(st, S, s->residualBlocks, W, 0);
      if (lp) {
        launch_mlp_forward(st, s->mlp->M, W.x, W.n, W.y, s->cus);
        launch_ws_apply(st, S, W);
      }
      if (lr) {
        launch_mlp_forward(st, s->mlpRigid->M, W.xR, W.nR, W.yR, s->cus);
        launch_ws_apply_rigid(st, S, W);
      }
      launch_residual(st, S, P, s->residualBlocks);
      launch_ws_reduce(st, S, s->residualBlocks, W, 1);
      launch_ws_select(st, S, W);
    }
    stage("learned-start");
  }
}

// `iterations` sweeps (every color, then the dual); the persistent mode runs
// them in one cooperative launch whose barrier counter starts at zero
hipError_t enqueue_iterations(avbd_solver* s, const DeviceState& S, const StepParams& P0, uint32_t iterations,
                              bool clearBarrier) {
  hipStream_t st = s->stream;
  StepParams P = P0;
  P.iterations = iterations;
  if (iterations == 0) return hipSuccess;
  if (s->sweepMode == 2) {
    if (clearBarrier) launch_fill_u32(st, S.counters + CTR_GBAR, 1, 0u);
    return launch_persistent_solve(st, S, P, s->dynamicColoring ? 0u : s->staticUsedColors, s->lanes, s->allMode,
                                   (uint32_t)s->persistentBlocks, 0u);
  }
  for (uint32_t it = 0; it < iterations; it++) {
    if (s->dynamicColoring) {
      for (uint32_t c = 0; c < kDynamicColorLaunches; c++)
        launch_primal_color(st, S, P, c, 0xFFFFFFFFu, 0u, s->lanes, kModeGeneral);
      launch_primal_tail(st, S, P, kDynamicColorLaunches, s->lanes);
    } else {
      for (uint32_t c = 0; c < s->staticUsedColors; c++) {
        launch_primal_color(st, S, P, c, s->staticColorCounts[c], s->staticColorStarts[c], s->lanes, s->colorMode[c]);
        // Planar-DAT after every color (GPUSolver.swift:4196-4229)
        if (s->datOn)
          launch_dat_color_pass(st, S, P, c, s->staticColorStarts[c], s->staticColorCounts[c], s->gridCap);
      }
    }
    launch_dual(st, S, P, s->gridCap);
    launch_ogc_refresh(st, S, P);
  }
  return hipSuccess;
}

void enqueue_post_solve(avbd_solver* s, const DeviceState& S, const StepParams& P, const StageFn& stage) {
  if (s->wsMode && (s->W.n || s->W.nR)) {
    const WarmState W = ws_params(s);
    launch_ws_target(s->stream, S, W);
    launch_ws_target_rigid(s->stream, S, W);
  }
  launch_dat_restore(s->stream, S, P);
  launch_finalize(s->stream, S, P);
  launch_viscosity(s->stream, S, P);
  stage("finalize");
}

// Enqueue one frame's kernels on the stream (eagerly or under capture).
hipError_t enqueue_frame(avbd_solver* s, const DeviceState& S, const StageFn& mark = nullptr) {
  auto stage = [&](const char* name) {
    if (mark) mark(name);
  };
  StepParams P = make_params(s);
  enqueue_pre_solve(s, S, P, stage);
  hipError_t e = enqueue_iterations(s, S, P, P.iterations, false);
  if (e != hipSuccess) return e;
  stage("solve");
  enqueue_post_solve(s, S, P, stage);
  return hipGetLastError();
}

void invalidate_graphs(avbd_solver* s) {
  for (int i = 0; i < 2; i++) {
    if (s->graphs[i]) hipGraphExecDestroy(s->graphs[i]);
    s->graphs[i] = nullptr;
    for (hipGraphExec_t* g : {&s->resPre[i], &s->resChunk[i], &s->resPost[i]}) {
      if (*g) hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  }
  s->graphsValid = false;
  s->resValid = false;
}

// captures what `body` enqueues on the solver stream into an executable graph
hipError_t capture_graph(avbd_solver* s, const std::function<hipError_t()>& body, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t eq = body();
  e = hipStreamEndCapture(s->stream, &g);
  if (eq != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return eq;
  }
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  return e;
}

hipError_t build_graphs(avbd_solver* s) {
  invalidate_graphs(s);
  DeviceState S = s->S;
  for (int p = 0; p < 2; p++) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t eq = enqueue_frame(s, S);
    e = hipStreamEndCapture(s->stream, &g);
    if (eq != hipSuccess) { if (g) hipGraphDestroy(g); return eq; }
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&s->graphs[p], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    std::swap(S.manifolds, S.prevManifolds);
    std::swap(S.soft, S.prevSoft);
  }
  s->graphsValid = true;
  s->graphBaseFrame = s->frameIndex + 1;
  return hipSuccess;
}

avbd_status retire_slot(avbd_solver* s, int slot) {
  if (!s->slotBusy[slot]) return AVBD_OK;
  s->slotBusy[slot] = false;
  int64_t frame = s->slotFrame[slot];
  hipError_t e = hipEventSynchronize(s->frameDone[slot]);
  if (e != hipSuccess) {
    char m[160];
    snprintf(m, sizeof(m), "frame %lld did not complete: %s", (long long)frame, hipGetErrorString(e));
    return latch(s, AVBD_FAIL_COMMAND_EXECUTION, frame, 0, 0, (int)e, m);
  }
  if (s->completionFailKind != 0 && (s->completionFailFrame < 0 || s->completionFailFrame == frame)) {
    const int kind = s->completionFailKind;
    s->completionFailKind = 0;  // one shot
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: completion failure injected for testing", (long long)frame);
    return latch(s, kind, frame, 0, 0, 0, m);
  }
  const uint32_t* c = s->pinnedCounters + slot * CTR_TOTAL;
  s->stats.frame_index = frame;
  s->stats.pair_candidates = (int32_t)c[CTR_PAIR_CANDIDATES];
  s->stats.num_pairs = (int32_t)c[CTR_PAIRS];
  s->stats.num_soft = (int32_t)c[CTR_SOFT];
  s->stats.soft_candidates = (int32_t)c[CTR_SOFT_CANDIDATES];
  s->stats.num_active_manifolds = (int32_t)c[CTR_ACTIVE_MANIFOLDS];
  s->stats.dat_pairs = (int32_t)c[CTR_DAT_PAIRS];
  s->stats.dat_vt_pairs = (int32_t)c[CTR_DAT_VT];
  s->stats.ws_accepted = (int32_t)c[CTR_WS_ACCEPT];
  s->stats.ws_residual_plain = as_float(c[CTR_WS_RES_PLAIN]);
  s->stats.ws_residual_learned = as_float(c[CTR_WS_RES_LEARNED]);
  s->stats.dat_ee_pairs = (int32_t)c[CTR_DAT_EE];
  s->stats.num_soft_rt = (int32_t)c[CTR_SOFT_RT];
  s->stats.dat_truncations = (int32_t)c[CTR_DAT_TRUNC];
  s->stats.dat_invalid_anchors = (int32_t)c[CTR_DAT_INVALID];
  s->stats.dat_vt_degeneracies = (int32_t)c[CTR_DAT_VT_DEGEN];
  s->stats.dat_ee_degeneracies = (int32_t)c[CTR_DAT_EE_DEGEN];
  s->stats.dat_nonfinite = (int32_t)c[CTR_DAT_NONFINITE];
  s->lastOgcTrunc = c[CTR_OGC_TRUNC];
  s->lastOgcRefresh = c[CTR_OGC_REFRESH];
  s->stats.color_conflicts = (int32_t)c[CTR_COLOR_CONFLICTS];
  if (s->dynamicColoring) {
    int maxUsed = -1;
    for (int k = 0; k < kMaxColors; k++) {
      s->stats.color_counts[k] = (int32_t)c[CTR_COLOR_BASE + k];
      if (c[CTR_COLOR_BASE + k] > 0) maxUsed = k;
    }
    s->stats.max_color_used = maxUsed;
    s->stats.num_colors = maxUsed + 1;
  }
  if (c[CTR_GBAR_FAIL] != 0) {
    return latch(s, AVBD_FAIL_BARRIER_TIMEOUT, frame, (int64_t)c[CTR_GBAR_FAIL], 0, 0,
                 "persistent sweep grid barrier exceeded its spin bound");
  }
  if ((int64_t)c[CTR_PAIR_CANDIDATES] > (int64_t)s->maxPairs) {
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: rigid-pair demand %u exceeds capacity %u", (long long)frame,
             c[CTR_PAIR_CANDIDATES], s->maxPairs);
    return latch(s, AVBD_FAIL_RIGID_PAIR_CAPACITY, frame, c[CTR_PAIR_CANDIDATES], s->maxPairs, 0, m);
  }
  if (s->nTris && (int64_t)c[CTR_SOFT_CANDIDATES] > (int64_t)s->maxSoft) {
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: soft-contact demand %u exceeds capacity %u", (long long)frame,
             c[CTR_SOFT_CANDIDATES], s->maxSoft);
    return latch(s, AVBD_FAIL_SOFT_CONTACT_CAPACITY, frame, c[CTR_SOFT_CANDIDATES], s->maxSoft, 0, m);
  }
  if (s->nTris && (int64_t)c[CTR_DAT_PAIRS] > (int64_t)s->maxDatPairs) {
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: Planar-DAT safety-pair demand %u exceeds capacity %u", (long long)frame,
             c[CTR_DAT_PAIRS], s->maxDatPairs);
    return latch(s, AVBD_FAIL_DAT_PAIR_CAPACITY, frame, c[CTR_DAT_PAIRS], s->maxDatPairs, 0, m);
  }
  if (s->nTris && c[CTR_DAT_OVERFLOW] > 0) {
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: %u Planar-DAT element AABBs exceeded the supported 4x4x4 grid span",
             (long long)frame, c[CTR_DAT_OVERFLOW]);
    return latch(s, AVBD_FAIL_DAT_GRID_SPAN, frame, c[CTR_DAT_OVERFLOW], 0, 0, m);
  }
  if (s->nTris && c[CTR_DAT_INVALID] > 0) {
    char m[200];
    snprintf(m, sizeof(m),
             "physics frame %lld: %u Planar-DAT anchors were invalid (V-T coincident %u, E-E degenerate %u, "
             "non-finite %u)",
             (long long)frame, c[CTR_DAT_INVALID], c[CTR_DAT_VT_DEGEN], c[CTR_DAT_EE_DEGEN], c[CTR_DAT_NONFINITE]);
    return latch(s, AVBD_FAIL_DAT_INVALID_ANCHOR, frame, c[CTR_DAT_INVALID], 0, 0, m);
  }
  if (c[CTR_COLOR_CONFLICTS] > 0) {
    char m[160];
    snprintf(m, sizeof(m), "physics frame %lld: dynamic coloring retained %u conflicting bodies", (long long)frame,
             c[CTR_COLOR_CONFLICTS]);
    return latch(s, AVBD_FAIL_UNRESOLVED_COLORING, frame, c[CTR_COLOR_CONFLICTS], 0, 0, m);
  }
  return AVBD_OK;
}

avbd_status sync_all(avbd_solver* s) {
  avbd_status first = AVBD_OK;
  // retire in frame order
  int order[2] = {0, 1};
  if (s->slotBusy[0] && s->slotBusy[1] && s->slotFrame[1] < s->slotFrame[0]) { order[0] = 1; order[1] = 0; }
  for (int i = 0; i < 2; i++) {
    avbd_status st = retire_slot(s, order[i]);
    if (st != AVBD_OK && first == AVBD_OK) first = st;
  }
  if (first != AVBD_OK) return first;
  return s->failed ? AVBD_ERR_RUNTIME_FAILURE : AVBD_OK;
}

// ---------------------------------------------------------------------------
// upload precompute (GPUSolver.swift:901-2081)
// ---------------------------------------------------------------------------
V3 rd3(const float* p, uint32_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
F4 rd4(const float* p, uint32_t i) { return F4{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }

bool body_is_dynamic(const avbd_scene_desc* d, uint32_t i) {
  if (d->body_mass && !isnan(d->body_mass[i])) return d->body_mass[i] > 0.0f;
  return d->body_density[i] > 0.0f;
}

// Pre-validates everything this backend cannot run.
avbd_status validate(const avbd_scene_desc* d) {
  if (!d || d->abi_version != AVBD_ABI_VERSION) return AVBD_ERR_INVALID_ARGUMENT;
  if (d->num_bodies == 0) return AVBD_ERR_INVALID_ARGUMENT;
  if (!d->body_size || !d->body_density || !d->body_friction || !d->body_position || !d->body_shape)
    return AVBD_ERR_INVALID_ARGUMENT;
  if (!(d->settings.dt > 0.0f) || d->settings.iterations < 0) return AVBD_ERR_INVALID_ARGUMENT;
  if (!(d->settings.collision_margin >= 0.0f) || !isfinite(d->settings.collision_margin))
    return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_bodies; i++)
    if (d->body_shape[i] > AVBD_SHAPE_CAPSULE) return AVBD_ERR_INVALID_ARGUMENT;
  if (d->num_colliders && (!d->collider_body || !d->collider_size || !d->collider_friction || !d->collider_shape))
    return AVBD_ERR_INVALID_ARGUMENT;
  uint64_t hullTotal = 0;
  for (uint32_t i = 0; i < d->num_colliders; i++) {
    if (d->collider_body[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
    uint32_t sh = d->collider_shape[i];
    bool enabled = !d->collider_enabled || d->collider_enabled[i];
    if (enabled && sh > AVBD_SHAPE_CAPSULE) return AVBD_ERR_INVALID_ARGUMENT;
    // convex hulls: at most 64 finite vertices (GPUSolver.swift:1024-1025)
    const uint32_t hn = d->collider_hull_count ? d->collider_hull_count[i] : 0u;
    if (hn > 64u || (hn && !d->hull_vertices)) return AVBD_ERR_INVALID_ARGUMENT;
    hullTotal += hn;
  }
  for (uint64_t k = 0; k < 3 * hullTotal; k++)
    if (!isfinite(d->hull_vertices[k])) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_joints; i++) {
    if (d->joint_body_b[i] < 0 || (uint32_t)d->joint_body_b[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
    if (d->joint_body_a[i] >= 0 && (uint32_t)d->joint_body_a[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t i = 0; i < d->num_springs; i++)
    if (d->spring_body_a[i] >= d->num_bodies || d->spring_body_b[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
  if (d->num_tets > 0 && (!d->tet_ids || !d->tet_mu || !d->tet_lambda)) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_tets * 4; i++)
    if (d->tet_ids[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
  if (d->num_tris > 0 && !d->tri_ids) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_tris * 3; i++)
    if (d->tri_ids[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_exclusions * 2; i++)
    if (d->exclusion_pairs[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
  if (d->num_spinners > 0 && (!d->spinner_body || !d->spinner_axis_omega)) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < d->num_spinners; i++) {
    if (d->spinner_body[i] >= d->num_bodies) return AVBD_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < 4; k++)
      if (!isfinite(d->spinner_axis_omega[4 * i + k])) return AVBD_ERR_INVALID_ARGUMENT;
  }
  return AVBD_OK;
}

// A spinner's per-frame rotation (x, y, z, w) by omega dt about its unit
// axis, in float as the reference's host computes it before each frame
// (GPUSolver.swift:4418-4427).
F4 spinner_rotation(F4 axisOmega, float dt) {
  const float angle = axisOmega.w * dt;
  const double half = 0.5 * (double)angle;
  const float sn = (float)std::sin(half), cs = (float)std::cos(half);
  return F4{axisOmega.x * sn, axisOmega.y * sn, axisOmega.z * sn, cs};
}

std::vector<F4> spinner_rotations(const avbd_solver* s, float dt) {
  std::vector<F4> dq(s->spinAxisOmega.size());
  for (size_t k = 0; k < dq.size(); k++) dq[k] = spinner_rotation(s->spinAxisOmega[k], dt);
  return dq;
}

// Boundary faces of the tet meshes (faces used by exactly one tet), wound
// outward against the tet's opposite vertex and sorted; they are collision
// triangles after the authored ones (GPUSolver.swift:384-390, 872-897).
std::vector<std::array<uint32_t, 3>> tet_boundary_faces(const avbd_scene_desc* d) {
  struct Face {
    std::array<uint32_t, 3> tri;
    uint32_t opp;
    int count;
  };
  std::map<std::array<uint32_t, 3>, Face> faces;
  static const int fi[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
  for (uint32_t t = 0; t < d->num_tets; t++) {
    const uint32_t* ids = d->tet_ids + 4 * t;
    for (int f = 0; f < 4; f++) {
      std::array<uint32_t, 3> tri{ids[fi[f][0]], ids[fi[f][1]], ids[fi[f][2]]};
      std::array<uint32_t, 3> k = tri;
      std::sort(k.begin(), k.end());
      auto it = faces.find(k);
      if (it != faces.end()) it->second.count++;
      else faces.emplace(k, Face{tri, ids[fi[f][3]], 1});
    }
  }
  std::vector<std::array<uint32_t, 3>> out;
  for (auto& kv : faces) {
    if (kv.second.count != 1) continue;
    std::array<uint32_t, 3> t = kv.second.tri;
    V3 pa = rd3(d->body_position, t[0]);
    V3 n = cross(rd3(d->body_position, t[1]) - pa, rd3(d->body_position, t[2]) - pa);
    if (dot(n, rd3(d->body_position, kv.second.opp) - pa) > 0.0f) std::swap(t[1], t[2]);
    out.push_back(t);
  }
  std::sort(out.begin(), out.end());
  return out;
}

// ---------------------------------------------------------------------------
// cloth surface topology and elements (GPUSolver.swift:391-443, 1289-1690)
// ---------------------------------------------------------------------------
struct MembraneHost {
  uint32_t ids[3];
  float dm[4];  // inverse rest shape, column major (a, b, c, d)
  float muA, lamA;
};
struct BendHost {
  uint32_t ids[4];  // edge ends, then the two opposite vertices
  float K[4];
  float s;
};
struct ClothHost {
  std::vector<U4> tris;
  std::vector<U2> edges;
  std::vector<uint32_t> nbrStart, nbrCount, nbrList, particles, group, selfCollision;
  std::vector<MembraneHost> membranes;
  std::vector<BendHost> bends;
  uint32_t elemHashSize = 64, maxDatPairs = 1, maxSoft = 1, softMapCapacity = 64;
  uint32_t numContactEdges = 0;  // edges of authored triangles (E-E contacts); tet-boundary edges follow
  bool pairStream = false;       // some V-T / E-E pair can be accepted: two surfaces, or a self-colliding one
  bool datOn = false;            // Planar-DAT limiter: a self-colliding (authored) surface exists
  bool isoOn = false;            // isotropic OGC limiter: tets, and two surfaces or an authored one
  // isotropic mode topology: triangle edge-adjacency, 2-rings, incident
  // edges, authored-vertex flags
  std::vector<U4> triAdj;
  std::vector<uint32_t> nbr2Start, nbr2Count, nbr2List, vEdgeStart, vEdgeCount, vEdgeList, clothVert;
  float elemMargin = 0, datQueryRadius = 0, elemCellSize = 0, surfaceContactCellSize = 0;
};

float body_pos_dist(const avbd_scene_desc* d, uint32_t a, uint32_t b) {
  V3 d3 = rd3(d->body_position, a) - rd3(d->body_position, b);
  return sqrtf(d3.x * d3.x + d3.y * d3.y + d3.z * d3.z);
}

// Static data of the isotropic mode (GPUSolver.swift:1336-1460): triangle
// adjacency over shared edges (filled in ascending edge-key order), sorted
// 2-rings (the 1-ring and its 1-rings, the body itself excluded), incident
// edges of every body in edge order, and the vertices of authored triangles.
void build_iso_topology(ClothHost& c, uint32_t nb, uint32_t na) {
  const uint32_t nr = (uint32_t)c.tris.size();
  std::map<uint64_t, std::pair<int64_t, int64_t>> edgeTris;
  for (uint32_t t = 0; t < nr; t++) {
    const U4 id = c.tris[t];
    const uint32_t ev[3][2] = {{id.x, id.y}, {id.y, id.z}, {id.x, id.z}};
    for (const auto& e : ev) {
      const uint64_t k = (uint64_t)std::min(e[0], e[1]) << 32 | std::max(e[0], e[1]);
      auto it = edgeTris.find(k);
      if (it == edgeTris.end()) edgeTris.emplace(k, std::make_pair((int64_t)t, (int64_t)-1));
      else it->second.second = t;
    }
  }
  c.triAdj.assign(std::max<uint32_t>(1, nr), U4{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu});
  std::vector<int> fill(std::max<uint32_t>(1, nr), 0);
  auto put = [&](uint32_t t, uint32_t o) {
    if (fill[t] == 0) c.triAdj[t].x = o;
    else if (fill[t] == 1) c.triAdj[t].y = o;
    else if (fill[t] == 2) c.triAdj[t].z = o;
    else return;
    fill[t]++;
  };
  for (const auto& kv : edgeTris) {
    if (kv.second.second < 0) continue;
    put((uint32_t)kv.second.first, (uint32_t)kv.second.second);
    put((uint32_t)kv.second.second, (uint32_t)kv.second.first);
  }
  c.nbr2Start.assign(nb, 0);
  c.nbr2Count.assign(nb, 0);
  c.nbr2List.clear();
  std::vector<uint32_t> two;
  for (uint32_t i = 0; i < nb; i++) {
    c.nbr2Start[i] = (uint32_t)c.nbr2List.size();
    if (c.nbrCount[i] == 0) continue;
    two.clear();
    for (uint32_t k = 0; k < c.nbrCount[i]; k++) {
      const uint32_t n1 = c.nbrList[c.nbrStart[i] + k];
      two.push_back(n1);
      for (uint32_t q = 0; q < c.nbrCount[n1]; q++) two.push_back(c.nbrList[c.nbrStart[n1] + q]);
    }
    std::sort(two.begin(), two.end());
    two.erase(std::unique(two.begin(), two.end()), two.end());
    two.erase(std::remove(two.begin(), two.end(), i), two.end());
    c.nbr2Count[i] = (uint32_t)two.size();
    c.nbr2List.insert(c.nbr2List.end(), two.begin(), two.end());
  }
  std::vector<std::vector<uint32_t>> ve(nb);
  for (uint32_t e = 0; e < (uint32_t)c.edges.size(); e++) {
    ve[c.edges[e].x].push_back(e);
    ve[c.edges[e].y].push_back(e);
  }
  c.vEdgeStart.assign(nb, 0);
  c.vEdgeCount.assign(nb, 0);
  c.vEdgeList.clear();
  for (uint32_t i = 0; i < nb; i++) {
    c.vEdgeStart[i] = (uint32_t)c.vEdgeList.size();
    c.vEdgeCount[i] = (uint32_t)ve[i].size();
    c.vEdgeList.insert(c.vEdgeList.end(), ve[i].begin(), ve[i].end());
  }
  c.clothVert.assign(nb, 0);
  for (uint32_t t = 0; t < na; t++) {
    c.clothVert[c.tris[t].x] = 1;
    c.clothVert[c.tris[t].y] = 1;
    c.clothVert[c.tris[t].z] = 1;
  }
}

// Collision triangles are the authored ones, then the tet boundary faces
// (GPUSolver.swift:1289-1334): membranes and hinges come from the authored
// triangles only, E-E contacts from their edges only (the tet-boundary edges
// follow them in the edge list and feed the Planar-DAT pair stream alone).
void build_cloth(const avbd_scene_desc* d, ClothHost& c, const std::vector<std::array<uint32_t, 3>>& boundary) {
  const uint32_t nb = d->num_bodies, na = d->num_tris;
  auto isParticle = [&](uint32_t i) { return d->body_is_particle && d->body_is_particle[i]; };
  auto sizeX = [&](uint32_t i) { return d->body_size[3 * i]; };
  std::vector<std::array<uint32_t, 3>> coll(na);
  for (uint32_t t = 0; t < na; t++) coll[t] = {d->tri_ids[3 * t], d->tri_ids[3 * t + 1], d->tri_ids[3 * t + 2]};
  for (const auto& f : boundary) coll.push_back(f);
  const uint32_t nr = (uint32_t)coll.size();
  c.tris.resize(nr);
  std::set<uint64_t> seen;
  std::vector<uint64_t> keys;  // unique edges in first-appearance order
  float maxElemR = 0.0f, maxPartR = 0.0f;
  for (uint32_t t = 0; t < nr; t++) {
    if (t == na) c.numContactEdges = (uint32_t)keys.size();
    uint32_t a = coll[t][0], b = coll[t][1], cc = coll[t][2];
    c.tris[t] = U4{a, b, cc, 0};
    uint32_t e[3][2] = {{a, b}, {b, cc}, {a, cc}};
    for (auto& uv : e) {
      uint64_t k = (uint64_t)std::min(uv[0], uv[1]) << 32 | std::max(uv[0], uv[1]);
      if (seen.insert(k).second) keys.push_back(k);
    }
    V3 pa = rd3(d->body_position, a), pb = rd3(d->body_position, b), pc = rd3(d->body_position, cc);
    V3 m = (pa + pb + pc) / 3.0f;
    float thick = fmaxf(sizeX(a), fmaxf(sizeX(b), sizeX(cc))) / 2.0f;
    maxPartR = fmaxf(maxPartR, thick);
    maxElemR = fmaxf(maxElemR, fmaxf(length(m - pa), fmaxf(length(m - pb), length(m - pc))) + thick);
  }
  if (nr == na) c.numContactEdges = (uint32_t)keys.size();
  c.edges.resize(keys.size());
  std::vector<std::vector<uint32_t>> nbr(nb);
  float edgeLenSum = 0.0f;
  for (size_t i = 0; i < keys.size(); i++) {
    uint32_t a = (uint32_t)(keys[i] >> 32), b = (uint32_t)(keys[i] & 0xFFFFFFFFu);
    c.edges[i] = U2{a, b};
    nbr[a].push_back(b);
    nbr[b].push_back(a);
    maxElemR = fmaxf(maxElemR, body_pos_dist(d, a, b) / 2.0f + fmaxf(sizeX(a), sizeX(b)) / 2.0f);
    maxPartR = fmaxf(maxPartR, fmaxf(sizeX(a), sizeX(b)) / 2.0f);
    edgeLenSum += body_pos_dist(d, a, b);
  }
  c.nbrStart.assign(nb, 0);
  c.nbrCount.assign(nb, 0);
  for (uint32_t i = 0; i < nb; i++) {
    std::sort(nbr[i].begin(), nbr[i].end());
    c.nbrStart[i] = (uint32_t)c.nbrList.size();
    c.nbrCount[i] = (uint32_t)nbr[i].size();
    c.nbrList.insert(c.nbrList.end(), nbr[i].begin(), nbr[i].end());
  }
  // connected surfaces (union-find over triangles), numbered by first
  // triangle; a surface holding an authored triangle self-collides, a tet
  // boundary does not (GPUSolver.swift:1476-1530)
  std::vector<uint32_t> parent(nb);
  for (uint32_t i = 0; i < nb; i++) parent[i] = i;
  std::function<uint32_t(uint32_t)> find = [&](uint32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (const U4& t : c.tris) {
    uint32_t r = find(t.x);
    parent[find(t.y)] = r;
    parent[find(t.z)] = r;
  }
  c.group.assign(nb, 0);
  c.selfCollision.assign(nb, 0);
  std::set<uint32_t> selfRoots;
  for (uint32_t t = 0; t < na; t++) selfRoots.insert(find(c.tris[t].x));
  std::map<uint32_t, uint32_t> gidx;
  std::vector<uint8_t> surface(nb, 0);
  for (const U4& t : c.tris) {
    uint32_t root = find(t.x);
    auto it = gidx.find(root);
    if (it == gidx.end()) it = gidx.emplace(root, (uint32_t)gidx.size() + 1).first;
    for (uint32_t v : {t.x, t.y, t.z})
      if (isParticle(v)) {
        c.group[v] = it->second;
        c.selfCollision[v] = selfRoots.count(root) ? 1u : 0u;
        surface[v] = 1;
      }
  }
  // The surface limiter follows the reference's automatic selection
  // (GPUSolver.swift:93-99): Planar-DAT, with its V-T / E-E pair stream and
  // the stream's contacts, for authored (self-colliding) surfaces without
  // tets; the isotropic bound wherever tets exist -- the scene format
  // carries no volumetric self-collision flag -- with its own tracked
  // contacts.  With one closed surface and no authored triangle every
  // candidate pair of the isotropic trackers is a same-solid pair, which they
  // skip: the mode is then left off, an identity on the results.
  c.datOn = !selfRoots.empty();
  c.pairStream = c.datOn;
  {
    uint32_t np0 = 0;
    for (uint32_t i = 0; i < nb; i++) np0 += surface[i];
    c.isoOn = d->num_tets > 0 && np0 > 0 && (gidx.size() >= 2 || na > 0);
  }
  if (c.isoOn) {
    c.datOn = false;
    c.pairStream = false;
  }
  for (uint32_t i = 0; i < nb; i++)
    if (surface[i]) c.particles.push_back(i);
  // membranes: rest 2x2 in the triangle's own frame, inverse column major
  std::map<uint64_t, std::vector<std::pair<uint32_t, float>>> edgeTris;  // edge -> (opposite, bend)
  for (uint32_t t = 0; t < na; t++) {
    float mu = d->tri_mu ? d->tri_mu[t] : 0.0f;
    if (!(mu > 0.0f)) continue;
    float lam = d->tri_lambda ? d->tri_lambda[t] : mu;
    float bend = d->tri_bend ? d->tri_bend[t] : 0.0f;
    uint32_t i0 = d->tri_ids[3 * t], i1 = d->tri_ids[3 * t + 1], i2 = d->tri_ids[3 * t + 2];
    double x0[3], d1[3], d2[3];
    for (int k = 0; k < 3; k++) {
      x0[k] = d->body_position[3 * i0 + k];
      d1[k] = (double)d->body_position[3 * i1 + k] - x0[k];
      d2[k] = (double)d->body_position[3 * i2 + k] - x0[k];
    }
    double n[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
    double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double area = nl / 2.0;
    if (!(area > 1e-10)) continue;
    double l1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    double e1[3] = {d1[0] / l1, d1[1] / l1, d1[2] / l1};
    double nn[3] = {n[0] / nl, n[1] / nl, n[2] / nl};
    double e2[3] = {nn[1] * e1[2] - nn[2] * e1[1], nn[2] * e1[0] - nn[0] * e1[2], nn[0] * e1[1] - nn[1] * e1[0]};
    double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    for (int k = 0; k < 3; k++) e2[k] /= l2;
    auto dotd = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double a = dotd(d1, e1), b = dotd(d1, e2), cc = dotd(d2, e1), dd = dotd(d2, e2);
    double det = a * dd - b * cc;
    if (!(fabs(det) > 1e-12)) continue;
    MembraneHost mh;
    mh.ids[0] = i0; mh.ids[1] = i1; mh.ids[2] = i2;
    mh.dm[0] = (float)(dd / det);
    mh.dm[1] = (float)(-b / det);
    mh.dm[2] = (float)(-cc / det);
    mh.dm[3] = (float)(a / det);
    mh.muA = (float)(mu * area);
    mh.lamA = (float)(lam * area);
    c.membranes.push_back(mh);
    if (bend > 0.0f) {
      uint32_t ev[3][3] = {{i0, i1, i2}, {i1, i2, i0}, {i0, i2, i1}};
      for (auto& e : ev) {
        uint64_t k = (uint64_t)std::min(e[0], e[1]) << 32 | std::max(e[0], e[1]);
        edgeTris[k].push_back({e[2], bend});
      }
    }
  }
  // hinges over edges shared by exactly two bending triangles, in edge-key
  // order: K spans the null space of {1, x, y} of the unfolded rest flaps,
  // scaled to the cotangent convention; s = 3 (k1 + k2) / 2 / (A1 + A2)
  for (auto& kv : edgeTris) {
    if (kv.second.size() != 2) continue;
    uint32_t v0 = (uint32_t)(kv.first >> 32), v1 = (uint32_t)(kv.first & 0xFFFFFFFFu);
    uint32_t opp2 = kv.second[0].first, opp3 = kv.second[1].first;
    float k1 = kv.second[0].second, k2 = kv.second[1].second;
    double p0[3], p1[3];
    for (int k = 0; k < 3; k++) { p0[k] = d->body_position[3 * v0 + k]; p1[k] = d->body_position[3 * v1 + k]; }
    auto dist = [&](uint32_t o, const double* p) {
      double s2 = 0;
      for (int k = 0; k < 3; k++) { double q = (double)d->body_position[3 * o + k] - p[k]; s2 += q * q; }
      return sqrt(s2);
    };
    double L = sqrt((p0[0] - p1[0]) * (p0[0] - p1[0]) + (p0[1] - p1[1]) * (p0[1] - p1[1]) +
                    (p0[2] - p1[2]) * (p0[2] - p1[2]));
    if (!(L > 1e-9)) continue;
    auto unfold = [&](uint32_t o, double side, double& x, double& y) {
      double l0 = dist(o, p0), l1 = dist(o, p1);
      x = (L * L + l0 * l0 - l1 * l1) / (2 * L);
      double h2 = l0 * l0 - x * x;
      if (!(h2 > 1e-12)) return false;
      y = side * sqrt(h2);
      return true;
    };
    double q[4][2] = {{0, 0}, {L, 0}, {0, 0}, {0, 0}};
    if (!unfold(opp2, 1.0, q[2][0], q[2][1]) || !unfold(opp3, -1.0, q[3][0], q[3][1])) continue;
    auto minor = [&](int c0, int c1, int c2) {
      int cs[3] = {c0, c1, c2};
      double m[3][3];
      for (int j = 0; j < 3; j++) { m[0][j] = 1.0; m[1][j] = q[cs[j]][0]; m[2][j] = q[cs[j]][1]; }
      return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
             m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    };
    double K[4] = {minor(1, 2, 3), -minor(0, 2, 3), minor(0, 1, 3), -minor(0, 1, 2)};
    auto cot = [&](int apex, int ia, int ib) {
      double ux = q[ia][0] - q[apex][0], uy = q[ia][1] - q[apex][1];
      double vx = q[ib][0] - q[apex][0], vy = q[ib][1] - q[apex][1];
      double cr = ux * vy - uy * vx;
      return fabs(cr) > 1e-12 ? (ux * vx + uy * vy) / fabs(cr) : 0.0;
    };
    double want = cot(0, 1, 2) + cot(1, 0, 2);
    if (!(fabs(K[2]) > 1e-12)) continue;
    double scale = want / K[2];
    double a1 = fabs((q[1][0] - q[0][0]) * q[2][1]) / 2, a2 = fabs((q[1][0] - q[0][0]) * q[3][1]) / 2;
    BendHost bh;
    bh.ids[0] = v0; bh.ids[1] = v1; bh.ids[2] = opp2; bh.ids[3] = opp3;
    for (int k = 0; k < 4; k++) bh.K[k] = (float)(K[k] * scale);
    bh.s = (float)(3.0 * ((double)k1 + (double)k2) / 2.0 / std::max(a1 + a2, 1e-9));
    c.bends.push_back(bh);
  }
  // grids, query radius and capacities (GPUSolver.swift:391-443, 1643-1682)
  float meanEdge = keys.empty() ? 0.2f : edgeLenSum / (float)keys.size();
  c.elemMargin = fminf(0.01f, fmaxf(0.002f, 0.5f * maxPartR));
  float forceReach = 2.0f * maxPartR + c.elemMargin;
  c.datQueryRadius = fmaxf(1.5f * forceReach, 0.75f * meanEdge);
  float reach = (2.0f * maxPartR + c.elemMargin) / 0.7f;
  c.surfaceContactCellSize = fmaxf(0.02f, fmaxf(fmaxf(reach, meanEdge * 1.05f), maxElemR + maxPartR + c.elemMargin));
  c.elemCellSize = fmaxf(0.02f, fmaxf(fmaxf(reach, meanEdge * 1.05f), 2.0f * maxElemR + c.datQueryRadius));
  c.elemHashSize = next_pow2(std::max<uint32_t>(64u, 8u * nr));
  uint64_t np = c.particles.size(), ne = keys.size();
  uint64_t pairCap = std::max<uint64_t>(1, 64 * np + 128 * ne);
  uint64_t budget = std::min<uint64_t>(pairCap, 8 * np + 8 * ne);
  c.maxDatPairs = (uint32_t)std::min<uint64_t>(pairCap, 0xFFFFFFF0ull);
  c.maxSoft = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, 4ull * nr + budget), kMaxSoftRecords - 1);
  if (c.isoOn) {
    // tracked-set capacity (GPUSolver.swift:439-442)
    c.maxSoft = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, 4 * np + 4ull * nr + 4ull * c.numContactEdges),
                                             kMaxSoftRecords - 1);
    build_iso_topology(c, nb, na);
  }
  c.softMapCapacity = next_pow2(std::max<uint32_t>(64u, 2u * c.maxSoft));
}

}  // namespace

extern "C" {

void avbd_settings_default(avbd_settings* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->dt = 1.0f / 60.0f;
  o->gravity = -10.0f;
  o->iterations = 10;
  o->collision_margin = 0.01f;
  o->friction_combine_mode = 0;
  o->alpha = 0.99f;
  o->beta_lin = 10000.0f;
  o->beta_ang = 100.0f;
  o->gamma = 0.999f;
  o->lambda_max = 1.0e6f;
}

void avbd_options_default(avbd_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->max_pairs_per_body = 16;
  o->lanes_per_body = 0;
  o->sweep_mode = 0;
  o->use_graph = 1;
  o->pair_capacity_override = 0;
}

const char* avbd_version(void) { return "avbd-mi355x 0.1 (gfx950, wave64)"; }

}  // extern "C"

namespace {

avbd_status build_static_and_upload(avbd_solver* s, const avbd_scene_desc* d);

// closes the running create phase under `name` (after the stream drains, so
// device work is charged to the phase that enqueued it)
void create_phase(avbd_solver* s, const char* name) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  const auto t = std::chrono::steady_clock::now();
  s->createPhases.emplace_back(name, std::chrono::duration<double, std::milli>(t - s->phaseT0).count());
  s->phaseT0 = t;
}
avbd_status static_coloring(avbd_solver* s);
void autotune(avbd_solver* s);
void release_solver(avbd_solver* s);

}  // namespace

// ---------------------------------------------------------------------------
// Staging memory of avbd_set_body_states_noalloc: one fixed array per GPU,
// resolved to its device address once per device (at the first create on
// that device) and shared by every solver on it until the process ends.
// ---------------------------------------------------------------------------
constexpr size_t kFixedScratchWords = size_t(1) << 20;  // 4 MiB
static __device__ __align__(256) uint32_t scratch[kFixedScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
uint32_t* g_scratchPtr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device` (the current device).  The name
// itself is a host-side stand-in, never passed to a kernel.
static hipError_t resolve_fixed_scratch(int device, uint32_t** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[device]) {
    uint32_t* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    g_scratchPtr[device] = ptr;
  }
  *out = g_scratchPtr[device];
  return hipSuccess;
}

extern "C" avbd_status avbd_create(const avbd_scene_desc* d, const avbd_options* opts, avbd_solver** out) {
  if (!out) return AVBD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  avbd_status v = validate(d);
  if (v != AVBD_OK) return v;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AVBD_ERR_NO_DEVICE;
  avbd_solver* s = new avbd_solver();
  s->phaseT0 = std::chrono::steady_clock::now();
  hipGetDevice(&s->device);
  if (resolve_fixed_scratch(s->device, &s->fixedScratch) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return AVBD_ERR_NO_DEVICE;
  }
  if (opts) s->opts = *opts;
  else avbd_options_default(&s->opts);
  if (s->opts.max_pairs_per_body < 0) s->opts.max_pairs_per_body = 16;  // 0 is a valid request (4096-pair floor)
  s->settings = d->settings;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return AVBD_ERR_NO_DEVICE;
  }
  create_phase(s, "stream");
  avbd_status st = build_static_and_upload(s, d);
  if (st != AVBD_OK) {
    release_solver(s);
    return st;
  }
  *out = s;
  live_add(s, HandleKind::Solver);
  st = static_coloring(s);
  create_phase(s, "static coloring");
  if (st == AVBD_ERR_RUNTIME_FAILURE) return st;  // failure latched; caller reads it then destroys
  if (st != AVBD_OK) {
    avbd_destroy(s);
    *out = nullptr;
    return st;
  }
  autotune(s);
  create_phase(s, "autotune");
  s->stats.lanes_per_body = s->lanes;
  s->stats.sweep_mode = s->sweepMode;
  s->stats.uses_dynamic_coloring = s->dynamicColoring ? 1 : 0;
  return AVBD_OK;
}

extern "C" void avbd_destroy(avbd_solver* s) {
  // a handle already released at process exit is not touched again
  if (s && live_remove(s)) release_solver(s);
}

namespace {

void release_solver(avbd_solver* s) {
  if (!s) return;
  if (s->stream) hipStreamSynchronize(s->stream);
  invalidate_graphs(s);
  for (void* p : s->allocs) hipFree(p);
  for (int i = 0; i < 2; i++)
    if (s->frameDone[i]) hipEventDestroy(s->frameDone[i]);
  if (s->pinnedCounters) hipHostFree(s->pinnedCounters);
  if (s->residualPartialHost) hipHostFree(s->residualPartialHost);
  if (s->convHost) hipHostFree(s->convHost);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

avbd_status build_static_and_upload(avbd_solver* s, const avbd_scene_desc* d) {
  const uint32_t nb = d->num_bodies;
  s->nb = nb;
  s->nc = d->num_colliders;
  s->nj = d->num_joints;
  s->ns = d->num_springs;
  s->nt = d->num_tets;
  s->dynamicColoring = d->num_tets == 0 && d->num_tris == 0;

  // ---- bodies (GPUSolver.swift:914-969)
  std::vector<F4> pl(nb), pa(nb), vl(nb), va(nb), pv(nb), pr(nb), sh(nb);
  std::vector<float> gs(nb);
  s->mass.resize(nb);
  s->inertia.resize(3 * nb);
  s->isParticle.assign(nb, 0);
  const float kPi = 3.14159265358979f;
  for (uint32_t i = 0; i < nb; i++) {
    V3 size = rd3(d->body_size, i);
    uint32_t shape = d->body_shape[i];
    if (shape == AVBD_SHAPE_SPHERE) size = v3s(size.x);
    float density = d->body_density[i];
    float m = 0.0f, rad = 0.0f;
    V3 mom = v3s(0);
    if (shape == AVBD_SHAPE_BOX) {
      m = density > 0.0f ? size.x * size.y * size.z * density : 0.0f;
      mom = v3((size.y * size.y + size.z * size.z) / 12.0f * m, (size.x * size.x + size.z * size.z) / 12.0f * m,
               (size.x * size.x + size.y * size.y) / 12.0f * m);
      V3 h = size * 0.5f;
      rad = sqrtf(h.x * h.x + h.y * h.y + h.z * h.z);
    } else if (shape == AVBD_SHAPE_SPHERE) {
      float r = size.x / 2.0f;
      m = density > 0.0f ? 4.0f / 3.0f * kPi * r * r * r * density : 0.0f;
      mom = v3s(0.4f * m * r * r);
      rad = r;
    } else if (shape == AVBD_SHAPE_TORUS) {
      float R = size.x, r = size.y;
      m = density > 0.0f ? 2.0f * kPi * kPi * R * r * r * density : 0.0f;
      float iDia = m * (R * R / 2.0f + 5.0f * r * r / 8.0f), iAxis = m * (R * R + 3.0f * r * r / 4.0f);
      mom = v3(iDia, iDia, iAxis);
      rad = R + r;
    } else {
      float L = size.x, r = size.y;
      m = density > 0.0f ? kPi * r * r * (L + 4.0f * r / 3.0f) * density : 0.0f;
      float iAxis = 0.5f * m * r * r, iPerp = m * (L * L / 12.0f + r * r / 4.0f);
      mom = v3(iPerp, iPerp, iAxis);
      rad = L / 2.0f + r;
    }
    if (d->body_mass && !isnan(d->body_mass[i])) {
      m = d->body_mass[i];
      mom = d->body_inertia ? rd3(d->body_inertia, i) : v3s(0);
    }
    bool particle = d->body_is_particle && d->body_is_particle[i];
    s->isParticle[i] = particle ? 1 : 0;
    V3 pos = rd3(d->body_position, i);
    F4 rot = d->body_rotation ? rd4(d->body_rotation, i) : F4{0, 0, 0, 1};
    V3 vel = d->body_velocity ? rd3(d->body_velocity, i) : v3s(0);
    pl[i] = f4(pos, m);
    pa[i] = rot;
    vl[i] = f4(vel, 0);
    va[i] = F4{0, 0, 0, 0};
    pv[i] = f4(vel, 0);
    float fr = d->body_friction[i];
    pr[i] = f4(mom, fr);
    sh[i] = f4(size, particle ? -rad : rad);
    gs[i] = d->body_gravity_scale ? d->body_gravity_scale[i] : 1.0f;
    s->mass[i] = m;
    s->inertia[3 * i] = mom.x;
    s->inertia[3 * i + 1] = mom.y;
    s->inertia[3 * i + 2] = mom.z;
  }

  create_phase(s, "bodies");
  // ---- kinematic spinners (GPUSolver.swift:628-637), grouped by body
  {
    std::vector<int32_t> slot(d->num_spinners ? nb : 0, -1);
    std::vector<std::vector<F4>> per;
    for (uint32_t k = 0; k < d->num_spinners; k++) {
      const uint32_t b = d->spinner_body[k];
      if (slot[b] < 0) {
        slot[b] = (int32_t)s->spinBodies.size();
        s->spinBodies.push_back(b);
        per.emplace_back();
      }
      per[slot[b]].push_back(rd4(d->spinner_axis_omega, k));
    }
    s->spinStart.assign(1, 0u);
    for (const auto& v : per) {
      s->spinAxisOmega.insert(s->spinAxisOmega.end(), v.begin(), v.end());
      s->spinStart.push_back((uint32_t)s->spinAxisOmega.size());
    }
  }
  // ---- colliders (GPUSolver.swift:971-1090)
  const uint32_t nc = d->num_colliders;
  std::vector<uint32_t> co(nc), ct(nc), cg(nc), csc(nc);
  std::vector<F4> cs(nc), cp(nc), cq(nc), cf(nc);
  // convex hull colliders: (first vertex, count) per collider, vertices collider-local
  std::vector<U2> hullRange(nc, U2{0u, 0u});
  std::vector<F4> hullVerts;
  std::vector<float> radii;
  std::vector<float> crad(nc);
  uint32_t enabledCount = 0;
  bool hasIsolated = false;
  for (uint32_t i = 0; i < nc; i++) {
    uint32_t b = d->collider_body[i];
    uint32_t shape = d->collider_shape[i];
    V3 size = rd3(d->collider_size, i);
    if (shape == AVBD_SHAPE_SPHERE) size = v3s(size.x);
    float r;
    if (shape == AVBD_SHAPE_SPHERE) r = size.x / 2.0f;
    else if (shape == AVBD_SHAPE_TORUS) r = size.x + size.y;
    else if (shape == AVBD_SHAPE_CAPSULE) r = size.x / 2.0f + size.y;
    else { V3 h = size * 0.5f; r = sqrtf(h.x * h.x + h.y * h.y + h.z * h.z); }
    const uint32_t hullN = d->collider_hull_count ? d->collider_hull_count[i] : 0u;
    if (hullN) {  // the hull's radius is its farthest vertex (GPUSolver.swift:994-997)
      hullRange[i] = U2{(uint32_t)hullVerts.size(), hullN};
      r = 0.0f;
      for (uint32_t k = 0; k < hullN; k++) {
        V3 v = rd3(d->hull_vertices, hullRange[i].x + k);
        hullVerts.push_back(f4(v, 0.0f));
        r = fmaxf(r, length(v));
      }
    }
    crad[i] = r;
    bool particle = d->body_is_particle && d->body_is_particle[b];
    bool enabled = !d->collider_enabled || d->collider_enabled[i];
    float fr = d->collider_friction[i];
    float dfr = d->collider_dynamic_friction ? d->collider_dynamic_friction[i] : fr;
    co[i] = b;
    cg[i] = d->collider_group ? d->collider_group[i] : 0u;
    csc[i] = (!d->collider_shared || d->collider_shared[i]) ? 1u : 0u;
    cs[i] = f4(size, particle ? -r : r);
    cp[i] = f4(d->collider_local_position ? rd3(d->collider_local_position, i) : v3s(0), fr);
    F4 lq = d->collider_local_rotation ? rd4(d->collider_local_rotation, i) : F4{0, 0, 0, 1};
    float qn = sqrtf(lq.x * lq.x + lq.y * lq.y + lq.z * lq.z + lq.w * lq.w);
    cq[i] = F4{lq.x / qn, lq.y / qn, lq.z / qn, lq.w / qn};
    cf[i] = F4{fr, dfr, 0, 0};
    uint32_t flags = hullN ? 4u : shape;
    if (particle) flags |= kShapeParticle;
    if (d->collider_world_round_anchor && d->collider_world_round_anchor[i]) flags |= kShapeWorldRound;
    ct[i] = flags;
    if (enabled) {
      enabledCount++;
      if (cg[i] != 0) hasIsolated = true;
      if (body_is_dynamic(d, b)) radii.push_back(r);
    }
  }
  std::sort(radii.begin(), radii.end());
  float median = radii.empty() ? 0.5f : radii[radii.size() / 2];
  float threshold = median * 4.0f;
  std::vector<uint32_t> hashed, globals;
  for (uint32_t i = 0; i < nc; i++) {
    bool enabled = !d->collider_enabled || d->collider_enabled[i];
    if (!enabled) continue;
    if (crad[i] > threshold || (hasIsolated && cg[i] == 0)) globals.push_back(i);
    else hashed.push_back(i);
  }
  float maxHashedR = 0.05f, maxR = 0.0f;
  for (uint32_t i : hashed) {
    maxHashedR = fmaxf(maxHashedR, fabsf(cs[i].w));
    maxR = fmaxf(maxR, fabsf(cs[i].w));
  }
  s->cellSize = maxHashedR * 2.0f;
  // hash cells at least twice the largest hashed diameter (the 0.05 m radius
  // floor under dense particles): candidates come from a search grid of
  // cells just over that diameter instead of the bucket walk
  const float fine = 2.0f * maxR * 1.001f;
  s->fineCellSize = (maxR > 0.0f && 2.0f * fine <= s->cellSize) ? fine : 0.0f;
  s->nHashed = (uint32_t)hashed.size();
  s->nGlobal = (uint32_t)globals.size();
  std::vector<uint32_t> hashedRigid;
  for (uint32_t i : hashed)
    if (!(d->body_is_particle && d->body_is_particle[co[i]])) hashedRigid.push_back(i);
  s->nHashedRigid = (uint32_t)hashedRigid.size();
  s->maxPairs = std::max<uint32_t>(4096u, enabledCount * (uint32_t)s->opts.max_pairs_per_body);
  if (s->opts.pair_capacity_override > 0) s->maxPairs = (uint32_t)s->opts.pair_capacity_override;
  s->mapCapacity = next_pow2(2 * s->maxPairs);
  s->gridHashSize = next_pow2(std::max<uint32_t>(64u, 2u * std::max<uint32_t>(1u, enabledCount)));

  create_phase(s, "colliders");
  // ---- joints (GPUSolver.swift:1092-1193)
  const uint32_t nj = d->num_joints;
  std::vector<JointRec> jr(nj);
  std::vector<F4> penInit(2 * (size_t)std::max<uint32_t>(1, nj), F4{0, 0, 0, 0});
  s->jointHeaders.resize(nj);
  for (uint32_t i = 0; i < nj; i++) {
    JointRec g;
    memset(&g, 0, sizeof(g));
    int32_t a = d->joint_body_a[i], b = d->joint_body_b[i];
    float kl = d->joint_stiffness_lin[i], ka = d->joint_stiffness_ang[i];
    float fracture = d->joint_fracture ? d->joint_fracture[i] : INFINITY;
    uint32_t flags = 0;
    if (isinf(kl)) flags |= 1;
    if (isinf(ka)) flags |= 2;
    if (isfinite(fracture)) flags |= 4;
    if (d->joint_fracture_linear && d->joint_fracture_linear[i]) flags |= 8;
    g.header = U4{a >= 0 ? (uint32_t)a : kWorld, (uint32_t)b, 0, flags};
    const float bigK = 3.0e10f;
    g.rA = f4(rd3(d->joint_ra, i), fminf(kl, bigK));
    g.rB = f4(rd3(d->joint_rb, i), fminf(ka, bigK));
    V3 sizeA = a >= 0 ? rd3(d->body_size, (uint32_t)a) : v3s(0);
    V3 sz = sizeA + rd3(d->body_size, (uint32_t)b);
    g.C0Lin = F4{0, 0, 0, dot(sz, sz)};
    g.C0Ang = F4{0, 0, 0, fminf(fracture, 3.0e18f)};
    F4 qA0 = (a >= 0 && d->body_rotation) ? rd4(d->body_rotation, (uint32_t)a) : F4{0, 0, 0, 1};
    F4 qB0 = d->body_rotation ? rd4(d->body_rotation, (uint32_t)b) : F4{0, 0, 0, 1};
    F4 rel = q_mul(q_inv(qA0), qB0);
    float rn = sqrtf(rel.x * rel.x + rel.y * rel.y + rel.z * rel.z + rel.w * rel.w);
    g.restRel = F4{rel.x / rn, rel.y / rn, rel.z / rn, rel.w / rn};
    bool hinge = d->joint_hinge_axis && !isnan(d->joint_hinge_axis[3 * i]);
    const float* mo = d->joint_motor ? d->joint_motor + 6 * i : nullptr;
    float motorTorque = mo ? mo[1] : 0.0f;
    float lo = d->joint_limits ? d->joint_limits[2 * i] : 1.0f;
    float hi = d->joint_limits ? d->joint_limits[2 * i + 1] : -1.0f;
    const float rate = d->joint_motor_rate ? d->joint_motor_rate[i] : 0.0f;
    if (!std::isfinite(rate)) return AVBD_ERR_INVALID_ARGUMENT;
    if (hinge) {
      V3 ax = normalize(rd3(d->joint_hinge_axis, i));
      g.hingeAxis = f4(ax, 1.0f);
      g.dynamics.x = mo ? mo[4] : 0.0f;
      if (motorTorque > 0.0f) {
        g.motor = F4{mo[0], mo[1], 0.0f, mo[2]};
        g.limits = F4{lo, hi, mo[3], 0.0f};
        int mode = (int)mo[5];
        g.header.w |= (mode == 1 ? kMotorExplicitPD : (mode == 2 ? kMotorVelocity : kMotorImplicitPD));
        if (rate != 0.0f) {
          if (mode != 0) return AVBD_ERR_INVALID_ARGUMENT;  // legacy rates need implicit position PD
          s->rateJoints.push_back(i);
          s->rateValues.push_back(rate);
        }
      }
      if (lo < hi && motorTorque == 0.0f) g.limits = F4{lo, hi, 0.0f, 0.0f};
    } else if (motorTorque > 0.0f) {
      return AVBD_ERR_INVALID_ARGUMENT;  // a powered joint must define a hinge axis
    }
    float mA = a >= 0 ? pl[a].w : 0.0f, mB = pl[b].w;
    float eff = 0.0f;
    bool any = false;
    if (mA > 0.0f) { eff = mA; any = true; }
    if (mB > 0.0f) { eff = any ? fminf(eff, mB) : mB; any = true; }
    float floorK = fminf(1.0e9f, fmaxf(1.0f, eff / fmaxf(s->settings.dt * s->settings.dt, 1.0e-12f)));
    if (isinf(kl)) g.penaltyLin = F4{floorK, floorK, floorK, floorK};
    if (isinf(ka)) g.penaltyAng = F4{floorK, floorK, floorK, floorK};
    if (kl > 0.0f && isfinite(kl)) { float k = fminf(kl, 1e9f); g.penaltyLin = F4{k, k, k, k}; }
    if (ka > 0.0f && isfinite(ka)) { float k = fminf(ka, 1e9f); g.penaltyAng = F4{k, k, k, k}; }
    jr[i] = g;
    s->jointHeaders[i] = g.header;
    // initialJointPenaltyLin / Ang: restored by setBodyStates
    penInit[i] = g.penaltyLin;
    penInit[nj + i] = g.penaltyAng;
  }

  create_phase(s, "joints");
  // ---- springs (GPUSolver.swift:1195-1212)
  const uint32_t ns = d->num_springs;
  std::vector<SpringRec> sr(ns);
  for (uint32_t i = 0; i < ns; i++) {
    SpringRec g;
    memset(&g, 0, sizeof(g));
    uint32_t a = d->spring_body_a[i], b = d->spring_body_b[i];
    bool hard = d->spring_hard && d->spring_hard[i];
    g.header = U4{a, b, hard ? 1u : 0u, 0};
    float rest = d->spring_rest[i];
    V3 rA = rd3(d->spring_ra, i), rB = rd3(d->spring_rb, i);
    if (rest < 0.0f) {
      V3 pA = xyz(pl[a]) + q_rotate(pa[a], rA), pB = xyz(pl[b]) + q_rotate(pa[b], rB);
      rest = length(pA - pB);
    }
    g.rA = f4(rA, d->spring_stiffness[i]);
    g.rB = f4(rB, rest);
    sr[i] = g;
    s->springEndsHost.push_back(U2{a, b});
  }

  create_phase(s, "springs");
  // ---- tets (GPUSolver.swift:1214-1236): the rest shapes (Dm^-1 rows, 6V,
  // mu V, lambda V) are computed on the device once the adjacency is there
  const uint32_t nt = d->num_tets;
  s->tetIdsHost.resize(nt);
  for (uint32_t t = 0; t < nt; t++) {
    const uint32_t* id = d->tet_ids + 4 * t;
    s->tetIdsHost[t] = U4{id[0], id[1], id[2], id[3]};
  }

  create_phase(s, "tets");
  // ---- exclusions (GPUSolver.swift:1265-1288): sorted unique (lo, hi)
  std::vector<uint64_t> ex;
  for (uint32_t i = 0; i < nj; i++) {
    int32_t a = d->joint_body_a[i], b = d->joint_body_b[i];
    if (a < 0) continue;
    uint64_t lo = (uint64_t)std::min(a, b), hi = (uint64_t)std::max(a, b);
    ex.push_back(lo << 32 | hi);
  }
  for (uint32_t i = 0; i < ns; i++) {
    uint64_t lo = std::min(d->spring_body_a[i], d->spring_body_b[i]);
    uint64_t hi = std::max(d->spring_body_a[i], d->spring_body_b[i]);
    ex.push_back(lo << 32 | hi);
  }
  for (uint32_t i = 0; i < d->num_exclusions; i++) {
    uint64_t a = d->exclusion_pairs[2 * i], b = d->exclusion_pairs[2 * i + 1];
    ex.push_back(std::min(a, b) << 32 | std::max(a, b));
  }
  std::sort(ex.begin(), ex.end());
  ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
  std::vector<U2> exr(ex.size());
  for (size_t i = 0; i < ex.size(); i++) exr[i] = U2{(uint32_t)(ex[i] >> 32), (uint32_t)(ex[i] & 0xFFFFFFFFu)};
  s->nExcl = (uint32_t)ex.size();
  create_phase(s, "exclusions");

  // ---- collision surfaces: authored triangles and tet boundary faces; the
  // particles of one connected surface share a soft collision group
  // (GPUSolver.swift:1476-1522)
  std::vector<uint32_t> clothGroup(nb, 0);
  ClothHost ch;
  if (d->num_tris > 0 || nt > 0) {
    std::vector<std::array<uint32_t, 3>> boundary;
    if (nt && !gpu_tet_boundary_faces(d->tet_ids, nt, d->body_position, nb, s->stream, boundary)) {
      (void)hipGetLastError();
      boundary = tet_boundary_faces(d);
    }
    create_phase(s, "tet boundary faces");
    build_cloth(d, ch, boundary);
    clothGroup = ch.group;
    s->nTris = (uint32_t)ch.tris.size();
    s->numContactEdges = ch.numContactEdges;
    s->datOn = ch.datOn;
    s->isoOn = ch.isoOn;
    s->pairStream = ch.pairStream;
    s->nEdges = (uint32_t)ch.edges.size();
    s->nParticles = (uint32_t)ch.particles.size();
    s->nMembranes = (uint32_t)ch.membranes.size();
    s->nBends = (uint32_t)ch.bends.size();
    s->elemHashSize = ch.elemHashSize;
    s->maxDatPairs = ch.maxDatPairs;
    if (s->opts.dat_pair_capacity_override > 0)
      s->maxDatPairs = std::min<uint32_t>(s->maxDatPairs, (uint32_t)s->opts.dat_pair_capacity_override);
    s->maxSoft = ch.maxSoft;
    if (s->opts.soft_capacity_override > 0)
      s->maxSoft = std::min<uint32_t>(s->maxSoft, (uint32_t)s->opts.soft_capacity_override);
    s->softMapCapacity = ch.softMapCapacity;
    s->elemCellSize = ch.elemCellSize;
    s->elemMargin = ch.elemMargin;
    s->datQueryRadius = ch.datQueryRadius;
    s->datQueryRadiusDerived = ch.datQueryRadius;
    s->surfaceContactCellSize = ch.surfaceContactCellSize;
  }

  create_phase(s, "collision surfaces");
  // ---- static topology CSR (per body: joints, springs, tets, membranes,
  // bends, each ascending), counted then filled in that order
  auto dyn = [&](uint32_t b) { return b != kWorld && pl[b].w > 0.0f; };
  auto zero3 = [](F4 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; };
  std::vector<uint32_t> sStart(nb), sCount(nb, 0), sList;
  for (int pass = 0; pass < 2; pass++) {
    std::vector<uint32_t> cur;
    if (pass == 1) {
      uint32_t run = 0;
      for (uint32_t b = 0; b < nb; b++) { sStart[b] = run; run += sCount[b]; }
      sList.assign(run, 0u);
      cur = sStart;
    }
    auto add = [&](uint32_t b, uint32_t e) {
      if (!dyn(b)) return;
      if (pass == 0) sCount[b]++;
      else sList[cur[b]++] = e;
    };
    for (uint32_t i = 0; i < nj; i++) {
      const JointRec& g = jr[i];
      if (g.rA.w == 0.0f && g.rB.w == 0.0f && g.motor.y == 0.0f) continue;  // inert slot
      add(g.header.x, (FK_JOINT << kAdjShift) | i);
      add(g.header.y, (FK_JOINT << kAdjShift) | i);
    }
    // particle-particle springs without offsets take a stamp-major slot
    for (uint32_t i = 0; i < ns; i++) {
      bool slot = s->isParticle[sr[i].header.x] && s->isParticle[sr[i].header.y] && zero3(sr[i].rA) && zero3(sr[i].rB);
      uint32_t e = ((slot ? FK_PSPRING : FK_SPRING) << kAdjShift) | i;
      add(sr[i].header.x, e);
      add(sr[i].header.y, e);
    }
    for (uint32_t t = 0; t < nt; t++) {
      const uint32_t* id = d->tet_ids + 4 * t;
      for (int k = 0; k < 4; k++) add(id[k], (FK_TET << kAdjShift) | t);
    }
    for (uint32_t m = 0; m < ch.membranes.size(); m++)
      for (int k = 0; k < 3; k++) add(ch.membranes[m].ids[k], (FK_MEMBRANE << kAdjShift) | m);
    for (uint32_t m = 0; m < ch.bends.size(); m++)
      for (int k = 0; k < 4; k++) add(ch.bends[m].ids[k], (FK_BEND << kAdjShift) | m);
  }
  s->sAdjCount = sCount;
  s->sAdjListHost = sList;
  // cloth slots at each membrane / bend / particle-spring entry's CSR
  // position (same-color partner bits are added once the palette is known)
  bool anySlot = false;
  for (uint32_t e : sList) {
    uint32_t kind = e >> kAdjShift;
    if (kind == FK_MEMBRANE || kind == FK_BEND || kind == FK_PSPRING) anySlot = true;
  }
  s->hasClothSlots = anySlot;
  std::vector<ClothSlot> cslots(std::max<size_t>(1, anySlot ? sList.size() : 1));
  memset(cslots.data(), 0, cslots.size() * sizeof(ClothSlot));
  for (uint32_t b = 0; b < nb && anySlot; b++) {
    for (uint32_t k = 0; k < sCount[b]; k++) {
      uint32_t e = sList[sStart[b] + k], kind = e >> kAdjShift, i = e & kAdjIndexMask;
      ClothSlot cs{};
      if (kind == FK_MEMBRANE) {
        const MembraneHost& m = ch.membranes[i];
        // shape-function gradients w_1 = (a, c), w_2 = (b, d), w_0 = -(w_1 + w_2)
        float w[3][2] = {{-(m.dm[0] + m.dm[1]), -(m.dm[2] + m.dm[3])}, {m.dm[0], m.dm[2]}, {m.dm[1], m.dm[3]}};
        uint32_t o[2];
        float ow[2][2];
        int n = 0;
        for (int q = 0; q < 3; q++)
          if (m.ids[q] != b && n < 2) { o[n] = m.ids[q]; ow[n][0] = w[q][0]; ow[n][1] = w[q][1]; n++; }
        cs.u = U4{o[0], o[1], as_uint(m.muA), as_uint(m.lamA)};
        cs.f = F4{ow[0][0], ow[0][1], ow[1][0], ow[1][1]};
      } else if (kind == FK_BEND) {
        const BendHost& h = ch.bends[i];
        uint32_t o[3];
        float ok[3], kself = 0.0f;
        int n = 0;
        for (int q = 0; q < 4; q++) {
          if (h.ids[q] == b) kself = h.K[q];
          else if (n < 3) { o[n] = h.ids[q]; ok[n] = h.K[q]; n++; }
        }
        cs.u = U4{o[0], o[1], o[2], as_uint(h.s)};
        cs.f = F4{ok[0], ok[1], ok[2], kself};
      } else if (kind == FK_PSPRING) {
        const SpringRec& r = sr[i];
        cs.u = U4{r.header.x == b ? r.header.y : r.header.x, i, r.header.z, 0u};
        cs.f = F4{r.rA.w, r.rB.w, 0.0f, 0.0f};
      } else {
        continue;
      }
      cslots[sStart[b] + k] = cs;
    }
  }
  s->clothSlotsHost = cslots;
  s->sAdjStartHost = sStart;

  create_phase(s, "static CSR and slots");
  // ---- device buffers
  DeviceState& S = s->S;
  hipError_t e = hipSuccess;
#define ALLOC(ptr, n) \
  if ((e = dev_alloc(s, &ptr, (n))) != hipSuccess) return AVBD_ERR_ALLOC;
  ALLOC(S.posLin, nb) ALLOC(S.posLo, nb) ALLOC(S.posAng, nb) ALLOC(S.initLin, nb) ALLOC(S.initAng, nb) ALLOC(S.inertLin, nb)
  ALLOC(S.inertAng, nb) ALLOC(S.velLin, nb) ALLOC(S.velAng, nb) ALLOC(S.prevVelLin, nb) ALLOC(S.props, nb)
  ALLOC(S.shape, nb) ALLOC(S.gravityScale, nb) ALLOC(S.clothGroup, nb) ALLOC(S.posLinSnap, nb)
  ALLOC(S.posAngSnap, nb)
  ALLOC(S.colliderOwner, nc) ALLOC(S.colliderShapeType, nc) ALLOC(S.colliderGroup, nc) ALLOC(S.colliderShared, nc)
  ALLOC(S.colliderShape, nc) ALLOC(S.colliderLocalPos, nc) ALLOC(S.colliderLocalRot, nc)
  ALLOC(S.colliderHullRange, std::max<uint32_t>(1, nc)) ALLOC(S.hullVerts, std::max<size_t>(1, hullVerts.size()))
  ALLOC(S.colliderFriction, nc)
  ALLOC(S.hashedIdx, std::max<size_t>(1, hashed.size())) ALLOC(S.globalIdx, std::max<size_t>(1, globals.size()))
  ALLOC(S.cellCount, s->gridHashSize) ALLOC(S.cellStart, s->gridHashSize) ALLOC(S.cellRigid, s->gridHashSize)
  ALLOC(S.cellBodies, std::max<size_t>(1, hashed.size())) ALLOC(S.bodyCellSlot, std::max<size_t>(1, hashed.size()))
  ALLOC(S.hashCenter, std::max<size_t>(1, hashed.size())) ALLOC(S.hashMeta, std::max<size_t>(1, hashed.size()))
  ALLOC(S.cellPos, std::max<size_t>(1, hashed.size())) ALLOC(S.cellMeta, std::max<size_t>(1, hashed.size()))
  ALLOC(S.colliderToHashed, std::max<uint32_t>(1, nc))
  {
    const bool fineOn = s->fineCellSize > 0.0f;
    ALLOC(S.fineCount, fineOn ? s->gridHashSize : 1u) ALLOC(S.fineStart, fineOn ? s->gridHashSize : 1u)
    ALLOC(S.fineBodies, fineOn ? std::max<size_t>(1, hashed.size()) : 1u)
    ALLOC(S.fineSlot, fineOn ? std::max<size_t>(1, hashed.size()) : 1u)
    ALLOC(S.largeCells, hashed.size() / 33 + 1)
  }
  uint32_t producers = s->nHashed + s->nGlobal;
  // pair producers: one lane each where that walk is short or hides its
  // latency -- every hashed producer without the search grid (small
  // buckets) or when there are many (kLaneMany), else the soft-group
  // particles of scenes with no hashed non-particle collider (globals
  // only); one wave each for the rest and the globals
  constexpr size_t kLaneMany = 32768;
  std::vector<uint32_t> bpLane, bpWave;
  for (size_t k = 0; k < hashed.size(); k++) {
    const bool lane = s->fineCellSize <= 0.0f || hashed.size() >= kLaneMany ||
                      (clothGroup[co[hashed[k]]] != 0u && s->nHashedRigid == 0);
    (lane ? bpLane : bpWave).push_back((uint32_t)k);
  }
  for (uint32_t g = 0; g < s->nGlobal; g++) bpWave.push_back(s->nHashed + g);
  s->numBpLane = (uint32_t)bpLane.size();
  s->numBpWave = (uint32_t)bpWave.size();
  ALLOC(S.bpLane, std::max<size_t>(1, bpLane.size())) ALLOC(S.bpWave, std::max<uint32_t>(1, producers))
  ALLOC(S.bpWaveCount, 1)
  ALLOC(S.pairCount, std::max<uint32_t>(1, producers)) ALLOC(S.pairStart, std::max<uint32_t>(1, producers))
  ALLOC(S.pairs, s->maxPairs) ALLOC(S.exclusions, std::max<size_t>(1, exr.size()))
  ALLOC(S.manifolds, s->maxPairs) ALLOC(S.prevManifolds, s->maxPairs)
  ALLOC(S.mapKeyA, s->mapCapacity) ALLOC(S.mapKeyB, s->mapCapacity) ALLOC(S.mapVal, s->mapCapacity)
  ALLOC(S.tets, std::max<uint32_t>(1, nt)) ALLOC(S.joints, std::max<uint32_t>(1, nj))
  ALLOC(s->jointPenInit, 2 * (size_t)std::max<uint32_t>(1, nj))
  ALLOC(S.springs, std::max<uint32_t>(1, ns))
  ALLOC(S.sAdjStart, nb) ALLOC(S.sAdjCount, nb) ALLOC(S.sAdjList, std::max<size_t>(1, sList.size()))
  ALLOC(S.tetSlots, std::max<size_t>(1, sList.size()))
  ALLOC(S.dAdjCount, nb) ALLOC(S.dAdjStart, nb) ALLOC(S.dAdjCursor, nb)
  S.dAdjCapacity = 2 * s->maxPairs + 4 * s->maxSoft;
  ALLOC(S.dAdjList, S.dAdjCapacity)
  ALLOC(S.colors, nb) ALLOC(S.colorsB, nb) ALLOC(S.colorList, nb) ALLOC(S.colorStart, kMaxColors + 2)
  ALLOC(S.colorWork, nb)
  uint32_t bb = (nb + 255) / 256;
  ALLOC(S.bodySlot, std::max<uint32_t>(nb, 2 * kMaxColors * bb))
  ALLOC(S.colorFlags, 32) ALLOC(S.counters, CTR_TOTAL)
  ALLOC(S.hashedRigid, std::max<size_t>(1, hashedRigid.size()))
  ALLOC(S.clothSlots, cslots.size())
  ALLOC(s->rateJointDev, std::max<size_t>(1, s->rateJoints.size()))
  ALLOC(s->rateValueDev, std::max<size_t>(1, s->rateValues.size()))
  ALLOC(s->spinBodyDev, std::max<size_t>(1, s->spinBodies.size()))
  ALLOC(s->spinStartDev, s->spinStart.size())
  ALLOC(s->spinDqDev, std::max<size_t>(1, s->spinAxisOmega.size()))
  if (!s->spinBodies.empty()) {
    ALLOC(S.spinVel, nb)
  }
  bool anyTorus = false;
  for (uint32_t i = 0; i < nc; i++) anyTorus = anyTorus || d->collider_shape[i] == AVBD_SHAPE_TORUS;
  if (anyTorus) {
    ALLOC(S.torusTrig, 64)
  }
  {
    const bool cl = s->nTris > 0;
    const uint32_t nElem = cl ? s->nTris + s->nEdges : 0u, nProd = cl ? s->nParticles + s->nEdges : 0u;
    ALLOC(S.tris, std::max<uint32_t>(1, s->nTris)) ALLOC(S.edges, std::max<uint32_t>(1, s->nEdges))
    ALLOC(S.nbrStart, cl ? nb : 1u) ALLOC(S.nbrCount, cl ? nb : 1u)
    ALLOC(S.nbrList, std::max<size_t>(1, ch.nbrList.size())) ALLOC(S.particleIdx, std::max<uint32_t>(1, s->nParticles))
    ALLOC(S.selfCollision, cl ? nb : 1u)
    ALLOC(S.elemCellCount, cl ? 2 * s->elemHashSize + 1 : 1u) ALLOC(S.elemCellStart, cl ? 2 * s->elemHashSize + 1 : 1u)
    ALLOC(S.elemCellCursor, cl ? 2 * s->elemHashSize : 1u) ALLOC(S.elemCells, std::max<uint64_t>(1, 64ull * nElem))
    ALLOC(S.elemBox, std::max<uint32_t>(2, 2 * nElem))
    ALLOC(S.datAnchor, cl ? nb : 1u) ALLOC(S.datT, cl ? nb : 1u)
    ALLOC(S.datCount, nProd + 1) ALLOC(S.datStart, nProd + 1)
    ALLOC(S.datPairs, cl ? s->maxDatPairs : 1u)
    uint32_t flagN = cl ? std::max(std::max(s->nTris, s->maxDatPairs), s->nParticles + s->numContactEdges) + 1 : 1u;
    ALLOC(S.softFlag, flagN) ALLOC(S.softStart, flagN)
    ALLOC(S.soft, cl ? s->maxSoft : 1u) ALLOC(S.prevSoft, cl ? s->maxSoft : 1u)
    ALLOC(S.softMapKeyA, cl ? s->softMapCapacity : 1u) ALLOC(S.softMapKeyB, cl ? s->softMapCapacity : 1u)
    ALLOC(S.softMapVal, cl ? s->softMapCapacity : 1u)
    ALLOC(S.velScratch, cl ? nb : 1u)
    const bool iso = s->isoOn;
    const uint32_t nce = std::max<uint32_t>(1, s->numContactEdges);
    ALLOC(S.triAdj, iso ? std::max<uint32_t>(1, s->nTris) : 1u)
    ALLOC(S.nbr2Start, iso ? nb : 1u) ALLOC(S.nbr2Count, iso ? nb : 1u)
    ALLOC(S.nbr2List, std::max<size_t>(1, ch.nbr2List.size()))
    ALLOC(S.vEdgeStart, iso ? nb : 1u) ALLOC(S.vEdgeCount, iso ? nb : 1u)
    ALLOC(S.vEdgeList, std::max<size_t>(1, ch.vEdgeList.size())) ALLOC(S.clothVert, iso ? nb : 1u)
    ALLOC(S.vtTrack, iso ? nb : 1u) ALLOC(S.vtPrev, iso ? nb : 1u)
    ALLOC(S.isoSlotLo, iso ? 2u * s->elemHashSize : 1u) ALLOC(S.isoSlotHi, iso ? 2u * s->elemHashSize : 1u)
    ALLOC(S.eeTrack, iso ? nce : 1u) ALLOC(S.eePrev, iso ? nce : 1u)
    ALLOC(S.ogcBound, iso ? nb : 1u) ALLOC(S.ogcPrev, iso ? nb : 1u) ALLOC(S.isoFrame, 1u)
  }
  uint32_t scanMax = std::max<uint32_t>(std::max<uint32_t>(s->gridHashSize, producers), std::max(nb, kMaxColors * bb));
  scanMax = std::max<uint32_t>(scanMax, std::max(2 * s->elemHashSize + 1, s->nParticles + s->nEdges + 1));
  scanMax = std::max<uint32_t>(scanMax, std::max(s->nTris, s->maxDatPairs) + 1);
  ALLOC(S.scanScratch, scan_scratch_words(scanMax) + 16)
  s->residualBlocks = bb;
  ALLOC(S.residualPartial, 2 * bb)
  s->devScratchWords = std::max<uint32_t>(2 * s->maxPairs, 64);
  ALLOC(s->devScratchU32, s->devScratchWords)
#undef ALLOC
  create_phase(s, "device allocation");

  // static metadata of the hashed colliders and the collider -> hashed index map
  std::vector<U4> hmeta(std::max<size_t>(1, hashed.size()), U4{0, 0, 0, 0});
  std::vector<uint32_t> c2h(std::max<uint32_t>(1, nc), 0xFFFFFFFFu);
  for (size_t k = 0; k < hashed.size(); k++) {
    uint32_t c = hashed[k], b = co[c];
    hmeta[k] = U4{b, cg[c], csc[c] | ((pl[b].w > 0.0f) ? 2u : 0u), clothGroup[b]};
    c2h[c] = (uint32_t)k;
  }
  if (upload(S.hashMeta, hmeta) || upload(S.colliderToHashed, c2h) || upload(S.bpLane, bpLane) ||
      upload(S.bpWave, bpWave))
    return AVBD_ERR_HIP;
  if (!s->spinBodies.empty()) {
    // surface angular velocity per body: the last spinner of a body sets it
    std::vector<F4> sv(nb, F4{0.0f, 0.0f, 0.0f, 0.0f});
    for (uint32_t k = 0; k < d->num_spinners; k++) {
      const F4 ao = rd4(d->spinner_axis_omega, k);
      sv[d->spinner_body[k]] = F4{ao.x * ao.w, ao.y * ao.w, ao.z * ao.w, 1.0f};
    }
    s->spinDt = s->settings.dt;
    if (upload(S.spinVel, sv) || upload(s->spinBodyDev, s->spinBodies) || upload(s->spinStartDev, s->spinStart) ||
        upload(s->spinDqDev, spinner_rotations(s, s->spinDt)))
      return AVBD_ERR_HIP;
  }
  if (anyTorus) {
    // cos / sin of the spine sample angles 2 pi i / 32, rounded to float
    std::vector<float> trig(64);
    for (int i = 0; i < 32; i++) {
      const double a = 6.283185307179586476925 * (double)i / 32.0;
      trig[2 * i] = (float)std::cos(a);
      trig[2 * i + 1] = (float)std::sin(a);
    }
    if (upload(S.torusTrig, trig)) return AVBD_ERR_HIP;
  }
  if (upload(S.posLin, pl) || upload(S.posAng, pa) || upload(S.velLin, vl) || upload(S.velAng, va) ||
      upload(S.prevVelLin, pv) || upload(S.props, pr) || upload(S.shape, sh) || upload(S.gravityScale, gs) ||
      upload(S.clothGroup, clothGroup) || upload(S.colliderOwner, co) || upload(S.colliderShapeType, ct) ||
      upload(S.colliderGroup, cg) || upload(S.colliderShared, csc) || upload(S.colliderShape, cs) ||
      upload(S.colliderHullRange, hullRange) || upload(S.hullVerts, hullVerts) ||
      upload(S.colliderLocalPos, cp) || upload(S.colliderLocalRot, cq) || upload(S.colliderFriction, cf) ||
      upload(S.hashedIdx, hashed) || upload(S.globalIdx, globals) || upload(S.exclusions, exr) ||
      upload(S.joints, jr) || upload(s->jointPenInit, penInit) || upload(S.springs, sr) ||
      upload(S.sAdjStart, sStart) ||
      upload(S.sAdjCount, sCount) || upload(S.sAdjList, sList) ||
      upload(S.hashedRigid, hashedRigid) || upload(S.clothSlots, cslots) ||
      upload(s->rateJointDev, s->rateJoints) || upload(s->rateValueDev, s->rateValues) || upload(S.tris, ch.tris) ||
      upload(S.edges, ch.edges) || upload(S.nbrStart, ch.nbrStart) || upload(S.nbrCount, ch.nbrCount) ||
      upload(S.nbrList, ch.nbrList) || upload(S.particleIdx, ch.particles) ||
      upload(S.selfCollision, ch.selfCollision) ||
      upload(S.posLinSnap, pl) ||
      upload(S.posAngSnap, pa))
    return AVBD_ERR_HIP;
  if (nt && gpu_tet_precompute(d->tet_ids, d->tet_mu, d->tet_lambda, nt, d->body_position, nb, S.sAdjStart,
                               S.sAdjCount, S.sAdjList, S.tets, S.tetSlots, s->stream) != hipSuccess)
    return AVBD_ERR_HIP;

  if (s->isoOn) {
    // isotropic mode: topology, empty tracked sets, cleared bounds
    const std::vector<U4> none4(std::max<uint32_t>(s->nb, s->numContactEdges),
                                U4{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu});
    const std::vector<U4> noneB(none4.begin(), none4.begin() + s->nb);
    const std::vector<U4> noneE(none4.begin(), none4.begin() + s->numContactEdges);
    const std::vector<uint32_t> far(s->nb, 0x7F7F7F7Fu);
    if (upload(S.triAdj, ch.triAdj) || upload(S.nbr2Start, ch.nbr2Start) || upload(S.nbr2Count, ch.nbr2Count) ||
        upload(S.nbr2List, ch.nbr2List) || upload(S.vEdgeStart, ch.vEdgeStart) ||
        upload(S.vEdgeCount, ch.vEdgeCount) || upload(S.vEdgeList, ch.vEdgeList) ||
        upload(S.clothVert, ch.clothVert) || upload(S.vtTrack, noneB) || upload(S.vtPrev, noneB) ||
        upload(S.eeTrack, noneE) || upload(S.eePrev, noneE) || upload(S.ogcBound, far))
      return AVBD_ERR_HIP;
  }

  create_phase(s, "uploads");
  if (hipHostMalloc((void**)&s->pinnedCounters, 2 * CTR_TOTAL * sizeof(uint32_t), hipHostMallocDefault) !=
      hipSuccess)
    return AVBD_ERR_ALLOC;
  memset(s->pinnedCounters, 0, 2 * CTR_TOTAL * sizeof(uint32_t));
  if (hipHostMalloc((void**)&s->residualPartialHost, 2 * bb * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return AVBD_ERR_ALLOC;
  // written by the device's residual checks, read by the host while kernels run
  if (hipHostMalloc((void**)&s->convHost, sizeof(ConvHost), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void**)&s->convDev, s->convHost, 0) != hipSuccess)
    return AVBD_ERR_ALLOC;
  memset(s->convHost, 0, sizeof(ConvHost));
  for (int i = 0; i < 2; i++)
    if (hipEventCreateWithFlags(&s->frameDone[i], hipEventDisableTiming) != hipSuccess) return AVBD_ERR_HIP;
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  s->gridCap = (uint32_t)std::max(256, 8 * cus);
  s->cus = (uint32_t)std::max(1, cus);
  create_phase(s, "runtime resources");
  return AVBD_OK;
}

// Static palette: Jacobi fixed point of the greedy rule on the GPU, then
// stable index-ordered buckets (GPUSolver.swift:1956-2050).
avbd_status static_coloring(avbd_solver* s) {
  DeviceState& S = s->S;
  StepParams P = make_params(s);
  hipStream_t st = s->stream;
  launch_fill_u32(st, S.colors, s->nb, 0u);
  uint32_t* in = S.colors;
  uint32_t* out = S.colorsB;
  uint32_t changed = 1;
  uint32_t rounds = 0;
  // The fixed point is reached within (longest increasing-index path + 1)
  // rounds; nb + 1 bounds it for any graph.
  while (changed && rounds <= s->nb + 1) {
    launch_fill_u32(st, S.counters + CTR_COLOR_CHANGED, 1, 0u);
    launch_static_color_pass(st, S, P, in, out);
    std::swap(in, out);
    rounds++;
    if ((rounds & 7u) == 0 || rounds < 8) {
      if (hipMemcpyAsync(&changed, S.counters + CTR_COLOR_CHANGED, 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return AVBD_ERR_HIP;
      if (hipStreamSynchronize(st) != hipSuccess) return AVBD_ERR_HIP;
    }
  }
  // the loop can also end on its round bound between two checks: the last
  // passes must finish before the colors are read (the stream does not
  // synchronize with the null stream)
  if (hipStreamSynchronize(st) != hipSuccess) return AVBD_ERR_HIP;
  std::vector<uint32_t> colors(s->nb);
  if (hipMemcpy(colors.data(), in, 4 * s->nb, hipMemcpyDeviceToHost) != hipSuccess) return AVBD_ERR_HIP;
  if (in != S.colors && hipMemcpy(S.colors, in, 4 * s->nb, hipMemcpyDeviceToDevice) != hipSuccess)
    return AVBD_ERR_HIP;
  uint32_t maxColor = 0;
  std::vector<std::vector<uint32_t>> buckets(kMaxColors);
  for (uint32_t v = 0; v < s->nb; v++) {
    if (!(s->mass[v] > 0.0f)) continue;
    if (colors[v] >= (uint32_t)kMaxColors) {
      char m[160];
      snprintf(m, sizeof(m), "greedy static solver coloring at body %u exhausted %d colors (next color %d)", v,
               kMaxColors, kMaxColors + 1);
      return latch(s, AVBD_FAIL_STATIC_COLOR_CAPACITY, -1, kMaxColors + 1, kMaxColors, 0, m);
    }
    maxColor = std::max(maxColor, colors[v]);
    buckets[colors[v]].push_back(v);
  }
  s->staticUsedColors = maxColor + 1;
  std::vector<uint32_t> clist, cstart;
  for (int c = 0; c < kMaxColors; c++) {
    cstart.push_back((uint32_t)clist.size());
    clist.insert(clist.end(), buckets[c].begin(), buckets[c].end());
  }
  cstart.push_back((uint32_t)clist.size());
  cstart.push_back(s->staticUsedColors);
  if (upload(S.colorStart, cstart) || upload(S.colorList, clist)) return AVBD_ERR_HIP;
  // per-color work records: one load gives the body and its static CSR range
  {
    std::vector<uint32_t> st(s->nb);
    uint32_t run = 0;
    for (uint32_t b = 0; b < s->nb; b++) { st[b] = run; run += s->sAdjCount[b]; }
    std::vector<U4> work(clist.size());
    for (size_t i = 0; i < clist.size(); i++) work[i] = U4{clist[i], st[clist[i]], s->sAdjCount[clist[i]], 0u};
    if (upload(S.colorWork, work)) return AVBD_ERR_HIP;
  }
  s->staticColorStarts.assign(cstart.begin(), cstart.begin() + kMaxColors);
  s->staticColorCounts.assign(kMaxColors, 0);
  s->colorMode.assign(kMaxColors, kModeGeneral);
  std::vector<uint32_t> sStart(s->nb);
  {
    uint32_t run = 0;
    for (uint32_t b = 0; b < s->nb; b++) { sStart[b] = run; run += s->sAdjCount[b]; }
  }
  bool first = true;
  s->allMode = kModeGeneral;
  bool uniform = !s->dynamicColoring;
  for (int c = 0; c < kMaxColors; c++) {
    s->staticColorCounts[c] = (uint32_t)buckets[c].size();
    s->stats.color_counts[c] = (int32_t)buckets[c].size();
    bool particles = true, tetOnly = true;
    for (uint32_t v : buckets[c]) {
      if (!s->isParticle[v]) { particles = false; break; }
      for (uint32_t k = 0; k < s->sAdjCount[v] && tetOnly; k++)
        if ((s->sAdjListHost[sStart[v] + k] >> kAdjShift) != FK_TET) tetOnly = false;
    }
    int mode = !particles ? kModeGeneral : (tetOnly ? kModeTet : kModeParticle);
    s->colorMode[c] = (uint8_t)mode;
    if (c < (int)s->staticUsedColors) {
      if (first) { s->allMode = mode; first = false; }
      else if (mode != s->allMode) uniform = false;
    }
  }
  if (!uniform) s->allMode = kModeGeneral;
  // cloth slots: flag partners of the same static color (snapshot reads)
  if (s->hasClothSlots) {
    std::vector<ClothSlot> cs = s->clothSlotsHost;
    auto mark = [&](uint32_t& id, uint32_t self) {
      if (s->mass[id] > 0.0f && colors[id] == colors[self]) id |= kSnapBit;
    };
    for (uint32_t b = 0; b < s->nb; b++) {
      for (uint32_t k = 0; k < s->sAdjCount[b]; k++) {
        uint32_t kind = s->sAdjListHost[sStart[b] + k] >> kAdjShift;
        ClothSlot& q = cs[sStart[b] + k];
        if (kind == FK_MEMBRANE) { mark(q.u.x, b); mark(q.u.y, b); }
        else if (kind == FK_BEND) { mark(q.u.x, b); mark(q.u.y, b); mark(q.u.z, b); }
        else if (kind == FK_PSPRING) mark(q.u.x, b);
      }
    }
    if (upload(S.clothSlots, cs)) return AVBD_ERR_HIP;
  }
  s->stats.max_color_used = (int32_t)s->staticUsedColors - 1;
  s->stats.num_colors = (int32_t)s->staticUsedColors;
  return AVBD_OK;
}

// Times primal sweeps / iterations for each lane split and sweep mode on the
// initial state, then restores every buffer the trials touched.
void autotune(avbd_solver* s) {
  DeviceState& S = s->S;
  hipStream_t st = s->stream;
  const uint32_t nb = s->nb;
  int wantLanes = s->opts.lanes_per_body;
  int wantMode = s->opts.sweep_mode;
  std::vector<int> laneCands = {4, 8, 16, 32};
  if (wantLanes == 4 || wantLanes == 8 || wantLanes == 16 || wantLanes == 32) laneCands = {wantLanes};
  std::vector<int> modeCands = {1, 2};
  if (wantMode == 1 || wantMode == 2) modeCands = {wantMode};
  // dynamic coloring with per-color launches would dispatch all 64 colors
  if (s->dynamicColoring && wantMode == 0) modeCands = {2, 1};
  // cloth: the Planar-DAT passes run between the colors' launches
  if (s->datOn || s->isoOn) modeCands = {1};
  if (laneCands.size() == 1 && modeCands.size() == 1) {
    s->lanes = laneCands[0];
    s->sweepMode = modeCands[0];
  }
  // persistent grid size per lane count
  auto pblocks = [&](int L) {
    int maxB = persistent_max_blocks(L, s->allMode);
    uint32_t maxCount = 0;
    for (uint32_t c : s->staticColorCounts) maxCount = std::max(maxCount, c);
    if (s->dynamicColoring) maxCount = nb;
    uint32_t want = (maxCount * (uint32_t)L + 255) / 256;
    want = (want + 7) / 8 * 8;
    if (want < 8) want = 8;
    if (maxB <= 0) return 0;
    return (int)std::min<uint32_t>(want, (uint32_t)maxB);
  };
  if (laneCands.size() == 1 && modeCands.size() == 1) {
    s->persistentBlocks = pblocks(s->lanes);
    if (s->sweepMode == 2 && s->persistentBlocks <= 0) s->sweepMode = 1;
    return;
  }
  // backup
  std::vector<std::pair<void*, size_t>> bufs = {
      {S.posLin, nb * 16}, {S.posLo, nb * 16}, {S.posAng, nb * 16}, {S.initLin, nb * 16}, {S.initAng, nb * 16},
      {S.inertLin, nb * 16}, {S.inertAng, nb * 16}, {S.posLinSnap, nb * 16}, {S.posAngSnap, nb * 16},
      {S.joints, std::max<uint32_t>(1, s->nj) * sizeof(JointRec)},
      {S.springs, std::max<uint32_t>(1, s->ns) * sizeof(SpringRec)}, {S.counters, CTR_TOTAL * 4},
      {S.dAdjCount, nb * 4}, {S.dAdjStart, nb * 4}, {S.colors, nb * 4}, {S.colorsB, nb * 4},
      {S.colorList, nb * 4}, {S.colorStart, (kMaxColors + 2) * 4}};
  std::vector<void*> backups;
  for (auto& b : bufs) {
    void* q = nullptr;
    if (hipMalloc(&q, b.second) != hipSuccess) { backups.push_back(nullptr); continue; }
    hipMemcpyAsync(q, b.first, b.second, hipMemcpyDeviceToDevice, st);
    backups.push_back(q);
  }
  // a plausible solve state: inertial target = start = current pose
  hipMemcpyAsync(S.initLin, S.posLin, nb * 16, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(S.inertLin, S.posLin, nb * 16, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(S.initAng, S.posAng, nb * 16, hipMemcpyDeviceToDevice, st);
  hipMemcpyAsync(S.inertAng, S.posAng, nb * 16, hipMemcpyDeviceToDevice, st);
  launch_fill_u32(st, S.counters, CTR_TOTAL, 0u);
  launch_fill_u32(st, S.dAdjCount, nb, 0u);
  launch_fill_u32(st, S.dAdjStart, nb, 0u);
  if (s->dynamicColoring) {
    StepParams P = make_params(s);
    launch_dynamic_coloring(st, S, P);
  }
  StepParams P = make_params(s);
  P.iterations = 2;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  float best = 3.4e38f;
  int bestL = 8, bestM = 1, bestBlocks = 0;
  for (int M : modeCands) {
    for (int L : laneCands) {
      int blocks = pblocks(L);
      if (M == 2 && blocks <= 0) continue;
      auto run = [&]() -> bool {
        if (M == 2) {
          launch_fill_u32(st, S.counters + CTR_GBAR, 2, 0u);
          return launch_persistent_solve(st, S, P, s->dynamicColoring ? 0u : s->staticUsedColors, L, s->allMode,
                                         (uint32_t)blocks, 0u) == hipSuccess;
        }
        for (uint32_t it = 0; it < P.iterations; it++) {
          if (s->dynamicColoring) {
            for (uint32_t c = 0; c < kDynamicColorLaunches; c++)
              launch_primal_color(st, S, P, c, 0xFFFFFFFFu, 0u, L, kModeGeneral);
            launch_primal_tail(st, S, P, kDynamicColorLaunches, L);
          } else {
            for (uint32_t c = 0; c < s->staticUsedColors; c++)
              launch_primal_color(st, S, P, c, s->staticColorCounts[c], s->staticColorStarts[c], L, s->colorMode[c]);
          }
          launch_dual(st, S, P, s->gridCap);
        }
        return hipGetLastError() == hipSuccess;
      };
      if (!run()) { hipGetLastError(); continue; }  // warm-up (and launchability check)
      hipEventRecord(e0, st);
      bool ok = true;
      for (int r = 0; r < 3 && ok; r++) ok = run();
      hipEventRecord(e1, st);
      if (!ok || hipEventSynchronize(e1) != hipSuccess) { hipGetLastError(); continue; }
      uint32_t fail = 0;
      hipMemcpy(&fail, S.counters + CTR_GBAR_FAIL, 4, hipMemcpyDeviceToHost);
      if (fail) continue;
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bestL = L; bestM = M; bestBlocks = blocks; }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  for (size_t i = 0; i < bufs.size(); i++) {
    if (!backups[i]) continue;
    hipMemcpyAsync(bufs[i].first, backups[i], bufs[i].second, hipMemcpyDeviceToDevice, st);
  }
  hipStreamSynchronize(st);
  for (void* q : backups)
    if (q) hipFree(q);
  s->lanes = bestL;
  s->sweepMode = bestM;
  s->persistentBlocks = bestBlocks ? bestBlocks : pblocks(bestL);
  if (s->sweepMode == 2 && s->persistentBlocks <= 0) s->sweepMode = 1;
  hipStreamSynchronize(st);
}

}  // namespace

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------
extern "C" avbd_status avbd_submit_step(avbd_solver* s) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  if (s->failed) return AVBD_ERR_RUNTIME_FAILURE;
  int64_t frame = s->frameIndex + 1;
  int slot = (int)(frame & 1);
  // throttle: at most two frames in flight; the slot's previous owner retires first
  if (s->slotBusy[slot]) {
    avbd_status st = retire_slot(s, slot);
    if (st != AVBD_OK) return st;
  }
  static const char* stages[] = {"broadphase", "narrowphase", "warmstart", "adjacency", "coloring", "solve",
                                 "finalize"};
  for (const char* stg : stages) {
    if (stage_denied(s, stg, frame)) {
      sync_all(s);
      char m[160];
      snprintf(m, sizeof(m), "physics frame %lld: encoder creation failed at %s", (long long)frame, stg);
      return latch(s, AVBD_FAIL_STAGE_DENIED, frame, 0, 0, 0, m);
    }
  }
  bool useGraph = s->opts.use_graph && !s->graphsDisabled && !s->profiling;
  if (useGraph && !s->graphsValid) {
    hipError_t e = build_graphs(s);
    if (e != hipSuccess) {
      hipGetLastError();
      invalidate_graphs(s);
      s->graphsDisabled = true;
      useGraph = false;
    }
  }
  if (useGraph) {
    // graphs[0] was captured with the buffer roles of frame graphBaseFrame
    HIP_OR_LATCH(s, hipGraphLaunch(s->graphs[(frame - s->graphBaseFrame) & 1], s->stream), frame);
  } else if (s->profiling) {
    // eager, with per-stage events
    hipStream_t st = s->stream;
    std::vector<hipEvent_t> ev;
    std::vector<std::string> names;
    auto mark = [&](const char* name) {
      hipEvent_t e;
      hipEventCreate(&e);
      hipEventRecord(e, st);
      ev.push_back(e);
      if (name) names.push_back(name);
    };
    mark(nullptr);
    HIP_OR_LATCH(s, enqueue_frame(s, s->S, mark), frame);
    HIP_OR_LATCH(s, hipStreamSynchronize(st), frame);
    s->profNames = names;
    s->profMs.assign(names.size(), 0.0f);
    for (size_t i = 0; i + 1 < ev.size(); i++) hipEventElapsedTime(&s->profMs[i], ev[i], ev[i + 1]);
    for (auto e : ev) hipEventDestroy(e);
  } else {
    HIP_OR_LATCH(s, enqueue_frame(s, s->S), frame);
  }
  std::swap(s->S.manifolds, s->S.prevManifolds);
  std::swap(s->S.soft, s->S.prevSoft);
  HIP_OR_LATCH(s,
               hipMemcpyAsync(s->pinnedCounters + slot * CTR_TOTAL, s->S.counters, CTR_TOTAL * 4,
                              hipMemcpyDeviceToHost, s->stream),
               frame);
  HIP_OR_LATCH(s, hipEventRecord(s->frameDone[slot], s->stream), frame);
  s->slotBusy[slot] = true;
  s->slotFrame[slot] = frame;
  s->frameIndex = frame;
  return AVBD_OK;
}

extern "C" avbd_status avbd_step(avbd_solver* s) { return avbd_submit_step(s); }

// Residual-terminated frames.  Iterations run in chunks of `check_every`
// sweeps; each chunk ends with the residual kernel and a one-block check
// that sets the frame's done flag on the device (r <= tol or the sweep cap)
// and publishes its progress to mapped host memory.  Every sweep kernel
// returns at entry once the flag is set, so chunks can be queued ahead
// without a host round trip per chunk: the host keeps a few chunks in
// flight, watches the progress record, and stops queueing at the first done
// check.  The result is that of the first chunk whose check passes, the
// same as checking on the host after every chunk.  Pre-solve, chunk and
// post-solve are captured graphs (even / odd frame roles) unless graphs are
// off; a final partial chunk (cap not a multiple of check_every) is
// enqueued directly.
namespace {

constexpr uint32_t kChunksInFlight = 3;

hipError_t enqueue_chunk(avbd_solver* s, const DeviceState& S, const StepParams& P, uint32_t n, float tol,
                         uint32_t maxIt) {
  StepParams Pc = P;
  Pc.convCheck = 1u;  // the chunk's kernels stop once a check has ended the frame
  hipError_t e = enqueue_iterations(s, S, Pc, n, true);
  if (e != hipSuccess) return e;
  launch_residual(s->stream, S, Pc, s->residualBlocks, true);
  launch_residual_check(s->stream, S, s->residualBlocks, tol, maxIt, n, s->convDev);
  return hipGetLastError();
}

hipError_t build_residual_graphs(avbd_solver* s, float tol, int32_t maxIt, int32_t check) {
  for (int i = 0; i < 2; i++)
    for (hipGraphExec_t* g : {&s->resPre[i], &s->resChunk[i], &s->resPost[i]}) {
      if (*g) hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  s->resValid = false;
  DeviceState S = s->S;
  const StepParams P = make_params(s);
  auto none = [](const char*) {};
  for (int p = 0; p < 2; p++) {
    hipError_t e = capture_graph(s, [&]() { enqueue_pre_solve(s, S, P, none); return hipGetLastError(); },
                                 &s->resPre[p]);
    if (e == hipSuccess)
      e = capture_graph(s, [&]() { return enqueue_chunk(s, S, P, (uint32_t)check, tol, (uint32_t)maxIt); },
                        &s->resChunk[p]);
    if (e == hipSuccess)
      e = capture_graph(s, [&]() { enqueue_post_solve(s, S, P, none); return hipGetLastError(); }, &s->resPost[p]);
    if (e != hipSuccess) return e;
    std::swap(S.manifolds, S.prevManifolds);
    std::swap(S.soft, S.prevSoft);
  }
  s->resValid = true;
  s->resCheck = check;
  s->resMax = maxIt;
  s->resTol = tol;
  s->resBase = s->frameIndex + 1;
  return hipSuccess;
}

}  // namespace

extern "C" avbd_status avbd_step_to_residual(avbd_solver* s, float tol, int32_t max_iterations, int32_t check_every,
                                             int32_t* iterations_used, float* residual) {
  if (!s || !(tol > 0.0f) || max_iterations <= 0 || check_every <= 0) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st0 = sync_all(s);
  if (st0 != AVBD_OK) return st0;
  const int64_t frame = s->frameIndex + 1;
  const int slot = (int)(frame & 1);
  hipStream_t st = s->stream;
  bool useGraph = s->opts.use_graph && !s->graphsDisabled && !s->profiling;
  if (useGraph && !(s->resValid && s->resCheck == check_every && s->resMax == max_iterations && s->resTol == tol)) {
    if (build_residual_graphs(s, tol, max_iterations, check_every) != hipSuccess) {
      (void)hipGetLastError();
      invalidate_graphs(s);
      s->graphsDisabled = true;
      useGraph = false;
    }
  }
  const int parity = useGraph ? (int)((frame - s->resBase) & 1) : 0;
  const DeviceState S = s->S;
  const StepParams P = make_params(s);
  auto none = [](const char*) {};
  // the device has retired every earlier check (sync_all above)
  memset(s->convHost, 0, sizeof(ConvHost));
  if (useGraph) HIP_OR_LATCH(s, hipGraphLaunch(s->resPre[parity], st), frame);
  else enqueue_pre_solve(s, S, P, none);
  const uint32_t full = (uint32_t)(max_iterations / check_every), tail = (uint32_t)(max_iterations % check_every);
  const uint32_t total = full + (tail ? 1u : 0u);
  uint32_t launched = 0, seen = 0;
  volatile ConvHost* h = s->convHost;
  for (;;) {
    while (launched < total && launched - seen < kChunksInFlight) {
      if (launched < full && useGraph) {
        HIP_OR_LATCH(s, hipGraphLaunch(s->resChunk[parity], st), frame);
      } else {
        const uint32_t n = launched < full ? (uint32_t)check_every : tail;
        HIP_OR_LATCH(s, enqueue_chunk(s, S, P, n, tol, (uint32_t)max_iterations), frame);
      }
      launched++;
    }
    // wait for the next check to land (or the stream to go idle / fail)
    for (uint32_t spin = 0;; spin++) {
      const uint32_t q = __atomic_load_n(&s->convHost->seq, __ATOMIC_ACQUIRE);
      if (q != seen) {
        seen = q;
        break;
      }
      if ((spin & 255u) == 255u) {
        hipError_t e = hipStreamQuery(st);
        if (e == hipSuccess) {
          seen = __atomic_load_n(&s->convHost->seq, __ATOMIC_ACQUIRE);
          break;
        }
        if (e != hipErrorNotReady) HIP_OR_LATCH(s, e, frame);
      }
      __builtin_ia32_pause();
    }
    if (h->done != 0u || seen >= total || (seen == launched && launched == total)) break;
  }
  if (useGraph) HIP_OR_LATCH(s, hipGraphLaunch(s->resPost[parity], st), frame);
  else enqueue_post_solve(s, S, P, none);
  HIP_OR_LATCH(s, hipGetLastError(), frame);
  std::swap(s->S.manifolds, s->S.prevManifolds);
  std::swap(s->S.soft, s->S.prevSoft);
  HIP_OR_LATCH(s,
               hipMemcpyAsync(s->pinnedCounters + slot * CTR_TOTAL, s->S.counters, CTR_TOTAL * 4,
                              hipMemcpyDeviceToHost, st),
               frame);
  HIP_OR_LATCH(s, hipEventRecord(s->frameDone[slot], st), frame);
  s->slotBusy[slot] = true;
  s->slotFrame[slot] = frame;
  s->frameIndex = frame;
  avbd_status out = sync_all(s);
  if (iterations_used) *iterations_used = (int32_t)s->pinnedCounters[slot * CTR_TOTAL + CTR_CONV_ITERS];
  if (residual) *residual = as_float(s->pinnedCounters[slot * CTR_TOTAL + CTR_CONV_RES]);
  return out;
}

extern "C" avbd_status avbd_synchronize(avbd_solver* s) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  return sync_all(s);
}

extern "C" avbd_status avbd_get_settings(const avbd_solver* s, avbd_settings* out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  *out = s->settings;
  return AVBD_OK;
}

extern "C" avbd_status avbd_set_settings(avbd_solver* s, const avbd_settings* in) {
  if (!s || !in) return AVBD_ERR_INVALID_ARGUMENT;
  if (!(in->dt > 0.0f) || in->iterations < 0 || !(in->collision_margin >= 0.0f)) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  s->settings = *in;
  if (!s->spinBodies.empty() && s->settings.dt != s->spinDt) {
    // the spinners' per-frame rotations follow the step size
    s->spinDt = s->settings.dt;
    if (upload(s->spinDqDev, spinner_rotations(s, s->spinDt))) return AVBD_ERR_HIP;
  }
  invalidate_graphs(s);
  return st == AVBD_ERR_RUNTIME_FAILURE ? st : AVBD_OK;
}

// ---------------------------------------------------------------------------
// readback / reset
// ---------------------------------------------------------------------------
namespace {
template <typename T>
avbd_status read_dev(avbd_solver* s, std::vector<T>& out, const T* src, size_t n) {
  out.resize(n);
  if (n == 0) return AVBD_OK;
  if (hipMemcpy(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return AVBD_ERR_HIP;
  return AVBD_OK;
}
}  // namespace

extern "C" avbd_status avbd_get_body_states(avbd_solver* s, const uint32_t* idx, uint32_t n, float* pos3,
                                            float* rot4, float* vel3, float* angvel3) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  (void)st;  // state stays readable after a failure, as the reference's shared buffers do
  std::vector<F4> pl, pa, vl, va;
  if (read_dev(s, pl, s->S.posLin, s->nb) || read_dev(s, pa, s->S.posAng, s->nb) ||
      read_dev(s, vl, s->S.velLin, s->nb) || read_dev(s, va, s->S.velAng, s->nb))
    return AVBD_ERR_HIP;
  for (uint32_t k = 0; k < n; k++) {
    uint32_t i = idx ? idx[k] : k;
    if (i >= s->nb) return AVBD_ERR_INVALID_ARGUMENT;
    if (pos3) { pos3[3 * k] = pl[i].x; pos3[3 * k + 1] = pl[i].y; pos3[3 * k + 2] = pl[i].z; }
    if (rot4) { rot4[4 * k] = pa[i].x; rot4[4 * k + 1] = pa[i].y; rot4[4 * k + 2] = pa[i].z; rot4[4 * k + 3] = pa[i].w; }
    if (vel3) { vel3[3 * k] = vl[i].x; vel3[3 * k + 1] = vl[i].y; vel3[3 * k + 2] = vl[i].z; }
    if (angvel3) { angvel3[3 * k] = va[i].x; angvel3[3 * k + 1] = va[i].y; angvel3[3 * k + 2] = va[i].z; }
  }
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_body_mass(avbd_solver* s, uint32_t n, float* mass, float* inertia3) {
  if (!s || n > s->nb) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; i++) {
    if (mass) mass[i] = s->mass[i];
    if (inertia3) for (int k = 0; k < 3; k++) inertia3[3 * i + k] = s->inertia[3 * i + k];
  }
  return AVBD_OK;
}

namespace {
// Clears the warm starts of every joint, spring and manifold (both frame
// buffers) incident to a flagged body (GPUSolver.swift:2433-2472).
__global__ void k_reset_incident(DeviceState S, StepParams P, const uint32_t* flag, const F4* initPenLin,
                                 const F4* initPenAng, ManifoldRec* other, uint32_t maxPairs) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  auto flagged = [&](uint32_t b) { return b != kWorld && b < P.numBodies && flag[b] != 0u; };
  if (gid < P.numJoints) {
    JointRec& j = S.joints[gid];
    if (flagged(j.header.y) || flagged(j.header.x)) {
      j.lambdaLin = F4{0, 0, 0, 0};
      j.lambdaAng = F4{0, 0, 0, 0};
      j.motor.z = 0.0f;
      j.dynamics.y = 0.0f;
      j.dynamics.z = 0.0f;
      j.penaltyLin = initPenLin[gid];
      j.penaltyAng = initPenAng[gid];
    }
  }
  if (gid < P.numSprings) {
    SpringRec& sp = S.springs[gid];
    if (flagged(sp.header.x) || flagged(sp.header.y)) sp.dual = F4{0, 0, 0, 0};
  }
  if (gid < maxPairs) {
    ManifoldRec* bufs[2] = {S.manifolds, other};
    for (int k = 0; k < 2; k++) {
      ManifoldRec& m = bufs[k][gid];
      if (flagged(m.header.x) || flagged(m.header.y)) {
        uint32_t* w = (uint32_t*)&m;
        for (uint32_t q = 0; q < sizeof(ManifoldRec) / 4; q++) w[q] = 0u;
      }
    }
  }
}
__global__ void k_write_states(DeviceState S, const uint32_t* idx, uint32_t n, const float* p, const float* r,
                               const float* v, const float* w) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t i = idx[k];
  float m = S.posLin[i].w;
  S.posLin[i] = F4{p[3 * k], p[3 * k + 1], p[3 * k + 2], m};
  S.posLo[i] = F4{0.0f, 0.0f, 0.0f, 0.0f};
  S.posAng[i] = F4{r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]};
  S.velLin[i] = F4{v[3 * k], v[3 * k + 1], v[3 * k + 2], 0.0f};
  S.velAng[i] = F4{w[3 * k], w[3 * k + 1], w[3 * k + 2], 0.0f};
  S.prevVelLin[i] = F4{v[3 * k], v[3 * k + 1], v[3 * k + 2], 0.0f};
}
}  // namespace

extern "C" avbd_status avbd_set_body_states(avbd_solver* s, const uint32_t* idx, uint32_t n, const float* pos3,
                                            const float* rot4, const float* vel3, const float* angvel3) {
  if (!s || !pos3 || !rot4 || !vel3 || !angvel3) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  std::vector<uint32_t> ids(n), flag(s->nb, 0u);
  for (uint32_t k = 0; k < n; k++) {
    uint32_t i = idx ? idx[k] : k;
    if (i >= s->nb) return AVBD_ERR_INVALID_ARGUMENT;
    ids[k] = i;
    flag[i] = 1u;
  }
  // one staging block: ids [n], flags [nb], pos / rot / vel / angvel [13 n]
  const size_t fl = (size_t)n * 13;
  const size_t words = (size_t)n + s->nb + fl;
  void* block = nullptr;
  if (hipMalloc(&block, 4 * words) != hipSuccess) {
    (void)hipGetLastError();
    return AVBD_ERR_ALLOC;
  }
  uint32_t* dIdx = (uint32_t*)block;
  uint32_t* dFlag = dIdx + n;
  float* dBuf = (float*)(dFlag + s->nb);
  std::vector<float> buf(fl);
  memcpy(buf.data(), pos3, 12 * (size_t)n);
  memcpy(buf.data() + 3 * (size_t)n, rot4, 16 * (size_t)n);
  memcpy(buf.data() + 7 * (size_t)n, vel3, 12 * (size_t)n);
  memcpy(buf.data() + 10 * (size_t)n, angvel3, 12 * (size_t)n);
  if (hipMemcpy(dIdx, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dFlag, flag.data(), 4 * (size_t)s->nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dBuf, buf.data(), 4 * fl, hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(block);
    return AVBD_ERR_HIP;
  }
  // the incident joints get back the penalties stored at upload
  // (initialJointPenaltyLin / Ang), whatever dt the settings hold now
  F4* dPen = s->jointPenInit;
  StepParams P = make_params(s);
  hipLaunchKernelGGL(k_write_states, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->S, dIdx, n, dBuf,
                     dBuf + 3 * n, dBuf + 7 * n, dBuf + 10 * n);
  uint32_t total = std::max(std::max(s->nj, s->ns), s->maxPairs);
  hipLaunchKernelGGL(k_reset_incident, dim3((total + 255) / 256), dim3(256), 0, s->stream, s->S, P, dFlag, dPen,
                     dPen + s->nj, s->S.prevManifolds, s->maxPairs);
  // written states invalidate the learned stage's correction history
  if (s->wsBuilt) {
    hipMemsetAsync(s->W.corr, 0, 16 * (size_t)s->nb, s->stream);
    hipMemsetAsync(s->W.corrPrev, 0, 16 * (size_t)s->nb, s->stream);
    hipMemsetAsync(s->W.corrAng, 0, 16 * (size_t)s->nb, s->stream);
  }
  hipError_t e = hipStreamSynchronize(s->stream);
  hipFree(block);
  if (e != hipSuccess) return latch(s, AVBD_FAIL_COMMAND_EXECUTION, s->frameIndex, 0, 0, (int)e, "set_body_states");
  return AVBD_OK;
}

namespace {
// body ids of a reset: the caller's list, or 0..n-1 when it passed none
__global__ void k_iota(uint32_t* out, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = k;
}
__global__ void k_flag_bodies(const uint32_t* idx, uint32_t n, uint32_t* flag) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flag[idx[k]] = 1u;
}
}  // namespace

// avbd_set_body_states with its staging (ids, per-body flags, the 13 floats
// per body) in the device's fixed scratch array instead of a fresh
// allocation; the same kernels do the write and the warm-start clearing.
extern "C" avbd_status avbd_set_body_states_noalloc(avbd_solver* s, const uint32_t* idx, uint32_t n,
                                                    const float* pos3, const float* rot4, const float* vel3,
                                                    const float* angvel3) {
  if (!s || !pos3 || !rot4 || !vel3 || !angvel3) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  if ((size_t)n * 14 + s->nb > kFixedScratchWords) return AVBD_ERR_TOO_LARGE;
  if (!s->fixedScratch) return AVBD_ERR_NO_DEVICE;
  if (idx)
    for (uint32_t k = 0; k < n; k++)
      if (idx[k] >= s->nb) return AVBD_ERR_INVALID_ARGUMENT;
  if (!idx && n > s->nb) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  hipStream_t q = s->stream;
  uint32_t* dIdx = s->fixedScratch;                 // [n]
  uint32_t* dFlag = dIdx + n;                       // [nb], cleared below
  float* dBuf = (float*)(dFlag + s->nb);            // pos [3n], rot [4n], vel [3n], angvel [3n]
  const size_t n3 = 12 * (size_t)n, n4 = 16 * (size_t)n;
  hipError_t e = hipSuccess;
  if (idx) e = hipMemcpyAsync(dIdx, idx, 4 * (size_t)n, hipMemcpyHostToDevice, q);
  else hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, q, dIdx, n);
  if (e == hipSuccess) e = hipMemcpyAsync(dBuf, pos3, n3, hipMemcpyHostToDevice, q);
  if (e == hipSuccess) e = hipMemcpyAsync(dBuf + 3 * (size_t)n, rot4, n4, hipMemcpyHostToDevice, q);
  if (e == hipSuccess) e = hipMemcpyAsync(dBuf + 7 * (size_t)n, vel3, n3, hipMemcpyHostToDevice, q);
  if (e == hipSuccess) e = hipMemcpyAsync(dBuf + 10 * (size_t)n, angvel3, n3, hipMemcpyHostToDevice, q);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(q);
    return AVBD_ERR_HIP;
  }
  // the flags keep the previous call's contents: clear them with a fill kernel
  launch_fill_u32(q, dFlag, s->nb, 0u);
  hipLaunchKernelGGL(k_flag_bodies, dim3((n + 255) / 256), dim3(256), 0, q, dIdx, n, dFlag);
  F4* dPen = s->jointPenInit;
  StepParams P = make_params(s);
  hipLaunchKernelGGL(k_write_states, dim3((n + 255) / 256), dim3(256), 0, q, s->S, dIdx, n, dBuf,
                     dBuf + 3 * n, dBuf + 7 * n, dBuf + 10 * n);
  uint32_t total = std::max(std::max(s->nj, s->ns), s->maxPairs);
  hipLaunchKernelGGL(k_reset_incident, dim3((total + 255) / 256), dim3(256), 0, q, s->S, P, dFlag, dPen,
                     dPen + s->nj, s->S.prevManifolds, s->maxPairs);
  if (s->wsBuilt) {
    launch_fill_u32(q, (uint32_t*)s->W.corr, 4 * s->nb, 0u);
    launch_fill_u32(q, (uint32_t*)s->W.corrPrev, 4 * s->nb, 0u);
    launch_fill_u32(q, (uint32_t*)s->W.corrAng, 4 * s->nb, 0u);
  }
  e = hipStreamSynchronize(q);
  if (e != hipSuccess) return latch(s, AVBD_FAIL_COMMAND_EXECUTION, s->frameIndex, 0, 0, (int)e, "set_body_states");
  return AVBD_OK;
}

namespace {
// field 0: motor.x (target angle / velocity), 1: motor.y (effort limit), 2: rA.xyz (world anchor)
__global__ void k_set_joint_field(DeviceState S, const uint32_t* j, const float* v, uint32_t n, int field) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  JointRec& r = S.joints[j[k]];
  if (field == 0) r.motor.x = v[k];
  else if (field == 1) r.motor.y = v[k];
  else r.rA = F4{v[3 * k], v[3 * k + 1], v[3 * k + 2], r.rA.w};
}
avbd_status set_joint_field(avbd_solver* s, const uint32_t* joints, uint32_t n, const float* v, int field) {
  const uint32_t width = field == 2 ? 3u : 1u;
  void* block = nullptr;  // joint ids [n], then values [width n]
  if (hipMalloc(&block, 4 * (size_t)n * (1 + width)) != hipSuccess) {
    (void)hipGetLastError();
    return AVBD_ERR_ALLOC;
  }
  uint32_t* dj = (uint32_t*)block;
  float* dv = (float*)(dj + n);
  hipError_t e = hipMemcpy(dj, joints, 4 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv, v, 4 * (size_t)width * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_set_joint_field, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->S, dj, dv, n, field);
    e = hipStreamSynchronize(s->stream);
  }
  hipFree(block);
  return e == hipSuccess ? AVBD_OK : AVBD_ERR_HIP;
}
// the preconditions of GPUSolver.swift:2496-2590 as argument errors; kept
// joints (mode 0 targets are no-ops) compacted into idx / val
avbd_status motor_updates(avbd_solver* s, const uint32_t* joints, uint32_t n, const float* v, int kind,
                          std::vector<uint32_t>& idx, std::vector<float>& val) {
  if (!s || (n && (!joints || !v))) return AVBD_ERR_INVALID_ARGUMENT;
  for (uint32_t k = 0; k < n; k++) {
    if (joints[k] >= s->nj || !isfinite(v[k])) return AVBD_ERR_INVALID_ARGUMENT;
    const uint32_t mode = s->jointHeaders[joints[k]].w & kMotorModeMask;
    if (kind == 0) {         // position targets: position-PD motors; unpowered joints ignore them
      if (mode == 0) continue;
      if (mode != kMotorImplicitPD && mode != kMotorExplicitPD) return AVBD_ERR_INVALID_ARGUMENT;
    } else if (kind == 1) {  // velocity commands: velocity motors only
      if (mode != kMotorVelocity) return AVBD_ERR_INVALID_ARGUMENT;
    } else {                 // effort limits: authored motors only, finite and non-negative
      if (mode == 0 || v[k] < 0.0f) return AVBD_ERR_INVALID_ARGUMENT;
    }
    idx.push_back(joints[k]);
    val.push_back(v[k]);
  }
  return AVBD_OK;
}
avbd_status set_motor(avbd_solver* s, const uint32_t* joints, uint32_t n, const float* v, int kind) {
  std::vector<uint32_t> idx;
  std::vector<float> val;
  avbd_status st = motor_updates(s, joints, n, v, kind, idx, val);
  if (st != AVBD_OK) return st;
  if (idx.empty()) return AVBD_OK;
  st = sync_all(s);
  if (st != AVBD_OK) return st;
  return set_joint_field(s, idx.data(), (uint32_t)idx.size(), val.data(), kind == 2 ? 1 : 0);
}
}  // namespace

extern "C" avbd_status avbd_set_motor_targets(avbd_solver* s, const uint32_t* joints, uint32_t n, const float* t) {
  return set_motor(s, joints, n, t, 0);
}
extern "C" avbd_status avbd_set_motor_velocities(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                                 const float* rad_per_s) {
  return set_motor(s, joints, n, rad_per_s, 1);
}
extern "C" avbd_status avbd_set_motor_torques(avbd_solver* s, const uint32_t* joints, uint32_t n, const float* t) {
  return set_motor(s, joints, n, t, 2);
}

extern "C" avbd_status avbd_set_joint_world_anchors(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                                    const float* point3) {
  if (!s || (n && (!joints || !point3))) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  for (uint32_t k = 0; k < n; k++)
    if (joints[k] >= s->nj) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  return set_joint_field(s, joints, n, point3, 2);
}

extern "C" avbd_status avbd_set_body_poses(avbd_solver* s, const uint32_t* idx, uint32_t n, const float* pos3,
                                           const float* rot4) {
  if (!s || (n && (!pos3 || !rot4))) return AVBD_ERR_INVALID_ARGUMENT;
  std::vector<float> zero(3 * (size_t)n, 0.0f);
  return avbd_set_body_states(s, idx, n, pos3, rot4, zero.data(), zero.data());
}

extern "C" avbd_status avbd_get_motor_states(avbd_solver* s, const uint32_t* joints, uint32_t n, float* angle,
                                             float* velocity) {
  if (!s || (n && !joints)) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  for (uint32_t k = 0; k < n; k++)
    if (joints[k] >= s->nj) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  (void)st;  // states stay readable after a failure
  std::vector<JointRec> jr;
  std::vector<F4> pa, va;
  if (read_dev(s, jr, s->S.joints, s->nj) || read_dev(s, pa, s->S.posAng, s->nb) ||
      read_dev(s, va, s->S.velAng, s->nb))
    return AVBD_ERR_HIP;
  for (uint32_t k = 0; k < n; k++) {
    const JointRec& j = jr[joints[k]];
    const uint32_t a = j.header.x, b = j.header.y;
    const F4 qA = a == kWorld ? F4{0, 0, 0, 1} : pa[a];
    const F4 qB = pa[b];
    // relative rotation of the child from its rest pose, shortest arc
    F4 r = q_mul(q_inv(q_mul(qA, j.restRel)), qB);
    if (r.w < 0.0f) r = F4{-r.x, -r.y, -r.z, -r.w};
    const V3 axis = xyz(j.hingeAxis);
    if (angle) angle[k] = 2.0f * atan2f(dot(xyz(r), axis), r.w);
    const V3 wB = xyz(va[b]), wA = a == kWorld ? v3s(0.0f) : xyz(va[a]);
    if (velocity) velocity[k] = dot(wB - wA, q_rotate(qB, axis));
  }
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_joint_diagnostics(avbd_solver* s, int32_t* broken, float* max_lambda_lin,
                                                  float* max_lambda_ang) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  std::vector<JointRec> jr;
  if (read_dev(s, jr, s->S.joints, s->nj)) return AVBD_ERR_HIP;
  int32_t nb = 0;
  float ml = 0.0f, ma = 0.0f;
  for (const JointRec& j : jr) {
    if (j.header.z != 0) {
      nb++;
      continue;
    }
    ml = fmaxf(ml, length(xyz(j.lambdaLin)));
    ma = fmaxf(ma, length(xyz(j.lambdaAng)));
  }
  if (broken) *broken = nb;
  if (max_lambda_lin) *max_lambda_lin = ml;
  if (max_lambda_ang) *max_lambda_ang = ma;
  return AVBD_OK;
}

// ---------------------------------------------------------------------------
// stats / debug
// ---------------------------------------------------------------------------
extern "C" avbd_status avbd_get_stats(avbd_solver* s, avbd_stats* out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  s->stats.lanes_per_body = s->lanes;
  s->stats.sweep_mode = s->sweepMode;
  s->stats.uses_dynamic_coloring = s->dynamicColoring ? 1 : 0;
  *out = s->stats;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_colors(avbd_solver* s, uint32_t n, uint32_t* colors) {
  if (!s || !colors || n > s->nb) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  if (hipMemcpy(colors, s->S.colors, 4 * n, hipMemcpyDeviceToHost) != hipSuccess) return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_cell_lists(avbd_solver* s, uint32_t* hash_size, uint32_t* num_hashed, float* cell_size,
                                           uint32_t* cell_start, uint32_t* cell_count, uint32_t* cell_bodies) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  if (hash_size) *hash_size = s->gridHashSize;
  if (num_hashed) *num_hashed = s->nHashed;
  if (cell_size) *cell_size = s->cellSize;
  if (cell_start && hipMemcpy(cell_start, s->S.cellStart, 4 * s->gridHashSize, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  if (cell_count && hipMemcpy(cell_count, s->S.cellCount, 4 * s->gridHashSize, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  if (cell_bodies && s->nHashed && hipMemcpy(cell_bodies, s->S.cellBodies, 4 * s->nHashed, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_pairs(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* pairs) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  uint32_t cnt = 0;
  if (hipMemcpy(&cnt, s->S.counters + CTR_PAIRS, 4, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  *n = cnt;
  uint32_t m = std::min(cnt, capacity);
  if (pairs && m && hipMemcpy(pairs, s->S.pairs, 8 * m, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_manifolds(avbd_solver* s, uint32_t capacity, uint32_t* n, void* records) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  uint32_t cnt = 0;
  if (hipMemcpy(&cnt, s->S.counters + CTR_PAIRS, 4, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  *n = cnt;
  uint32_t m = std::min(cnt, capacity);
  // the frame just retired wrote into what is now prevManifolds
  if (records && m && hipMemcpy(records, s->S.prevManifolds, sizeof(ManifoldRec) * m, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_cloth_info(avbd_solver* s, avbd_cloth_info* info, uint32_t* edges,
                                           uint32_t* particles) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  if (info) {
    info->num_tris = s->nTris;
    info->num_edges = s->nEdges;
    info->num_particles = s->nParticles;
    info->num_membranes = s->nMembranes;
    info->num_bends = s->nBends;
    info->elem_hash_size = s->elemHashSize;
    info->max_soft = s->maxSoft;
    info->soft_map_capacity = s->softMapCapacity;
    info->max_dat_pairs = s->maxDatPairs;
    info->elem_cell_size = s->elemCellSize;
    info->elem_margin = s->elemMargin;
    info->dat_query_radius = s->datQueryRadius;
    info->surface_contact_cell_size = s->surfaceContactCellSize;
    info->num_contact_edges = s->numContactEdges;
    info->pair_stream = s->pairStream ? 1u : 0u;
    info->dat_on = s->datOn ? 1u : 0u;
    info->iso_on = s->isoOn ? 1u : 0u;
  }
  if (edges && s->nEdges && hipMemcpy(edges, s->S.edges, 8 * (size_t)s->nEdges, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  if (particles && s->nParticles &&
      hipMemcpy(particles, s->S.particleIdx, 4 * (size_t)s->nParticles, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_motor_targets(avbd_solver* s, uint32_t n, float* targets) {
  if (!s || (n && !targets) || n > s->nj) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  std::vector<JointRec> j(n);
  if (n && hipMemcpy(j.data(), s->S.joints, sizeof(JointRec) * n, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  for (uint32_t i = 0; i < n; i++) targets[i] = j[i].motor.x;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_collision_tris(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* tris) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  *n = s->nTris;
  uint32_t m = std::min(s->nTris, capacity);
  if (!tris || !m) return AVBD_OK;
  std::vector<U4> tmp(m);
  if (hipMemcpy(tmp.data(), s->S.tris, sizeof(U4) * m, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  for (uint32_t i = 0; i < m; i++) {
    tris[3 * i] = tmp[i].x;
    tris[3 * i + 1] = tmp[i].y;
    tris[3 * i + 2] = tmp[i].z;
  }
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_dat_pairs(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* pairs) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  *n = 0;
  if (!s->pairStream) return AVBD_OK;
  uint32_t cnt = 0;
  // the stream of the accepted predictor pose (its demand is the working count)
  if (hipMemcpy(&cnt, s->S.datStart + s->nParticles + s->nEdges, 4, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  cnt = std::min(cnt, s->maxDatPairs);
  *n = cnt;
  uint32_t m = std::min(cnt, capacity);
  if (pairs && m) {
    std::vector<U4> tmp(m);
    if (hipMemcpy(tmp.data(), s->S.datPairs, sizeof(U4) * m, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
    for (uint32_t i = 0; i < m; i++) {
      pairs[3 * i] = tmp[i].x;
      pairs[3 * i + 1] = tmp[i].y;
      pairs[3 * i + 2] = tmp[i].z;
    }
  }
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_soft_contacts(avbd_solver* s, uint32_t capacity, uint32_t* n, void* records) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  *n = 0;
  if (!s->nTris) return AVBD_OK;
  uint32_t cnt = 0;
  if (hipMemcpy(&cnt, s->S.counters + CTR_SOFT, 4, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  *n = cnt;
  uint32_t m = std::min(cnt, capacity);
  // the frame just retired wrote into what is now prevSoft
  if (records && m && hipMemcpy(records, s->S.prevSoft, sizeof(SoftRec) * m, hipMemcpyDeviceToHost))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_create_profile(avbd_solver* s, uint32_t index, const char** name, float* ms) {
  if (!s || index >= s->createPhases.size()) return AVBD_ERR_INVALID_ARGUMENT;
  if (name) *name = s->createPhases[index].first;
  if (ms) *ms = (float)s->createPhases[index].second;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_iso_state(avbd_solver* s, uint32_t* vt_track, uint32_t* ee_track, float* bound,
                                          float* anchor, uint32_t* counts) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  if (!s->isoOn) return AVBD_ERR_UNSUPPORTED;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  const size_t nb = s->nb, ne = s->numContactEdges;
  if (vt_track && hipMemcpy(vt_track, s->S.vtTrack, 16 * nb, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  if (ee_track && ne && hipMemcpy(ee_track, s->S.eeTrack, 16 * ne, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  if (bound && hipMemcpy(bound, s->S.ogcBound, 4 * nb, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
  if (anchor) {
    std::vector<F4> tmp(nb);
    if (hipMemcpy(tmp.data(), s->S.ogcPrev, 16 * nb, hipMemcpyDeviceToHost)) return AVBD_ERR_HIP;
    for (size_t i = 0; i < nb; i++) {
      anchor[3 * i] = tmp[i].x;
      anchor[3 * i + 1] = tmp[i].y;
      anchor[3 * i + 2] = tmp[i].z;
    }
  }
  if (counts) {
    counts[0] = s->lastOgcTrunc;
    counts[1] = s->lastOgcRefresh;
  }
  return AVBD_OK;
}

// Runs one of the persistence-map lookup kernels over host query pairs.
static avbd_status map_query(avbd_solver* s, const uint32_t* pairs, uint32_t n, int32_t* out, bool soft) {
  if (!s || (n && (!pairs || !out))) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  sync_all(s);
  U2* dq = nullptr;
  int32_t* dout = nullptr;
  if (hipMalloc(&dq, 8ull * n) != hipSuccess) return AVBD_ERR_ALLOC;
  if (hipMalloc(&dout, 4ull * n) != hipSuccess) {
    hipFree(dq);
    return AVBD_ERR_ALLOC;
  }
  hipError_t e = hipMemcpy(dq, pairs, 8ull * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    StepParams P = make_params(s);
    if (soft) {
      if (P.softMapCapacity && s->S.softMapKeyA) launch_soft_map_find(s->stream, s->S, P, dq, dout, n);
      else e = hipMemset(dout, 0xFF, 4ull * n);  // no element contacts: every key is absent
    } else {
      launch_map_find(s->stream, s->S, P, dq, dout, n);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dout, 4ull * n, hipMemcpyDeviceToHost);
  hipFree(dq);
  hipFree(dout);
  return e == hipSuccess ? AVBD_OK : AVBD_ERR_HIP;
}

extern "C" avbd_status avbd_map_find(avbd_solver* s, const uint32_t* pairs, uint32_t n, int32_t* out) {
  return map_query(s, pairs, n, out, false);
}

extern "C" avbd_status avbd_soft_map_find(avbd_solver* s, const uint32_t* keys, uint32_t n, int32_t* out) {
  return map_query(s, keys, n, out, true);
}

extern "C" avbd_status avbd_max_constraint_error(avbd_solver* s, float* out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  // the last frame's manifolds are in prevManifolds after the swap
  DeviceState S = s->S;
  std::swap(S.manifolds, S.prevManifolds);
  std::swap(S.soft, S.prevSoft);
  StepParams P = make_params(s);
  launch_fill_u32(s->stream, s->devScratchU32, 1, 0u);
  launch_diag_error(s->stream, S, P, s->devScratchU32, s->gridCap);
  uint32_t bits = 0;
  if (hipMemcpyAsync(&bits, s->devScratchU32, 4, hipMemcpyDeviceToHost, s->stream) ||
      hipStreamSynchronize(s->stream))
    return AVBD_ERR_HIP;
  *out = as_float(bits);
  return AVBD_OK;
}

extern "C" avbd_status avbd_residual(avbd_solver* s, float* out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  DeviceState S = s->S;
  std::swap(S.manifolds, S.prevManifolds);
  std::swap(S.soft, S.prevSoft);
  StepParams P = make_params(s);
  launch_residual(s->stream, S, P, s->residualBlocks);
  if (hipMemcpyAsync(s->residualPartialHost, S.residualPartial, 16 * (size_t)s->residualBlocks,
                     hipMemcpyDeviceToHost, s->stream) ||
      hipStreamSynchronize(s->stream))
    return AVBD_ERR_HIP;
  double g2 = 0.0, w2 = 0.0;
  for (uint32_t b = 0; b < s->residualBlocks; b++) {
    g2 += s->residualPartialHost[2 * b];
    w2 += s->residualPartialHost[2 * b + 1];
  }
  *out = w2 > 0.0 ? (float)sqrt(g2 / w2) : (float)sqrt(g2);
  return AVBD_OK;
}

extern "C" avbd_status avbd_last_failure(const avbd_solver* s, avbd_failure* out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  if (s->failed) *out = s->failure;
  else { memset(out, 0, sizeof(*out)); out->frame = -1; }
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_profile(avbd_solver* s, int32_t enable, char* names, uint32_t names_cap, float* ms,
                                        uint32_t* n_stages) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  sync_all(s);
  s->profiling = enable != 0;
  if (n_stages) *n_stages = (uint32_t)s->profNames.size();
  if (names && names_cap) {
    std::string joined;
    for (size_t i = 0; i < s->profNames.size(); i++) {
      if (i) joined += ",";
      joined += s->profNames[i];
    }
    snprintf(names, names_cap, "%s", joined.c_str());
  }
  if (ms)
    for (size_t i = 0; i < s->profMs.size(); i++) ms[i] = s->profMs[i];
  return AVBD_OK;
}

extern "C" avbd_status avbd_time_primal(avbd_solver* s, int32_t reps, float* avg_launch_us, double* sweep_bytes,
                                        int32_t* launches_per_sweep) {
  if (!s || reps <= 0) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  DeviceState S = s->S;
  std::swap(S.manifolds, S.prevManifolds);  // the last frame's contacts
  std::swap(S.soft, S.prevSoft);
  StepParams P = make_params(s);
  // algorithmic bytes of one sweep: per body fixed + per stamp record model
  std::vector<uint32_t> dc, dstart, dlist;
  if (read_dev(s, dc, S.dAdjCount, s->nb) || read_dev(s, dstart, S.dAdjStart, s->nb)) return AVBD_ERR_HIP;
  uint32_t dtotal = 0;
  for (uint32_t b = 0; b < s->nb; b++) dtotal = std::max(dtotal, dstart[b] + dc[b]);
  if (read_dev(s, dlist, S.dAdjList, dtotal)) return AVBD_ERR_HIP;
  std::vector<ManifoldRec> mf;
  uint32_t npairs = 0;
  hipMemcpy(&npairs, S.counters + CTR_PAIRS, 4, hipMemcpyDeviceToHost);
  if (read_dev(s, mf, (const ManifoldRec*)S.manifolds, npairs)) return AVBD_ERR_HIP;
  double bytes = 0.0;
  std::vector<uint32_t> sStart(s->nb);
  {
    uint32_t run = 0;
    for (uint32_t b = 0; b < s->nb; b++) { sStart[b] = run; run += s->sAdjCount[b]; }
  }
  for (uint32_t b = 0; b < s->nb; b++) {
    if (!(s->mass[b] > 0.0f)) continue;
    // particles: posLin r/w 32 + inertLin 16 + shape 16 + CSR 8; rigid adds posAng r/w 32, inertAng 16, props 16
    bool rigid = !s->isParticle[b];
    bytes += rigid ? 136.0 : 72.0;
    for (uint32_t k = 0; k < s->sAdjCount[b]; k++) {
      uint32_t e = s->sAdjListHost[sStart[b] + k], kind = e >> kAdjShift;
      if (kind == FK_TET) bytes += 4 + 64 + 4 * 16;
      else if (kind == FK_JOINT) bytes += 4 + 224 + 2 * 32 + 2 * 16;
      else if (kind == FK_SPRING || kind == FK_PSPRING) bytes += 4 + 64 + 2 * 32;
      else if (kind == FK_MEMBRANE) bytes += 4 + 48 + 3 * 16;
      else if (kind == FK_BEND) bytes += 4 + 48 + 4 * 16;
    }
    for (uint32_t k = 0; k < dc[b]; k++) {
      uint32_t e = dlist[dstart[b] + k];
      if ((e >> kAdjShift) == FK_SOFT) {
        bytes += 4 + 112 + 4 * 16;  // record + the four participant positions
        continue;
      }
      uint32_t m = e & kAdjIndexMask;
      uint32_t nct = m < mf.size() ? mf[m].header.z : 0;
      bytes += 4 + 64 + 80.0 * nct + 2 * 32 + 2 * 32;
    }
  }
  // the `reps` sweeps are captured in one hipGraph and replayed (warm-up
  // replay, then one timed replay), so the events bracket device-side
  // dispatch only, as the frame's own graph does
  // a dynamic palette runs its first colors one launch each and the rest in
  // one ordered tail launch, as the frame does
  auto sweeps = [&]() {
    for (int r = 0; r < reps; r++) {
      if (s->dynamicColoring) {
        for (uint32_t c = 0; c < kDynamicColorLaunches; c++)
          launch_primal_color(s->stream, S, P, c, 0xFFFFFFFFu, 0u, s->lanes, kModeGeneral);
        launch_primal_tail(s->stream, S, P, kDynamicColorLaunches, s->lanes);
      } else {
        for (uint32_t c = 0; c < s->staticUsedColors; c++)
          launch_primal_color(s->stream, S, P, c, s->staticColorCounts[c], s->staticColorStarts[c], s->lanes,
                              s->colorMode[c]);
      }
    }
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t he = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
  if (he == hipSuccess) {
    sweeps();
    he = hipStreamEndCapture(s->stream, &graph);
  }
  if (he == hipSuccess) he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (he != hipSuccess) {
    hipGetLastError();
    return AVBD_ERR_HIP;
  }
  // a residual-terminated frame leaves its end flag set until the next
  // frame starts, and every sweep kernel returns at entry while it is set:
  // clear it for the timed sweeps and put it back afterwards
  uint32_t convDone = 0;
  he = hipMemcpy(&convDone, s->S.counters + CTR_CONV_DONE, 4, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemsetAsync(s->S.counters + CTR_CONV_DONE, 0, 4, s->stream);
  if (he != hipSuccess) {
    hipGraphExecDestroy(exec);
    return AVBD_ERR_HIP;
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipGraphLaunch(exec, s->stream);
  hipEventRecord(e0, s->stream);
  hipGraphLaunch(exec, s->stream);
  hipEventRecord(e1, s->stream);
  he = hipEventSynchronize(e1);
  if (he == hipSuccess) he = hipMemcpy(s->S.counters + CTR_CONV_DONE, &convDone, 4, hipMemcpyHostToDevice);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipGraphExecDestroy(exec);
  if (he != hipSuccess) return AVBD_ERR_HIP;
  int32_t launches = s->dynamicColoring ? (int32_t)kDynamicColorLaunches + 1 : (int32_t)s->staticUsedColors;
  if (avg_launch_us) *avg_launch_us = ms * 1000.0f / (float)(reps * launches);
  if (sweep_bytes) *sweep_bytes = bytes;
  if (launches_per_sweep) *launches_per_sweep = launches;
  return AVBD_OK;
}

// ---------------------------------------------------------------------------
// learned warm start (no reference counterpart)
// ---------------------------------------------------------------------------
namespace {
// dynamic particles in index order and their topological 1-ring: every body
// sharing a tet, membrane, bend or spring with the particle
avbd_status build_warm_state(avbd_solver* s) {
  std::vector<uint32_t> idx;
  for (uint32_t i = 0; i < s->nb; i++)
    if (s->isParticle[i] && s->mass[i] > 0.0f) idx.push_back(i);
  const uint32_t n = (uint32_t)idx.size();
  std::vector<uint32_t> rStart(n), rCount(n), rList;
  std::vector<uint32_t> nb;
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t b = idx[k];
    nb.clear();
    const uint32_t start = s->sAdjStartHost[b], cnt = s->sAdjCount[b];
    for (uint32_t q = 0; q < cnt; q++) {
      const uint32_t e = s->sAdjListHost[start + q], kind = e >> kAdjShift, id = e & kAdjIndexMask;
      if (kind == FK_TET) {
        const U4 t = s->tetIdsHost[id];
        nb.insert(nb.end(), {t.x, t.y, t.z, t.w});
      } else if (kind == FK_SPRING || kind == FK_PSPRING) {
        nb.push_back(s->springEndsHost[id].x);
        nb.push_back(s->springEndsHost[id].y);
      } else if (kind == FK_MEMBRANE) {
        const ClothSlot& c = s->clothSlotsHost[start + q];
        nb.push_back(c.u.x & ~kSnapBit);
        nb.push_back(c.u.y & ~kSnapBit);
      } else if (kind == FK_BEND) {
        const ClothSlot& c = s->clothSlotsHost[start + q];
        nb.push_back(c.u.x & ~kSnapBit);
        nb.push_back(c.u.y & ~kSnapBit);
        nb.push_back(c.u.z & ~kSnapBit);
      }
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    rStart[k] = (uint32_t)rList.size();
    for (uint32_t j : nb)
      if (j != b && j < s->nb) rList.push_back(j);
    rCount[k] = (uint32_t)rList.size() - rStart[k];
  }
  // dynamic rigid bodies and their joint partners
  std::vector<uint32_t> ridx;
  for (uint32_t i = 0; i < s->nb; i++)
    if (!s->isParticle[i] && s->mass[i] > 0.0f) ridx.push_back(i);
  const uint32_t nR = (uint32_t)ridx.size();
  std::vector<uint32_t> qStart(nR), qCount(nR), qList;
  for (uint32_t k = 0; k < nR; k++) {
    const uint32_t b = ridx[k];
    nb.clear();
    const uint32_t start = s->sAdjStartHost[b], cnt = s->sAdjCount[b];
    for (uint32_t q = 0; q < cnt; q++) {
      const uint32_t e = s->sAdjListHost[start + q];
      if ((e >> kAdjShift) != FK_JOINT) continue;
      const U4 h = s->jointHeaders[e & kAdjIndexMask];
      const uint32_t o = h.x == b ? h.y : h.x;
      if (o != kWorld && o != b && o < s->nb) nb.push_back(o);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    qStart[k] = (uint32_t)qList.size();
    qList.insert(qList.end(), nb.begin(), nb.end());
    qCount[k] = (uint32_t)qList.size() - qStart[k];
  }
  WarmState W{};
  W.n = n;
  W.nR = nR;
  hipError_t e;
  __bf16* xb = nullptr;
  __bf16* xr = nullptr;
  const size_t m = std::max<uint32_t>(n, 1);
  if ((e = dev_alloc(s, &W.idx, m)) || (e = dev_alloc(s, &W.ringStart, m)) || (e = dev_alloc(s, &W.ringCount, m)) ||
      (e = dev_alloc(s, &W.ringList, std::max<size_t>(rList.size(), 1))) || (e = dev_alloc(s, &W.guess, m)) ||
      (e = dev_alloc(s, &W.corr, s->nb)) || (e = dev_alloc(s, &W.corrPrev, s->nb)) || (e = dev_alloc(s, &xb, m * kMlpInPad)) ||
      (e = dev_alloc(s, &W.y, m * kMlpOut)) || (e = dev_alloc(s, &W.logX, m * kMlpIn)) ||
      (e = dev_alloc(s, &W.logT, m * kMlpOut)) || (e = dev_alloc(s, &W.res, 4)) ||
      (e = dev_alloc(s, &W.pred, s->nb)) || (e = dev_alloc(s, &W.pred2, s->nb)))
    return AVBD_ERR_ALLOC;
  const size_t mr = std::max<uint32_t>(nR, 1);
  if ((e = dev_alloc(s, &W.ridx, mr)) || (e = dev_alloc(s, &W.rRingStart, mr)) ||
      (e = dev_alloc(s, &W.rRingCount, mr)) || (e = dev_alloc(s, &W.rRingList, std::max<size_t>(qList.size(), 1))) ||
      (e = dev_alloc(s, &W.rGuessLin, mr)) || (e = dev_alloc(s, &W.rGuessAng, mr)) ||
      (e = dev_alloc(s, &W.corrAng, s->nb)) || (e = dev_alloc(s, &xr, mr * kMlpInPad)) ||
      (e = dev_alloc(s, &W.yR, mr * kMlpOutRigid)) || (e = dev_alloc(s, &W.logXR, mr * kMlpIn)) ||
      (e = dev_alloc(s, &W.logTR, mr * kMlpOutRigid)))
    return AVBD_ERR_ALLOC;
  W.x = xb;
  W.xR = xr;
  if (upload(W.ridx, ridx) || upload(W.rRingStart, qStart) || upload(W.rRingCount, qCount) ||
      upload(W.rRingList, qList))
    return AVBD_ERR_HIP;
  if (upload(W.idx, idx) || upload(W.ringStart, rStart) || upload(W.ringCount, rCount) || upload(W.ringList, rList))
    return AVBD_ERR_HIP;
  s->W = W;
  s->wsBuilt = true;
  return AVBD_OK;
}
}  // namespace

extern "C" avbd_status avbd_set_warm_start(avbd_solver* s, avbd_mlp* m, int32_t mode) {
  const int32_t flags = mode & 0xFF, passes = (mode >> AVBD_WS_SMOOTH_SHIFT) & 0xF;
  if (!s || mode < 0 || (mode >> (AVBD_WS_SMOOTH_SHIFT + 4)) != 0 ||
      flags > (AVBD_WS_LEARNED | AVBD_WS_LOG | AVBD_WS_FORCE))
    return AVBD_ERR_INVALID_ARGUMENT;
  if ((flags & AVBD_WS_LEARNED) && !m) return AVBD_ERR_INVALID_ARGUMENT;
  if ((flags & AVBD_WS_FORCE) && !(flags & AVBD_WS_LEARNED)) return AVBD_ERR_INVALID_ARGUMENT;
  if (flags == 0) mode = 0;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  if (mode && !s->wsBuilt) {
    st = build_warm_state(s);
    if (st != AVBD_OK) return st;
  }
  // a fresh attachment starts without correction history
  if (mode && !s->wsMode &&
      (hipMemset(s->W.corr, 0, 16 * (size_t)s->nb) != hipSuccess ||
       hipMemset(s->W.corrPrev, 0, 16 * (size_t)s->nb) != hipSuccess ||
       hipMemset(s->W.corrAng, 0, 16 * (size_t)s->nb) != hipSuccess))
    return AVBD_ERR_HIP;
  s->wsMode = mode;
  if (!mode) {
    s->mlp = s->mlpRigid = nullptr;
  } else if (m) {
    if (m->M.nOut == (uint32_t)kMlpOutRigid) s->mlpRigid = m;
    else s->mlp = m;
  }
  if (!m || m->M.nOut != (uint32_t)kMlpOutRigid) s->W.smoothPasses = (uint32_t)passes;
  invalidate_graphs(s);
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_warm_start_samples_rigid(avbd_solver* s, uint32_t capacity, uint32_t* n, float* x,
                                                         float* t, uint32_t* idx) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  if (!(s->wsMode & AVBD_WS_LOG)) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  *n = s->W.nR;
  const size_t c = std::min(capacity, s->W.nR);
  if (c == 0) return AVBD_OK;
  if ((x && hipMemcpy(x, s->W.logXR, c * kMlpIn * 4, hipMemcpyDeviceToHost)) ||
      (t && hipMemcpy(t, s->W.logTR, c * kMlpOutRigid * 4, hipMemcpyDeviceToHost)) ||
      (idx && hipMemcpy(idx, s->W.ridx, c * 4, hipMemcpyDeviceToHost)))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_get_warm_start_samples(avbd_solver* s, uint32_t capacity, uint32_t* n, float* x, float* t,
                                                   uint32_t* idx) {
  if (!s || !n) return AVBD_ERR_INVALID_ARGUMENT;
  if (!(s->wsMode & AVBD_WS_LOG)) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  *n = s->W.n;
  const size_t c = std::min(capacity, s->W.n);
  if (c == 0) return AVBD_OK;
  if ((x && hipMemcpy(x, s->W.logX, c * kMlpIn * 4, hipMemcpyDeviceToHost)) ||
      (t && hipMemcpy(t, s->W.logT, c * kMlpOut * 4, hipMemcpyDeviceToHost)) ||
      (idx && hipMemcpy(idx, s->W.idx, c * 4, hipMemcpyDeviceToHost)))
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_test_fail_completion(avbd_solver* s, int32_t kind, int64_t frame) {
  if (!s || kind < 0 || kind > AVBD_FAIL_DAT_INVALID_ANCHOR) return AVBD_ERR_INVALID_ARGUMENT;
  s->completionFailKind = kind;
  s->completionFailFrame = frame;
  return AVBD_OK;
}

// ---------------------------------------------------------------------------
// speculation snapshots (GPUSolver.swift:2948-3057): the temporal state of a
// rigid scene -- body poses / velocities / inertial targets, joint and
// spring multipliers, both manifold buffers, the pair map, the persistent
// colors, the counters -- copied device to device into one HBM block, so a
// rollout can fork and rewind without crossing PCIe
// ---------------------------------------------------------------------------
struct avbd_snapshot {
  const avbd_solver* owner = nullptr;
  void* block = nullptr;
  size_t bytes = 0;
  int64_t frameIndex = 0;
  avbd_stats stats{};
};

namespace {

// the snapshot's buffers, in a fixed order, with the manifold / soft roles
// of S (the solver's current roles, or the roles a restore will commit)
std::vector<std::pair<void*, size_t>> speculation_buffers(const avbd_solver* s, const DeviceState& S) {
  const size_t nb = s->nb;
  std::vector<std::pair<void*, size_t>> b = {
      {S.posLin, 16 * nb}, {S.posLo, 16 * nb}, {S.posAng, 16 * nb}, {S.initLin, 16 * nb}, {S.initAng, 16 * nb},
      {S.inertLin, 16 * nb}, {S.inertAng, 16 * nb}, {S.velLin, 16 * nb}, {S.velAng, 16 * nb},
      {S.prevVelLin, 16 * nb}, {S.posLinSnap, 16 * nb}, {S.posAngSnap, 16 * nb},
      {S.joints, sizeof(JointRec) * std::max<size_t>(1, s->nj)},
      {S.springs, sizeof(SpringRec) * std::max<size_t>(1, s->ns)},
      {S.manifolds, sizeof(ManifoldRec) * (size_t)s->maxPairs},
      {S.prevManifolds, sizeof(ManifoldRec) * (size_t)s->maxPairs},
      {S.mapKeyA, 4 * (size_t)s->mapCapacity}, {S.mapKeyB, 4 * (size_t)s->mapCapacity},
      {S.mapVal, 4 * (size_t)s->mapCapacity}, {S.colors, 4 * nb}, {S.colorsB, 4 * nb},
      {S.counters, 4 * (size_t)CTR_TOTAL}};
  if (s->W.corr) {  // the learned warm start's correction history
    b.push_back({s->W.corr, 16 * nb});
    b.push_back({s->W.corrPrev, 16 * nb});
    b.push_back({s->W.corrAng, 16 * nb});
  }
  return b;
}

size_t spec_align(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" avbd_status avbd_capture_speculation_snapshot(avbd_solver* s, avbd_snapshot** out) {
  if (!s || !out) return AVBD_ERR_INVALID_ARGUMENT;
  if (s->nTris || s->nt) return AVBD_ERR_UNSUPPORTED;  // rigid scenes only, as the reference
  if (*out && (*out)->owner != s) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  const auto bufs = speculation_buffers(s, s->S);
  size_t bytes = 0;
  for (const auto& b : bufs) bytes += spec_align(b.second);
  avbd_snapshot* snap = *out;
  bool fresh = false;
  if (!snap) {
    snap = new avbd_snapshot();
    snap->owner = s;
    fresh = true;
  } else if (!live_contains(snap)) {
    return AVBD_ERR_INVALID_ARGUMENT;
  }
  if (snap->bytes != bytes) {  // grows when a learned warm start was attached since
    if (snap->block) hipFree(snap->block);
    snap->block = nullptr;
    snap->bytes = 0;
    if (hipMalloc(&snap->block, bytes) != hipSuccess) {
      hipGetLastError();
      if (fresh) delete snap;
      return AVBD_ERR_ALLOC;
    }
    snap->bytes = bytes;
  }
  size_t off = 0;
  for (const auto& b : bufs) {
    if (hipMemcpyAsync((char*)snap->block + off, b.first, b.second, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
      return AVBD_ERR_HIP;
    off += spec_align(b.second);
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) return AVBD_ERR_HIP;
  snap->frameIndex = s->frameIndex;
  snap->stats = s->stats;
  if (fresh) live_add(snap, HandleKind::Snapshot);
  *out = snap;
  return AVBD_OK;
}

extern "C" avbd_status avbd_restore_speculation_snapshot(avbd_solver* s, const avbd_snapshot* snap) {
  if (!s || !snap || snap->owner != s) return AVBD_ERR_INVALID_ARGUMENT;
  if (s->nTris || s->nt) return AVBD_ERR_UNSUPPORTED;
  if (!live_contains(snap)) return AVBD_ERR_INVALID_ARGUMENT;
  avbd_status st = sync_all(s);
  if (st != AVBD_OK) return st;
  // manifold and soft buffers alternate roles every frame (the captured
  // graphs bake the roles by frame parity): the restore copies into the
  // roles of the snapshot's frame and commits those roles only after every
  // copy has completed, so a rejected or failed restore leaves the solver
  // as it was
  DeviceState R = s->S;
  if (((s->frameIndex - snap->frameIndex) & 1) != 0) {
    std::swap(R.manifolds, R.prevManifolds);
    std::swap(R.soft, R.prevSoft);
  }
  const auto bufs = speculation_buffers(s, R);
  size_t bytes = 0;
  for (const auto& b : bufs) bytes += spec_align(b.second);
  if (bytes != snap->bytes) return AVBD_ERR_INVALID_ARGUMENT;  // a warm start attached or detached since
  size_t off = 0;
  for (const auto& b : bufs) {
    if (hipMemcpyAsync(b.first, (const char*)snap->block + off, b.second, hipMemcpyDeviceToDevice, s->stream) !=
        hipSuccess)
      return AVBD_ERR_HIP;
    off += spec_align(b.second);
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) return AVBD_ERR_HIP;
  s->S.manifolds = R.manifolds;
  s->S.prevManifolds = R.prevManifolds;
  s->S.soft = R.soft;
  s->S.prevSoft = R.prevSoft;
  s->frameIndex = snap->frameIndex;
  s->stats = snap->stats;
  return AVBD_OK;
}

namespace {

void release_snapshot(avbd_snapshot* snap) {
  if (snap->block) hipFree(snap->block);
  delete snap;
}

}  // namespace

extern "C" void avbd_destroy_speculation_snapshot(avbd_snapshot* snap) {
  if (snap && live_remove(snap)) release_snapshot(snap);
}

namespace avbd {

void release_network(avbd_mlp* m);  // mlp_api.hip

void release_all_at_exit() {
  LiveRegistry& r = live_registry();
  std::vector<LiveHandle> h;
  {
    std::lock_guard<std::mutex> g(r.mu);
    h.swap(r.live);
    r.exiting = true;
  }
  // snapshots and solvers before the networks a solver may reference
  for (const LiveHandle& x : h)
    if (x.kind == HandleKind::Snapshot) release_snapshot((avbd_snapshot*)x.p);
  for (const LiveHandle& x : h)
    if (x.kind == HandleKind::Solver) release_solver((avbd_solver*)x.p);
  for (const LiveHandle& x : h)
    if (x.kind == HandleKind::Network) release_network((avbd_mlp*)x.p);
}

}  // namespace avbd

extern "C" avbd_status avbd_test_set_dat_query_radius(avbd_solver* s, float r) {
  if (!s || !(r == r)) return AVBD_ERR_INVALID_ARGUMENT;
  s->datQueryRadius = r > 0.0f ? r : s->datQueryRadiusDerived;
  invalidate_graphs(s);  // captured launches hold the old parameters
  return AVBD_OK;
}

extern "C" avbd_status avbd_test_deny_stage(avbd_solver* s, const char* stage, int64_t frame) {
  if (!s) return AVBD_ERR_INVALID_ARGUMENT;
  s->deniedStage = stage ? stage : "";
  s->deniedFrame = frame;
  return AVBD_OK;
}
