// state.hpp -- the device-side view of one solver: every buffer pointer the
// kernels touch, passed by value as a kernel argument (scalar kernarg loads).
#pragma once
#include "common.hpp"

namespace avbd {

struct alignas(8) U2 {
  uint32_t x, y;
};

struct DeviceState {
  // bodies (one F4 per body, each field its own array)
  F4 *posLin, *posAng, *initLin, *initAng, *inertLin, *inertAng;
  // low part of the current linear position: the iterate is posLin + posLo
  // exactly (posLin holds it rounded to f32).  Zeroed at the frame start;
  // the primal carries the sum of its corrections here, so the elastic and
  // inertial terms see differences finer than one f32 ulp of the position
  F4* posLo;
  F4 *velLin, *velAng, *prevVelLin;
  F4 *props;          // xyz principal moment, w friction
  F4 *shape;          // xyz size, w = +radius (rigid) / -radius (particle)
  float* gravityScale;
  uint32_t* clothGroup;
  // colliders
  uint32_t *colliderOwner, *colliderShapeType, *colliderGroup, *colliderShared;
  F4 *colliderShape, *colliderLocalPos, *colliderLocalRot, *colliderFriction;
  U2* colliderHullRange;      // convex hulls: (first vertex, count), count 0 = primitive
  F4* hullVerts;              // hull vertices, collider-local
  float* torusTrig;           // cos, sin of the 32 torus spine sample angles (torus colliders only)
  F4* spinVel;                // kinematic spinners: per body axis * omega, w = 1 if spinning; null if none
  // broad phase
  uint32_t *hashedIdx, *globalIdx;
  uint32_t* hashedRigid;      // hashed colliders whose owner is not a particle
  uint32_t *cellCount, *cellStart, *cellBodies, *cellRigid;
  // per hashed collider (index = position in hashedIdx): this frame's center
  // and |radius|, and static metadata (owner, group, shared | dynamic << 1,
  // soft group of the owner); both re-gathered into cell order each frame so
  // a candidate test is one level of contiguous loads
  F4* hashCenter;
  U4* hashMeta;
  uint32_t* colliderToHashed;
  F4* cellPos;
  U4* cellMeta;
  U2* bodyCellSlot;
  // candidate search grid (cells of 2 x the largest hashed radius, same
  // hash): hashed-collider indices per bucket, unordered
  uint32_t *fineCount, *fineStart, *fineBodies;
  U2* fineSlot;
  uint32_t* largeCells;       // buckets of more than 32 colliders (sorted by a workgroup each)
  uint32_t *pairCount, *pairStart;
  // pair producers by path: bpLane one lane each; bpWave one wave each,
  // its static part (built at create) followed by the lane producers
  // deferred this frame, bpWaveCount[0] entries in all
  uint32_t *bpLane, *bpWave, *bpWaveCount;
  U2* pairs;
  U2* exclusions;
  // contacts
  ManifoldRec *manifolds, *prevManifolds;
  uint32_t *mapKeyA, *mapKeyB, *mapVal;
  // elements
  TetRec* tets;
  JointRec* joints;
  SpringRec* springs;
  // adjacency: static topology CSR (built once) + per-frame contact CSR
  uint32_t *sAdjStart, *sAdjCount, *sAdjList;
  TetSlot* tetSlots;          // parallel to sAdjList (tet entries populated)
  uint32_t *dAdjCount, *dAdjStart, *dAdjCursor, *dAdjList;
  // colors
  uint32_t *colors, *colorsB, *colorList, *colorStart, *bodySlot;
  U4* colorWork;              // static palette, color order: (body, static CSR start, count, 0)
  // start-of-iteration pose snapshot: contact partners that share the
  // body's color are read from here, so a same-color contact is a Jacobi
  // coupling with a schedule-independent result
  F4 *posLinSnap, *posAngSnap;
  // cloth surfaces
  ClothSlot* clothSlots;      // parallel to sAdjList (membrane / bend / particle-spring entries)
  U4* tris;                   // collision triangles (vertex ids in xyz)
  U2* edges;                  // unique surface edges (ids ascending)
  uint32_t *nbrStart, *nbrCount, *nbrList;  // sorted topological 1-ring of every body
  uint32_t* particleIdx;      // surface particles, ascending
  uint32_t* selfCollision;    // per body: its surface self-collides
  // element grid: counts / starts / cursors per (bucket, triangles | edges), interleaved
  uint32_t *elemCellCount, *elemCellStart, *elemCellCursor, *elemCells;
  F4* elemBox;                // per element (triangles, then edges): padded box lo, hi; lo.w = surface group bits (0: mixed)
  F4* datAnchor;              // Planar-DAT anchor pose of the color passes (accepted predictor pose)
  uint32_t* datT;             // per body: step fraction of the running pass (f32 bits, 1.0 between passes)
  uint32_t *datCount, *datStart;  // per producer: surface particles, then edges (+1 total)
  U4* datPairs;               // (kind, owner, other, 0), producer-major, other ascending
  uint32_t *softFlag, *softStart;  // emission counts / offsets (RT per triangle, DAT per pair)
  SoftRec *soft, *prevSoft;
  uint32_t *softMapKeyA, *softMapKeyB, *softMapVal;
  F4* velScratch;             // unsmoothed velocities for the 1-ring viscosity gather
  // isotropic OGC limiter
  U4* triAdj;                 // per collision triangle: its edge neighbours (ascending edge order, 0xFFFFFFFF: none)
  uint32_t *nbr2Start, *nbr2Count, *nbr2List;  // sorted topological 2-ring of every body (itself excluded)
  uint32_t *vEdgeStart, *vEdgeCount, *vEdgeList;  // incident edges of every body, in edge order
  uint32_t* clothVert;        // per body: 1 for a vertex of an authored triangle
  U4 *vtTrack, *vtPrev;       // per body: tracked triangles of this frame / of the previous one
  // per element-grid slot (bucket triangles | edges) this frame: the least
  // and greatest solid key of its entries (a closed surface's group when the
  // element lies on it and no authored triangle touches it, else all ones)
  uint32_t *isoSlotLo, *isoSlotHi;
  U4 *eeTrack, *eePrev;       // per contact edge: tracked edges of this frame / of the previous one
  uint32_t* ogcBound;         // per body: squared clearance bound (f32 bits, atomic min)
  F4* ogcPrev;                // per body: the pose the bounds hold around
  uint32_t* isoFrame;         // frames stepped so far (the trackers' reseed stagger)
  // misc
  uint32_t* colorFlags;       // per-pass changed flags of the dynamic coloring
  uint32_t* counters;
  uint32_t* scanScratch;      // block sums for the scan
  double* residualPartial;    // per-block residual partial sums (|g|^2, |m g_grav|^2)
  uint32_t dAdjCapacity;
};

}  // namespace avbd
