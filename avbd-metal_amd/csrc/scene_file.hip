// scene_file.hip -- versioned on-disk form of the scene wire format
// (avbd_scene_desc): fixtures and large scenes load without rebuilding them.
//
// Layout (little-endian): the 8-byte magic "AVBDSCEN", the file version,
// the descriptor ABI version it was written from, the 15 settings fields
// (4 bytes each), the nine counts (bodies, colliders, joints, springs,
// tets, triangles, exclusions, hull vertices, spinners), the number of arrays, then
// per present array: its id, element size, element count and data padded
// to 8 bytes.  Absent (NULL) arrays are not written and read back as NULL.
// The reader checks every id, element size and count against the counts,
// so a file describes exactly one valid descriptor layout.
#include <stdio.h>
#include <string.h>

#include <cstdint>
#include <vector>

#include "../../include/avbd.h"

namespace {

constexpr char kMagic[8] = {'A', 'V', 'B', 'D', 'S', 'C', 'E', 'N'};
constexpr uint32_t kFileVersion = AVBD_SCENE_FILE_VERSION;

enum Dim : uint32_t { kBodies, kColliders, kJoints, kSprings, kTets, kTris, kExclusions, kHullVerts, kSpinners, kDims };

struct Field {
  uint32_t id;
  size_t offset;   // of the pointer member in avbd_scene_desc
  uint32_t elem;   // bytes per element
  Dim dim;
  uint32_t width;  // elements per item of the dimension
};

#define F(id, member, type, dim, width) {id, offsetof(avbd_scene_desc, member), sizeof(type), dim, width}
const Field kFields[] = {
    F(1, body_size, float, kBodies, 3),
    F(2, body_density, float, kBodies, 1),
    F(3, body_friction, float, kBodies, 1),
    F(4, body_dynamic_friction, float, kBodies, 1),
    F(5, body_position, float, kBodies, 3),
    F(6, body_rotation, float, kBodies, 4),
    F(7, body_velocity, float, kBodies, 3),
    F(8, body_shape, uint32_t, kBodies, 1),
    F(9, body_is_particle, uint8_t, kBodies, 1),
    F(10, body_mass, float, kBodies, 1),
    F(11, body_inertia, float, kBodies, 3),
    F(12, body_gravity_scale, float, kBodies, 1),
    F(20, collider_body, uint32_t, kColliders, 1),
    F(21, collider_size, float, kColliders, 3),
    F(22, collider_friction, float, kColliders, 1),
    F(23, collider_dynamic_friction, float, kColliders, 1),
    F(24, collider_local_position, float, kColliders, 3),
    F(25, collider_local_rotation, float, kColliders, 4),
    F(26, collider_shape, uint32_t, kColliders, 1),
    F(27, collider_group, uint32_t, kColliders, 1),
    F(28, collider_shared, uint8_t, kColliders, 1),
    F(29, collider_enabled, uint8_t, kColliders, 1),
    F(30, collider_world_round_anchor, uint8_t, kColliders, 1),
    F(31, collider_hull_count, uint32_t, kColliders, 1),
    F(32, hull_vertices, float, kHullVerts, 3),
    F(40, joint_body_a, int32_t, kJoints, 1),
    F(41, joint_body_b, int32_t, kJoints, 1),
    F(42, joint_ra, float, kJoints, 3),
    F(43, joint_rb, float, kJoints, 3),
    F(44, joint_stiffness_lin, float, kJoints, 1),
    F(45, joint_stiffness_ang, float, kJoints, 1),
    F(46, joint_fracture, float, kJoints, 1),
    F(47, joint_fracture_linear, uint8_t, kJoints, 1),
    F(48, joint_hinge_axis, float, kJoints, 3),
    F(49, joint_motor, float, kJoints, 6),
    F(50, joint_limits, float, kJoints, 2),
    F(51, joint_motor_rate, float, kJoints, 1),
    F(60, spring_body_a, uint32_t, kSprings, 1),
    F(61, spring_body_b, uint32_t, kSprings, 1),
    F(62, spring_ra, float, kSprings, 3),
    F(63, spring_rb, float, kSprings, 3),
    F(64, spring_stiffness, float, kSprings, 1),
    F(65, spring_rest, float, kSprings, 1),
    F(66, spring_hard, uint8_t, kSprings, 1),
    F(70, tet_ids, uint32_t, kTets, 4),
    F(71, tet_mu, float, kTets, 1),
    F(72, tet_lambda, float, kTets, 1),
    F(80, tri_ids, uint32_t, kTris, 3),
    F(81, tri_mu, float, kTris, 1),
    F(82, tri_lambda, float, kTris, 1),
    F(83, tri_bend, float, kTris, 1),
    F(90, exclusion_pairs, uint32_t, kExclusions, 2),
    F(100, spinner_body, uint32_t, kSpinners, 1),
    F(101, spinner_axis_omega, float, kSpinners, 4),
};
#undef F

const void*& member(avbd_scene_desc* d, const Field& f) { return *(const void**)((char*)d + f.offset); }
const void* member(const avbd_scene_desc* d, const Field& f) { return *(const void* const*)((const char*)d + f.offset); }

// the fifteen settings fields in declaration order, all four bytes wide
static_assert(sizeof(avbd_settings) == 15 * 4, "settings layout");

void counts_of(const avbd_scene_desc* d, uint64_t n[kDims]) {
  n[kBodies] = d->num_bodies;
  n[kColliders] = d->num_colliders;
  n[kJoints] = d->num_joints;
  n[kSprings] = d->num_springs;
  n[kTets] = d->num_tets;
  n[kTris] = d->num_tris;
  n[kExclusions] = d->num_exclusions;
  uint64_t hv = 0;
  if (d->collider_hull_count)
    for (uint32_t c = 0; c < d->num_colliders; c++) hv += d->collider_hull_count[c];
  n[kHullVerts] = hv;
  n[kSpinners] = d->num_spinners;
}

// what avbd_scene_read hands out: the descriptor, then its arrays' storage
struct Owned {
  avbd_scene_desc desc;
  uint64_t* block;  // 8-byte aligned storage of every array
};

}  // namespace

extern "C" avbd_status avbd_scene_write(const avbd_scene_desc* d, const char* path) {
  if (!d || !path || d->abi_version != AVBD_ABI_VERSION) return AVBD_ERR_INVALID_ARGUMENT;
  FILE* f = fopen(path, "wb");
  if (!f) return AVBD_ERR_INVALID_ARGUMENT;
  uint64_t n[kDims];
  counts_of(d, n);
  bool ok = fwrite(kMagic, 1, 8, f) == 8;
  const uint32_t head[2] = {kFileVersion, AVBD_ABI_VERSION};
  ok = ok && fwrite(head, 4, 2, f) == 2;
  ok = ok && fwrite(&d->settings, 4, 15, f) == 15;
  uint32_t cnt[kDims];
  for (int k = 0; k < kDims; k++) cnt[k] = (uint32_t)n[k];
  ok = ok && fwrite(cnt, 4, kDims, f) == (size_t)kDims;
  uint32_t present = 0;
  for (const Field& fd : kFields)
    if (member(d, fd) && n[fd.dim] > 0) present++;
  ok = ok && fwrite(&present, 4, 1, f) == 1;
  static const uint8_t pad[8] = {0};
  for (const Field& fd : kFields) {
    const void* p = member(d, fd);
    if (!p || n[fd.dim] == 0 || !ok) continue;
    const uint64_t count = n[fd.dim] * fd.width, bytes = count * fd.elem;
    const uint32_t rec[2] = {fd.id, fd.elem};
    ok = fwrite(rec, 4, 2, f) == 2 && fwrite(&count, 8, 1, f) == 1 && fwrite(p, 1, bytes, f) == bytes &&
         fwrite(pad, 1, (8 - bytes % 8) % 8, f) == (8 - bytes % 8) % 8;
  }
  ok = fclose(f) == 0 && ok;
  return ok ? AVBD_OK : AVBD_ERR_INVALID_ARGUMENT;
}

extern "C" avbd_status avbd_scene_read(const char* path, avbd_scene_desc** out) {
  if (!path || !out) return AVBD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return AVBD_ERR_INVALID_ARGUMENT;
  uint64_t fileBytes = 0;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long e = ftell(f);
    fileBytes = e > 0 ? (uint64_t)e : 0;
  }
  rewind(f);
  auto fail = [&]() {
    fclose(f);
    return AVBD_ERR_INVALID_ARGUMENT;
  };
  char magic[8];
  uint32_t head[2];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kMagic, 8) != 0 || fread(head, 4, 2, f) != 2) return fail();
  if (head[0] != kFileVersion) return fail();
  Owned* o = new Owned();
  memset(o, 0, sizeof(*o));
  avbd_scene_desc& d = o->desc;
  d.abi_version = AVBD_ABI_VERSION;
  uint32_t cnt[kDims], present = 0;
  if (fread(&d.settings, 4, 15, f) != 15 || fread(cnt, 4, kDims, f) != (size_t)kDims || fread(&present, 4, 1, f) != 1) {
    delete o;
    return fail();
  }
  d.num_bodies = cnt[kBodies];
  d.num_colliders = cnt[kColliders];
  d.num_joints = cnt[kJoints];
  d.num_springs = cnt[kSprings];
  d.num_tets = cnt[kTets];
  d.num_tris = cnt[kTris];
  d.num_exclusions = cnt[kExclusions];
  d.num_spinners = cnt[kSpinners];
  // read every record, then check the hull vertex count against the counts
  struct Rec {
    const Field* fd;
    size_t at;  // word offset in the block
  };
  std::vector<Rec> recs;
  std::vector<uint64_t> block;
  bool seen[256] = {false};
  for (uint32_t r = 0; r < present; r++) {
    uint32_t rec[2];
    uint64_t count = 0;
    if (fread(rec, 4, 2, f) != 2 || fread(&count, 8, 1, f) != 1) { delete o; return fail(); }
    const Field* fd = nullptr;
    for (const Field& x : kFields)
      if (x.id == rec[0]) fd = &x;
    if (!fd || rec[1] != fd->elem || rec[0] >= 256 || seen[rec[0]] ||
        count != (uint64_t)cnt[fd->dim] * fd->width || count == 0) {
      delete o;
      return fail();
    }
    seen[rec[0]] = true;
    const uint64_t bytes = count * fd->elem, words = (bytes + 7) / 8;
    if (words * 8 > fileBytes) { delete o; return fail(); }  // longer than the file itself
    const size_t at = block.size();
    block.resize(at + words, 0);
    if (fread(block.data() + at, 1, words * 8, f) != words * 8) { delete o; return fail(); }
    recs.push_back({fd, at});
  }
  fclose(f);
  o->block = new uint64_t[block.size() ? block.size() : 1];
  if (!block.empty()) memcpy(o->block, block.data(), 8 * block.size());
  for (const Rec& r : recs) member(&d, *r.fd) = o->block + r.at;
  uint64_t n[kDims];
  counts_of(&d, n);
  if (n[kHullVerts] != cnt[kHullVerts] || (cnt[kHullVerts] > 0 && !d.hull_vertices)) {
    delete[] o->block;
    delete o;
    return AVBD_ERR_INVALID_ARGUMENT;
  }
  static_assert(offsetof(Owned, desc) == 0, "the descriptor opens the owned record");
  *out = &o->desc;
  return AVBD_OK;
}

extern "C" void avbd_scene_free(avbd_scene_desc* d) {
  if (!d) return;
  Owned* o = reinterpret_cast<Owned*>(d);
  delete[] o->block;
  delete o;
}
