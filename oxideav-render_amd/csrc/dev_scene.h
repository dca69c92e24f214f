// Device-resident scene, path-state and queue layouts of the oxr path
// tracer (shared by the kernels and the host runtime).
#pragma once

#include <stdint.h>

#include "dev_tex.h"

namespace oxr {

// Per-material constants (PreparedMaterial + the untextured reads of
// TraceScene::material_oriented, oxideav-render src/prepare.rs:623-668,
// src/trace.rs:724-922), resolved and clamped on the host. 160 bytes.
struct alignas(16) DevMaterial {
  float base[4];
  float metallic, roughness, ior, specular;
  float emissive[3];
  uint32_t flags;  // MAT_DS | MAT_UNLIT
  float spec_color[3];
  float transmission;
  float thickness;
  float att_color[3];
  float att_dist;
  float clearcoat, cc_rough, sheen_rough;
  float sheen[3];
  float iridescence;
  float irid_ior, irid_thick, dt, dispersion;
  float dt_color[3];
  float aniso;      // KHR_materials_anisotropy strength (needs tangents)
  float aniso_rot;  // its rotation
  float pad[3];
};
static_assert(sizeof(DevMaterial) == 160, "DevMaterial is 160 bytes");

constexpr uint32_t MAT_DS = 1u;
constexpr uint32_t MAT_UNLIT = 2u;

// Punctual light (PreparedLight, src/prepare.rs:693-714) with the
// spot-cone constants of PreparedLight::sample precomputed. 64 bytes.
struct alignas(16) DevLight {
  uint32_t kind;
  float base[3];  // colour x intensity
  float pos[3];
  uint32_t has_range;
  float dir[3];
  float range;
  float spot_scale, spot_offset;
  uint32_t index;  // position in the scene's light list, zero-power lights included (the shadow-ray coin id)
  uint32_t pad1;
};
static_assert(sizeof(DevLight) == 64, "DevLight is 64 bytes");

struct Node4Q;
struct TriRec;

struct DevScene {
  const Node4Q* nodes;
  const TriRec* tris;
  const float* pos;         // 9 per global triangle
  const float* nrm;         // 9 per global triangle (or null)
  const uint32_t* tri_info; // material id | TI_HAS_NORMALS
  const DevMaterial* mats;
  const DevLight* lights;
  const uint32_t* l_tris;
  const float* l_pmf;
  const float* l_cdf;
  const uint32_t* l_lookup;
  uint32_t num_tris, num_mats, num_lights, num_emissive;
  uint32_t has_blend;
  uint32_t has_tex;        // textures, vertex colours or tangents: the TEX shading variant
  uint32_t has_alpha_tex;  // MASK / BLEND alpha varies over triangles: the ALPHA traversal variant
  uint32_t has_ext;        // shading variant: 0 no extension lobe, 1 extension lobes but no iridescence, 2 iridescence
  DevTextures tex;
};

constexpr uint32_t TI_HAS_NORMALS = 0x80000000u;
constexpr uint32_t TI_MAT_MASK = 0x7fffffffu;

// Equirectangular environment map (EnvironmentMap, src/pathtrace.rs:
// 1557-1741): texels (row 0 = +Y), the luminance x sin(theta) CDFs and the
// per-texel mass. img == nullptr: no map (the constant ambient sky).
struct DevEnv {
  const float4* img;
  const float* marginal;     // h + 1
  const float* conditional;  // h * (w + 1)
  const float* mass;         // w * h
  uint32_t w, h;
  float intensity;
  uint32_t uniform;
};

// Camera (src/camera.rs:61-91) + the integrator constants.
struct DevFrame {
  float eye[3], forward[3], side[3], up[3];
  float half_w, half_h;
  uint32_t projection;
  uint32_t width, height;
  uint32_t max_bounces, rr_start, seed, strategy;
  float clamp, ambient;
  uint32_t nslots;  // NEE slots per path: punctual lights + 1 area
  float p_env;      // Integrator::p_env: 0, 1/2 or 1
  DevEnv env;
  float spread;     // 2 half_h / height (src/pathtrace.rs:2537-2540): one pixel row on the image plane
  // Render window: accumulator pixel i is frame pixel (win_x0 + i % win_w,
  // win_y0 + i / win_w); the whole frame by default.
  uint32_t win_x0, win_y0, win_w;
};

// Path flags (meta.z >> 24).
constexpr uint32_t PF_UNCOVERED = 1u;
constexpr uint32_t PF_PENDING = 2u;
constexpr uint32_t PF_DONE = 4u;

constexpr uint32_t HIT_MISS = 0xffffffffu;
constexpr uint32_t HIT_BACK = 0x80000000u;
constexpr uint32_t NO_SKIP = 0xffffffffu;

// One render job of a wavefront batch: its scene, camera/options and
// accumulator. A batch mixes paths of several jobs (continuous batching);
// every path carries its job index.
struct DevJob {
  DevScene sc;
  DevFrame fr;
  double* accum;
  double* sumsq;
};

// A batch segment: paths [path0, path0 + npx * S) of job `job` cover
// pixels [px0, px0 + npx) x samples [first, first + S), pixel-major
// within a sample.
struct Seg {
  uint32_t job, path0, npx, px0, first, S;
};

constexpr int MAX_SEGS = 64;
struct SegTable {
  uint32_t n;
  Seg s[MAX_SEGS];
};

// Device control block: queue counters, work counters, error, stats.
// A counter alone in its 256-byte block. Device-scope atomics on one
// address (and on one line) serialise device-wide, about 11 ns apiece on
// MI355X; counters that one kernel bumps concurrently (the two append
// queues, the traversal segments) therefore live on lines of their own.
struct alignas(256) PadCounter {
  uint32_t v;
  uint32_t pad_[63];
};
constexpr uint32_t PAD_COUNTER_WORDS = sizeof(PadCounter) / sizeof(uint32_t);
struct alignas(256) PadCounter64 {
  unsigned long long v;
  unsigned long long pad_[31];
};

struct Ctrl {
  PadCounter qcount[2];
  PadCounter scount;
  PadCounter work[16];  // traversal work counters: closest [0, 8), any-hit [8, 16), one per queue segment
  PadCounter64 trav[4];  // counting refines: node and triangle fetches of closest-hit [0, 1] and occlusion [2, 3] rays
  uint32_t error;
  unsigned long long rays_closest;
  unsigned long long rays_shadow;
};

constexpr uint32_t ERR_STACK = 1u;

}  // namespace oxr
