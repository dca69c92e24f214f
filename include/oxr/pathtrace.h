/*
 * oxr — MI355X (gfx950) path tracer behind oxideav-render's PathTrace
 * backend. C ABI of liboxr.so.
 *
 * Every entry point returns OXR_OK (0) or a negative status; the text of
 * the last failure on the calling thread is oxr_last_error(). Inputs are
 * copied (or consumed) during the call; outputs go into caller buffers.
 * A handle may move between threads but must not be used by two threads
 * at once (the reference's `Renderer: Send` + `&mut self` contract,
 * oxideav-render src/lib.rs:136-167).
 *
 * Reference interfaces replaced (oxideav-render 0.0.6):
 *   oxr_pt_create / oxr_pt_destroy  -> PathTracer::new / Drop
 *                                      (src/pathtrace.rs:2420-2433)
 *   oxr_pt_upload_scene             -> the scene half of PathTracer::sync:
 *                                      TraceScene::new + EmissiveLights::build
 *                                      (src/pathtrace.rs:2484-2497,
 *                                       src/trace.rs:320-351,
 *                                       src/pathtrace.rs:1479-1543)
 *   oxr_pt_set_camera               -> Camera::resolve result held in State
 *                                      (src/pathtrace.rs:2488, :2499)
 *   oxr_pt_set_options              -> the options half of PathTracer::sync
 *                                      and its radiance_key reset rule
 *                                      (src/pathtrace.rs:2474-2508, :2402-2418)
 *   oxr_pt_reset                    -> PathTracer::reset (src/pathtrace.rs:2463)
 *   oxr_pt_refine                   -> PathTracer::refine (src/pathtrace.rs:2531-2591)
 *   oxr_pt_samples                  -> PathTracer::samples (src/pathtrace.rs:2511)
 *   oxr_pt_read_accum               -> the accumulator `accum: Vec<[f64;4]>`
 *                                      (src/pathtrace.rs:2377)
 *   oxr_pt_resolve_hdr / _rgba8     -> PathTracer::hdr / image
 *                                      (src/pathtrace.rs:2687-2734)
 *   oxr_sobol_owen_4d               -> sobol_owen_4d (src/pathtrace.rs:388-398)
 *   oxr_trace_closest               -> TraceScene::closest_hit_filtered
 *                                      (src/trace.rs:405-422)
 *   oxr_trace_any                   -> TraceScene::occluded_filtered
 *                                      (src/trace.rs:432-450)
 *   oxr_bsdf_eval / oxr_bsdf_sample -> Bsdf::new + Bsdf::eval / Bsdf::sample
 *                                      (src/pathtrace.rs:853-1279)
 */
#ifndef OXR_PATHTRACE_H
#define OXR_PATHTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------- */
#define OXR_OK 0
#define OXR_ERR_INVALID (-1)   /* bad argument (null pointer, size, range)  */
#define OXR_ERR_DEVICE (-2)    /* HIP runtime failure (Error::Backend)      */
#define OXR_ERR_STATE (-3)     /* call out of order (e.g. trace before scene)*/
#define OXR_ERR_MEMORY (-4)    /* device allocation failed / over budget    */

/* ---- scene data contract (a prepared, world-space triangle soup) ---- */

/* Alpha modes (oxideav_mesh3d::AlphaMode). */
#define OXR_ALPHA_OPAQUE 0u
#define OXR_ALPHA_MASK 1u
#define OXR_ALPHA_BLEND 2u

/* Material flags. */
#define OXR_MAT_DOUBLE_SIDED 1u
#define OXR_MAT_UNLIT 2u

/* One glTF metallic-roughness material with its KHR extension scalars,
 * already resolved to the values TraceScene::material_oriented reads when
 * no texture is bound (src/trace.rs:700-923; PreparedMaterial,
 * src/prepare.rs:433-476). The library applies the same finite/clamp
 * rules as material_oriented at upload. */
typedef struct oxr_material {
  float base_color[4];          /* factor, linear RGBA                      */
  float metallic;
  float roughness;
  float emissive[3];            /* emissiveFactor x emissive_strength       */
  uint32_t flags;               /* OXR_MAT_*                                */
  uint32_t alpha_mode;          /* OXR_ALPHA_*                              */
  float alpha_cutoff;
  float ior;                    /* KHR_materials_ior (1.5 default)          */
  float specular;               /* KHR_materials_specular factor (1)        */
  float specular_color[3];      /* (1,1,1)                                  */
  float transmission;           /* KHR_materials_transmission (0)           */
  float thickness;              /* KHR_materials_volume (0 = thin)          */
  float attenuation_color[3];   /* (1,1,1)                                  */
  float attenuation_distance;   /* +inf = none                              */
  float clearcoat;              /* KHR_materials_clearcoat (0)              */
  float clearcoat_roughness;
  float sheen_color[3];         /* KHR_materials_sheen (0,0,0)              */
  float sheen_roughness;
  float anisotropy;             /* KHR_materials_anisotropy strength (0)    */
  float anisotropy_rotation;
  float iridescence;            /* KHR_materials_iridescence factor (0)     */
  float iridescence_ior;        /* 1.3                                      */
  float iridescence_thickness;  /* thicknessMax without a texture (400)    */
  float diffuse_transmission;   /* KHR_materials_diffuse_transmission (0)   */
  float diffuse_transmission_color[3];
  float dispersion;             /* KHR_materials_dispersion (0)             */
} oxr_material;

/* Punctual light kinds (KHR_lights_punctual; src/prepare.rs:680-720). */
#define OXR_LIGHT_DIRECTIONAL 0u
#define OXR_LIGHT_POINT 1u
#define OXR_LIGHT_SPOT 2u

typedef struct oxr_light {
  uint32_t kind;
  float color[3];
  float intensity;              /* lux (directional) or candela             */
  float position[3];
  float direction[3];           /* unit direction the light travels         */
  float range;                  /* <= 0 or non-finite: infinite             */
  float inner_cone_angle;       /* radians                                  */
  float outer_cone_angle;       /* radians                                  */
} oxr_light;

/* Textures: a decoded image plus the glTF sampler of the texture that
 * references it (PreparedTexture, src/texture.rs:384-470). The library
 * builds the linear-light mip chains of src/texture.rs:323-381 at upload
 * (an sRGB chain for colour slots, a linear chain for data slots). */
#define OXR_TEXFMT_RGBA8 0u     /* 4 bytes per texel                        */
#define OXR_TEXFMT_RGBA32F 1u   /* 4 floats per texel, taken as linear      */
#define OXR_FILTER_NEAREST 0u   /* magnification filters                    */
#define OXR_FILTER_LINEAR 1u
#define OXR_MIN_NEAREST 0u      /* minification filters (glTF)              */
#define OXR_MIN_LINEAR 1u
#define OXR_MIN_NEAREST_MIP_NEAREST 2u
#define OXR_MIN_LINEAR_MIP_NEAREST 3u
#define OXR_MIN_NEAREST_MIP_LINEAR 4u
#define OXR_MIN_LINEAR_MIP_LINEAR 5u
#define OXR_WRAP_REPEAT 0u
#define OXR_WRAP_CLAMP_TO_EDGE 1u
#define OXR_WRAP_MIRRORED_REPEAT 2u

typedef struct oxr_texture {
  uint32_t width, height;
  uint32_t format;              /* OXR_TEXFMT_*                              */
  const void* pixels;           /* width x height RGBA, row-major, row 0 at
                                   v = 0                                     */
  uint32_t mag_filter;          /* OXR_FILTER_* (the effective filter: glTF
                                   leaves it open; the default is linear)    */
  uint32_t min_filter;          /* OXR_MIN_* (default linear-mip-linear)     */
  uint32_t wrap_s, wrap_t;      /* OXR_WRAP_*                                */
} oxr_texture;

/* A material's reference to a texture (TextureBinding, src/prepare.rs:
 * 386-431): the texture, its TEXCOORD set and KHR_texture_transform.
 * The transform maps uv to (T R S) uv with T = offset, R = rotation by
 * `rotation` radians as in KHR_texture_transform, S = scale. */
typedef struct oxr_texture_binding {
  int32_t texture;              /* index into oxr_scene_desc.textures; -1 =
                                   no texture                                */
  uint32_t uv_set;              /* 0 or 1                                    */
  uint32_t has_transform;
  float offset[2];
  float rotation;
  float scale[2];
} oxr_texture_binding;

/* The texture slots TraceScene::material_oriented samples (src/trace.rs:
 * 717-895), one entry per material. Colour slots are sampled in sRGB, the
 * others as linear data; the comment names the channel that is used. The
 * occlusion texture has no slot: the path tracer ignores it
 * (src/pathtrace.rs:241-243). */
typedef struct oxr_material_textures {
  oxr_texture_binding base_color;         /* sRGB RGBA x base colour     */
  oxr_texture_binding metallic_roughness; /* G x roughness, B x metallic */
  oxr_texture_binding normal;             /* tangent-space normal map    */
  oxr_texture_binding emissive;           /* sRGB RGB x emissive         */
  oxr_texture_binding specular;           /* A x specular factor         */
  oxr_texture_binding specular_color;     /* sRGB RGB x specular colour  */
  oxr_texture_binding transmission;       /* R x transmission            */
  oxr_texture_binding thickness;          /* G x thickness               */
  oxr_texture_binding clearcoat;          /* R x clearcoat               */
  oxr_texture_binding clearcoat_roughness;/* G x clearcoat roughness     */
  oxr_texture_binding clearcoat_normal;   /* clearcoat normal map        */
  oxr_texture_binding sheen_color;        /* sRGB RGB x sheen colour     */
  oxr_texture_binding sheen_roughness;    /* A x sheen roughness         */
  oxr_texture_binding anisotropy;         /* RG direction, B x strength  */
  oxr_texture_binding iridescence;        /* R x iridescence             */
  oxr_texture_binding iridescence_thickness; /* G: thickness min..max    */
  oxr_texture_binding diffuse_transmission;  /* A x factor               */
  oxr_texture_binding diffuse_transmission_color; /* sRGB RGB x colour   */
  float normal_scale;           /* normalTexture.scale (1)                   */
  float clearcoat_normal_scale; /* clearcoatNormalTexture.scale (1)          */
  float iridescence_thickness_min; /* iridescenceThicknessMinimum (100); the
                                      maximum is oxr_material's value        */
  float pad;
} oxr_material_textures;

/* Triangle flags. */
#define OXR_TRI_HAS_NORMALS 1u
#define OXR_TRI_HAS_UV0 2u      /* TEXCOORD_0 present for the triangle's item */
#define OXR_TRI_HAS_UV1 4u
#define OXR_TRI_HAS_TANGENTS 8u
#define OXR_TRI_HAS_COLORS 16u

typedef struct oxr_scene_desc {
  uint32_t num_tris;
  const float* positions;       /* num_tris x 9: v0,v1,v2 (CCW front)      */
  const float* normals;         /* num_tris x 9 vertex normals, or NULL     */
  const uint32_t* tri_flags;    /* num_tris OXR_TRI_*, or NULL (each
                                   OXR_TRI_HAS_* set where its array below
                                   is given)                                 */
  const uint32_t* material_ids; /* num_tris, index into materials          */
  uint32_t num_materials;
  const oxr_material* materials;
  uint32_t num_lights;
  const oxr_light* lights;      /* punctual lights (PreparedScene::lights)  */
  /* Textured materials (all may be NULL / 0: an untextured scene). */
  uint32_t num_textures;
  const oxr_texture* textures;
  const oxr_material_textures* material_textures; /* num_materials, or NULL */
  const float* uv0;             /* num_tris x 6: TEXCOORD_0 of v0, v1, v2    */
  const float* uv1;             /* num_tris x 6: TEXCOORD_1                  */
  const float* tangents;        /* num_tris x 12: xyz + handedness w per
                                   vertex                                    */
  const float* colors;          /* num_tris x 12: COLOR_0 RGBA per vertex    */
} oxr_scene_desc;

/* Resolved camera (src/camera.rs:61-91): eye + orthonormal basis + the
 * half extents of the view plane at unit distance. */
#define OXR_PROJ_PERSPECTIVE 0u
#define OXR_PROJ_ORTHOGRAPHIC 1u

typedef struct oxr_camera {
  float eye[3];
  float forward[3];
  float side[3];
  float up[3];
  float half_w;
  float half_h;
  uint32_t projection;
} oxr_camera;

/* LightStrategy (src/options.rs:51-64). */
#define OXR_STRATEGY_MIS 0u
#define OXR_STRATEGY_LIGHT_ONLY 1u
#define OXR_STRATEGY_BSDF_ONLY 2u

/* POD mirror of PathTraceOptions (src/options.rs:67-146) plus the
 * RenderOptions fields the estimator reads: width, height, ambient. */
typedef struct oxr_pt_options {
  uint32_t width;
  uint32_t height;
  uint32_t samples_per_pixel;   /* 1..=2^20                                  */
  uint32_t max_bounces;         /* <= 1024                                   */
  uint32_t rr_start;            /* default 3                                 */
  float clamp;                  /* 0 = off                                   */
  uint32_t seed;
  uint32_t strategy;            /* OXR_STRATEGY_*                            */
  float ambient;                /* constant sky radiance                     */
} oxr_pt_options;

/* Display options for the resolve (RenderOptions::background, tone_map,
 * exposure; src/pathtrace.rs:2593-2734, src/hdr.rs:44-66). */
#define OXR_TONEMAP_CLAMP 0u
#define OXR_TONEMAP_REINHARD 1u
#define OXR_TONEMAP_ACES 2u

typedef struct oxr_display {
  uint8_t background[4];        /* sRGB bytes + alpha                       */
  uint32_t tone_map;
  float exposure;
} oxr_display;

/* Ray and hit records of the parity entry points (also the device
 * layout of the ray queues): 32-byte ray, 16-byte hit. */
typedef struct oxr_ray {
  float o[3];
  float t_min;
  float d[3];
  float t_max;
} oxr_ray;

#define OXR_HIT_MISS 0xffffffffu
#define OXR_HIT_BACK 0x80000000u  /* bit 31 of prim: back-face hit         */

typedef struct oxr_hit {
  float t;
  uint32_t prim;                /* global triangle id | OXR_HIT_BACK, or
                                   OXR_HIT_MISS                              */
  float u;                      /* barycentric weight of v1                 */
  float v;                      /* barycentric weight of v2 (v0: 1-u-v)     */
} oxr_hit;

/* ---- progressive path tracer ---------------------------------------- */
typedef struct oxr_pt oxr_pt;

const char* oxr_last_error(void);

/* Number of visible HIP devices (0 when none: then every device call
 * below fails with OXR_ERR_DEVICE). */
int oxr_device_count(int* out);

int oxr_pt_create(int device, oxr_pt** out);
int oxr_pt_destroy(oxr_pt* h);

/* Upload a scene: builds the SAH BVH (binned SAH, collapsed to 4-wide
 * nodes), the emissive-triangle power table and the constant tables, and
 * keeps everything resident on the device. Resets the accumulation. */
int oxr_pt_upload_scene(oxr_pt* h, const oxr_scene_desc* scene);

/* Camera and options. Each resets the accumulation only when a value the
 * estimator reads changed (the radiance_key rule); *reset_out (may be
 * NULL) tells whether it did. */
int oxr_pt_set_camera(oxr_pt* h, const oxr_camera* cam, int* reset_out);
int oxr_pt_set_options(oxr_pt* h, const oxr_pt_options* opts, int* reset_out);
int oxr_pt_reset(oxr_pt* h);
/* Render window: refine traces only the frame pixels x in [x0, x0 + w),
 * y in [y0, y0 + h), with the samples (pixel seeds, sample indices,
 * camera rays) those pixels have in the whole-frame render, and the
 * accumulator, read_accum and the resolves hold just those pixels
 * (w x h, row-major). w = h = 0 restores the whole frame. A change resets
 * the accumulation. No counterpart in the reference (its refine always
 * covers the frame); it serves crop renders and convergence studies. */
int oxr_pt_set_window(oxr_pt* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hgt);
/* PathTracer::set_environment (src/pathtrace.rs:2449-2454): install an
 * equirectangular HDR environment (rgba: width*height*4 floats, scene-linear,
 * row 0 = +Y) scaled by `intensity`, or remove it (rgba = NULL). It replaces
 * the constant `ambient` sky. Resets the accumulation. */
int oxr_pt_set_environment(oxr_pt* h, const float* rgba, uint32_t width, uint32_t height, float intensity);
/* Parity entry points: EnvironmentMap::{pdf, radiance} at unit directions
 * (3n floats; le 3n) and EnvironmentMap::sample of (u0, u1) pairs (2n floats;
 * dirs 3n, pdf n, ok n) on the installed map (src/pathtrace.rs:1677-1736). */
int oxr_pt_env_eval(oxr_pt* h, uint32_t n, const float* dirs, float* pdf, float* le);
int oxr_pt_env_sample(oxr_pt* h, uint32_t n, const float* u, float* dirs, float* pdf, uint8_t* ok);

/* Texture parity entry points (the tests of the texture layer).
 * oxr_pt_tex_sample: PreparedTexture::sample_lod (src/texture.rs:417-447)
 * of texture `tex` in sRGB (srgb != 0) or linear space at n (u, v) pairs
 * (2n floats) and levels of detail (n): 4 floats each. OXR_ERR_INVALID
 * when no material slot uses the texture in that space.
 * oxr_pt_material_eval: TraceScene::material_oriented (src/trace.rs:
 * 700-923) at n hits (triangle g[i], barycentrics 3n) with the ray cone
 * (width[i], dir 3n; src/trace.rs:518-552) where width > 0 and level 0
 * otherwise, and the double-sided flip (flip[i]; normals reversed back as
 * the integrator uses them); then Integrator::emission (src/pathtrace.rs:
 * 1951-1966) and the any-hit alpha (src/trace.rs:492-506). 43 floats per
 * hit: the 35 of oxr_material_sample, the base alpha, the emissive colour,
 * the emission, the alpha. */
int oxr_pt_tex_sample(oxr_pt* h, uint32_t tex, int srgb, uint32_t n, const float* uv, const float* lod, float* out);
int oxr_pt_material_eval(oxr_pt* h, uint32_t n, const uint32_t* g, const float* bary, const float* width,
                         const float* dir, const uint8_t* flip, float* out);

/* Add n samples per pixel, continuing the global sample index; a no-op
 * before the first upload. refine(a); refine(b) == refine(a+b) bit for
 * bit. Returns after the device finished. */
int oxr_pt_refine(oxr_pt* h, uint32_t n);
int oxr_pt_samples(oxr_pt* h, uint32_t* out);

/* Accumulator readback: accum = W*H*4 doubles (sum r, g, b over covered
 * samples, covered count), sum_sq = W*H*3 doubles of the squared
 * radiance (may be NULL). */
int oxr_pt_read_accum(oxr_pt* h, double* accum, double* sum_sq);

/* Resolve on the device: W*H*4 floats (hdr) or bytes (rgba8). */
int oxr_pt_resolve_hdr(oxr_pt* h, const oxr_display* disp, float* out);
int oxr_pt_resolve_rgba8(oxr_pt* h, const oxr_display* disp, uint8_t* out);

/* Counters of the last refine: paths traced, closest-hit rays, shadow
 * rays, per-kernel device time (ms) when timing was enabled, and the BVH
 * nodes and triangles the closest-hit rays fetched when traversal counting
 * was enabled. */
typedef struct oxr_pt_stats {
  uint64_t paths;
  uint64_t closest_rays;
  uint64_t shadow_rays;
  double ms_closest;
  double ms_anyhit;
  double ms_shade;
  double ms_other;
  uint32_t launches_closest;
  uint32_t launches_anyhit;
  uint32_t graph_replays;   /* batches whose bounce loop ran as a hipGraph */
  uint32_t graph_captures;  /* of which (re)captured the graph first      */
  uint64_t closest_nodes;   /* node fetches of the closest-hit rays        */
  uint64_t closest_tris;    /* triangle fetches of the closest-hit rays    */
  uint64_t anyhit_nodes;    /* node fetches of the occlusion rays          */
  uint64_t anyhit_tris;     /* triangles the occlusion rays tested         */
} oxr_pt_stats;
/* enable: bit 0 (OXR_TIMING_EVENTS) times every kernel launch with HIP
 * events; bit 1 (OXR_TIMING_COUNTS) counts the closest-hit and occlusion
 * traversal work of single-scene refines (counting variants of the
 * kernels). Either runs the bounce loop eagerly instead of as a graph. */
#define OXR_TIMING_EVENTS 1
#define OXR_TIMING_COUNTS 2
int oxr_pt_set_timing(oxr_pt* h, int enable);
int oxr_pt_stats_get(oxr_pt* h, oxr_pt_stats* out);

/* Autotuner (no counterpart in the reference): times refine(spp) of the
 * synced scene/camera/options under each candidate launch configuration —
 * paths per wavefront batch in {2M, 4M, 8M, 16M, 32M} (within the device
 * memory cap) x shading-kernel occupancy in {1, 2} waves per SIMD x the
 * traversal queue whole or split into 8 per-XCD segments — keeps the
 * fastest (each candidate is warmed once and timed twice, the faster run
 * counting); then, at that configuration, times the occlusion kernel's
 * refill threshold at 32 idle lanes against the engine's 48 (unless
 * OXR_REFILL_ANY fixes it), and resets the accumulation. */
typedef struct oxr_tune_result {
  uint32_t path_budget;     /* chosen paths per batch                  */
  uint32_t shade_waves;     /* chosen shading occupancy                */
  uint32_t trials;          /* configurations timed                    */
  uint32_t xcd_segments;    /* chosen traversal queue split (1 or 8)   */
  double paths_per_s;       /* of the chosen configuration             */
  double worst_paths_per_s; /* of the slowest candidate                */
  uint32_t refill_closest;  /* idle lanes at which a traversal wave    */
  uint32_t refill_any;      /* takes new rays (closest hit, occlusion) */
} oxr_tune_result;
int oxr_pt_autotune(oxr_pt* h, uint32_t spp, oxr_tune_result* out);

/* ---- serving: continuous batching of render jobs ---------------------
 *
 * No counterpart in the reference, which renders one scene per call. A
 * server keeps scenes (BVH, tables) and job accumulators resident on one
 * device. Jobs are admitted FIFO while their accumulator fits the memory
 * budget. Each oxr_srv_step runs one wavefront round whose ray queues mix
 * the paths of every active job in shared launches. A job's result is
 * bit-identical to rendering it alone with oxr_pt_refine(spp): each
 * sample depends only on (scene, options, pixel, sample index), and each
 * pixel accumulates its samples in index order. */
typedef struct oxr_srv oxr_srv;

#define OXR_JOB_QUEUED 0u
#define OXR_JOB_ACTIVE 1u
#define OXR_JOB_DONE 2u

typedef struct oxr_job_status {
  uint32_t state;               /* OXR_JOB_*                                */
  uint32_t samples;             /* samples accumulated so far               */
  uint32_t target;              /* options.samples_per_pixel                */
  uint32_t rounds;              /* rounds the job took part in              */
  double t_submit_ms;           /* host clock (ms since server creation)    */
  double t_admit_ms;
  double t_done_ms;
} oxr_job_status;

typedef struct oxr_srv_stats {
  uint64_t rounds;
  uint64_t paths;
  uint64_t jobs_done;
  uint32_t active;
  uint32_t queued;
  uint64_t bytes_resident;      /* scenes + accumulators + wavefront buffers */
  double busy_ms;               /* wall time inside oxr_srv_step            */
  uint32_t graph_captures;      /* bounce loops captured as HIP graphs      */
  uint32_t graph_replays;       /* rounds whose bounce loop replayed one    */
  uint64_t single_scene_rounds; /* rounds of one scene (single-scene kernels) */
  uint32_t scenes;              /* resident scenes                          */
  uint32_t pad;
} oxr_srv_stats;

/* mem_budget: bytes the server may keep resident: scenes, the shared
 * wavefront buffers and job accumulators (0 = 1/4 of the device). A job is
 * admitted while its accumulator fits (or when nothing else is active). */
int oxr_srv_create(int device, uint64_t mem_budget, oxr_srv** out);
int oxr_srv_destroy(oxr_srv* s);
int oxr_srv_add_scene(oxr_srv* s, const oxr_scene_desc* scene, uint32_t* scene_id);
/* A scene prepared on the host only (validation, BVH, light tables,
 * materials): no device call, so any thread may build one while the
 * server's owner keeps stepping; the object keeps its own copies of the
 * scene's arrays. flags: OXR_BUILD_FAST builds a linear (Morton-order)
 * BVH with single-triangle leaves, about twice as fast to build as the
 * default binned SAH and a few percent slower to trace: for scenes that
 * arrive with a request. */
typedef struct oxr_prebuilt oxr_prebuilt;
#define OXR_BUILD_FAST 1u
int oxr_scene_prebuild(const oxr_scene_desc* scene, uint32_t flags, oxr_prebuilt** out);
int oxr_prebuilt_destroy(oxr_prebuilt* p);
/* Uploads a prebuilt scene to the server's device (owner thread). */
int oxr_srv_add_prebuilt(oxr_srv* s, const oxr_prebuilt* p, uint32_t* scene_id);
/* Frees a scene as soon as no queued or active job uses it; its id is not
 * reused and no longer accepted. */
int oxr_srv_remove_scene(oxr_srv* s, uint32_t scene_id);
int oxr_srv_submit(oxr_srv* s, uint32_t scene_id, const oxr_camera* cam, const oxr_pt_options* opts,
                   uint32_t* job_id);
/* Admit what fits, then run one round over the active jobs (no-op when
 * none). active/done (may be NULL): jobs active after the round, jobs that
 * finished in it. */
int oxr_srv_step(oxr_srv* s, uint32_t* active, uint32_t* done);
int oxr_srv_job_status(oxr_srv* s, uint32_t job_id, oxr_job_status* out);
int oxr_srv_read_accum(oxr_srv* s, uint32_t job_id, double* accum, double* sum_sq);
/* Frees a finished (or queued) job's accumulator and slot. */
int oxr_srv_release(oxr_srv* s, uint32_t job_id);
int oxr_srv_stats_get(oxr_srv* s, oxr_srv_stats* out);

/* ---- parity entry points --------------------------------------------- */

/* out[4*i..4*i+4] = sobol_owen_4d(index[i], pattern_seed[i]) on device. */
int oxr_sobol_owen_4d(int device, uint32_t n, const uint32_t* index,
                      const uint32_t* pattern_seed, float* out);

/* Closest hit / occlusion of n host rays against the uploaded scene, all
 * candidates accepted (no MASK or BLEND alpha filter), as
 * TraceScene::closest_hit / occluded (src/trace.rs:396-450). flags bit 0:
 * cull back faces of single-sided materials (the primary-ray rule). */
int oxr_trace_closest(oxr_pt* h, uint32_t n, const oxr_ray* rays, uint32_t flags,
                      oxr_hit* hits);
int oxr_trace_any(oxr_pt* h, uint32_t n, const oxr_ray* rays, uint8_t* occluded);

/* The same closest-hit query on rays already resident in device memory
 * (`stream` is a hipStream_t or NULL for the handle's stream). It shares
 * the handle's traversal buffers with refine, so on another stream it is
 * ordered after the handle's queued work and before its later work. The
 * call only enqueues; oxr_trace_device_check reports its outcome. */
int oxr_trace_closest_device(oxr_pt* h, uint32_t n, const void* d_rays,
                             void* d_hits, void* stream);
/* Waits for the handle's stream and reports a traversal stack overflow of
 * the last oxr_trace_closest_device call (OXR_ERR_DEVICE; its hits are
 * then misses). Call it before the next refine or trace on the handle. */
int oxr_trace_device_check(oxr_pt* h);
/* oxr_trace_closest_device without allocating: no hipMalloc, hipFree,
 * event or other allocating call. Its work counter and traversal-stack
 * spill live in a fixed 64 MiB device array per GPU (set up when the
 * library first uses the device), not in the handle's buffers, so it is
 * ordered only on `stream` (NULL: the handle's stream) and may be captured
 * into a HIP graph. Same hits as oxr_trace_closest_device. All calls on one
 * GPU share the array: calls that overlap on one GPU race. A tree deeper
 * than the array's spill capacity is rejected (OXR_ERR_INVALID). No
 * counterpart in the reference: it is for callers that cannot allocate
 * (HIP graph capture, tight serving loops). */
int oxr_trace_closest_device_noalloc(oxr_pt* h, uint32_t n, const void* d_rays,
                                     void* d_hits, void* stream);

/* Flattened BVH statistics: node and leaf counts, SAH cost, depth. */
typedef struct oxr_bvh_info {
  uint32_t num_nodes;
  uint32_t num_leaves;
  uint32_t num_tris;
  uint32_t max_depth;
  double sah_cost;
  double build_ms;
} oxr_bvh_info;
int oxr_pt_bvh_info(oxr_pt* h, oxr_bvh_info* out);

/* Copy of the flattened device BVH (for the oracle's traversal count):
 * nodes = num_nodes * 64 bytes (4-wide nodes with 8-bit quantised child
 * boxes), tris = num_tris * 48 bytes. */
int oxr_pt_bvh_export(oxr_pt* h, void* nodes, void* tris);

/* Per-point material inputs of Bsdf::new (the MaterialSample fields the
 * BSDF reads, src/trace.rs:201-267). */
typedef struct oxr_material_sample {
  float base_color[3];
  float metallic;
  float roughness;
  float normal[3];
  float ior;
  float specular;
  float specular_color[3];
  float transmission;
  float thickness;
  float clearcoat;
  float clearcoat_roughness;
  float clearcoat_normal[3];
  float sheen_color[3];
  float sheen_roughness;
  float anisotropy;
  float anisotropy_direction[3];
  float iridescence;
  float iridescence_ior;
  float iridescence_thickness;
  float diffuse_transmission;
  float diffuse_transmission_color[3];
} oxr_material_sample;

/* Bsdf::new(ms[i], v[i], ng[i], front[i]) then eval(l[i]) -> f (3), pdf.
 * Arrays of 3 floats per item. */
int oxr_bsdf_eval(int device, uint32_t n, const oxr_material_sample* ms,
                  const float* v, const float* ng, const uint8_t* front,
                  const float* l, float* f, float* pdf);
/* ... then sample(u[i]) -> l (3), f (3), pdf; ok[i] = 0 when None. */
int oxr_bsdf_sample(int device, uint32_t n, const oxr_material_sample* ms,
                    const float* v, const float* ng, const uint8_t* front,
                    const float* u, float* l, float* f, float* pdf, uint8_t* ok);

#ifdef __cplusplus
}
#endif

#endif /* OXR_PATHTRACE_H */
