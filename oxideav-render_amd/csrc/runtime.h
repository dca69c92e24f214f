// Internal interface between the host runtime (api.cpp) and the kernel
// translation unit (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_guide.h"
#include "dev_scene.h"

namespace oxr {

// Wavefront buffers for one batch of paths (capacity `cap` paths).
struct PathBuffers {
  float4* qray[2] = {nullptr, nullptr};  // 2 float4 per ray: (o, aux), (d, t_max)
  uint32_t* qpid[2] = {nullptr, nullptr};
  uint4* hits = nullptr;
  float4* sray = nullptr;  // shadow rays: (o, skip gid), (d, t_max)
  uint32_t* sidx = nullptr;
  float4* beta = nullptr;  // (beta, prev_pdf)
  float4* rad = nullptr;
  float4* prev = nullptr;
  float4* sigma = nullptr;  // (sigma, in-medium flag)
  uint4* meta = nullptr;    // (x | y << 16, sample, k | hero << 16 | flags << 24, pixel seed)
  float4* nee = nullptr;    // cap * nslots contributions
  uint8_t* vis = nullptr;   // cap * nslots: 0 none, 1 pending, 2 visible
  float4* out = nullptr;    // per-sample radiance + covered flag
  uint32_t* pjob = nullptr; // job index per path
  uint32_t cap = 0;
  uint32_t nslots = 0;
};

namespace launch {
hipError_t upload_constants(const uint32_t* sobol128, const float* sheen512, const float* srgb255);
int trace_blocks_per_cu(bool any, bool alpha = false);
int trace_block_threads();
int stack_lds_entries();
// work: nseg counters (closest: Ctrl::work, any-hit: Ctrl::work + 8);
// nseg = 8 splits the queue into per-XCD segments (1 when the grid is not
// a multiple of it). Bits 8-15 of nseg: the idle lanes at which a wave
// takes new rays (0 = 64).
// path: rays of the wavefront (else host rays of the parity entry points);
// multi: the launch mixes scenes (each ray's scene from its job), else all
// rays trace `sc`. asc non-null: the textured-alpha variant (MASK / BLEND
// alpha varying over triangles); single-scene launches read that scene's
// device copy through it.
void closest(hipStream_t s, int grid, const DevScene& sc, const DevJob* jobs, const uint32_t* pjob, bool path,
             bool multi, const float4* rays, const uint32_t* qpid, const uint4* meta, uint32_t fixed_count,
             const uint32_t* count_ptr, uint32_t* work, uint32_t cull_all, uint4* hits, uint32_t* spill,
             uint32_t spill_cap, uint32_t* err, uint32_t nseg = 1, const DevScene* asc = nullptr,
             unsigned long long* counts = nullptr);
void anyhit(hipStream_t s, int grid, const DevScene& sc, const DevJob* jobs, const uint32_t* pjob, bool path,
            bool multi, const float4* rays, const uint32_t* sidx, const uint4* meta, uint32_t nslots, uint32_t fixed_count,
            const uint32_t* count_ptr, uint32_t* work, uint8_t* out, uint32_t* spill, uint32_t spill_cap,
            uint32_t* err, uint32_t nseg = 1, const DevScene* asc = nullptr, unsigned long long* counts = nullptr);
void camera(hipStream_t s, const DevJob* jobs, const SegTable& segs, uint32_t P, const PathBuffers& pb, Ctrl* ctrl);
void prep(hipStream_t s, Ctrl* ctrl, uint32_t cur, uint32_t count_shadow_prev);
// guide: null for the plain estimator.
// Texture parity entry points: sample_lod of a chain at (u, v, lod)
// triples; material_oriented + emission + alpha at hits (43 floats each).
void tex_sample(hipStream_t s, const DevTextures& T, uint32_t chain, uint32_t n, const float* uv, const float* lod,
                float4* out);
void material_eval(hipStream_t s, const DevScene& sc, uint32_t n, const uint32_t* g, const float* bary,
                   const float* width, const float* dir, const uint8_t* flip, float* out);
// tex: the TEX variant (textures, vertex colours or tangents in a scene
// of the batch).
// multi: the batch mixes jobs (the server's rounds).
// ext: a material of the batch has an extension lobe (clearcoat, sheen,
// iridescence, transmission, diffuse transmission or anisotropy).
// area: a scene of the batch has emissive triangles, or a job has an
// environment map. kb: the bounce of every entry of the launch's queue
// (bounce 0 reads no path state but the metadata).
void shade(hipStream_t s, int grid, int waves, const DevJob* jobs, uint32_t nslots, const PathBuffers& pb, Ctrl* ctrl,
           uint32_t cur, const DevGuideUse* guide, bool tex, bool multi, int ext, bool area, uint32_t kb);
// Training samples (2 float4 each) of the guide-recorded paths of a batch,
// appended at *count (entries past cap are dropped).
void guide_collect(hipStream_t s, uint32_t P, const DevGuideUse& gu, const PathBuffers& pb, float4* samp,
                   uint32_t* count, uint32_t cap);
void finalize(hipStream_t s, const DevJob* jobs, const SegTable& segs, uint32_t P, uint32_t nslots,
              const PathBuffers& pb);
void reduce(hipStream_t s, uint32_t npx, uint32_t S, const float4* out, double* accum, double* sumsq);
void fill_f64(hipStream_t s, double* p, size_t n);
void zero_ctrl(hipStream_t s, Ctrl* c);
void resolve(hipStream_t s, bool rgba8, uint32_t npx, uint32_t samples, const double* accum, const float bg[4],
             const uint8_t bgb[4], uint32_t tone_map, float exposure, void* out);
void sobol(hipStream_t s, uint32_t n, const uint32_t* idx, const uint32_t* seed, float4* out);
void env(hipStream_t s, const DevEnv& e, uint32_t n, int do_sample, const float* in, float* dirs, float* pdf,
         float* le, uint8_t* ok);
void bsdf(hipStream_t s, uint32_t n, int do_sample, const float* ms, const float* v, const float* ng,
          const uint8_t* front, const float* in, float* l, float* f, float* pdf, uint8_t* ok);
}  // namespace launch

}  // namespace oxr
