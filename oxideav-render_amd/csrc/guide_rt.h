// Internal interface between the guide's host runtime and its kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_guide.h"

namespace oxr {
namespace glaunch {
// Inference over `fixed_count` queries, or *count_ptr when non-null
// (device counter; `cap` bounds the launch). Positions come from `pos`
// (3 floats each) when non-null, else from the ray queue: o + t d.
void infer(hipStream_t s, const DevGuideNet& n, uint32_t fixed_count, const uint32_t* count_ptr, uint32_t cap,
           const float* pos, const float4* qray, const uint4* hits, GuideLobes* out, float* raw);
// Loss gradient of `count` samples (count % 32 == 0): the grid part by
// atomics into ggrid (which must hold zeros) and, for the levels gc
// covers, into gc's copies (zero on entry, zero again on return) folded
// into ggrid; the MLP part into gmlp.
void train_grad(hipStream_t s, const DevGuideNet& n, uint32_t count, const float4* samp, float* inv_wsum,
                float* ggrid, const GradCopies& gc, float* gmlp, uint16_t* xs, uint16_t* ds, float4* part,
                float* bpart, uint32_t nsplit, float* loss);
void adam(hipStream_t s, uint64_t n_grid, uint64_t n_total, float* p, float* m, float* v, float* grad, float lr,
          float b1, float b2, float eps, float c1, float c2);
void pack(hipStream_t s, const float* mlp, uint4* frag);
void fill(hipStream_t s, float* p, uint64_t n);
// dst[i] = src[start + i * stride], i < count (samples of 2 float4).
void gather(hipStream_t s, const float4* src, uint32_t start, uint32_t stride, uint32_t count, float4* dst);
void pdf_sample(hipStream_t s, uint32_t n, const GuideLobes* lobes, const float* dir, const float* u, float* dir_out,
                float* pdf_out);
uint32_t wtiles();
size_t xs_group();  // bf16 elements of stored activations per 16 samples
size_t ds_group();  // bf16 elements of stored deltas per 16 samples
}  // namespace glaunch
}  // namespace oxr
