/*
 * liboxr.so — learned path guiding for the HIP path tracer (gfx950).
 *
 * oxideav-render has no path guiding: its PathTrace estimator samples
 * directions from the BSDF only (src/pathtrace.rs:2281-2322). This adds an
 * optional proposal on top of it. The estimator stays unbiased, and with no
 * guide attached the tracer is exactly the reference estimator.
 *
 * The guide has two parts:
 *
 *  - Encoding: a multiresolution hash grid over scene positions.
 *    OXR_GUIDE_LEVELS levels, OXR_GUIDE_FEATURES features each, 2^log2_table
 *    entries per level.
 *  - Network: an MLP with OXR_GUIDE_HIDDEN units per hidden layer.
 *    - Layout: 32 -> 64 -> 64 -> 64 -> 48, ReLU between layers.
 *    - Precision: BF16 operands, FP32 accumulation (MFMA on gfx950).
 *    - Output: an OXR_GUIDE_LOBES von Mises-Fisher mixture over world
 *      directions.
 *
 * It is trained online on the tracer's own radiance samples. The loss is
 * the cross-entropy of the mixture against the incident radiance:
 *
 *   L = sum_i w_i * -log q(w_i | x_i) / sum_i w_i,
 *   w_i = luminance(L_i) / p_i.
 *
 * At eligible vertices (no transmission, not a near-mirror metal), the
 * shading kernel uses one-sample MIS. It draws the next direction from
 * the guide with probability `mix`, else from the BSDF, and weights the
 * sample by the mixture pdf:
 *
 *   p = mix * q + (1 - mix) * p_bsdf.
 *
 * NEE's MIS weights use the same p, so the estimator remains unbiased.
 * Russian roulette on a guided path decides on the throughput the BSDF
 * sampler alone would have produced (beta * p / p_bsdf per guided vertex):
 * a guided direction into a bright region has a small f / p by design, and
 * throughput roulette would end most such paths and scale up the rest.
 *
 * Parameters are one f32 array in this order:
 *
 *   grid[levels][2^log2_table][2]
 *   W1[64][32] b1[64]  W2[64][64] b2[64]  W3[64][64] b3[64]  W4[48][64] b4[48]
 *
 * W is [out][in]. Output o of the last layer is decoded as follows, with
 * t = o / 16, g = (o % 16) / 4, e = 4 * t + o % 4. If e < 10, output o is
 * component e % 5 of lobe 2 * g + e / 5; otherwise it is unused. A lobe's
 * components are (m.x, m.y, m.z, s, logit), and they map to the lobe as:
 *
 *   mu = m / |m|
 *   kappa = exp(clamp(s, -3, 8))
 *   pi = softmax(logit)
 */
#ifndef OXR_GUIDE_H
#define OXR_GUIDE_H

#include "oxr/pathtrace.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OXR_GUIDE_LEVELS 16
#define OXR_GUIDE_FEATURES 2
#define OXR_GUIDE_HIDDEN 64
#define OXR_GUIDE_OUTPUTS 48
#define OXR_GUIDE_LOBES 8
/* Largest oxr_guide_config.bounces: vertices past it sample the BSDF alone. */
#define OXR_GUIDE_MAX_BOUNCES 4
/* f32 count of the MLP part of the parameter array. */
#define OXR_GUIDE_MLP_PARAMS (64 * 32 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 48 * 64 + 48)

typedef struct oxr_guide oxr_guide;

typedef struct oxr_guide_config {
  uint32_t log2_table;   /* entries per level = 2^log2_table, 10..22 (19) */
  uint32_t base_res;     /* coarsest grid resolution (16) */
  uint32_t max_res;      /* finest grid resolution (2048) */
  uint32_t seed;         /* parameter initialisation */
  float aabb_min[3];     /* position normalisation box; min == max: the */
  float aabb_max[3];     /* scene's bounds when attached to a tracer     */
  float lr;              /* Adam step size (5e-3) */
  float beta1, beta2;    /* Adam moments (0.9, 0.99) */
  float eps;             /* Adam epsilon (1e-15) */
  float mix;             /* probability of sampling the guide at an eligible vertex (0.5) */
  uint32_t batch;        /* samples per online training step (65536) */
  uint32_t steps;        /* online training steps after each refine pass (4) */
  uint32_t record_shift; /* online samples come from paths with id % 2^shift == 0 (3) */
  uint32_t max_steps;    /* online training stops after this many Adam steps; 0 = never (0) */
  uint32_t bounces;      /* eligible vertices at bounce k < bounces use the guide, 1..OXR_GUIDE_MAX_BOUNCES (1) */
} oxr_guide_config;

typedef struct oxr_guide_stats {
  uint64_t steps;         /* Adam steps taken */
  uint64_t samples;       /* training samples consumed */
  uint64_t collected;     /* online samples collected (all passes) */
  float last_loss;        /* loss of the last step */
  float pad;
  double train_ms;        /* device time of training (when timing is on) */
  double infer_ms;        /* device time of inference inside refine */
} oxr_guide_stats;

/* Fills `cfg` with the defaults above. */
int oxr_guide_default_config(oxr_guide_config* cfg);
int oxr_guide_create(int device, const oxr_guide_config* cfg, oxr_guide** out);
int oxr_guide_destroy(oxr_guide* g);
/* f32 counts of the grid part and the MLP part of the parameter array. */
int oxr_guide_num_params(oxr_guide* g, uint64_t* n_grid, uint64_t* n_mlp);
/* Copy parameters out / in (either pointer may be null). Setting them
 * resets the Adam state. */
int oxr_guide_get_params(oxr_guide* g, float* grid, float* mlp);
int oxr_guide_set_params(oxr_guide* g, const float* grid, const float* mlp);
/* Inference on host positions (3n floats). `raw` (48n, may be null): the
 * network outputs; `lobes` (40n, may be null): per lobe (mu.xyz, kappa, pi). */
int oxr_guide_infer(oxr_guide* g, uint32_t n, const float* pos, float* raw, float* lobes);
/* Mixture pdf at unit directions `dir` (3n), and sampling from the mixture
 * with `u` (3n uniforms: lobe, then the two of the vMF draw): dir_out (3n),
 * pdf_out (n). */
int oxr_guide_pdf(oxr_guide* g, uint32_t n, const float* pos, const float* dir, float* pdf);
int oxr_guide_sample(oxr_guide* g, uint32_t n, const float* pos, const float* u, float* dir_out, float* pdf_out);
/* Loss and gradient of a host batch (pos 3n, dir 3n, weight n) at the
 * current parameters; no update. grad_grid / grad_mlp as the parameter
 * array (either may be null). */
int oxr_guide_gradient(oxr_guide* g, uint32_t n, const float* pos, const float* dir, const float* weight,
                       float* grad_grid, float* grad_mlp, double* loss);
/* `steps` Adam steps on a host batch (the whole batch each step); loss of
 * the last step in *loss (may be null). */
int oxr_guide_train(oxr_guide* g, uint32_t n, const float* pos, const float* dir, const float* weight,
                    uint32_t steps, double* loss);
int oxr_guide_stats_get(oxr_guide* g, oxr_guide_stats* out);
/* The training samples the last online pass collected (inspection and
 * tests): up to `max` samples of 8 floats each into `out` (may be null),
 * (position.xyz, weight, direction.xyz, pdf of the proposal that drew the
 * direction); *n receives the number available. */
int oxr_guide_collected(oxr_guide* g, uint32_t max, float* out, uint32_t* n);

/* Attach a guide to a tracer (null detaches). With `train` != 0 every
 * refine pass collects training samples and runs cfg.steps Adam steps after
 * the pass. The guide must live on the tracer's device and outlive the
 * attachment; attaching resets the accumulator. A guide whose box is empty
 * takes the scene's bounds. */
int oxr_pt_set_guide(oxr_pt* h, oxr_guide* g, int train);

#ifdef __cplusplus
}
#endif

#endif /* OXR_GUIDE_H */
