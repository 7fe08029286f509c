/* oracle.h - CPU restatement of the reference rendering hot path (TEST
 * INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and the
 * bench's cpu_baseline leg; never by the product library).
 *
 * Function                 reference                                 pinned by
 * oracle_philox            (replaces cuRAND, core/render.cuh:82-91)  Random123 KAT vectors
 * oracle_uniforms          draw -> (0,1] float                       range/uniformity tests
 * oracle_math              shared log/exp/sincos/atan/acos, HG      libm within ulp bounds; closed forms
 * oracle_ray_box           core/vector.cu:41-75                      geometric closed forms
 * oracle_hg_sample_dir     core/vector.cu:211-237                    chi^2 of cos vs HG pdf; unit length
 * oracle_density           core/render.cuh:67-71, volume.cu:685-703  brute-force trilinear (numpy)
 * oracle_max_density       core/volume.cu:1219-1224                  numpy max
 * oracle_render            core/render.cuh:93-258, volume.cu:147-179 closed forms: empty volume = sky,
 *                                                                    albedo 0 = Beer-Lambert, single
 *                                                                    scattering quadrature (statistical)
 * oracle_render_frame      core/volume.cu:57-124,181-287,469-547      constant image passes the filter;
 *                                                                    low variance passes through; brute
 *                                                                    force on small images (numpy);
 *                                                                    running mean = oracle_render
 * oracle_tone_pack         core/render.cuh:640-732, volume.cu:248-263 Gamma / ACES closed forms (numpy)
 * oracle_histo_distance    core/volume.cu:187-207                    chi2(h,h) = 0, scale invariance, hand values
 * oracle_get_samples       core/render.cuh:276-315                   transparent medium = sky; albedo 0 =
 *                                                                    HG-averaged escape probability;
 *                                                                    single scattering = FP64 quadrature
 * oracle_mrpnn_grad        training loss and gradient (derived from    FP64 finite differences of an
 *                          core/radiancePredict.cu:370-486)          independent numpy forward
 * oracle_mrpnn_grad_bf16   the same with BF16 product operands        rounding = numpy BF16 (tests/lowp.py);
 *                                                                    the BF16 numpy forward's predictions;
 *                                                                    finite differences of its loss
 * oracle_reproject         core/volume.cu:125-141                    static camera maps to itself; pinhole
 *                                                                    projection into a moved camera (numpy)
 * oracle_frame_from_normal core/render.cuh:355-369                   orthonormal, first axis = input, handedness,
 *                                                                    second axis uniform about the first (KS)
 * oracle_render_mrpnn      core/render.cuh:469-523, volume.cu:147-179 zero network = single-scatter FP64
 *                                                                    quadrature per channel (light colour, albedo);
 *                                                                    channel ratios; static-camera MRPNN frames =
 *                                                                    closed-form blend of 1-frame renders
 * oracle_training_points   dataGen.cu:174-216 (MeanFreePathSample)    empty volume = no points; dense medium =
 *                                                                    uniform origins on the r = 0.5 sphere,
 *                                                                    keep / jitter / view-direction rates
 * oracle_radiances         core/volume.cu:17-34,329-402 (GetRadiances) PT: ray missing the box = sky, empty
 *                                                                    volume = sky, albedo 0 = Beer-Lambert,
 *                                                                    one bounce = single-scatter quadrature;
 *                                                                    MRPNN: equals oracle_render_mrpnn's
 *                                                                    composition on the same first vertices
 * oracle_shadow_terms      core/render.cuh:317-341 (GetTr/GetTrs)   Tr * HG: homogeneous box = exp(-sigma d) * HG
 *                                                                    within Monte Carlo error; empty = HG
 * oracle_env / set_hdri    core/render.cuh:57-59 (SkyBox)            constant map = constant sky; texel values at
 *                                                                    the equirectangular texel centres; wrap
 *                                                                    seam; axis directions' (u, v)
 * oracle_env_light         core/volume.cu:417-432 (Render)           rate formula; Roberts2 R2 points; colour =
 *                                                                    clamp-sampled map x 4 exp / rate
 * Later sections of oracle.cpp list their own pins.
 */
#ifndef MRP_ORACLE_H_
#define MRP_ORACLE_H_

#include <stdint.h>

#include "../include/mrp.h"

#ifdef __cplusplus
extern "C" {
#endif

int oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
int oracle_set_hdri(const float* rgba, int w, int h, float exposure);
int oracle_env(const float* dirs, int n, mrp_vec3 sky, float* out);
int oracle_env_light(int n, float choose, mrp_vec3 light_dir, mrp_vec3 light_color, float* out6);
int oracle_normalize3(const float* v, float* out);
int oracle_uniforms(uint32_t id, uint32_t sample, uint32_t tag, uint64_t seed, int n, float* out);
int oracle_frame_from_normal(const float* v, int n, uint32_t sample, uint32_t tag, uint64_t seed, float* out,
                             int* draws);
int oracle_math(int fn, const float* a, const float* b, float* out, int n);
int oracle_ray_box(const float* p, const float* d, float* offset, float* distance_out, int n);
int oracle_hg_sample_dir(const float* e0, const float* e1, const float* v, float g, float* out, int n);
int oracle_density(const float* grid, int res, const float* pos, float* out, int n);
int oracle_generate_cloud(const mrp_cloud_params* p, int res, float* out);
int oracle_max_density(const float* grid, int res, float* out);
int oracle_mips(const float* grid, int res, float* out);
int oracle_tr_mips(const float* mips, mrp_vec3 light_dir, float alpha, float* out);
int oracle_hg_lut(float g, uint64_t seed, float* out);
int oracle_mrpnn_features(const float* mips, const float* trs, const float* lut, int n, const float* pos,
                          const float* fx, const float* fy, const float* fz, const int* active, mrp_vec3 light_dir,
                          mrp_vec3 light_x, float alpha, float g, float* feat, float* gamma);
int oracle_mrpnn_forward(const float* w, int n, const float* feat, const float* gamma, float g, mrp_vec3 srp,
                         float* out);
int oracle_render_mrpnn(const float* grid, int res, float max_density, const float* mips, const float* trs,
                        const float* lut, const float* weights, const mrp_render_params* p, float* target,
                        double* sum64, int threads);
// One view of a batched serving call (mrp_serve_views): MRPNN samples, the
// histogram-variance fallback to delta tracking; see oracle.cpp.
int oracle_serve_view(const float* grid, int res, float max_density, const float* mips, const float* trs,
                      const float* lut, const float* weights, const mrp_render_params* p, int spp, float threshold,
                      int fallback_spp, float* rgb, float* histo, uint8_t* flags, uint64_t* fallback_lookups,
                      int threads);
int oracle_render(const float* grid, int res, float max_density, const mrp_render_params* p, float* target,
                  double* sum64, uint64_t* lookups_out, uint64_t* scatters_out, int threads, int row_begin,
                  int row_end);

int oracle_render_frame(const float* grid, int res, float max_density, const float* mips, const float* trs,
                        const float* lut, const float* weights, const mrp_frame_params* fp, uint32_t call_index,
                        int flip, int have_last, mrp_vec3 last_ori, mrp_vec3 last_up, mrp_vec3 last_right,
                        int last_predict, float* target, mrp_histogram* histo, uint32_t* bgra,
                        uint64_t* lookups_out, int threads);
int oracle_denoise_pack(const float* target, mrp_histogram* histo, int W, int H, int frame, int checker,
                        int checkerboard, int denoise, int predict, int tone, float exposure, uint32_t* bgra);
int oracle_tone_pack(const float* rgb, int n, float exposure, int tone, uint32_t* out);
int oracle_mrpnn_grad(const float* w, int n, const float* feat, const float* aux, double* grad, double* loss_mean,
                      double* pred);
int oracle_mrpnn_grad_bf16(const float* w, int n, const float* feat, const float* aux, double* grad,
                           double* loss_mean, double* pred);
int oracle_bf16_round(const double* in, int n, double* out);
int oracle_get_samples(const float* grid, int res, float max_density, int n, const mrp_sample_query* q,
                       mrp_vec3 light_color, mrp_vec3 sky, int multi, int spp, uint64_t seed, float* out,
                       uint64_t* lookups_out, int threads);
int oracle_training_points(const float* grid, int res, float max_density, int n_rays, float alpha, float g,
                           uint64_t seed, float* pos, float* view, uint64_t* lookups_out, int threads);
int oracle_radiances(const float* grid, int res, float max_density, const float* mips, const float* trs,
                     const float* lut, const float* weights, int n, const float* ori, const float* dir,
                     const mrp_render_params* p, float* out, uint64_t* lookups_out, int threads);
int oracle_shadow_terms(const float* grid, int res, float max_density, int n, const float* pos, const float* dir,
                        mrp_vec3 light_dir, float alpha, float g, int spp, uint64_t seed, float* out,
                        uint64_t* lookups_out, int threads);
int oracle_reproject(const float* pts, int n, mrp_vec3 lori, mrp_vec3 lup, mrp_vec3 lright, int W, int H, int* out);
int oracle_histo_distance(const mrp_histogram* a, const mrp_histogram* b, float* dist, float* weight);

#ifdef __cplusplus
}
#endif

#endif
