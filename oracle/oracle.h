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
int oracle_normalize3(const float* v, float* out);
int oracle_uniforms(uint32_t id, uint32_t sample, uint32_t tag, uint64_t seed, int n, float* out);
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
int oracle_render(const float* grid, int res, float max_density, const mrp_render_params* p, float* target,
                  double* sum64, uint64_t* lookups_out, uint64_t* scatters_out, int threads, int row_begin,
                  int row_end);

#ifdef __cplusplus
}
#endif

#endif
