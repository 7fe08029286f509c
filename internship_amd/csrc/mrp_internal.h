// mrp_internal.h - state behind the opaque mrp_volume handle and shared host helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mrp.h"
#include "mrp_medium.h"

namespace mrp {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define MRP_HIP(call)                                           \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return ::mrp::hip_fail(e_, #call); \
    } while (0)

constexpr int kNumMips = 9;      // density mips 256^3 .. 1^3
constexpr int kNumTrMips = 8;    // transmittance mips 128^3 .. 1^3
constexpr int kLutSize = 512;    // HG lookup table is kLutSize x kLutSize
constexpr int kNumQueues = 8;    // one work queue per XCD

struct Volume {
    int res = 0;
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t nn_events[9] = {};  // stage timing of MRPNN frames, created on first use
    // pinned host words the device copies into without a blocking copy: [0] the
    // last MRPNN chunk's descriptor count, [1..4] the lookup counters
    unsigned long long* host_words = nullptr;
    bool nn_pending = false;       // the last chunk's count and times await mrpnn_finish
    float* grid = nullptr;           // padded (res+2)^3, border zeros
    float4* cells = nullptr;         // corner-expanded cells (res+1)^3 x 2 float4, the tracer's layout
    float4* quads = nullptr;         // the padded grid as 16-byte quads in 1x4x2 bricks (quad_count), the first free flight's layout
    uint32_t* occ = nullptr;         // brick occupancy bits of the cells (Occupancy, occupancy_words)
    float max_density = 0.00001f;
    bool has_data = false;
    bool updated = false;
    float* mips[kNumMips] = {};      // padded (256>>m + 2)^3
    float* tr_mips[kNumTrMips] = {}; // unpadded (128>>m)^3
    bool tr_valid = false;
    mrp_vec3 tr_light{0, 0, 0};
    float tr_alpha = -1.0f;
    float* hg_lut = nullptr;         // kLutSize^2, x (cos) fastest
    // corner-expanded copies read by the descriptor gather (one 32/16-byte read per fetch)
    float4* mip_cells[kNumMips] = {};   // border cells of the padded mips, (r+1)^3 x 32 B
    float4* tr_cells[kNumTrMips] = {};  // clamp cells of the TR mips, (r+2)^3 x 32 B
    float4* lut_cells = nullptr;        // bilinear clamp cells of the HG table, 514^2 x 16 B
    // FP16 copies of the cells (8 corners in 16 B), every level in one buffer
    // (level m at mip16_offset / tr16_offset cells), read by the BF16/FP8
    // renderer's gather when the volume opts in to FP16 cells
    uint4* mip_cells16 = nullptr;
    uint4* tr_cells16 = nullptr;
    bool gather_cells16 = false;  // the tiled gather reads the FP16 cells (mrp_volume_set_gather_cells(16))
    bool lut_valid = false;
    float lut_g = -100.0f;
    uint64_t lut_seed = 0;
    const float* nn_weights = nullptr;       // FP32 MRPNN weights in HBM (owned by an mrp_network)
    const void* nn_packed = nullptr;         // BF16 MFMA-packed weights (owned by an mrp_network)
    const void* nn_packed_fp8 = nullptr;     // FP8-E4M3 MFMA-packed weights (owned by an mrp_network)
    // scratch
    uint32_t* queues = nullptr;              // kNumQueues counters
    unsigned long long* counters = nullptr;  // [lookups, paths, scatters, spare]
    void* samples = nullptr;                 // per-(sample, pixel) float4 results
    size_t samples_bytes = 0;
    size_t sample_budget = (size_t)4 << 30;
    void* nn_scratch = nullptr;              // MRPNN stages: vertices, features, pool, predictions
    size_t nn_scratch_bytes = 0;
    // batched serving (this volume runs the batch): view tables, MRPNN samples,
    // histograms and the fallback list; stage events
    void* serve_scratch = nullptr;
    size_t serve_scratch_bytes = 0;
    hipEvent_t serve_events[5] = {};
    // the environment map (`SetHDRI`, core/volume.cu:1426-1458): RGBA FP32
    // texels in HBM for the kernels' `SkyBox`, a host copy for the light choice
    float* hdri = nullptr;
    int hdri_w = 0, hdri_h = 0;
    float hdri_exp = 1.0f;
    std::vector<float> hdri_host;
    // progressive-frame state kept between mrp_render_frame calls, like the
    // reference's globals `flip`, `rand_cpu`, `lori/lup/lright` and
    // `last_predict` (core/volume.cu:50-56,404-407,469-547)
    int frame_flip = 0;
    uint32_t frame_calls = 0;
    bool frame_have_last = false;
    bool frame_last_predict = true;
    mrp_vec3 frame_last_ori{0, 0, 0}, frame_last_up{0, 0, 0}, frame_last_right{0, 0, 0};
};

// Which pixels a launch renders and which camera row convention it uses.
struct PixelMap {
    int checker = -1;  // -1: every pixel; 0/1: the checkerboard half of checker_col
    int rows_up = 0;   // 1: progressive camera (row 0 at the bottom)
    // a ray list instead of camera pixels (width = n rays, height 1): ray i
    // starts at ray_ori[3i..] along ray_dir[3i..] (HBM)
    const float* ray_ori = nullptr;
    const float* ray_dir = nullptr;
};

inline int mip_res(int m) { return 256 >> m; }
inline int tr_mip_res(int m) { return 128 >> m; }

// Ensure the per-volume sample buffer holds at least `bytes`.
int ensure_samples(Volume* v, size_t bytes);

// Kernel entry points implemented in the .hip files (host launchers).
int check_params(const Volume* v, const mrp_render_params* p);
int launch_accumulate(Volume* v, const void* samples, size_t P, int spp_begin, int n, float* target, double* sum64);
int mrpnn_dev(Volume* v, const mrp_render_params* p, float* target, double* sum64, mrp_trace_stats* st);
// One batch of per-(sample, pixel) float4 results (rgb, first-hit distance or
// -1) in v->samples[(s - s_begin) * P + pixel] for the pixels `map` selects.
int trace_samples(Volume* v, const mrp_render_params* p, PixelMap map, int s_begin, int n);
// Device time (HIP events) of each MRPNN stage, summed over calls.
struct NNTimes {
    double first_vertex = 0.0, compact = 0.0, shadow = 0.0, gather = 0.0, network = 0.0;
};
// With defer_last, the last chunk's count and stage times are left pending
// (no host synchronisation): after synchronising the stream, the caller adds
// them with mrpnn_finish.
// With `fold` set and one sample per pixel (n == 1, every pixel), the combine
// stage folds each pixel's sample straight into the running mean (and FP64
// sums) instead of writing it to v->samples.
struct SampleFold {
    float* target = nullptr;
    double* sum64 = nullptr;
    float rate = 0.0f;  // 1 / (1 + the sample's index)
};
int mrpnn_samples(Volume* v, const mrp_render_params* p, PixelMap map, int s_begin, int n, NNTimes* times,
                  uint64_t* n_active, bool defer_last = false, SampleFold fold = SampleFold{});
int mrpnn_finish(Volume* v, NNTimes* times, uint64_t* n_active);
// The MRPNN samples [0, spp) of n_views images in one pass of the stages
// (view k on vols[k] with ps[k], every view of the same size, width * height a
// multiple of 64): out[(k * spp + s) * P + pixel] = (radiance, first-vertex
// distance or -1).  vols[0]'s network, stream and scratch run them; tables_dev
// holds mrpnn_view_table_bytes(n_views) bytes of device memory.
int mrpnn_views(Volume* const* vols, const mrp_render_params* ps, int n_views, int spp, void* tables_dev,
                float4* out, NNTimes* times, uint64_t* n_active);
size_t mrpnn_view_table_bytes(int n_views);
// Denoise (past frame 0) and BGRA8 pack of n_images images of one size stored
// one after another: target 3 floats, histo kHistoFloats floats, bgra one word per pixel;
// max_count bounds every histogram's sample count.
int launch_denoise_images(hipStream_t stream, int width, int height, int n_images, int denoise, int tone,
                          float exposure, const float* target, const float* histo, uint32_t* bgra, int max_count);
// `NNPredict<MRPNN>` on n given rays (one sample each): out_dev n x float4
// (radiance, first-vertex distance or -1).
int mrpnn_rays(Volume* v, int n, const float* ori, const float* dir, const mrp_render_params* p, float* out,
               mrp_trace_stats* stats);
int clear_counters(Volume* v);
int read_counters(Volume* v, unsigned long long out[4]);
int launch_update(Volume* v);
int launch_update_tr(Volume* v, vec3 light_dir, float alpha);
int ensure_tr(Volume* v, mrp_vec3 light_dir, float alpha);       // cached Update_TR
int ensure_hg_lut(Volume* v, float g, uint64_t seed);           // cached UpdateHGLut
int launch_update_hg_lut(Volume* v, float g, uint64_t seed);
// The volume's environment (the map if one is set, else the constant sky).
EnvMap env_map(const Volume* v, vec3 sky);
// The light of MRPNN sample n with an environment map: the sun or a map
// direction, chosen by a host draw on Philox stream (0, n, tag 6, seed)
// (`Render`, core/volume.cu:417-432,471-486); unchanged without a map.
void env_light_of_sample(const Volume* v, int n, uint64_t seed, mrp_vec3* light_dir, mrp_vec3* light_color);

}  // namespace mrp

struct mrp_volume {
    mrp::Volume v;
};
