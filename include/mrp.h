/* mrp.h - C ABI of the MI355X (gfx950) MRPNN rendering hot path.
 *
 * The reference exposes this path as the C++ classes `VolumeRender`
 * (core/volume.hpp:45-146) and `Camera` (core/camera.hpp:10-41), backed by
 * process-global CUDA textures and `__device__` symbols.  This library keeps
 * the same operations but puts all state in an opaque handle and passes every
 * per-call setting in a plain struct, so that several volumes can live in one
 * process and nothing crosses the boundary except pointers, sizes and PODs.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative MRP_E* code on
 *     failure; mrp_last_error() returns the message of the last failure on the
 *     calling thread.  (The reference prints CUDA errors and continues,
 *     core/volume.cu:15.)
 *   - Calls are synchronous unless documented otherwise: work is queued on the
 *     volume's own HIP stream and the call waits for it, as the reference
 *     synchronises before returning (core/volume.cu:455,539).
 *   - Grids are FP32, index (x*res + y)*res + z (z fastest), voxel centres at
 *     (i+0.5)/res, the layout of `VolumeRender::datas` (core/volume.cu:938-940).
 *   - Positions are in the box [-0.5, 0.5]^3.
 *   - "dev" pointers are device (HBM) pointers; "host" pointers are host memory.
 */
#ifndef MRP_H_
#define MRP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRP_OK 0
#define MRP_E_INVALID -1   /* bad argument */
#define MRP_E_HIP -2       /* HIP runtime error */
#define MRP_E_STATE -3     /* call order violated (e.g. render before update) */
#define MRP_E_NOMEM -4     /* allocation failed */
#define MRP_E_IO -5        /* file error */

typedef struct mrp_vec3 {
    float x, y, z;
} mrp_vec3; /* layout of CUDA float3 (12 bytes) */

/* Per-pixel histogram of the progressive renderer: 3 x 10 bins of r/(r+1),
 * sample count, running mean of l^2 and of l = r+g+b (compressed).
 * Same 132-byte POD as `Histogram` (core/volume.hpp:12-19). */
#define MRP_HISTO_SIZE 10
typedef struct mrp_histogram {
    float bin[MRP_HISTO_SIZE * 3];
    float total_sample_num;
    float x2;
    float x;
} mrp_histogram;

/* `VolumeRender::RenderType` (core/volume.hpp:88-90). RPNN is not provided. */
#define MRP_RENDER_PT 0
#define MRP_RENDER_MRPNN 2

/* Everything one offline render needs.  Fields mirror the arguments of
 * `VolumeRender::Render(int2 size, float3 ori, float3 up, float3 right,
 * float3 lightDir, RenderType rt, float g, float alpha, float3 lightColor,
 * int multiScatter, int sampleNum)` (core/volume.hpp:141) plus the device
 * globals its setters write (`scatter_rate`, `tr_scale`, `enviroment_exp`,
 * core/render.cuh:14-16). */
typedef struct mrp_render_params {
    int width, height;          /* image size (the reference is square-only) */
    mrp_vec3 ori, up, right;    /* camera: dir = normalize(-ori + right(2u-1) + up(2v-1)) */
    mrp_vec3 light_dir;         /* direction towards the light (normalised inside) */
    mrp_vec3 light_color;
    mrp_vec3 sky;               /* constant environment radiance returned for escaping rays */
    mrp_vec3 scatter_rate;      /* per-bounce albedo multiplier (SetScatterRate) */
    float g;                    /* Henyey-Greenstein asymmetry */
    float alpha;                /* extinction scale: sigma = alpha * density */
    float tr_scale;             /* step scale of the shadow-ray tracker (SetTrScale), 1 = unbiased */
    int multi_scatter;          /* maximum number of scattering events per path */
    int spp_begin;              /* first sample (frame) index */
    int spp_count;              /* number of samples (frames) */
    uint64_t seed;              /* Philox key */
    int render_type;            /* MRP_RENDER_PT or MRP_RENDER_MRPNN */
    int mlp_precision;          /* MRPNN only: 0 = FP32, 1 = BF16 MFMA, 2 = FP8 MFMA */
} mrp_render_params;

/* Counters and timings of the last trace (filled when a pointer is given). */
typedef struct mrp_trace_stats {
    uint64_t lookups;       /* trilinear density lookups (8 FP32 corners each) */
    uint64_t paths;         /* (pixel, sample) paths traced */
    uint64_t scatters;      /* scattering events (MRPNN: first vertices that scattered = descriptors evaluated) */
    uint64_t max_path_steps; /* loop iterations of the longest path (latency tail) */
    double kernel_ms;       /* device time of the tracing kernel(s) (MRPNN: of the network), HIP events */
    double reduce_ms;       /* device time of the accumulation kernel(s) (MRPNN: compaction + descriptor gather) */
    double total_ms;        /* device time of the whole call */
    double first_vertex_ms; /* MRPNN only: device time of the first-vertex kernel (free flights and frames) */
    int launches;           /* tracing kernel launches */
    int grid_blocks;        /* persistent workgroups per launch */
    double shadow_ms;       /* MRPNN only: shadow rays of the first vertices that scatter */
    double compact_ms;      /* MRPNN only: compaction of those vertices and the count's read-back */
    double gather_ms;       /* MRPNN only: the descriptor gather kernel */
    uint64_t shadow_lookups; /* MRPNN only: the part of `lookups` made by the shadow rays */
} mrp_trace_stats;

/* ------------------------------------------------------------------ errors */
const char* mrp_last_error(void);
int mrp_device_count(int* count);

/* ------------------------------------------------ device memory plumbing
 * For hosts without their own HIP allocator (the Python tests and bench). */
int mrp_malloc(void** dev, size_t bytes);
int mrp_free(void* dev);
int mrp_memcpy_h2d(void* dev_dst, const void* host_src, size_t bytes);
int mrp_memcpy_d2h(void* host_dst, const void* dev_src, size_t bytes);
int mrp_memset_zero(void* dev, size_t bytes);
int mrp_memcpy_d2d(void* dev_dst, const void* dev_src, size_t bytes);
/* The training data loader's shuffle (no counterpart in the reference, whose
 * trainer shuffles a CSV on the host, dataGen.cu:335-337): dst row r = src row
 * idx_dev[r] for r < n_rows, rows of row_bytes (a multiple of 16) in HBM; an
 * index >= n_src_rows gives a row of zeros.  Queued on the null stream
 * (asynchronous). */
int mrp_gather_rows(void* dst_dev, const void* src_dev, const uint32_t* idx_dev, int n_rows, int row_bytes,
                    int n_src_rows);

/* ------------------------------------------------------------ measurement
 * No counterpart in the reference: the GPU's ceiling for random reads of
 * `width` bytes (16, 32 or 128) at hashed positions in a `buffer_bytes`
 * buffer (a power of two, 1 MiB to 16 GiB), `iters` reads per thread, 4.2M
 * threads; best of two timed launches.  It is the access shape of a trilinear
 * cell lookup, the bound the tracer's lookup rate is read against. */
int mrp_measure_random_reads(int width, uint64_t buffer_bytes, int iters, double* reads_per_s);

/* ------------------------------------------------------------------ volume */
typedef struct mrp_volume mrp_volume;

/* `VolumeRender(int resolution)` + `MallocMemory` (core/volume.cu:748-773,831-833). */
int mrp_volume_create(int resolution, mrp_volume** out);
/* `~VolumeRender` (core/volume.cu:913-936). */
int mrp_volume_destroy(mrp_volume* vol);
int mrp_volume_resolution(const mrp_volume* vol, int* res);
/* Copy res^3 FP32 densities in; the result of `SetDatas`/`SetData`
 * (core/volume.cu:938-944).  `grid_is_device` selects host or HBM source. */
int mrp_volume_set_data(mrp_volume* vol, const float* grid, int grid_is_device);
/* `VolumeRender::Update` (core/volume.cu:1203-1316): upload, majorant
 * (max density, floor 1e-5) and the 9-level density mip pyramid 256^3..1^3,
 * built on the GPU with the intended semantics (level 0 = trilinear resample
 * of the grid at 256^3 voxel centres, level m+1 = 2x2x2 mean of level m). */
int mrp_volume_update(mrp_volume* vol);
/* `max_density` after Update (core/volume.cu:1219-1226). */
int mrp_volume_max_density(const mrp_volume* vol, float* out);
/* Copy density mip `mip` ((256>>mip)^3 FP32, z fastest) to host memory. */
int mrp_volume_get_mip(const mrp_volume* vol, int mip, float* host_out);
/* No counterpart in the reference (whose tracer reads every density): the
 * brick occupancy Update builds for the trackers.  Cells c in [-1, res-1]^3
 * (corners c..c+1 of the zero-bordered grid) are grouped in bricks of
 * 2^shift cells per axis, u = c + 1, brick (ux>>shift, uy>>shift, uz>>shift),
 * n = (res >> shift) + 1 per axis, z fastest; bit b of word b/32 is set when
 * some corner of the brick's cells is not zero.  A lookup in a clear brick is
 * an exact zero and reads no grid memory.  Copies ceil(n^3/32) words to
 * host_out (NULL: sizes only). */
int mrp_volume_get_occupancy(const mrp_volume* vol, uint32_t* host_out, size_t max_words, int* shift, int* n);
/* `VolumeRender::Update_TR` (core/volume.cu:1320-1424), GPU path: 8
 * transmittance mips (128>>m)^3 marched 128 steps toward the light through
 * density mip m+1 with extinction alpha/1.73^(m+1). */
int mrp_volume_update_tr(mrp_volume* vol, mrp_vec3 light_dir, float alpha);
int mrp_volume_get_tr_mip(const mrp_volume* vol, int mip, float* host_out);
/* `VolumeRender::UpdateHGLut` + `Fill_Hg` (core/volume.cu:1082-1183): 512x512
 * cone-averaged HG table, 1024 scrambled-Halton samples per texel. */
int mrp_volume_update_hg_lut(mrp_volume* vol, float g, uint64_t seed);
int mrp_volume_get_hg_lut(const mrp_volume* vol, float* host_out /* 512*512 */);

/* `SetHDRI` + `SetEnvExp` (core/volume.cu:1426-1462): an environment map for
 * `SkyBox` (core/render.cuh:57-59) in place of the constant sky: width x
 * height RGBA FP32 texels (row-major, row 0 at the top: v = acos(dir.y)/pi,
 * u = atan2(-dir.z, dir.x)/(2 pi) + 1/2), read bilinearly with wrap
 * addressing and scaled by exposure.  Escaping paths of every renderer and
 * `GetSample` read it, and MRPNN renders choose per sample (per call for
 * progressive frames) between the sun and a direction of the map as the
 * reference's Render does (core/volume.cu:417-432,471-486): with probability
 * 4 exposure / (4 exposure + max(lightColor)) the light becomes the direction
 * of the R2 point of the sample index mapped to the sphere with the map's
 * colour there times 4 exposure / probability, else the sun's colour is
 * divided by 1 - probability (the draw: Philox stream (0, sample, tag 6,
 * seed)).  width = 0 (or rgba_host NULL) removes the map. */
int mrp_volume_set_hdri(mrp_volume* vol, const float* rgba_host, int width, int height, float exposure);
/* Radiance RGBE ".hdr" file (the reference's `load_hdr_float4`,
 * core/hdr_loader.h:249) into width x height RGBA FP32 texels (alpha 1):
 * with rgba_out NULL only width and height are returned; max_texels bounds
 * the output.  Flat and run-length-encoded scanlines; host only. */
int mrp_read_hdr(const char* path, float* rgba_out, size_t max_texels, int* width, int* height);

/* ------------------------------------------------------- synthetic inputs */
/* Seeded Perlin-fBm cloud (this project's own generator; the reference's
 * procedural volume, genall.cu:32-67, is a non-compiling stub):
 * density = clamp((v - threshold)/(1 - threshold), 0, 1) * sphere falloff,
 * v = (fBm + 1)/2 at voxel centres.  Writes res^3 FP32 (z fastest) to `out`
 * (HBM if out_is_device, else host memory).  Same bits as the CPU oracle. */
typedef struct mrp_cloud_params {
    uint64_t seed;
    int octaves;
    float base_freq, lacunarity, gain, threshold;
} mrp_cloud_params;
int mrp_generate_cloud(const mrp_cloud_params* p, int res, float* out, int out_is_device);

/* ------------------------------------------------------------------ render */
/* Offline render to a host image: `VolumeRender::Render(int2 size, ...)`
 * (core/volume.cu:410-467).  rgb_host receives width*height*3 FP32, row-major
 * (row 0 first), the running mean over samples [spp_begin, spp_begin+spp_count).
 * stats may be NULL. */
int mrp_render_offline(mrp_volume* vol, const mrp_render_params* p, float* rgb_host, mrp_trace_stats* stats);

/* Same computation, everything resident in HBM.
 *   target_dev : width*height*3 FP32 running mean, updated in place (must hold
 *                the mean of samples [0, spp_begin) on entry; zeros if spp_begin = 0)
 *   sum64_dev  : optional width*height*3 FP64 per-pixel sums, accumulated
 *                (+=) over the samples of this call (the high-spp oracle image)
 * This is the hot path the throughput benchmark times. */
int mrp_render_offline_dev(mrp_volume* vol, const mrp_render_params* p, float* target_dev, double* sum64_dev,
                           mrp_trace_stats* stats);

/* One progressive frame: `VolumeRender::Render(float3* target, Histogram*
 * histo, unsigned* target2, int2 size, ..., int randseed, RenderType rt,
 * int toneType, bool denoise)` (core/volume.cu:469-547), all buffers
 * caller-owned in HBM (the reference's GUI passes device pointers and a mapped
 * pixel buffer, GUI.hpp:138-168):
 *   target_dev : width*height*3 FP32; PT: running mean; MRPNN: the frame
 *                blended with the reprojected previous frame
 *   histo_dev  : width*height histograms; PT: per-pixel statistics (cleared
 *                per pixel at its first frame); MRPNN: the previous frame
 *   bgra_dev   : width*height 0xAARRGGBB after exposure and tone curve
 * r.spp_begin is the frame number (`randseed`: 0 restarts the running means),
 * r.spp_count must be 1.  The volume keeps the state the reference keeps in
 * globals: the checkerboard parity (toggles every call), a call counter that
 * selects the random stream (`rand_cpu`), the previous camera (reprojection)
 * and whether the previous frame was MRPNN (histograms are cleared when
 * switching to MRPNN).  Rows follow the reference's progressive camera:
 * v = (row + xi)/height, row 0 at the bottom. */
typedef struct mrp_frame_params {
    mrp_render_params r;
    int tone;          /* 0 none, 1 gamma 2.2, 2 ACES */
    int denoise;       /* PT: histogram-distance filter of noisy pixels */
    int checkerboard;  /* `SetCheckboard`: render one checkerboard half per call */
    float exposure;    /* `SetExposure` */
} mrp_frame_params;
int mrp_render_frame(mrp_volume* vol, const mrp_frame_params* p, float* target_dev, mrp_histogram* histo_dev,
                     uint32_t* bgra_dev, mrp_trace_stats* stats);
/* Forget the progressive state (parity 0, call counter 0, no previous camera). */
int mrp_volume_reset_frames(mrp_volume* vol);
/* The progressive state: checkerboard parity of the next call and calls made. */
int mrp_volume_frame_state(const mrp_volume* vol, int* flip, uint32_t* calls);

/* ------------------------------------------------------ batched serving */
/* Several scenes and cameras rendered as one batch: the MRPNN image of every
 * view (`NNPredict<MRPNN>`, core/render.cuh:469-523, as mrp_render_offline
 * with render_type MRP_RENDER_MRPNN), a delta-tracking fallback for the pixels
 * whose sample statistics say the prediction is noisy, and the histogram
 * denoise and BGRA8 pack of the frame path (`Denoise<true>`,
 * core/volume.cu:208-264).  The reference renders one scene per call; this
 * entry point has no counterpart there and serves callers that batch scenes
 * (a render service).
 *   volumes[k], views[k] : view k's volume (updated, no environment map) and
 *                parameters; every view has the same width and height
 *                (width*height a multiple of 64), and the same g, albedo and
 *                mlp_precision (one network evaluates the batch: volumes[0]'s);
 *                views of one volume share its light, alpha and seed.
 *                spp_begin and spp_count are ignored; multi_scatter bounds the
 *                fallback paths.
 *   per pixel:  samples 0 .. mrpnn_spp-1 are MRPNN samples (Philox stream
 *                (pixel, s, 0, seed), as mrp_render_offline), folded into the
 *                running mean and the pixel histogram of the frame path; if the
 *                histogram's |E[l^2] - E[l]^2| exceeds var_threshold, the pixel
 *                is re-estimated from scratch (mean and histogram) with the
 *                delta-tracking samples mrpnn_spp .. mrpnn_spp+fallback_spp-1
 *                (`CalculateRadiance`, core/volume.cu:17-34), all fallback
 *                pixels of all views in one persistent launch.
 *   rgb_dev    : n_views*width*height*3 FP32, the estimate before the denoise
 *   bgra_dev   : n_views*width*height 0xAARRGGBB after denoise (if set),
 *                exposure and tone curve (NULL: no pack)
 *   flags_dev  : optional n_views*width*height bytes, 1 for fallback pixels
 * stats may be NULL. */
typedef struct mrp_serve_params {
    int mrpnn_spp;        /* MRPNN samples per pixel, >= 1 (2 or more give the variance a meaning) */
    float var_threshold;  /* fallback when the pixel's histogram variance exceeds this */
    int fallback_spp;     /* delta-tracking samples of a fallback pixel, >= 1 */
    int denoise;          /* 1: `Denoise<true>` on the estimate before the pack */
    int tone;             /* 0 none, 1 gamma 2.2, 2 ACES */
    float exposure;
} mrp_serve_params;
typedef struct mrp_serve_stats {
    uint64_t mrpnn_lookups;     /* first free flights and shadow rays */
    uint64_t fallback_lookups;  /* delta-tracking paths of the fallback */
    uint64_t descriptors;       /* first vertices that scattered */
    uint64_t fallback_pixels;
    float total_ms;             /* device time of the call (HIP events) */
    float mrpnn_ms, accumulate_ms, fallback_ms, denoise_ms;
} mrp_serve_stats;
int mrp_serve_views(mrp_volume* const* volumes, const mrp_render_params* views, int n_views,
                    const mrp_serve_params* sp, float* rgb_dev, uint32_t* bgra_dev, uint8_t* flags_dev,
                    mrp_serve_stats* stats);

/* Tone curve on host colours (n RGB triples): `Camera::None/Gamma/ACES`
 * (core/camera.cu:66-160, core/render.cuh:640-732); tone 0 none, 1 gamma,
 * 2 ACES.  The same arithmetic as the frame path's device tone map. */
int mrp_tone_map_host(const float* rgb_host, size_t n, int tone, float* out_host);

/* Cap (bytes, default 4 GiB) of the per-volume buffer that holds the
 * per-(sample, pixel) results of one batch; larger sample counts are split
 * into batches. */
int mrp_set_sample_buffer_budget(mrp_volume* vol, size_t bytes);

/* ------------------------------------------------------ training targets */
/* One query of `GetSamples` (core/volume.cu:550-626): a scattering point, the
 * direction the first HG lobe is drawn about, the unit light direction, and the
 * medium (extinction scale, HG g, albedo). 48 bytes. */
typedef struct mrp_sample_query {
    mrp_vec3 pos;
    mrp_vec3 dir;
    mrp_vec3 light_dir;
    float alpha, g, scatter;
} mrp_sample_query;
/* `GetSample` (core/render.cuh:276-315) for n queries in HBM: the mean over
 * spp paths of the multiple-scattering radiance reaching the point (next-event
 * light at every vertex, albedo 1.001*scatter per bounce, sky on escape), the
 * high-order part the network is trained to predict.  out_dev: n x 3 FP32.
 * Path k of query j draws from Philox stream (j, k, tag 1). */
int mrp_get_samples(mrp_volume* vol, int n, const mrp_sample_query* q_dev, mrp_vec3 light_color, mrp_vec3 sky,
                    int multi_scatter, int spp, uint64_t seed, float* out_dev, mrp_trace_stats* stats);

/* Training points: the sampling of `GetDesiredCountSample` /
 * `MeanFreePathSample` (dataGen.cu:174-216) on the GPU.  n_rays rays start
 * uniformly on the sphere of radius 0.5, heading along normalize(r +
 * normalize(-origin)) for a uniform direction r, and visit up to 4 successive
 * delta-tracking vertices (extinction scale alpha, HG g); a vertex is kept if
 * it is the first, lies more than 1/8 from the previous one, or with
 * probability 0.1; a kept point moves by a uniform offset in the ball of
 * radius 1/128 with probability 0.5, and its view direction is the ray's
 * first direction with probability 0.75, else its arrival direction.  Ray r
 * draws from Philox stream (r, 0, tag 2, seed).  Outputs in HBM, 4 slots per
 * ray: pos_dev 4 n_rays float4 (position, 1 if the slot holds a point else
 * 0), view_dev 4 n_rays float4 (unit view direction, 0).  lookups (may be
 * NULL): trilinear density lookups made. */
int mrp_sample_training_points(mrp_volume* vol, int n_rays, float alpha, float g, uint64_t seed, float* pos_dev,
                               float* view_dev, uint64_t* lookups);

/* `GetRadiances` (core/volume.cu:329-402, `CalculateRadianceMulti`,
 * core/volume.cu:17-34) on n given rays (HBM, n x 3 FP32 origins and
 * directions).  The render parameters give the light, medium, sky and albedo;
 * their camera and image fields are ignored.  render_type PT: p->spp_count
 * paths per ray (`CalculateRadiance`'s sampleNum, core/render.cuh:200-258),
 * one after another on Philox stream (ray, 0, tag 4, seed).  MRPNN:
 * `NNPredict` once per ray (the reference ignores sampleNum there) on stream
 * (ray, 0, tag 0, seed), through the renderer's stages with p->mlp_precision.
 * out_dev: n x 4 FP32 (radiance, distance to the first scattering vertex or
 * -1).  stats may be NULL. */
int mrp_get_radiances(mrp_volume* vol, int n, const float* ori_dev, const float* dir_dev,
                      const mrp_render_params* p, float* out_dev, mrp_trace_stats* stats);
/* `GetTr` / `GetTrs` (core/volume.cu:302-323,636-673, `ShadowTerm_TR(s)`,
 * core/render.cuh:317-341) on n points (HBM, n x 3 FP32 positions and
 * directions): out_dev[i] = the mean over spp estimates of Tr(pos -> light) *
 * HG(dir . light, g) (dir as given), drawn one after another on Philox stream
 * (i, 0, tag 5, seed).  The reference's GetTrs scales this by lightColor * pi
 * and GetTr rounds sampleNum up to a multiple of 32; the C++ facade does
 * both.  lookups may be NULL. */
int mrp_get_shadow_terms(mrp_volume* vol, int n, const float* pos_dev, const float* dir_dev, mrp_vec3 light_dir,
                         float alpha, float g, int spp, uint64_t seed, float* out_dev, uint64_t* lookups);

/* ---------------------------------------------------------------- network */
/* Trained MRPNN weights: 49,673 FP32 values, the 146 arrays of
 * core/NNWeight.cuh concatenated in header order (row-major [out][in]).
 * The handle keeps an FP32 copy and a BF16 MFMA-operand packing in HBM.
 * Rendering with render_type = MRP_RENDER_MRPNN uses the network set on the
 * volume; the reference compiles its weights in (`#include "NNWeight.cuh"`,
 * core/radiancePredict.cu:305). */
typedef struct mrp_network mrp_network;
int mrp_network_create(const float* weights_host, size_t count, mrp_network** out);
int mrp_network_destroy(mrp_network* net);
int mrp_volume_set_network(mrp_volume* vol, const mrp_network* net);

/* Descriptor gather on caller-given points (the gather of `RadiancePredict`,
 * core/radiancePredict.cu:319-368), all pointers in HBM:
 *   pos_act  n x float4 (position, active flag)   frame_* n x float4 (X, Y, Z axes in .xyz)
 *   feat     n x 576 FP32 [log-density 192 | transmittance 192 | log-phase 192]
 *   gamma    n FP32, acos(X . LX)
 * Builds/refreshes the transmittance mips (light_dir, alpha) and the HG table
 * (g, seed) first, as `Render` does (core/volume.cu:434-437). */
int mrp_mrpnn_features(mrp_volume* vol, int n, const float* pos_act_dev, const float* frame_x_dev,
                       const float* frame_y_dev, const float* frame_z_dev, mrp_vec3 light_dir, float alpha, float g,
                       uint64_t seed, float* feat_dev, float* gamma_dev);
/* The gather as the BF16/FP8 renderer runs it (the same part of
 * `RadiancePredict`, core/radiancePredict.cu:319-368), on caller-given points:
 *   tiles  ceil(n/32) x 36,864 bytes: the 576 features of each descriptor
 *          rounded to BF16 and laid out as the MFMA network's B operands
 *          (tiles of 32 descriptors; per stream k (density, transmittance,
 *          phase) and 16-feature chunk c, the 16-byte chunk of lane 32 kg + d
 *          holds features 16c + 8h + 4kg + e, h, e in 0..1 / 0..3, of
 *          descriptor d); items past n in the last tile are not written
 *   pool   n x 72 FP32: the squeeze-excitation blocks' per-segment means
 *          (0..35) and maxima (36..71), from the FP32 features (`AVGMAX`,
 *          core/radiancePredict.cu:21-33)
 *   gamma  n FP32
 * The density and transmittance lookups read the cell format chosen with
 * mrp_volume_set_gather_cells (FP32 unless set). */
int mrp_mrpnn_features_tiled(mrp_volume* vol, int n, const float* pos_act_dev, const float* frame_x_dev,
                             const float* frame_y_dev, const float* frame_z_dev, mrp_vec3 light_dir, float alpha,
                             float g, uint64_t seed, void* tiles_dev, float* pool_dev, float* gamma_dev);
/* Cell format of the BF16/FP8 renderers' gather: 32 (default) reads the FP32
 * corners of the density and transmittance mips, as the reference's FP32
 * textures hold them (core/volume.cu:1289-1298,1412-1417); 16 reads FP16
 * copies (one 16-byte load per lookup instead of two). The FP32 renderer always
 * reads FP32 cells. No counterpart in the reference. */
int mrp_volume_set_gather_cells(mrp_volume* vol, int bits);
/* The network on n descriptors (HBM): out[n][3] = max(exp(P_c)-1, 0)*srp_c,
 * srp = (scatter_rate/1.001)^4 per channel.  precision 0 = FP32, 1 = BF16 MFMA,
 * 2 = FP8-E4M3 MFMA (power-of-two scales: per layer for the weights, per layer
 * and descriptor for the activations; FP32 accumulation, FP32 bias). */
int mrp_mlp_forward(const mrp_network* net, int n, const float* feat_dev, const float* gamma_dev, float g,
                    mrp_vec3 srp, int precision, float* out_dev);

/* The MFMA network (precision 1 = BF16, 2 = FP8) on descriptors in the tiled
 * layout of mrp_mrpnn_features_tiled, with their pooled statistics: the
 * second half of `RadiancePredict` (core/radiancePredict.cu:370-486) as the
 * renderer runs it. out_dev: n x 3 FP32. */
int mrp_mlp_forward_tiled(const mrp_network* net, int n, const void* tiles_dev, const float* pool_dev,
                          const float* gamma_dev, float g, mrp_vec3 srp, int precision, float* out_dev);

/* mrp_mlp_forward without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The BF16/FP8
 * networks keep the descriptors' pooled statistics in a fixed 18 MiB device
 * array of the library (one per GPU, held until the process ends) instead of a
 * buffer allocated per call; n is limited to 65,536 descriptors (larger n
 * returns MRP_E_INVALID).  Calls on one GPU share that array: calls that
 * overlap on one GPU race.  Same kernels and results as mrp_mlp_forward. */
int mrp_mlp_forward_noalloc(const mrp_network* net, int n, const float* feat_dev, const float* gamma_dev, float g,
                            mrp_vec3 srp, int precision, float* out_dev);

/* --------------------------------------------------------------- training */
/* Training of the MRPNN network on the GPU (forward, backward, Adam), derived
 * from the inference network (core/radiancePredict.cu:370-486); the
 * reference trains with Adam, lr 1e-3 (train/train.py:536-546).  A sample is
 * the gather's 576 features (feat_dev, batch x 576 FP32) plus aux_dev
 * (batch x float4: gamma = acos(X . LX), g, srp = (scatter/1.001)^4, target).
 * The prediction is p = max(exp(P)-1, 0) * srp for one channel; the loss is
 * the batch mean of (p - target)^2 / (target^2 + 0.01), reduced in FP64.
 * Weight gradients are FP64; master weights and Adam moments are FP32. */
typedef struct mrp_trainer mrp_trainer;
int mrp_trainer_create(const float* weights_host, size_t count, int max_batch, mrp_trainer** out);
int mrp_trainer_destroy(mrp_trainer* t);
/* Arithmetic of the forward and backward products (train/train.py:220-254
 * trains in FP32): 0 = FP32 (default), 1 = BF16 matrix cores - BF16 copies of
 * the FP32 master weights made before each gradient, activations rounded to
 * BF16 at each product, FP32 accumulation and records; weight gradients,
 * loss and Adam are the same for both. */
int mrp_trainer_set_precision(mrp_trainer* t, int precision);
/* Floats of per-sample activation record (HBM: record_floats * max_batch * 4 bytes). */
int mrp_trainer_record_floats(const mrp_trainer* t, int* out);
/* Loss and gradient of the current weights on a batch (no update); grad_host
 * (49,673 doubles) may be NULL. */
int mrp_trainer_gradient(mrp_trainer* t, int batch, const float* feat_dev, const float* aux_dev, double* grad_host,
                         double* loss_out);
/* Gradient + one Adam step (beta1 0.9, beta2 0.999, eps 1e-8). loss_out: the
 * batch loss before the step. */
int mrp_trainer_step(mrp_trainer* t, int batch, const float* feat_dev, const float* aux_dev, float lr,
                     double* loss_out);
int mrp_trainer_get_weights(const mrp_trainer* t, float* host_out);
/* Predictions p of the last gradient/step call (batch floats). */
int mrp_trainer_predictions(const mrp_trainer* t, int batch, float* host_out);

#ifdef __cplusplus
}
#endif

#endif /* MRP_H_ */
