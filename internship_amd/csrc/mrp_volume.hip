// mrp_volume.hip - per-volume and per-light precompute on the GPU.
//
//   set_data   : host/HBM grid -> padded HBM grid (zero border)
//   update     : majorant (max density) + 9-level density mip pyramid
//                (`VolumeRender::Update`, core/volume.cu:1203-1316, with the
//                intended semantics of its broken mip loop)
//   update_tr  : 8 transmittance mips (`Fill_TR`/`Update_TR`, core/volume.cu:1320-1424)
//   hg lut     : cone-averaged HG table (`Fill_Hg`/`UpdateHGLut`, core/volume.cu:1082-1183)
//
// All of these are HBM-streaming kernels with one thread per output voxel; the
// grid and mips stay resident in HBM (a 512^3 grid and its pyramid take about
// 0.6 GB of the 288 GB) so no host round trip happens between stages.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "mrp_hglut.h"
#include "mrp_internal.h"
#include "mrp_procedural.h"

namespace mrp {

__global__ void fill_zero_kernel(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = 0.0f;
}

// src: res^3 unpadded (z fastest) -> dst: padded interior.
__global__ void pad_kernel(const float* __restrict__ src, int res, float* __restrict__ dst) {
    const size_t n = (size_t)res * res * res;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % res);
        const int y = (int)((i / res) % res);
        const int x = (int)(i / ((size_t)res * res));
        dst[padded_index(res, x, y, z)] = src[i];
    }
}

// padded -> unpadded res^3
__global__ void unpad_kernel(const float* __restrict__ src, int res, float* __restrict__ dst) {
    const size_t n = (size_t)res * res * res;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % res);
        const int y = (int)((i / res) % res);
        const int x = (int)(i / ((size_t)res * res));
        dst[i] = src[padded_index(res, x, y, z)];
    }
}

// cells[(x,y,z)] = the 8 corners (x..x+1, y..y+1, z..z+1) of the padded grid,
// for x,y,z in [-1, res-1].
__global__ void build_cells_kernel(const float* __restrict__ grid, int res, float4* __restrict__ cells) {
    const int c = res + 1;
    const size_t n = (size_t)c * c * c;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t P = (size_t)res + 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % c) - 1;
        const int y = (int)((i / c) % c) - 1;
        const int x = (int)(i / ((size_t)c * c)) - 1;
        const float* g = grid + padded_index(res, x, y, z);
        cells[2 * i] = make_float4(g[0], g[1], g[P], g[P + 1]);
        cells[2 * i + 1] = make_float4(g[P * P], g[P * P + 1], g[P * P + P], g[P * P + P + 1]);
    }
}

// quads = the padded grid in the quad layout (quad_x/y/z); slots past the
// last quad in the last bricks are 0.
__global__ void build_quads_kernel(const float* __restrict__ grid, int res, QuadGeom g, size_t n,
                                   float4* __restrict__ quads) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint32_t nbz = g.sy / kQuadBrick, nby = g.sx / g.sy;
    const size_t P = (size_t)res + 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t w = (uint32_t)(i % kQuadBrick);
        const size_t b = i / kQuadBrick;
        const uint32_t bz = (uint32_t)(b % nbz), by = (uint32_t)((b / nbz) % nby), ux = (uint32_t)(b / ((size_t)nbz * nby));
        const uint32_t uy = 4u * by + (w >> 1), uz = 2u * bz + (w & 1u);
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ux <= (uint32_t)res && uy <= (uint32_t)res + 1u && uz <= (uint32_t)res) {
            const float* c = grid + padded_index(res, (int)ux - 1, (int)uy - 1, (int)uz - 1);
            q = make_float4(c[0], c[1], c[P * P], c[P * P + 1]);
        }
        quads[i] = q;
    }
}

// Clamp cells of an unpadded res^3 grid (trilinear_clamp's corners) for cell
// coordinates in [-1, res].
__global__ void build_clamp_cells_kernel(const float* __restrict__ src, int res, float4* __restrict__ cells) {
    const int c = res + 2;
    const size_t n = (size_t)c * c * c;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t R = (size_t)res;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % c) - 1, y = (int)((i / c) % c) - 1, x = (int)(i / ((size_t)c * c)) - 1;
        const int hi = res - 1;
        const int x0 = x < 0 ? 0 : (x > hi ? hi : x), x1 = x + 1 > hi ? hi : x + 1;
        const int y0 = y < 0 ? 0 : (y > hi ? hi : y), y1 = y + 1 > hi ? hi : y + 1;
        const int z0 = z < 0 ? 0 : (z > hi ? hi : z), z1 = z + 1 > hi ? hi : z + 1;
#define MRP_AT(a, b, d) src[((size_t)(a) * R + (size_t)(b)) * R + (size_t)(d)]
        cells[2 * i] = make_float4(MRP_AT(x0, y0, z0), MRP_AT(x0, y0, z1), MRP_AT(x0, y1, z0), MRP_AT(x0, y1, z1));
        cells[2 * i + 1] = make_float4(MRP_AT(x1, y0, z0), MRP_AT(x1, y0, z1), MRP_AT(x1, y1, z0), MRP_AT(x1, y1, z1));
#undef MRP_AT
    }
}

// FP16 copy of n cells (two float4 each) as one uint4 each.
__global__ void cells_half_kernel(const float4* __restrict__ cells, size_t n, uint4* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 lo = cells[2 * i], hi = cells[2 * i + 1];
    out[i] = make_uint4(pack_half2(lo.x, lo.y), pack_half2(lo.z, lo.w), pack_half2(hi.x, hi.y), pack_half2(hi.z, hi.w));
}

// Bilinear clamp cells of a w x h table (x fastest) for cell coordinates in [-1, w] x [-1, h].
__global__ void build_lut_cells_kernel(const float* __restrict__ lut, int w, int h, float4* __restrict__ cells) {
    const int cw = w + 2, ch = h + 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cw * ch) return;
    const int x = i % cw - 1, y = i / cw - 1;
    const int x0 = x < 0 ? 0 : (x > w - 1 ? w - 1 : x), x1 = x + 1 > w - 1 ? w - 1 : x + 1;
    const int y0 = y < 0 ? 0 : (y > h - 1 ? h - 1 : y), y1 = y + 1 > h - 1 ? h - 1 : y + 1;
    cells[i] = make_float4(lut[y0 * w + x0], lut[y0 * w + x1], lut[y1 * w + x0], lut[y1 * w + x1]);
}

// max over the interior of a padded grid; result as float bits in *out
// (initialised to bits(1e-5f) > 0, so unsigned ordering equals float ordering).
__global__ __launch_bounds__(256) void max_kernel(const float* __restrict__ src, int res, unsigned* out) {
    __shared__ float red[256];
    const size_t n = (size_t)res * res * res;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    float m = 0.00001f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % res);
        const int y = (int)((i / res) % res);
        const int x = (int)(i / ((size_t)res * res));
        m = fmax_(m, src[padded_index(res, x, y, z)]);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax_(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, f2u(red[0]));
}

// dst (res r, padded) voxel (x,y,z) = trilinear border sample of src (res s,
// padded) at the voxel centre uv = (2i+1)/(2r): `Sample(source, source_res,
// float3{u,v,w})` as used by the mip loop (core/volume.cu:1238-1241).
__global__ void resample_kernel(const float* __restrict__ src, int s, float* __restrict__ dst, int r) {
    const size_t n = (size_t)r * r * r;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const float inv = 1.0f / (float)(2 * r);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % r);
        const int y = (int)((i / r) % r);
        const int x = (int)(i / ((size_t)r * r));
        const float u = (float)(2 * x + 1) * inv, v = (float)(2 * y + 1) * inv, w = (float)(2 * z + 1) * inv;
        dst[padded_index(r, x, y, z)] = trilinear_border_uv(src, s, u, v, w);
    }
}

// dst (res r) = mean of the 2x2x2 block of src (res 2r), summed dx, dy, dz
// nested as in core/volume.cu:1259-1275, divided by 8.
__global__ void box_kernel(const float* __restrict__ src, float* __restrict__ dst, int r) {
    const size_t n = (size_t)r * r * r;
    const int s = 2 * r;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % r);
        const int y = (int)((i / r) % r);
        const int x = (int)(i / ((size_t)r * r));
        float sum = 0.0f;
        for (int dx = 0; dx < 2; dx++)
            for (int dy = 0; dy < 2; dy++)
                for (int dz = 0; dz < 2; dz++) sum += src[padded_index(s, 2 * x + dx, 2 * y + dy, 2 * z + dz)];
        dst[padded_index(r, x, y, z)] = sum / 8.0f;
    }
}

// TR mip voxel: march 128 steps from the voxel centre to the box exit along
// the light through density mip (same resolution), T = exp(-sum * a * step).
__global__ void fill_tr_kernel(const float* __restrict__ mip, int r, float* __restrict__ tr, vec3 light,
                               float alpha_m) {
    const size_t n = (size_t)r * r * r;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % r);
        const int y = (int)((i / r) % r);
        const int x = (int)(i / ((size_t)r * r));
        const vec3 uv = v3(((float)x + 0.5f) / (float)r, ((float)y + 0.5f) / (float)r, ((float)z + 0.5f) / (float)r);
        const vec3 ori = uv - 0.5f;
        const float dis = ray_box_distance(ori, light);
        const float step = dis / 128.0f;
        vec3 lp = ori;
        float shadow = 0.0f;
        for (int k = 0; k < 128; k++) {
            lp = lp + light * step;
            shadow = shadow + trilinear_border(mip, r, lp);
        }
        tr[i] = expf_(-shadow * alpha_m * step);
    }
}

__global__ __launch_bounds__(64) void fill_hg_kernel(float* __restrict__ lut, const float* __restrict__ tables,
                                                     float g, float blend, uint32_t seed_lo, uint32_t seed_hi) {
    const int i = blockIdx.x * 8 + (threadIdx.x & 7);
    const int j = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (i >= kHgLutSize || j >= kHgLutSize) return;
    const float* halton = tables;
    const float angle = tables[4 * kHgLutSamples + j];
    u32x4 c;
    c.v[0] = (uint32_t)(j * kHgLutSize + i);
    c.v[1] = 0u;
    c.v[2] = 0u;
    c.v[3] = 2u;
    const float bn = u01_(philox4x32_10(c, seed_lo, seed_hi).v[0]);
    lut[j * kHgLutSize + i] = hg_lut_texel(i, angle, g, blend, bn, halton);
}

__global__ void cloud_kernel(const int* __restrict__ perm, CloudParams c, int res, float* __restrict__ out) {
    __shared__ int sp[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) sp[i] = perm[i];
    __syncthreads();
    const size_t n = (size_t)res * res * res;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int z = (int)(i % res);
        const int y = (int)((i / res) % res);
        const int x = (int)(i / ((size_t)res * res));
        out[i] = cloud_density(sp, c, res, x, y, z);
    }
}

// Brick occupancy (Occupancy, mrp_medium.h): one workgroup per brick scans
// the corners of the brick's cells, padded indices [b << shift, (b + 1) <<
// shift] per axis (cell u has its corners at u and u + 1), clipped to the
// padded grid, and sets the brick's bit if any corner is not zero (a -0 counts
// as zero: it blends to +-0 like +0).
__global__ __launch_bounds__(256) void occupancy_kernel(const float* __restrict__ grid, int res, uint32_t shift,
                                                        uint32_t n, uint32_t* __restrict__ bits) {
    const uint32_t b = blockIdx.x;
    const uint32_t bz = b % n, by = (b / n) % n, bx = b / (n * n);
    const uint32_t P = (uint32_t)res + 2u, B = 1u << shift;
    const uint32_t x0 = bx * B, y0 = by * B, z0 = bz * B;
    const uint32_t ex = min(x0 + B, P - 1u) + 1u - x0, ey = min(y0 + B, P - 1u) + 1u - y0,
                   ez = min(z0 + B, P - 1u) + 1u - z0;  // extents (>= 1: every brick starts inside the grid)
    const uint32_t cnt = ex * ey * ez;
    int any = 0;
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
        const uint32_t z = z0 + i % ez, y = y0 + (i / ez) % ey, x = x0 + i / (ez * ey);
        any |= grid[((size_t)x * P + y) * P + z] != 0.0f;
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) atomicOr(&bits[b >> 5], 1u << (b & 31u));
}

__global__ void clear_words_kernel(uint32_t* __restrict__ w, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 0u;
}

static unsigned blocks_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int launch_update(Volume* v) {
    const int res = v->res;
    // majorant
    unsigned* dmax = nullptr;
    MRP_HIP(hipMalloc(&dmax, sizeof(unsigned)));
    const unsigned init = f2u(0.00001f);
    MRP_HIP(hipMemcpyAsync(dmax, &init, sizeof(unsigned), hipMemcpyHostToDevice, v->stream));
    max_kernel<<<blocks_for((size_t)res * res * res) > 4096 ? 4096 : blocks_for((size_t)res * res * res), 256, 0,
                 v->stream>>>(v->grid, res, dmax);
    unsigned bits = 0;
    MRP_HIP(hipMemcpyAsync(&bits, dmax, sizeof(unsigned), hipMemcpyDeviceToHost, v->stream));
    MRP_HIP(hipStreamSynchronize(v->stream));
    hipFree(dmax);
    v->max_density = u2f(bits);
    build_cells_kernel<<<blocks_for(cell_count(res)), 256, 0, v->stream>>>(v->grid, res, v->cells);
    build_quads_kernel<<<blocks_for(quad_count(res)), 256, 0, v->stream>>>(v->grid, res, quad_geom(res),
                                                                            quad_count(res), v->quads);
    {
        const uint32_t on = occupancy_n(res);
        const size_t words = occupancy_words(res);
        clear_words_kernel<<<(unsigned)((words + 255) / 256), 256, 0, v->stream>>>(v->occ, words);
        occupancy_kernel<<<on * on * on, 256, 0, v->stream>>>(v->grid, res, occupancy_shift(res), on, v->occ);
    }

    // source chain: halve (by trilinear resampling) while above 2x256
    int r = 256;
    while (r * 2 < res) r <<= 1;
    const float* src = v->grid;
    int src_res = res;
    std::vector<float*> temps;
    for (; r > 256; r >>= 1) {
        float* tmp = nullptr;
        MRP_HIP(hipMalloc(&tmp, padded_count(r) * sizeof(float)));
        temps.push_back(tmp);
        fill_zero_kernel<<<blocks_for(padded_count(r)), 256, 0, v->stream>>>(tmp, padded_count(r));
        resample_kernel<<<blocks_for((size_t)r * r * r), 256, 0, v->stream>>>(src, src_res, tmp, r);
        src = tmp;
        src_res = r;
    }
    for (int m = 0; m < kNumMips; ++m)
        fill_zero_kernel<<<blocks_for(padded_count(mip_res(m))), 256, 0, v->stream>>>(v->mips[m],
                                                                                      padded_count(mip_res(m)));
    resample_kernel<<<blocks_for((size_t)256 * 256 * 256), 256, 0, v->stream>>>(src, src_res, v->mips[0], 256);
    for (int m = 1; m < kNumMips; ++m) {
        const int rm = mip_res(m);
        box_kernel<<<blocks_for((size_t)rm * rm * rm), 256, 0, v->stream>>>(v->mips[m - 1], v->mips[m], rm);
    }
    for (int m = 0; m < kNumMips; ++m) {
        build_cells_kernel<<<blocks_for(cell_count(mip_res(m))), 256, 0, v->stream>>>(v->mips[m], mip_res(m),
                                                                                      v->mip_cells[m]);
        const size_t n = cell_count(mip_res(m));
        cells_half_kernel<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(v->mip_cells[m], n,
                                                                               v->mip_cells16 + mip16_offset(m));
    }
    MRP_HIP(hipGetLastError());
    MRP_HIP(hipStreamSynchronize(v->stream));
    for (float* t : temps) hipFree(t);
    v->updated = true;
    v->tr_valid = false;
    return MRP_OK;
}

// 1.73^(m+1) by repeated FP32 multiplication (the reference uses powf).
static float tr_alpha_scale(int m) {
    float a = 1.73f;
    for (int k = 0; k < m; ++k) a *= 1.73f;
    return a;
}

int launch_update_tr(Volume* v, vec3 light, float alpha) {
    for (int m = 0; m < kNumTrMips; ++m) {
        const int r = tr_mip_res(m);
        const float am = alpha / tr_alpha_scale(m);
        fill_tr_kernel<<<blocks_for((size_t)r * r * r), 256, 0, v->stream>>>(v->mips[m + 1], r, v->tr_mips[m], light,
                                                                             am);
        build_clamp_cells_kernel<<<blocks_for(clamp_cell_count(r)), 256, 0, v->stream>>>(v->tr_mips[m], r,
                                                                                         v->tr_cells[m]);
        const size_t n = clamp_cell_count(r);
        cells_half_kernel<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(v->tr_cells[m], n,
                                                                               v->tr_cells16 + tr16_offset(m));
    }
    MRP_HIP(hipGetLastError());
    MRP_HIP(hipStreamSynchronize(v->stream));
    return MRP_OK;
}

int launch_update_hg_lut(Volume* v, float g, uint64_t seed) {
    std::vector<float> tables(4 * kHgLutSamples + kHgLutSize);
    halton_table(tables.data());
    for (int j = 0; j < kHgLutSize; ++j) {
        const float aa = ((float)j + 0.5f) / (float)kHgLutSize;
        tables[4 * kHgLutSamples + j] = (float)((double)aa * 3.1415926535 * 60.0 / 180.0);
    }
    float* dt = nullptr;
    MRP_HIP(hipMalloc(&dt, tables.size() * sizeof(float)));
    MRP_HIP(hipMemcpyAsync(dt, tables.data(), tables.size() * sizeof(float), hipMemcpyHostToDevice, v->stream));
    dim3 grid(kHgLutSize / 8, kHgLutSize / 8);
    fill_hg_kernel<<<grid, 64, 0, v->stream>>>(v->hg_lut, dt, g, hg_lut_blend(g), (uint32_t)seed,
                                               (uint32_t)(seed >> 32));
    const int nc = (int)lut_cell_count(kHgLutSize, kHgLutSize);
    build_lut_cells_kernel<<<(nc + 255) / 256, 256, 0, v->stream>>>(v->hg_lut, kHgLutSize, kHgLutSize, v->lut_cells);
    MRP_HIP(hipGetLastError());
    MRP_HIP(hipStreamSynchronize(v->stream));
    hipFree(dt);
    return MRP_OK;
}

int ensure_tr(Volume* v, mrp_vec3 light_dir, float alpha) {
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() has not run");
    if (!(alpha >= 0.0f)) return fail(MRP_E_INVALID, "alpha must be >= 0");
    const vec3 l = v3(light_dir.x, light_dir.y, light_dir.z);
    if (!(dot(l, l) > 0.0f)) return fail(MRP_E_INVALID, "light_dir must be non-zero");
    if (v->tr_valid && v->tr_light.x == light_dir.x && v->tr_light.y == light_dir.y && v->tr_light.z == light_dir.z &&
        v->tr_alpha == alpha)
        return MRP_OK;
    const int rc = launch_update_tr(v, normalize(l), alpha);
    if (rc) return rc;
    v->tr_valid = true;
    v->tr_light = light_dir;
    v->tr_alpha = alpha;
    return MRP_OK;
}

int ensure_hg_lut(Volume* v, float g, uint64_t seed) {
    if (!(g > -1.0f && g < 1.0f)) return fail(MRP_E_INVALID, "g must be in (-1, 1)");
    if (v->lut_valid && v->lut_g == g && v->lut_seed == seed) return MRP_OK;
    const int rc = launch_update_hg_lut(v, g, seed);
    if (rc) return rc;
    v->lut_valid = true;
    v->lut_g = g;
    v->lut_seed = seed;
    return MRP_OK;
}

}  // namespace mrp

using namespace mrp;

extern "C" {

int mrp_generate_cloud(const mrp_cloud_params* p, int res, float* out, int out_is_device) {
    if (!p || !out) return fail(MRP_E_INVALID, "NULL argument");
    if (res < 1 || res > 2048) return fail(MRP_E_INVALID, "resolution out of range");
    if (p->octaves < 0 || p->octaves > 16 || !(p->threshold < 1.0f)) return fail(MRP_E_INVALID, "bad cloud params");
    int perm[512];
    cloud_permutation(p->seed, perm);
    CloudParams c{p->seed, p->octaves, p->base_freq, p->lacunarity, p->gain, p->threshold};
    const size_t n = (size_t)res * res * res;
    int* dperm = nullptr;
    float* dst = out;
    float* tmp = nullptr;
    MRP_HIP(hipMalloc(&dperm, sizeof(perm)));
    hipError_t e = hipMemcpy(dperm, perm, sizeof(perm), hipMemcpyHostToDevice);
    if (e == hipSuccess && !out_is_device) {
        e = hipMalloc(&tmp, n * sizeof(float));
        dst = tmp;
    }
    if (e == hipSuccess) {
        cloud_kernel<<<blocks_for(n), 256>>>(dperm, c, res, dst);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && !out_is_device) e = hipMemcpy(out, tmp, n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dperm);
    if (tmp) hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "generate_cloud");
    return MRP_OK;
}

int mrp_volume_set_data(mrp_volume* h, const float* grid, int grid_is_device) {
    if (!h || !grid) return fail(MRP_E_INVALID, "NULL argument");
    Volume* v = &h->v;
    MRP_HIP(hipSetDevice(v->device));
    const size_t n = (size_t)v->res * v->res * v->res;
    const float* src = grid;
    float* tmp = nullptr;
    if (!grid_is_device) {
        MRP_HIP(hipMalloc(&tmp, n * sizeof(float)));
        MRP_HIP(hipMemcpyAsync(tmp, grid, n * sizeof(float), hipMemcpyHostToDevice, v->stream));
        src = tmp;
    }
    fill_zero_kernel<<<blocks_for(padded_count(v->res)), 256, 0, v->stream>>>(v->grid, padded_count(v->res));
    pad_kernel<<<blocks_for(n), 256, 0, v->stream>>>(src, v->res, v->grid);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(v->stream);
    if (tmp) hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "set_data");
    v->has_data = true;
    v->updated = false;
    v->tr_valid = false;
    return MRP_OK;
}

int mrp_volume_update(mrp_volume* h) {
    if (!h) return fail(MRP_E_INVALID, "NULL volume");
    if (!h->v.has_data) return fail(MRP_E_STATE, "mrp_volume_set_data() has not run");
    MRP_HIP(hipSetDevice(h->v.device));
    return launch_update(&h->v);
}

int mrp_volume_get_mip(const mrp_volume* h, int mip, float* host_out) {
    if (!h || !host_out) return fail(MRP_E_INVALID, "NULL argument");
    if (mip < 0 || mip >= kNumMips) return fail(MRP_E_INVALID, "mip out of range");
    const Volume* v = &h->v;
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() has not run");
    MRP_HIP(hipSetDevice(v->device));
    const int r = mip_res(mip);
    const size_t n = (size_t)r * r * r;
    float* tmp = nullptr;
    MRP_HIP(hipMalloc(&tmp, n * sizeof(float)));
    unpad_kernel<<<blocks_for(n), 256, 0, v->stream>>>(v->mips[mip], r, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, tmp, n * sizeof(float), hipMemcpyDeviceToHost, v->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v->stream);
    hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "get_mip");
    return MRP_OK;
}

int mrp_volume_get_occupancy(const mrp_volume* h, uint32_t* host_out, size_t max_words, int* shift, int* n) {
    if (!h || !shift || !n) return fail(MRP_E_INVALID, "NULL argument");
    const Volume* v = &h->v;
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() has not run");
    *shift = (int)occupancy_shift(v->res);
    *n = (int)occupancy_n(v->res);
    if (!host_out) return MRP_OK;  // size query
    const size_t words = occupancy_words(v->res);
    if (max_words < words) return fail(MRP_E_INVALID, "output buffer holds fewer than ceil(n^3/32) words");
    MRP_HIP(hipSetDevice(v->device));
    MRP_HIP(hipStreamSynchronize(v->stream));
    MRP_HIP(hipMemcpy(host_out, v->occ, words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return MRP_OK;
}

int mrp_volume_update_tr(mrp_volume* h, mrp_vec3 light_dir, float alpha) {
    if (!h) return fail(MRP_E_INVALID, "NULL volume");
    MRP_HIP(hipSetDevice(h->v.device));
    return ensure_tr(&h->v, light_dir, alpha);
}

int mrp_volume_get_tr_mip(const mrp_volume* h, int mip, float* host_out) {
    if (!h || !host_out) return fail(MRP_E_INVALID, "NULL argument");
    if (mip < 0 || mip >= kNumTrMips) return fail(MRP_E_INVALID, "mip out of range");
    const Volume* v = &h->v;
    if (!v->tr_valid) return fail(MRP_E_STATE, "mrp_volume_update_tr() has not run");
    MRP_HIP(hipSetDevice(v->device));
    const size_t r = (size_t)tr_mip_res(mip);
    MRP_HIP(hipMemcpy(host_out, v->tr_mips[mip], r * r * r * sizeof(float), hipMemcpyDeviceToHost));
    return MRP_OK;
}

int mrp_volume_update_hg_lut(mrp_volume* h, float g, uint64_t seed) {
    if (!h) return fail(MRP_E_INVALID, "NULL volume");
    MRP_HIP(hipSetDevice(h->v.device));
    return ensure_hg_lut(&h->v, g, seed);
}

int mrp_volume_get_hg_lut(const mrp_volume* h, float* host_out) {
    if (!h || !host_out) return fail(MRP_E_INVALID, "NULL argument");
    const Volume* v = &h->v;
    if (!v->lut_valid) return fail(MRP_E_STATE, "mrp_volume_update_hg_lut() has not run");
    MRP_HIP(hipSetDevice(v->device));
    MRP_HIP(hipMemcpy(host_out, v->hg_lut, (size_t)kLutSize * kLutSize * sizeof(float), hipMemcpyDeviceToHost));
    return MRP_OK;
}

}  // extern "C"
