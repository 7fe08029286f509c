// mrp_medium.h - leaf functions of the delta-tracking medium model, shared by
// the gfx950 kernels and the CPU oracle (same source, same bits).
//
// Each function states the reference function whose arithmetic it follows.
#pragma once

#include "mrp_math.h"

namespace mrp {

// Density grids live in HBM as FP32, z fastest, with a one-voxel zero border
// on every face: voxel (x,y,z) of an N^3 grid is stored at
// ((x+1)*(N+2) + (y+1))*(N+2) + (z+1).  The border makes the eight corner
// reads of a trilinear lookup branch-free for every cell index in [-1, N-1],
// which is exactly the set of cells a border-addressed texture can touch with
// a non-zero corner.
struct Grid {
    const float* data;
    int res;
};

// Majorant constants of the trackers.  The reference divides by the
// majorant SMax = maxDensity * alpha at every step (`t -= log(1-xi)/SMax`,
// `S/SMax`, core/render.cuh:101,109,145,158); here the three reciprocals are
// formed once and every step multiplies (a rounding-level difference; the
// reference's own -use_fast_math build divides approximately).
struct Majorant {
    float inv_smax;     // 1/SMax: free-flight distance scale
    float inv_smax_tr;  // tr_scale/SMax: shadow-ray distance scale
    float ratio;        // alpha/SMax: density -> collision probability
};
MRP_HD Majorant make_majorant(float max_density, float alpha, float tr_scale) {
    const float smax = max_density * alpha;
    Majorant m;
    m.inv_smax = 1.0f / smax;
    m.inv_smax_tr = tr_scale / smax;
    m.ratio = alpha / smax;
    return m;
}
MRP_HD size_t padded_index(int res, int x, int y, int z) {
    const size_t p = (size_t)res + 2;
    return ((size_t)(x + 1) * p + (size_t)(y + 1)) * p + (size_t)(z + 1);
}
MRP_HD size_t padded_count(int res) {
    const size_t p = (size_t)res + 2;
    return p * p * p;
}

// Trilinear fetch at texture coordinate uv in [0,1]^3 with texture-border
// (zero) addressing and exact FP32 weights.  Follows the host sampler
// `Sample(float*, int, float3 uv)` (core/volume.cu:685-703): texel coordinate
// uv*res - 0.5, lerp along z, then y, then x.  Returns 0 outside the support.
MRP_HD float trilinear_border_uv(const float* data, int res, float ux, float uy, float uz) {
    const float fr = (float)res;
    const float tx = ux * fr - 0.5f;
    const float ty = uy * fr - 0.5f;
    const float tz = uz * fr - 0.5f;
    if (!(tx >= -1.0f && tx < fr && ty >= -1.0f && ty < fr && tz >= -1.0f && tz < fr)) return 0.0f;
    const float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
    const float wx = tx - fx, wy = ty - fy, wz = tz - fz;
    const int x = (int)fx, y = (int)fy, z = (int)fz;
    const size_t P = (size_t)res + 2;
    const float* c = data + padded_index(res, x, y, z);
    const float c000 = c[0], c001 = c[1];
    const float c010 = c[P], c011 = c[P + 1];
    const float c100 = c[P * P], c101 = c[P * P + 1];
    const float c110 = c[P * P + P], c111 = c[P * P + P + 1];
    return lerp(lerp(lerp(c000, c001, wz), lerp(c010, c011, wz), wy),
                lerp(lerp(c100, c101, wz), lerp(c110, c111, wz), wy), wx);
}

// Corner-expanded cell layout of the tracer's density grid: cell (x,y,z) for
// x,y,z in [-1, res-1] holds the eight corners of its trilinear neighbourhood
// as two float4 (c000 c001 c010 c011 | c100 c101 c110 c111), cells z fastest.
// A lookup is then one 32-byte aligned read from a single cache line instead
// of four 8-byte reads from four lines (z-pair, y+1, x+1, x+1/y+1) of a
// linear grid; the price is 8x the footprint (4.3 GB for 512^3).
MRP_HD size_t cell_index(int res, int x, int y, int z) {
    const size_t c = (size_t)res + 1;
    return ((size_t)(x + 1) * c + (size_t)(y + 1)) * c + (size_t)(z + 1);
}

// Occupancy of the tracer's trilinear cells, by bricks: bit b of `bits`
// (bricks z fastest, n per axis) is set when some cell u = cell + 1 in brick
// b = u >> shift has a corner that is not zero.  A cell in a clear brick has
// eight zero corners, so its trilinear value is an exact zero, and the
// trackers use a density only through d * ratio > u (false for +-0 and
// u in (0,1]), d < 0 (false) and 1 - max(0, d * ratio) (exactly 1): a lookup
// there is answered without reading the grid, with the same draws, results
// and lookup count.  Over the 512^3 benchmark cloud at 1080p about 72 % of
// the first free flights' lookups and 45 % of the shadow rays' fall in clear
// 16^3 bricks.
// The kernels read the bits from LDS: staged once per workgroup (4.5 KB at
// 512^3), a test costs an LDS read; from global memory the extra dependent
// read before each cell load made the path tracer slower (36.5 against
// 33.1 ms per 16-spp 1080p launch) while at most halving its HBM traffic.
struct Occupancy {
    const uint32_t* bits;  // nullptr: every cell counts as occupied
    uint32_t shift, n;
};
MRP_HD uint32_t occupancy_shift(int res) { return res >= 256 ? 4u : (res >= 64 ? 3u : 2u); }
MRP_HD uint32_t occupancy_n(int res) { return ((uint32_t)res >> occupancy_shift(res)) + 1u; }
MRP_HD size_t occupancy_words(int res) {
    const size_t n = occupancy_n(res);
    return (n * n * n + 31) / 32;
}
#if defined(__HIPCC__)
// Copies the bits of `o` into the workgroup's LDS array `lds` (occupancy_words
// words) and returns an Occupancy that reads them there.  Every thread of the
// workgroup must call it (it ends with a barrier).
__device__ __forceinline__ Occupancy stage_occupancy(const Occupancy& o, int res, uint32_t* lds) {
    if (!o.bits) return o;  // uniform
    const uint32_t nw = (uint32_t)occupancy_words(res);
    for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) lds[w] = o.bits[w];
    __syncthreads();
    return Occupancy{lds, o.shift, o.n};
}
#endif
MRP_HD size_t cell_count(int res) {
    const size_t c = (size_t)res + 1;
    return c * c * c;
}

// Density at box position p from the cell layout: the same arithmetic as
// trilinear_border (same corners, same weights, same blend order).
#if defined(__HIPCC__)
// a * b + c with v_mad_u32_u24 (full rate; a, b < 2^24).  Left to itself the
// compiler turns a multiply-add of small values into the quarter-rate 64-bit
// v_mad_u64_u32.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// 32-bit cell indices on the device: resolutions are at most 1024
// (mrp_volume_create), so twice the largest index (2 * 1025^3) stays below
// 2^32 and every factor is under 2^24.
__device__ __forceinline__ uint32_t cell_index32(int res, int x, int y, int z) {
    const uint32_t c = (uint32_t)res + 1u;
    return mad_u24(mad_u24((uint32_t)(x + 1), c, (uint32_t)(y + 1)), c, (uint32_t)(z + 1));
}
__device__ __forceinline__ uint32_t clamp_cell_index32(int res, int x, int y, int z) {
    const uint32_t c = (uint32_t)res + 2u;
    return mad_u24(mad_u24((uint32_t)(x + 1), c, (uint32_t)(y + 1)), c, (uint32_t)(z + 1));
}
// Outside the cells the result is 0; branch-free (cell (-1,-1,-1) is read and
// discarded), so lanes of a wave need no divergent branch.
// Whether the cell with floor coordinates (fx, fy, fz) in [-1, res-1] may be
// nonzero (its brick's bit); a uniform test of the pointer skips the mask.
__device__ __forceinline__ bool cell_occupied(const Occupancy& o, float fx, float fy, float fz) {
    if (!o.bits) return true;
    const uint32_t ux = (uint32_t)(fx + 1.0f) >> o.shift, uy = (uint32_t)(fy + 1.0f) >> o.shift,
                   uz = (uint32_t)(fz + 1.0f) >> o.shift;
    const uint32_t b = mad_u24(mad_u24(ux, o.n, uy), o.n, uz);
    return (o.bits[b >> 5] >> (b & 31u)) & 1u;
}
__device__ __forceinline__ float trilinear_cells(const float4* __restrict__ cells, int res, vec3 p,
                                                 const Occupancy& occ = Occupancy{nullptr, 0u, 0u}) {
    const float fr = (float)res;
    const float tx = (p.x + 0.5f) * fr - 0.5f;
    const float ty = (p.y + 0.5f) * fr - 0.5f;
    const float tz = (p.z + 0.5f) * fr - 0.5f;
    const bool inr = tx >= -1.0f && tx < fr && ty >= -1.0f && ty < fr && tz >= -1.0f && tz < fr;
    // a point outside, or in a clear brick, reads cell (-1,-1,-1) and discards
    // it; the brick test sees only in-range coordinates
    const float gx = inr ? floorf(tx) : -1.0f, gy = inr ? floorf(ty) : -1.0f, gz = inr ? floorf(tz) : -1.0f;
    const bool in = inr && cell_occupied(occ, gx, gy, gz);
    const float fx = in ? gx : -1.0f, fy = in ? gy : -1.0f, fz = in ? gz : -1.0f;
    const float wx = tx - fx, wy = ty - fy, wz = tz - fz;
    const float4* c = cells + 2u * cell_index32(res, (int)fx, (int)fy, (int)fz);
    const float4 lo = c[0], hi = c[1];
    const float r = lerp(lerp(lerp(lo.x, lo.y, wz), lerp(lo.z, lo.w, wz), wy),
                         lerp(lerp(hi.x, hi.y, wz), lerp(hi.z, hi.w, wz), wy), wx);
    return in ? r : 0.0f;
}

// trilinear_cells in two halves, so that several lookups can be in flight:
// issue() starts the two loads, finish() blends (the same arithmetic; a point
// outside the cells gives +0).
struct CellsFetch {
    struct Token {
        float4 lo, hi;
        float wx, wy, wz;
        bool in;
    };
    const float4* __restrict__ cells;
    int res;
    Occupancy occ;
    // branch-free: a point outside, or in a clear brick, reads cell (-1,-1,-1)
    // (one line that stays in cache) and finish() discards it
    __device__ __forceinline__ Token issue(vec3 p) const {
        const float fr = (float)res;
        const float tx = (p.x + 0.5f) * fr - 0.5f;
        const float ty = (p.y + 0.5f) * fr - 0.5f;
        const float tz = (p.z + 0.5f) * fr - 0.5f;
        Token k;
        const bool inr = tx >= -1.0f && tx < fr && ty >= -1.0f && ty < fr && tz >= -1.0f && tz < fr;
        const float gx = inr ? floorf(tx) : -1.0f, gy = inr ? floorf(ty) : -1.0f, gz = inr ? floorf(tz) : -1.0f;
        k.in = inr && cell_occupied(occ, gx, gy, gz);  // the brick test sees only in-range coordinates
        const float fx = k.in ? gx : -1.0f, fy = k.in ? gy : -1.0f, fz = k.in ? gz : -1.0f;
        k.wx = tx - fx;
        k.wy = ty - fy;
        k.wz = tz - fz;
        const float4* c = cells + 2u * cell_index32(res, (int)fx, (int)fy, (int)fz);
        k.lo = c[0];
        k.hi = c[1];
        return k;
    }
    __device__ __forceinline__ float finish(const Token& k) const {
        const float4 lo = k.lo, hi = k.hi;
        const float r = lerp(lerp(lerp(lo.x, lo.y, k.wz), lerp(lo.z, lo.w, k.wz), k.wy),
                             lerp(lerp(hi.x, hi.y, k.wz), lerp(hi.z, hi.w, k.wz), k.wy), k.wx);
        return k.in ? r : 0.0f;
    }
};
#endif

// Quad layout of the tracer's density grid: quad (x,y,z), for x,z in
// [-1, res-1] and y in [-1, res], holds the corners (x..x+1) x (z..z+1) of
// row y as one float4 (d[x][y][z], d[x][y][z+1], d[x+1][y][z], d[x+1][y][z+1])
// of the padded grid; a lookup reads quads y and y+1 (two 16-byte loads,
// as for a cell).  Quads sit in bricks of 1 x 4 x 2 (x, y, z), 8 quads = one
// 128-byte line, bricks ordered z fastest, so a lookup's two quads share a
// line 3 times in 4, and the lines hold 4-fold rather than 8-fold copies.
// Over a 1080p frame of the 512^3 cloud the first free flights touch 53 bytes
// of distinct lines per lookup within one 8 x 8 pixel block (cells: 68).  The
// L2 keeps little of that reuse (its misses stay at cells' level), but the
// second quad's line is usually the first's: the first free flight takes
// 0.52-0.53 ms on quads against 0.56 ms on cells.  Shadow rays, which run
// mostly along +y, are slower on quads (0.29 against 0.26 ms) and read cells.
// The footprint is half the cells' (2.2 GB at 512^3).
// With u = coordinate + 1: ux, uz in [0, res], uy in [0, res+1].
constexpr uint32_t kQuadBrick = 8u;  // quads per brick
struct QuadGeom {
    uint32_t sx, sy;  // quads per step of ux and per brick step of uy (a brick step of uz is kQuadBrick)
};
MRP_HD QuadGeom quad_geom(int res) {
    QuadGeom g;
    g.sy = (((uint32_t)res + 2u) / 2u) * kQuadBrick;
    g.sx = (((uint32_t)res + 5u) / 4u) * g.sy;
    return g;
}
MRP_HD size_t quad_count(int res) { return ((size_t)res + 1) * quad_geom(res).sx; }
MRP_HD uint32_t quad_x(QuadGeom g, uint32_t ux) { return ux * g.sx; }
MRP_HD uint32_t quad_y(QuadGeom g, uint32_t uy) { return (uy >> 2) * g.sy + ((uy & 3u) << 1); }
MRP_HD uint32_t quad_z(uint32_t uz) { return (uz >> 1) * kQuadBrick + (uz & 1u); }

#if defined(__HIPCC__)
// CellsFetch on the quad layout: the same corners, weights and blend.
struct QuadsFetch {
    struct Token {
        float4 q0, q1;  // rows y and y+1: (c.00 c.01 c.10 c.11) over (x, z)
        float wx, wy, wz;
        bool in;
    };
    const float4* __restrict__ quads;
    int res;
    QuadGeom g;
    Occupancy occ;
    // branch-free: a point outside, or in a clear brick, reads cell (-1,-1,-1)
    // and finish() discards it
    __device__ __forceinline__ Token issue(vec3 p) const {
        const float fr = (float)res;
        const float tx = (p.x + 0.5f) * fr - 0.5f;
        const float ty = (p.y + 0.5f) * fr - 0.5f;
        const float tz = (p.z + 0.5f) * fr - 0.5f;
        Token k;
        const bool inr = tx >= -1.0f && tx < fr && ty >= -1.0f && ty < fr && tz >= -1.0f && tz < fr;
        const float gx = inr ? floorf(tx) : -1.0f, gy = inr ? floorf(ty) : -1.0f, gz = inr ? floorf(tz) : -1.0f;
        k.in = inr && cell_occupied(occ, gx, gy, gz);  // the brick test sees only in-range coordinates
        const float fx = k.in ? gx : -1.0f, fy = k.in ? gy : -1.0f, fz = k.in ? gz : -1.0f;
        k.wx = tx - fx;
        k.wy = ty - fy;
        k.wz = tz - fz;
        const uint32_t ux = (uint32_t)((int)fx + 1), uy = (uint32_t)((int)fy + 1), uz = (uint32_t)((int)fz + 1);
        const uint32_t xz = quad_x(g, ux) + quad_z(uz);
        k.q0 = quads[xz + quad_y(g, uy)];
        k.q1 = quads[xz + quad_y(g, uy + 1u)];
        return k;
    }
    __device__ __forceinline__ float finish(const Token& k) const {
        // corners c000 c001 c010 c011 | c100 c101 c110 c111 as in CellsFetch
        const float r = lerp(lerp(lerp(k.q0.x, k.q0.y, k.wz), lerp(k.q1.x, k.q1.y, k.wz), k.wy),
                             lerp(lerp(k.q0.z, k.q0.w, k.wz), lerp(k.q1.z, k.q1.w, k.wz), k.wy), k.wx);
        return k.in ? r : 0.0f;
    }
};
#endif

// Clamp-addressed cells: cell (x,y,z) for x,y,z in [-1, res] holds the eight
// clamp-to-edge corners of its neighbourhood in the order of trilinear_clamp,
// (res+2)^3 cells; used for the transmittance mips by the descriptor gather.
MRP_HD size_t clamp_cell_index(int res, int x, int y, int z) {
    const size_t c = (size_t)res + 2;
    return ((size_t)(x + 1) * c + (size_t)(y + 1)) * c + (size_t)(z + 1);
}
MRP_HD size_t clamp_cell_count(int res) {
    const size_t c = (size_t)res + 2;
    return c * c * c;
}
// Bilinear clamp cells of a W x H table: cell (x,y) for x in [-1, W], y in
// [-1, H] holds (d[y0][x0], d[y0][x1], d[y1][x0], d[y1][x1]), (W+2)(H+2) cells.
MRP_HD size_t lut_cell_count(int w, int h) { return ((size_t)w + 2) * ((size_t)h + 2); }

#if defined(__HIPCC__)
// trilinear_clamp on clamp-addressed cells: the same coordinates, weights and
// blend order, one 32-byte read.
__device__ __forceinline__ float trilinear_clamp_cells(const float4* __restrict__ cells, int res, vec3 uv) {
    const float fr = (float)res;
    float tx = uv.x * fr - 0.5f, ty = uv.y * fr - 0.5f, tz = uv.z * fr - 0.5f;
    tx = tx >= -1.0f ? fmin_(tx, fr) : -1.0f;
    ty = ty >= -1.0f ? fmin_(ty, fr) : -1.0f;
    tz = tz >= -1.0f ? fmin_(tz, fr) : -1.0f;
    const float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
    const float wx = tx - fx, wy = ty - fy, wz = tz - fz;
    const float4* c = cells + 2u * clamp_cell_index32(res, (int)fx, (int)fy, (int)fz);
    const float4 lo = c[0], hi = c[1];
    return lerp(lerp(lerp(lo.x, lo.y, wz), lerp(lo.z, lo.w, wz), wy), lerp(lerp(hi.x, hi.y, wz), lerp(hi.z, hi.w, wz), wy),
                wx);
}
// FP16 cells: the same corners rounded to FP16 (round to nearest even,
// saturated at +-65504), 16 bytes per cell; the same coordinates, weights and
// blend order as the FP32 cells.  Level m of the border cells of the density
// mips (256 >> m)^3 starts at mip16_offset(m) cells, level m of the clamp cells
// of the transmittance mips (128 >> m)^3 at tr16_offset(m).
// (m is at most 9 / 8; a fixed-bound loop with selects, so a per-thread level
// costs no divergent loop)
MRP_HD uint32_t mip16_offset(int m) {
    uint32_t o = 0u;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const uint32_t c = (uint32_t)(256 >> k) + 1u;
        o += k < m ? c * c * c : 0u;
    }
    return o;
}
MRP_HD uint32_t tr16_offset(int m) {
    uint32_t o = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t c = (uint32_t)(128 >> k) + 2u;
        o += k < m ? c * c * c : 0u;
    }
    return o;
}
__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    a = fmax_(-65504.0f, fmin_(a, 65504.0f));
    b = fmax_(-65504.0f, fmin_(b, 65504.0f));
    return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)a) |
           ((uint32_t)__builtin_bit_cast(unsigned short, (_Float16)b) << 16);
}
__device__ __forceinline__ float half_lo(uint32_t w) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }
__device__ __forceinline__ float blend_cell16(uint4 q, float wx, float wy, float wz) {
    return lerp(lerp(lerp(half_lo(q.x), half_hi(q.x), wz), lerp(half_lo(q.y), half_hi(q.y), wz), wy),
                lerp(lerp(half_lo(q.z), half_hi(q.z), wz), lerp(half_lo(q.w), half_hi(q.w), wz), wy), wx);
}
__device__ __forceinline__ float trilinear_cells16(const uint4* __restrict__ cells, int res, vec3 p) {
    const float fr = (float)res;
    const float tx = (p.x + 0.5f) * fr - 0.5f;
    const float ty = (p.y + 0.5f) * fr - 0.5f;
    const float tz = (p.z + 0.5f) * fr - 0.5f;
    const bool in = tx >= -1.0f && tx < fr && ty >= -1.0f && ty < fr && tz >= -1.0f && tz < fr;
    const float fx = in ? floorf(tx) : -1.0f, fy = in ? floorf(ty) : -1.0f, fz = in ? floorf(tz) : -1.0f;
    const float r = blend_cell16(cells[cell_index32(res, (int)fx, (int)fy, (int)fz)], tx - fx, ty - fy, tz - fz);
    return in ? r : 0.0f;
}
__device__ __forceinline__ float trilinear_clamp_cells16(const uint4* __restrict__ cells, int res, vec3 uv) {
    const float fr = (float)res;
    float tx = uv.x * fr - 0.5f, ty = uv.y * fr - 0.5f, tz = uv.z * fr - 0.5f;
    tx = tx >= -1.0f ? fmin_(tx, fr) : -1.0f;
    ty = ty >= -1.0f ? fmin_(ty, fr) : -1.0f;
    tz = tz >= -1.0f ? fmin_(tz, fr) : -1.0f;
    const float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
    return blend_cell16(cells[clamp_cell_index32(res, (int)fx, (int)fy, (int)fz)], tx - fx, ty - fy, tz - fz);
}

// bilinear_clamp on LUT cells: one 16-byte read.
__device__ __forceinline__ float bilinear_cells(const float4* __restrict__ cells, int w, int h, float u, float v) {
    float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
    tx = tx >= -1.0f ? fmin_(tx, (float)w) : -1.0f;
    ty = ty >= -1.0f ? fmin_(ty, (float)h) : -1.0f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float wx = tx - fx, wy = ty - fy;
    const float4 q = cells[(uint32_t)((int)fy + 1) * ((uint32_t)w + 2u) + (uint32_t)((int)fx + 1)];
    return lerp(lerp(q.x, q.y, wx), lerp(q.z, q.w, wx), wy);
}
#endif

// Density at box position p in [-0.5,0.5]^3: uv = p + 0.5 as in `Density`
// (core/render.cuh:67-71) and `MipDensityDynamic` (core/renderResources.cu:42-67).
MRP_HD float trilinear_border(const float* data, int res, vec3 p) {
    return trilinear_border_uv(data, res, p.x + 0.5f, p.y + 0.5f, p.z + 0.5f);
}

// Trilinear fetch with clamp-to-edge addressing on an unpadded N^3 grid (z
// fastest).  Follows `SampleClamp` (core/volume.cu:704-722), the CPU twin of
// the clamp-addressed transmittance textures (core/volume.cu:1412-1417).
MRP_HD float trilinear_clamp(const float* data, int res, vec3 uv) {
    const float fr = (float)res;
    float tx = uv.x * fr - 0.5f, ty = uv.y * fr - 0.5f, tz = uv.z * fr - 0.5f;
    // Clamp the coordinate first so that the integer conversion is defined for
    // any input (NaN maps to the low edge); inside [-1, res] this changes no value.
    tx = tx >= -1.0f ? fmin_(tx, fr) : -1.0f;
    ty = ty >= -1.0f ? fmin_(ty, fr) : -1.0f;
    tz = tz >= -1.0f ? fmin_(tz, fr) : -1.0f;
    const float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
    const float wx = tx - fx, wy = ty - fy, wz = tz - fz;
    const int x = (int)fx, y = (int)fy, z = (int)fz;
    const int hi = res - 1;
    const int x0 = x < 0 ? 0 : (x > hi ? hi : x), x1 = x + 1 < 0 ? 0 : (x + 1 > hi ? hi : x + 1);
    const int y0 = y < 0 ? 0 : (y > hi ? hi : y), y1 = y + 1 < 0 ? 0 : (y + 1 > hi ? hi : y + 1);
    const int z0 = z < 0 ? 0 : (z > hi ? hi : z), z1 = z + 1 < 0 ? 0 : (z + 1 > hi ? hi : z + 1);
    const size_t R = (size_t)res;
#define MRP_AT(a, b, c) data[((size_t)(a) * R + (size_t)(b)) * R + (size_t)(c)]
    const float v = lerp(lerp(lerp(MRP_AT(x0, y0, z0), MRP_AT(x0, y0, z1), wz),
                              lerp(MRP_AT(x0, y1, z0), MRP_AT(x0, y1, z1), wz), wy),
                         lerp(lerp(MRP_AT(x1, y0, z0), MRP_AT(x1, y0, z1), wz),
                              lerp(MRP_AT(x1, y1, z0), MRP_AT(x1, y1, z1), wz), wy),
                         wx);
#undef MRP_AT
    return v;
}

// Bilinear fetch with clamp addressing on a W x H table stored row-major
// (x fastest), normalised coordinates: the HG lookup table read by the
// descriptor gather (`tex2D(_HGLut, u, v)`, core/radiancePredict.cu:353-354,
// texture state core/volume.cu:1173-1178).
MRP_HD float bilinear_clamp(const float* data, int w, int h, float u, float v) {
    float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
    tx = tx >= -1.0f ? fmin_(tx, (float)w) : -1.0f;
    ty = ty >= -1.0f ? fmin_(ty, (float)h) : -1.0f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float wx = tx - fx, wy = ty - fy;
    const int x = (int)fx, y = (int)fy;
    const int x0 = x < 0 ? 0 : (x > w - 1 ? w - 1 : x), x1 = x + 1 < 0 ? 0 : (x + 1 > w - 1 ? w - 1 : x + 1);
    const int y0 = y < 0 ? 0 : (y > h - 1 ? h - 1 : y), y1 = y + 1 < 0 ? 0 : (y + 1 > h - 1 ? h - 1 : y + 1);
    return lerp(lerp(data[y0 * w + x0], data[y0 * w + x1], wx), lerp(data[y1 * w + x0], data[y1 * w + x1], wx), wy);
}

// 1/x per component with 0 -> 1e9: `inv` (core/vector.cu:19).
MRP_HD vec3 inv3(vec3 a) {
    return v3(a.x == 0.0f ? 999999999.0f : 1.0f / a.x, a.y == 0.0f ? 999999999.0f : 1.0f / a.y,
              a.z == 0.0f ? 999999999.0f : 1.0f / a.z);
}

// Entry distance into the slab box of half-size 0.4999, -1 on a miss:
// `RayBoxOffset` (core/vector.cu:41-58).
MRP_HD float ray_box_offset(vec3 p, vec3 dir) {
    const vec3 id = inv3(dir);
    const vec3 bmax = v3(0.4999f, 0.4999f, 0.4999f);
    const vec3 to_axis = (-p) * id;
    const vec3 to_face = bmax * id;
    const vec3 d0 = to_axis + to_face, d1 = to_axis - to_face;
    const vec3 tmin = vmin(d0, d1), tmax = vmax(d0, d1);
    const float tmi = fmax_(tmin.x, fmax_(tmin.y, tmin.z));
    const float tma = fmin_(tmax.x, fmin_(tmax.y, tmax.z));
    return tma >= tmi ? fmax_(tmi, 0.0f) : -1.0f;
}

// Exit distance from the unit box: `RayBoxDistance` (core/vector.cu:60-75).
MRP_HD float ray_box_distance(vec3 p, vec3 dir) {
    const vec3 id = inv3(dir);
    const vec3 bmax = v3(0.5f, 0.5f, 0.5f);
    const vec3 to_axis = (-p) * id;
    const vec3 to_face = bmax * id;
    const vec3 d0 = to_axis + to_face, d1 = to_axis - to_face;
    const vec3 tmax = vmax(d0, d1);
    return fmin_(tmax.x, fmin_(tmax.y, tmax.z));
}

// Henyey-Greenstein phase value without the 1/pi factor, as the reference
// defines it: (1-g^2) / (1+g^2-2g cos)^1.5 / 4 (core/vector.cu:205-209).
// d^1.5 is formed as d*sqrt(d).
MRP_HD float hg_phase(float cos_t, float g) {
    const float g2 = g * g;
    const float d = (1.0f + g2) - 2.0f * g * cos_t;
    return ((1.0f - g2) / (d * sqrtf(d))) * 0.25f;
}

// Inverse-CDF HG cosine: `SampleHeneyGreenstein` (core/vector.cu:211-218).
MRP_HD float hg_sample_cos(float s, float g) {
    if (fabs_(g) < 0.0001f) return s * 2.0f - 1.0f;
    const float g2 = g * g;
    const float t0 = (1.0f - g2) / ((1.0f - g) + 2.0f * g * s);
    return ((1.0f + g2) - t0 * t0) / (2.0f * g);
}

// HG direction about v from two uniforms (e0 -> cos, e1 -> azimuth):
// `SampleHenyeyGreenstein` (core/vector.cu:220-237).
MRP_HD vec3 hg_sample_dir(float e0, float e1, vec3 v, float g) {
    const float cos_t = hg_sample_cos(e0, g);
    const float phi = 6.28318548f * e1;
    const float sin_t = sqrtf(fmax_(0.0f, 1.0f - cos_t * cos_t));
    vec3 v2;
    if (fabs_(v.z) > 0.707106769f) {
        v2 = normalize(cross(v, v3(0.0f, 1.0f, 0.0f)));
    } else {
        v2 = normalize(cross(v, v3(0.0f, 0.0f, 1.0f)));
    }
    const vec3 w3 = cross(v2, v);
    float sp, cp;
    sincosf_(phi, &sp, &cp);
    return (w3 * (sin_t * cp)) + (v2 * (sin_t * sp)) + (v * cos_t);
}

// Uniform direction on the sphere: `UniformSampleSphere` (core/vector.cu:190-203).
MRP_HD vec3 uniform_sphere(float ex, float ey) {
    const float phi = 6.28318548f * ex;
    const float cos_t = 1.0f - 2.0f * ey;
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    float sp, cp;
    sincosf_(phi, &sp, &cp);
    return v3(sin_t * cp, sin_t * sp, cos_t);
}

// Pinhole ray through pixel (row, col) with jitter (xi_u, xi_v), the offline
// camera of `RenderCamera` (core/volume.cu:158-163): u = 1-(col+xi)/W,
// v = 1-(row+xi)/H, dir = normalize(forward + right(2u-1) + up(2v-1)).
MRP_HD vec3 camera_dir(vec3 forward, vec3 right, vec3 up, int row, int col, int w, int h, float xi_u, float xi_v) {
    const float u = 1.0f - ((float)col + xi_u) / (float)w;
    const float v = 1.0f - ((float)row + xi_v) / (float)h;
    const vec3 d = forward + (right * (u * 2.0f - 1.0f)) + (up * (v * 2.0f - 1.0f));
    return normalize(d);
}

// The same ray with the row convention selectable: the progressive camera of
// `RenderCamera(target, histo, ...)` (core/volume.cu:69-70) takes
// v = (row + xi)/h (row 0 at the bottom, the GL convention of its pixel
// buffer) where the offline camera takes v = 1 - (row + xi)/h.
MRP_HD vec3 camera_dir_rows(vec3 forward, vec3 right, vec3 up, int row, int col, int w, int h, float xi_u, float xi_v,
                            int rows_up) {
    const float u = 1.0f - ((float)col + xi_u) / (float)w;
    const float vv = ((float)row + xi_v) / (float)h;
    const float v = rows_up ? vv : 1.0f - vv;
    const vec3 d = forward + (right * (u * 2.0f - 1.0f)) + (up * (v * 2.0f - 1.0f));
    return normalize(d);
}

// Checkerboard rendering renders, in each call, the pixels whose row + column
// + flip is even (`RenderCamera`, core/volume.cu:61-62): logical column k of
// row r is image column 2k + ((r + flip) & 1).  checker < 0: every pixel.
MRP_HD int checker_col(int row, int col_l, int checker) {
    return checker < 0 ? col_l : 2 * col_l + ((row + checker) & 1);
}
MRP_HD bool checker_rendered(int row, int col, int checker) {
    return checker < 0 || ((row + col + checker) & 1) == 0;
}

// Work order of a w x h pixel grid in 8 x 8 blocks: index l in [0, w*h) ->
// (row, col).  Bands of 8 rows (the last may be lower) are split into blocks
// of 8 columns (the last may be narrower), each walked row by row, so 64
// consecutive items form one block: a wave's rays leave the camera close
// together.  A bijection for every w, h > 0.
MRP_HD void block_order(uint32_t l, uint32_t w, uint32_t h, int* row, int* col) {
    const uint32_t band = l / (8u * w);
    const uint32_t hb = h - 8u * band < 8u ? h - 8u * band : 8u;
    const uint32_t lp = l - band * 8u * w;
    const uint32_t bx = lp / (8u * hb);
    const uint32_t r = lp - bx * 8u * hb;
    const uint32_t wb = w - 8u * bx < 8u ? w - 8u * bx : 8u;
    *row = (int)(8u * band + r / wb);
    *col = (int)(8u * bx + r % wb);
}

// Random orthonormal frame whose first axis is v1: `GetMatrixFromNormal`
// (core/render.cuh:355-369).  The rejection loop is bounded (64 tries, then a
// fixed helper axis) so that degenerate input cannot hang a wavefront.
template <class RNG>
MRP_HD void frame_from_normal(RNG* rng, vec3 v1, vec3* a1, vec3* a2, vec3* a3) {
    v1 = normalize(v1);
    vec3 v2 = v3(0.0f, 0.0f, 0.0f), v3_ = v3(0.0f, 0.0f, 0.0f);
    bool ok = false;
    for (int it = 0; it < 64 && !ok; ++it) {
        const float ex = rng_next(rng);
        const float ey = rng_next(rng);
        const vec3 r = uniform_sphere(ex, ey);
        if (fabs_(dot(r, v1)) < 0.01f) continue;
        v2 = normalize(cross(v1, r));
        v3_ = cross(v2, v1);
        ok = true;
    }
    if (!ok) {
        const vec3 h = fabs_(v1.x) < 0.9f ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f);
        v2 = normalize(cross(v1, h));
        v3_ = cross(v2, v1);
    }
    *a1 = v1;
    *a2 = v2;
    *a3 = v3_;
}

// ------------------------------------------------------------- environment
// The environment radiance `SkyBox` (core/render.cuh:57-59): an equirectangular
// map, u = atan2(-dir.z, dir.x) / (2 pi) + 1/2, v = acos(dir.y) / pi, read as a
// normalised, linear-filtered, wrap-addressed texture (bilinear on texel
// centres, exact FP32 weights), times the exposure; the constant sky when no
// map is set.  Pixels are RGBA FP32, row v, column u.
struct EnvMap {
    const float* px;  // w * h texels of 4 floats (RGBA)
    int w, h;
    float exposure;
    vec3 sky;
};
MRP_HD float env_u(vec3 d) { return atan2f_(-d.z, d.x) * 0.159154937f + 0.5f; }
MRP_HD float env_v(vec3 d) { return acosf_(fmax_(fmin_(d.y, 1.0f), -1.0f)) * 0.318309873f; }
MRP_HD vec3 env_texel(const EnvMap& e, int x, int y) {
    const float* q = e.px + 4 * ((size_t)y * (size_t)e.w + (size_t)x);
    return v3(q[0], q[1], q[2]);
}
MRP_HD int wrap_index(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}
MRP_HD vec3 env_radiance(const EnvMap& e, vec3 dir) {
    if (!e.px) return e.sky;
    const float tx = env_u(dir) * (float)e.w - 0.5f, ty = env_v(dir) * (float)e.h - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float wx = tx - fx, wy = ty - fy;
    const int x0 = wrap_index((int)fx, e.w), x1 = wrap_index((int)fx + 1, e.w);
    const int y0 = wrap_index((int)fy, e.h), y1 = wrap_index((int)fy + 1, e.h);
    const vec3 a = env_texel(e, x0, y0), b = env_texel(e, x1, y0), c = env_texel(e, x0, y1),
               d = env_texel(e, x1, y1);
    const vec3 top = v3(lerp(a.x, b.x, wx), lerp(a.y, b.y, wx), lerp(a.z, b.z, wx));
    const vec3 bot = v3(lerp(c.x, d.x, wx), lerp(c.y, d.y, wx), lerp(c.z, d.z, wx));
    return v3(lerp(top.x, bot.x, wy), lerp(top.y, bot.y, wy), lerp(top.z, bot.z, wy)) * e.exposure;
}
// `Image_host::Sample` (core/volume.hpp:23-41): bilinear on texel centres with
// clamp addressing, no exposure (the light colour of a sampled direction).
MRP_HD vec3 env_sample_clamp(const EnvMap& e, float u, float v) {
    const float tx = u * (float)e.w - 0.5f, ty = v * (float)e.h - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float wx = tx - fx, wy = ty - fy;
    auto cl = [](int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); };
    const int x0 = cl((int)fx, e.w), x1 = cl((int)fx + 1, e.w), y0 = cl((int)fy, e.h), y1 = cl((int)fy + 1, e.h);
    const vec3 a = env_texel(e, x0, y0), b = env_texel(e, x1, y0), c = env_texel(e, x0, y1),
               d = env_texel(e, x1, y1);
    const vec3 top = v3(lerp(a.x, b.x, wx), lerp(a.y, b.y, wx), lerp(a.z, b.z, wx));
    const vec3 bot = v3(lerp(c.x, d.x, wx), lerp(c.y, d.y, wx), lerp(c.z, d.z, wx));
    return v3(lerp(top.x, bot.x, wy), lerp(top.y, bot.y, wy), lerp(top.z, bot.z, wy));
}
// `Roberts2` (core/vector.cu:174-178): the R2 low-discrepancy point n,
// frac((1/g, 1/g^2) n + 1/2) with g the plastic number.
MRP_HD void roberts2(int n, float* x, float* y) {
    const float g = 1.32471795724474602596f;  // FP32, as the reference declares it
    const float a1 = 1.0f / g, a2 = 1.0f / (g * g);
    const float px = a1 * (float)n + 0.5f, py = a2 * (float)n + 0.5f;
    *x = px - floorf(px);
    *y = py - floorf(py);
}
// The light of MRPNN sample (or progressive call) n with an environment map
// (`Render`, core/volume.cu:417-432,471-486): with probability rate = 4 exp /
// (4 exp + max(lightColor)) the sun is replaced by the map: the direction is
// the uniform-sphere image of Roberts2(n) and the colour the map there times
// 4 exp / rate; otherwise the sun's colour is divided by 1 - rate.  choose is
// the uniform draw of that decision.
MRP_HD void env_light(const EnvMap& e, int n, float choose, vec3* light_dir, vec3* light_color) {
    const float m = fmax_(light_color->x, fmax_(light_color->y, light_color->z));
    const float rate = e.exposure * 4.0f / (e.exposure * 4.0f + m);
    if (choose < rate) {
        float rx, ry;
        roberts2(n, &rx, &ry);
        const vec3 dir = uniform_sphere(rx, ry);
        *light_dir = dir;
        *light_color = env_sample_clamp(e, env_u(dir), env_v(dir)) * (e.exposure * 4.0f / rate);
    } else {
        *light_color = *light_color / (1.0f - rate);
    }
}

// Cube root of x in (0, 1] (the radius of a uniform point in a ball).
MRP_HD float cbrt01(float x) { return expf_(logf_(x) * (1.0f / 3.0f)); }

// Training points of one random ray, the sampling of `GetDesiredCountSample`
// and `MeanFreePathSample` (dataGen.cu:174-216): the ray starts on the sphere
// of radius 0.5, uniform, heading along normalize(r + normalize(-ori)) for a
// uniform direction r; up to 4 successive delta-tracking vertices (free
// flight, then an HG direction, `next_vertex`) are visited, and vertex i is
// kept if i = 0, if it lies more than 1/8 from the previous one, or with
// probability 0.1.  A kept point is moved by a uniform offset in the ball of
// radius 1/128 with probability 0.5, and its view direction is the ray's
// first direction with probability 0.75, else the direction it arrived
// along.  Draw order per ray: origin (2), r (2), then per vertex the free
// flight's draws, keep, jitter and view (3), and the offset (3) when moved.
// The reference draws with one host generator and reuses the step's draw as
// the acceptance test in its free flight (dataGen.cu:150-165); here every
// ray has its own stream and the free flight is the renderer's.
struct TrainPoint {
    vec3 pos, view;
    int kept;
};
template <class RNG, class NextVertex>
MRP_HD int training_ray_points(RNG* rng, const NextVertex& next_vertex, TrainPoint out[4]) {
    for (int i = 0; i < 4; ++i) {
        out[i].pos = v3(0.0f, 0.0f, 0.0f);
        out[i].view = v3(0.0f, 0.0f, 0.0f);
        out[i].kept = 0;
    }
    const float a0 = rng_next(rng);
    const float a1 = rng_next(rng);
    const vec3 ori = uniform_sphere(a0, a1) * 0.5f;
    const float b0 = rng_next(rng);
    const float b1 = rng_next(rng);
    const vec3 dir = normalize(uniform_sphere(b0, b1) + normalize(-ori));
    const float off = ray_box_offset(ori, dir);
    if (off < 0.0f) return 0;
    vec3 sp = ori + dir * off, rd = dir;
    int kept = 0;
    for (int i = 0; i < 4; ++i) {
        vec3 np, nd;
        if (!next_vertex(rng, sp, rd, ray_box_distance(sp, rd), &np, &nd)) break;
        const vec3 d = sp - np;
        const float u_keep = rng_next(rng);
        const float u_jit = rng_next(rng);
        const float u_view = rng_next(rng);
        if (i == 0 || dot(d, d) > 1.0f / 64.0f || u_keep > 0.9f) {
            vec3 p = np;
            if (u_jit > 0.5f) {
                const float j0 = rng_next(rng);
                const float j1 = rng_next(rng);
                const float j2 = rng_next(rng);
                p = np + uniform_sphere(j0, j1) * (cbrt01(j2) * (1.0f / 128.0f));
            }
            out[i].pos = p;
            out[i].view = u_view > 0.25f ? dir : rd;
            out[i].kept = 1;
            ++kept;
        }
        sp = np;
        rd = nd;
    }
    return kept;
}

// `CalculateRadiance` (core/render.cuh:200-258) for one given ray with
// `sampleNum` = spp paths drawn one after another from the same stream: the
// mean radiance and the distance to the first scattering vertex of the last
// path that scattered (-1 if none).  next_vertex / transmittance are the
// caller's free flight and `Tr` (the renderer's cells on the GPU, the grid in
// the oracle).
template <class RNG, class NextVertex, class Transmit>
MRP_HD vec3 radiance_paths(RNG* rng, const NextVertex& next_vertex, const Transmit& transmittance, vec3 ori,
                           vec3 dir, vec3 light_dir, vec3 light_color, const EnvMap& env, vec3 scatter_rate, float g,
                           int multi_scatter, int spp, float* first_t) {
    dir = normalize(dir);
    light_dir = normalize(light_dir);
    float t = -1.0f;
    const float dis = ray_box_offset(ori, dir);
    if (dis < 0.0f) {
        *first_t = t;
        return env_radiance(env, dir);
    }
    vec3 res = v3(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < spp; ++i) {
        vec3 sample_pos = ori + dir * dis;
        vec3 ray_dir = dir;
        vec3 csr = v3(1.0f, 1.0f, 1.0f);
        for (int k = 0; k < multi_scatter; ++k) {
            vec3 np, nd;
            if (!next_vertex(rng, sample_pos, ray_dir, ray_box_distance(sample_pos, ray_dir), &np, &nd)) {
                res = res + env_radiance(env, ray_dir) * csr;
                break;
            }
            if (k == 0) t = distance(ori, np);
            sample_pos = np;
            csr = csr * scatter_rate;
            const float phase = hg_phase(dot(ray_dir, light_dir), g);
            const float tr = transmittance(rng, sample_pos, light_dir, ray_box_distance(sample_pos, light_dir));
            res = res + light_color * csr * (tr * phase);
            ray_dir = nd;
        }
    }
    *first_t = t;
    return res / (float)spp;
}

// `ShadowTerm_TR` / `ShadowTerm_TRs` (core/render.cuh:317-341): the mean over
// spp estimates, drawn one after another from one stream, of Tr(pos -> light)
// * HG(dir . light, g); dir is used as given (the reference's GetTrs passes
// an offset direction that need not be unit).
template <class RNG, class Transmit>
MRP_HD float shadow_term_paths(RNG* rng, const Transmit& transmittance, vec3 pos, vec3 dir, vec3 light_dir, float g,
                               int spp) {
    light_dir = normalize(light_dir);
    const float dis = ray_box_distance(pos, light_dir);
    const float phase = hg_phase(dot(dir, light_dir), g);
    float res = 0.0f;
    for (int i = 0; i < spp; ++i) res += transmittance(rng, pos, light_dir, dis) * phase;
    return res / (float)spp;
}

}  // namespace mrp
