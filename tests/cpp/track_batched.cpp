// The batched trackers (several lookups in flight, reading their draws from
// the stream's ring of generated words) against the sequential loops that
// state the reference semantics (one draw at a time from the plain stream),
// on random rays through a density field with negative regions, for several
// batch widths and step caps: the result, the lookup count, the stream
// position and the draws that follow must be identical; a seek back over the
// free flight's last two draws must give them again.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mrp_medium.h"
#include "mrp_track.h"

using namespace mrp;

struct Field {
    struct Token {
        vec3 p;
    };
    float scale, bias;
    Token issue(vec3 p) const { return Token{p}; }
    float finish(Token t) const {
        const vec3 p = t.p;
        if (p.x < -0.5f || p.y < -0.5f || p.z < -0.5f || p.x > 0.5f || p.y > 0.5f || p.z > 0.5f) return 0.0f;
        return scale * (std::sin(7.0f * p.x) * std::cos(5.0f * p.y) * std::sin(3.0f * p.z + 1.0f)) + bias;
    }
};

static bool same(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

template <int B>
static int run(const Field& F, int max_steps, int rays, uint64_t seed) {
    int bad = 0;
    for (int r = 0; r < rays; ++r) {
        Rng g = rng_make((uint32_t)r, 7u, 3u, seed);
        const vec3 pos = v3(rng_next(&g) - 0.5f, rng_next(&g) - 0.5f, rng_next(&g) - 0.5f);
        const vec3 dir = normalize(v3(rng_next(&g) - 0.5f, rng_next(&g) - 0.5f, rng_next(&g) - 0.5f));
        const float dis = 0.2f + 1.5f * rng_next(&g);
        const float inv_smax = 1.0f / (8.0f + 200.0f * rng_next(&g)), ratio = 0.5f + 2.0f * rng_next(&g);
        uint32_t wb[kRingWords];
        {
            Rng a = rng_make((uint32_t)r, 1u, 0u, seed);
            RingRng<1> b = ring_make<1>(wb, (uint32_t)r, 1u, 0u, seed);
            vec3 pa = v3(0, 0, 0), pb = v3(0, 0, 0);
            uint32_t la = 0, lb = 0;
            const bool ra = free_flight_seq(&a, F, inv_smax, ratio, pos, dir, dis, max_steps, &pa, &la);
            const bool rb = free_flight_batched<B>(&b, F, inv_smax, ratio, pos, dir, dis, max_steps, &pb, &lb);
            if (ra != rb || la != lb || (ra && (a.n != b.n || !same(pa.x, pb.x) || !same(pa.y, pb.y) || !same(pa.z, pb.z)))) {
                if (bad < 5) std::printf("free flight B=%d cap=%d ray %d: %d/%d lk %u/%u n %u/%u\n", B, max_steps, r, ra, rb, la, lb, a.n, b.n);
                ++bad;
            }
            if (ra) {
                Rng c = a;  // the free flight's two direction draws, again
                rng_seek(&c, a.n - 2u);
                rng_seek(&b, b.n - 2u);
                for (int k = 0; k < 2; ++k)
                    if (rng_next(&c) != rng_next(&b)) ++bad;
                if (rng_next(&a) != rng_next(&b)) ++bad;  // the stream continues identically
            }
        }
        {
            Rng a = rng_make((uint32_t)r, 2u, 0u, seed);
            RingRng<1> b = ring_make<1>(wb, (uint32_t)r, 2u, 0u, seed);
            if (r & 1) {  // a stream resumed part-way, as the shadow rays resume the first vertex's
                const uint32_t start = 5u + (uint32_t)r % 11u;
                rng_seek(&a, start);
                rng_seek(&b, start);
            }
            uint32_t la = 0, lb = 0;
            const float ta = transmittance_seq(&a, F, inv_smax, ratio, pos, dir, dis, max_steps, &la);
            const float tb = transmittance_batched<B>(&b, F, inv_smax, ratio, pos, dir, dis, max_steps, &lb);
            if (!same(ta, tb) || la != lb || a.n != b.n) {
                if (bad < 5)
                    std::printf("transmittance B=%d cap=%d ray %d: %.9g/%.9g lk %u/%u n %u/%u\n", B, max_steps, r, ta, tb,
                                la, lb, a.n, b.n);
                ++bad;
            }
            if (rng_next(&a) != rng_next(&b)) ++bad;  // the stream continues identically
        }
    }
    return bad;
}

// block_order must visit every pixel of a w x h grid exactly once.
static int block_order_bad() {
    int bad = 0;
    const uint32_t ws[] = {1, 3, 7, 8, 9, 17, 64, 960, 961}, hs[] = {1, 5, 8, 13, 16, 90};
    for (uint32_t w : ws)
        for (uint32_t h : hs) {
            static unsigned char seen[961 * 90];
            std::memset(seen, 0, sizeof(seen));
            for (uint32_t l = 0; l < w * h; ++l) {
                int row, col;
                block_order(l, w, h, &row, &col);
                if (row < 0 || col < 0 || (uint32_t)row >= h || (uint32_t)col >= w || seen[row * w + col]++) ++bad;
            }
        }
    int row, col;  // 64 consecutive items of an aligned full block are one 8 x 8 block
    block_order(64u * 5u + 17u, 64u, 16u, &row, &col);
    if (row != 2 || col != 40 + 1) ++bad;
    return bad;
}

int main() {
    if (const int b = block_order_bad()) {
        std::printf("block_order: %d bad pixels\n", b);
        return 1;
    }
    const Field fields[3] = {{0.8f, 0.3f}, {0.3f, 0.05f}, {1.5f, 0.9f}};  // negative regions, thin, dense
    const int caps[5] = {1, 3, 7, 40, 10000};
    long bad = 0, cases = 0;
    for (const Field& F : fields)
        for (int cap : caps) {
            const int rays = cap == 10000 ? 20000 : 4000;
            bad += run<1>(F, cap, rays, 11) + run<2>(F, cap, rays, 12) + run<3>(F, cap, rays, 13) +
                   run<4>(F, cap, rays, 14) + run<5>(F, cap, rays, 15) + run<6>(F, cap, rays, 16);
            cases += 6L * rays;
        }
    std::printf("%ld rays x 2 trackers, %ld mismatches\n", cases, bad);
    return bad == 0 ? 0 : 1;
}
