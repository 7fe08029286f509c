// mrp_track.h - delta tracking and ratio tracking with several density
// lookups in flight per thread.
//
// In both trackers the next tentative position depends only on the random
// draws (t -= log(1 - u) / majorant); a density value only decides whether
// the walk stops, or moves t when it is negative (t -= d).  So a thread can
// draw the next kTrackBatch steps ahead, issue all their lookups, and then
// consume the results in order.  The counter-based stream makes the look-ahead
// free: the draws are taken from a copy of the stream, and the committed state
// is set back (rng_seek) to exactly the draws the one-step-at-a-time loop
// would have consumed.  Results, draw counts and lookup counts are identical
// to the sequential loops below, which state the reference semantics:
//   free_flight_seq   : `DeterminateNextVertex` with an unused direction
//                       (core/render.cuh:138-172) as NNPredict's first vertex uses it;
//   transmittance_seq : `Tr` (core/render.cuh:93-136): ratio tracking while
//                       tr > 0.5, then delta tracking while tr > 0.001.
// Both cap the walk at max_steps loop tests (10000 in the reference); the
// counter is advanced by every loop test, as `loop_num++ < 10000 && ...` does.
//
// Fetch: issue(vec3) -> Fetch::Token starts a lookup, finish(Token) -> float
// completes it (the kernels split a trilinear cell fetch this way).
#pragma once

#include "mrp_math.h"

namespace mrp {

template <class Fetch>
MRP_HD bool free_flight_seq(Rng* rng, const Fetch& F, float inv_smax, float ratio, vec3 pos, vec3 dir, float dis,
                            int max_steps, vec3* npos, uint32_t* lk) {
    float t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < max_steps) {
        float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * inv_smax;
        if (t > dis) return false;
        rk = rng_next(rng);
        const float d = F.finish(F.issue(pos + (dir * t)));
        *lk += 1u;
        if (d * ratio > rk) break;
        if (d < 0.0f) t -= d;
    }
    (void)rng_next(rng);  // the HG direction draws (direction unused)
    (void)rng_next(rng);
    *npos = (dir * t) + pos;
    return true;
}

template <class Fetch>
MRP_HD float transmittance_seq(Rng* rng, const Fetch& F, float inv_smax_tr, float ratio, vec3 pos, vec3 dir,
                               float dis, int max_steps, uint32_t* lk) {
    float tr = 1.0f, t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < max_steps && tr > 0.5f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * inv_smax_tr;
        if (t > dis) break;
        const float d = F.finish(F.issue(pos + (dir * t)));
        *lk += 1u;
        tr *= 1.0f - fmax_(0.0f, d * ratio);
        if (d < 0.0f) t -= d;
    }
    while (loop_num++ < max_steps && tr > 0.001f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * inv_smax_tr;
        if (t > dis) break;
        const float d = F.finish(F.issue(pos + (dir * t)));
        *lk += 1u;
        if (rng_next(rng) < fmax_(0.0f, d * ratio)) {
            tr = 0.0f;
            break;
        }
        if (d < 0.0f) t -= d;
    }
    return tr;
}

// free_flight_seq with up to B lookups in flight.
template <int B, class Fetch>
MRP_HD bool free_flight_batched(Rng* rng, const Fetch& F, float inv_smax, float ratio, vec3 pos, vec3 dir, float dis,
                                int max_steps, vec3* npos, uint32_t* lk) {
    float t = 0.0f;
    int steps = 0;  // loop tests passed so far (iterations started)
    for (;;) {
        if (steps >= max_steps) break;  // the loop test fails: accept at t
        const uint32_t n0 = rng->n;
        Rng sp = *rng;
        float ts[B], ua[B];
        typename Fetch::Token tok[B];
        int m = 0;
        bool escape = false;
        float tc = t;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (escape || steps + j >= max_steps) break;
            const float rk = rng_next(&sp);
            tc -= logf_(1.0f - rk) * inv_smax;
            if (tc > dis) {
                escape = true;
                break;
            }
            ua[j] = rng_next(&sp);
            ts[j] = tc;
            tok[j] = F.issue(pos + (dir * tc));
            ++m;
        }
        // (no array is indexed with a run-time value, so ts/ua/tok stay in registers)
        int event = 0;  // 1: accepted at step j, 2: negative density at step j
        int j = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (event || q >= m) break;
            const float d = F.finish(tok[q]);
            *lk += 1u;
            j = q;
            if (d * ratio > ua[q]) {
                event = 1;
                t = ts[q];
            } else if (d < 0.0f) {
                event = 2;
                t = ts[q] - d;
            }
        }
        if (event == 1) {
            rng_seek(rng, n0 + 2u * (uint32_t)(j + 1));
            break;
        }
        if (event == 2) {
            steps += j + 1;
            rng_seek(rng, n0 + 2u * (uint32_t)(j + 1));
            continue;
        }
        steps += m;
        if (escape) {
            rng_seek(rng, n0 + 2u * (uint32_t)m + 1u);
            return false;
        }
        t = tc;     // the last speculated position (t itself when m == 0)
        *rng = sp;  // exactly 2m draws consumed
    }
    (void)rng_next(rng);
    (void)rng_next(rng);
    *npos = (dir * t) + pos;
    return true;
}

// transmittance_seq with up to B lookups in flight.
template <int B, class Fetch>
MRP_HD float transmittance_batched(Rng* rng, const Fetch& F, float inv_smax_tr, float ratio, vec3 pos, vec3 dir,
                                   float dis, int max_steps, uint32_t* lk) {
    float tr = 1.0f, t = 0.0f;
    int loop_num = 0;  // the reference's counter
    // ---- ratio tracking: one draw per step
    for (;;) {
        if (!(loop_num < max_steps && tr > 0.5f)) {
            ++loop_num;  // the failing test still increments
            break;
        }
        const uint32_t n0 = rng->n;
        Rng sp = *rng;
        float ts[B];
        typename Fetch::Token tok[B];
        int m = 0;
        bool escape = false;
        float tc = t;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (escape || loop_num + j >= max_steps) break;
            const float rk = rng_next(&sp);
            tc -= logf_(1.0f - rk) * inv_smax_tr;
            if (tc > dis) {
                escape = true;
                break;
            }
            ts[j] = tc;
            tok[j] = F.issue(pos + (dir * tc));
            ++m;
        }
        // commit step by step; the loop test precedes each step
        int done = 0;  // steps committed
        bool stop = false, restart = false;
        float tdone = t;  // position of the last committed step
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (stop || restart || q >= m) break;
            if (!(tr > 0.5f)) {  // test of step q fails: leave ratio tracking before its draw
                stop = true;
                break;
            }
            const float d = F.finish(tok[q]);
            *lk += 1u;
            tr *= 1.0f - fmax_(0.0f, d * ratio);
            done = q + 1;
            tdone = ts[q];
            if (d < 0.0f) {
                t = ts[q] - d;
                restart = true;
            }
        }
        loop_num += done;
        if (restart) {
            rng_seek(rng, n0 + (uint32_t)done);
            continue;
        }
        t = tdone;
        if (stop) {
            ++loop_num;
            rng_seek(rng, n0 + (uint32_t)done);
            break;
        }
        if (escape) {
            // the escaping step's test: it runs only if the walk is still in ratio tracking
            if (loop_num < max_steps && tr > 0.5f) return tr;  // t > dis: delta tracking breaks at once too
            ++loop_num;
            rng_seek(rng, n0 + (uint32_t)done);
            break;
        }
        *rng = sp;  // exactly m draws consumed
    }
    // ---- delta tracking: two draws per step; tr stays until a collision zeroes it
    if (!(tr > 0.001f)) return tr;
    for (;;) {
        if (loop_num >= max_steps) return tr;
        const uint32_t n0 = rng->n;
        Rng sp = *rng;
        float ts[B], ua[B];
        typename Fetch::Token tok[B];
        int m = 0;
        bool escape = false;
        float tc = t;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (escape || loop_num + j >= max_steps) break;
            const float rk = rng_next(&sp);
            tc -= logf_(1.0f - rk) * inv_smax_tr;
            if (tc > dis) {
                escape = true;
                break;
            }
            ts[j] = tc;
            tok[j] = F.issue(pos + (dir * tc));
            ua[j] = rng_next(&sp);
            ++m;
        }
        bool restart = false;
        int done = 0;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (restart || q >= m) break;
            const float d = F.finish(tok[q]);
            *lk += 1u;
            done = q + 1;
            if (ua[q] < fmax_(0.0f, d * ratio)) return 0.0f;
            if (d < 0.0f) {
                t = ts[q] - d;
                restart = true;
            }
        }
        loop_num += done;
        if (restart) {
            rng_seek(rng, n0 + 2u * (uint32_t)done);
            continue;
        }
        if (escape) return tr;
        t = tc;  // the last speculated position (t itself when m == 0)
        *rng = sp;
    }
}

}  // namespace mrp
