// mrp_track.h - delta tracking and ratio tracking with several density
// lookups in flight per thread.
//
// In both trackers the next tentative position depends only on the random
// draws (t -= log(1 - u) / majorant); a density value only decides whether
// the walk stops, or moves t when it is negative (t -= d).  So a thread can
// draw the next B steps ahead, issue all their lookups, and then consume the
// results in order.  The counter-based stream makes the look-ahead free: a
// round makes the words of its window in the stream's ring at one site
// (rng_fill), reads them by index, and then sets the stream to exactly the
// draws the one-step-at-a-time loop would have consumed.  Results, draw
// counts and lookup counts are identical to the sequential loops below, which
// state the reference semantics:
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

template <class R, class Fetch>
MRP_HD bool free_flight_seq(R* rng, const Fetch& F, float inv_smax, float ratio, vec3 pos, vec3 dir, float dis,
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

template <class R, class Fetch>
MRP_HD float transmittance_seq(R* rng, const Fetch& F, float inv_smax_tr, float ratio, vec3 pos, vec3 dir,
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
template <int B, uint32_t kStride, class Fetch>
MRP_HD bool free_flight_batched(RingRng<kStride>* rng, const Fetch& F, float inv_smax, float ratio, vec3 pos,
                                vec3 dir, float dis, int max_steps, vec3* npos, uint32_t* lk) {
    static_assert(2 * B + 3 <= (int)kRingWords, "a round's draws must stay in the ring");
    float t = 0.0f;
    int steps = 0;  // loop tests passed so far (iterations started)
    for (;;) {
        if (steps >= max_steps) break;  // the loop test fails: accept at t
        const uint32_t n0 = rng->n;
        rng_fill(rng, n0 + 2u * (uint32_t)B);  // step j draws words n0 + 2j (distance) and n0 + 2j + 1
        float ts[B], ua[B];
        typename Fetch::Token tok[B];
        int m = 0;
        bool escape = false;
        float tc = t;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (escape || steps + j >= max_steps) break;
            tc -= logf_(1.0f - rng_word(*rng, n0 + 2u * (uint32_t)j)) * inv_smax;
            if (tc > dis) {
                escape = true;
                break;
            }
            ua[j] = rng_word(*rng, n0 + 2u * (uint32_t)j + 1u);
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
            rng->n = n0 + 2u * (uint32_t)(j + 1);
            break;
        }
        if (event == 2) {
            steps += j + 1;
            rng->n = n0 + 2u * (uint32_t)(j + 1);
            continue;
        }
        steps += m;
        if (escape) {
            rng->n = n0 + 2u * (uint32_t)m + 1u;
            return false;
        }
        t = tc;                             // the last speculated position (t itself when m == 0)
        rng->n = n0 + 2u * (uint32_t)m;  // exactly 2m draws consumed
    }
    rng->n += 2u;  // the HG direction draws (direction unused)
    *npos = (dir * t) + pos;
    return true;
}

// State of a batched transmittance walk between rounds.
struct TrWalk {
    float t, tr;
    int loop_num;  // the reference's loop counter (advanced by every loop test)
    int phase;     // 1: ratio tracking, 2: delta tracking, 0: finished (tr is the result)
};
MRP_HD TrWalk tr_walk_start() {
    TrWalk w;
    w.t = 0.0f;
    w.tr = 1.0f;
    w.loop_num = 0;
    w.phase = 1;
    return w;
}

// One round of transmittance_seq with up to B lookups in flight: the walk
// `w` advances by one window of steps (or moves to its next phase), and the
// stream is left at exactly the draws the one-step loops would have used.
// Both phases share one speculation block (a delta-tracking step draws a
// second number), so lanes of a wave in different phases issue their
// lookups together and wait for memory once per round.
template <int B, uint32_t kStride, class Fetch>
MRP_HD void transmittance_round(RingRng<kStride>* rng, const Fetch& F, float inv_smax_tr, float ratio, vec3 pos,
                                vec3 dir, float dis, int max_steps, TrWalk* w, uint32_t* lk) {
    static_assert(2 * B + 3 <= (int)kRingWords, "a round's draws must stay in the ring");
    // the loop test at the top of each phase
    if (w->phase == 1 && !(w->loop_num < max_steps && w->tr > 0.5f)) {
        ++w->loop_num;  // the failing test still increments
        w->phase = w->tr > 0.001f ? 2 : 0;
    }
    if (w->phase == 2 && w->loop_num >= max_steps) w->phase = 0;
    if (w->phase == 0) return;
    const bool delta = w->phase == 2;  // delta tracking: two draws per step
    const uint32_t n0 = rng->n;
    const uint32_t per = delta ? 2u : 1u;  // step j draws words n0 + per * j (distance) and, delta, the next
    rng_fill(rng, n0 + per * (uint32_t)B);
    float ts[B], ua[B];
    typename Fetch::Token tok[B];
    int m = 0;
    bool escape = false;
    float tc = w->t;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        if (escape || w->loop_num + j >= max_steps) break;
        tc -= logf_(1.0f - rng_word(*rng, n0 + per * (uint32_t)j)) * inv_smax_tr;
        if (tc > dis) {
            escape = true;
            break;
        }
        ts[j] = tc;
        tok[j] = F.issue(pos + (dir * tc));
        ua[j] = delta ? rng_word(*rng, n0 + 2u * (uint32_t)j + 1u) : 0.0f;
        ++m;
    }
    if (!delta) {
        // ---- ratio tracking: commit step by step; the loop test precedes each step
        int done = 0;
        bool stop = false, restart = false;
        float tdone = w->t;
        float tr = w->tr;
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
                w->t = ts[q] - d;
                restart = true;
            }
        }
        w->tr = tr;
        w->loop_num += done;
        if (restart) {
            rng->n = n0 + (uint32_t)done;
            return;
        }
        w->t = tdone;
        if (stop) {
            ++w->loop_num;
            rng->n = n0 + (uint32_t)done;
            w->phase = tr > 0.001f ? 2 : 0;
            return;
        }
        if (escape) {
            // the escaping step's test: it runs only if the walk is still in ratio tracking
            if (w->loop_num < max_steps && tr > 0.5f) {
                // the step draws and t > dis ends ratio tracking; the delta-tracking
                // loop's test then passes unless the cap is reached (tr > 0.5 >
                // 0.001), and its step draws once more and ends too (t only grows)
                ++w->loop_num;
                rng->n = n0 + (uint32_t)done + 1u + (w->loop_num < max_steps ? 1u : 0u);
                w->phase = 0;
                return;
            }
            ++w->loop_num;
            rng->n = n0 + (uint32_t)done;
            w->phase = tr > 0.001f ? 2 : 0;
            return;
        }
        rng->n = n0 + (uint32_t)m;  // exactly m draws consumed
        return;
    }
    // ---- delta tracking: tr stays until a collision zeroes it
    bool restart = false, hit = false;
    int done = 0;
#pragma unroll
    for (int q = 0; q < B; ++q) {
        if (restart || hit || q >= m) break;
        const float d = F.finish(tok[q]);
        *lk += 1u;
        done = q + 1;
        if (ua[q] < fmax_(0.0f, d * ratio)) {
            hit = true;
        } else if (d < 0.0f) {
            w->t = ts[q] - d;
            restart = true;
        }
    }
    w->loop_num += done;
    if (hit) {
        w->tr = 0.0f;
        w->phase = 0;
        rng->n = n0 + 2u * (uint32_t)done;  // two draws per step, the last one the collision test
        return;
    }
    if (restart) {
        rng->n = n0 + 2u * (uint32_t)done;
        return;
    }
    if (escape) {
        w->phase = 0;
        rng->n = n0 + 2u * (uint32_t)done + 1u;  // done == m steps, then the escaping step's draw
        return;
    }
    w->t = tc;  // the last speculated position (t itself when m == 0)
    rng->n = n0 + 2u * (uint32_t)m;  // exactly 2m draws consumed
}

// transmittance_seq with up to B lookups in flight.
template <int B, uint32_t kStride, class Fetch>
MRP_HD float transmittance_batched(RingRng<kStride>* rng, const Fetch& F, float inv_smax_tr, float ratio, vec3 pos,
                                   vec3 dir, float dis, int max_steps, uint32_t* lk) {
    TrWalk w = tr_walk_start();
    while (w.phase != 0) transmittance_round<B>(rng, F, inv_smax_tr, ratio, pos, dir, dis, max_steps, &w, lk);
    return w.tr;
}

}  // namespace mrp
