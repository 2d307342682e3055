// bsm_inputs.hip -- host-side synthetic input streams of the workloads.
//
// The reference feeds areas from encoders and input generators
// (src/encode/, src/subcortex/input_generator.rs); the workloads here use
// seeded synthetic sparse distributed frames instead.  An SDR stream over n
// axons keeps exactly k = round(density * n) axons active (state 255); each
// next frame keeps round(overlap * k) of the previous frame's active axons (a
// uniform subset) and draws the rest uniformly from outside the previous
// active set.  Draws come from the engine's xorshift128 stream (Marsaglia
// 2003) seeded by splitmix64(seed).  Host code only: the serving loop fills the
// frames of every session that steps in a cycle with one call.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "bsm_internal.h"

struct bsm_sdr {
    uint32_t n, k, keep;
    uint32_t rng[4];
    std::vector<uint32_t> active;   // the current frame's active axons
    std::vector<uint8_t> member;    // membership of `active`
    std::vector<uint32_t> drop;     // scratch: axons leaving the active set
};

namespace bsm {

// Uniform in [0, n) by multiply-high of one 32-bit draw (bias below 2^-32 * n).
static inline uint32_t draw_below(uint32_t st[4], uint32_t n) {
    return (uint32_t)(((uint64_t)rng_next(st) * n) >> 32);
}

// Advances the stream: partial Fisher-Yates picks the kept subset into the
// front of `active`, then fresh axons outside the previous set fill the rest.
static void sdr_advance(bsm_sdr* s) {
    const uint32_t k = s->k, keep = s->keep;
    for (uint32_t i = 0; i < keep; i++) {
        const uint32_t j = i + draw_below(s->rng, k - i);
        const uint32_t t = s->active[i];
        s->active[i] = s->active[j];
        s->active[j] = t;
    }
    // axons leaving the set stay excluded from the fresh draws (they were in
    // the previous frame), so they are unmarked only afterwards
    s->drop.assign(s->active.begin() + keep, s->active.end());
    for (uint32_t i = keep; i < k; i++) {
        uint32_t c;
        do {
            c = draw_below(s->rng, s->n);
        } while (s->member[c]);
        s->member[c] = 1;
        s->active[i] = c;
    }
    for (uint32_t c : s->drop) s->member[c] = 0;
}

// Runs f(i) for i in [0, n), split over up to 8 threads when n is large (the
// streams are independent, so the frames do not depend on the split).
template <class F>
static void for_streams(uint32_t n, F&& f) {
    const uint32_t nt = n >= 64 ? std::min<uint32_t>(8, n / 16) : 1;
    if (nt <= 1) {
        for (uint32_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (uint32_t w = 1; w < nt; w++)
        pool.emplace_back([&, w] {
            for (uint32_t i = n * w / nt; i < n * (w + 1) / nt; i++) f(i);
        });
    for (uint32_t i = 0; i < n / nt; i++) f(i);
    for (std::thread& th : pool) th.join();
}

}  // namespace bsm

using namespace bsm;

extern "C" {

int bsm_sdr_create(uint32_t n_axons, double density, double overlap, uint64_t seed,
                   bsm_sdr** out) {
    if (!out) return fail("null out");
    if (n_axons == 0) return fail("n_axons must be >= 1");
    if (!(density > 0.0 && density <= 1.0) || !(overlap >= 0.0 && overlap <= 1.0))
        return fail("density must be in (0, 1] and overlap in [0, 1]");
    uint32_t k = (uint32_t)std::llround(density * n_axons);
    if (k < 1) k = 1;
    const uint32_t keep = (uint32_t)std::llround(overlap * k);
    if (n_axons - k < k - keep) return fail("density too high for the requested overlap");
    bsm_sdr* s = new bsm_sdr();
    s->n = n_axons;
    s->k = k;
    s->keep = keep;
    rng_seed(s->rng, seed);
    s->member.assign(n_axons, 0);
    s->active.resize(k);
    for (uint32_t i = 0; i < k; i++) {
        uint32_t c;
        do {
            c = draw_below(s->rng, n_axons);
        } while (s->member[c]);
        s->member[c] = 1;
        s->active[i] = c;
    }
    *out = s;
    return 0;
}

int bsm_sdr_destroy(bsm_sdr* s) {
    delete s;
    return 0;
}

int bsm_sdr_info(bsm_sdr* s, uint32_t* n_axons, uint32_t* k) {
    if (!s) return fail("null stream");
    if (n_axons) *n_axons = s->n;
    if (k) *k = s->k;
    return 0;
}

int bsm_sdr_next(bsm_sdr* s, uint8_t* frame) {
    if (!s || !frame) return fail("null arg");
    std::memset(frame, 0, s->n);
    for (uint32_t c : s->active) frame[c] = 255;
    sdr_advance(s);
    return 0;
}

int bsm_sdr_frames(bsm_sdr* s, uint32_t n_frames, uint8_t* frames) {
    if (!s || (!frames && n_frames)) return fail("null arg");
    for (uint32_t f = 0; f < n_frames; f++)
        if (bsm_sdr_next(s, frames + (size_t)f * s->n)) return 1;
    return 0;
}

int bsm_sdr_next_packed(bsm_sdr* const* s, uint32_t n, uint8_t* buf, const uint64_t* ofs) {
    if ((!s || !buf || !ofs) && n) return fail("null arg");
    for (uint32_t i = 0; i < n; i++)
        if (!s[i]) return fail("null stream");
    for_streams(n, [&](uint32_t i) {
        uint8_t* frame = buf + ofs[i];
        std::memset(frame, 0, s[i]->n);
        for (uint32_t c : s[i]->active) frame[c] = 255;
        sdr_advance(s[i]);
    });
    return 0;
}

int bsm_sdr_next_indices(bsm_sdr* s, uint32_t* idx) {
    if (!s || !idx) return fail("null arg");
    std::memcpy(idx, s->active.data(), (size_t)s->k * sizeof(uint32_t));
    sdr_advance(s);
    return 0;
}

int bsm_sdr_next_packed_sdr(bsm_sdr* const* s, uint32_t n, uint8_t* buf, const uint64_t* ofs) {
    if ((!s || !buf || !ofs) && n) return fail("null arg");
    for (uint32_t i = 0; i < n; i++) {
        if (!s[i]) return fail("null stream");
        if (ofs[i] & 3) return fail("SDR entry offsets must be 4-byte aligned");
    }
    for_streams(n, [&](uint32_t i) {
        uint32_t* e = (uint32_t*)(buf + ofs[i]);
        e[0] = s[i]->k;
        std::memcpy(e + 1, s[i]->active.data(), (size_t)s[i]->k * sizeof(uint32_t));
        sdr_advance(s[i]);
    });
    return 0;
}

}  // extern "C"
