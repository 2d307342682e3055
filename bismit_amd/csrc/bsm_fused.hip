// bsm_fused.hip -- fused gfx950 kernels of the per-cycle pipeline and the
// plain per-area orchestrator (CorticalArea::cycle, src/cortex/cortical_area.rs:716-789).
//
// Stages per area per cycle:
//   F1 kf_tuft      : tft_set_syn_flags + layer_cycle_syns_wow_vec4 + den_cycle_tft
//                     (+ tft_cycle for pyramidal tufts, + ssc_cycle for the SSC soma)
//                     cl/syns.cl:23-33,285-388, cl/bismit.cl:695-884
//   F2 kf_ssc_post  : inhib_simple | inhib_passthrough + ssc_mtp_simple
//                     cl/control.cl:28-217, cl/bismit.cl:888-915
//   F3 kf_pyr_post  : pyr_cycle + inhib_intra_column | pyr_output + tft_dst_mtp
//                     cl/bismit.cl:961-1022,1195-1268, cl/control.cl:298-351
//
// Every kernel is a persistent grid-stride loop over a flattened work-item list
// (tiles of columns), so one launch can step any number of areas (batched form)
// or one area (plain form).
//
// Synapse states are double-buffered: F1 writes st[par] and never reads the
// old state; SYN_PREV_ACTIVE of the flags is "st[par^1] != 0", read by the
// distal LTP of F3 and materialised only at readback.  F1 therefore streams
// per synapse the offsets (1 B packed, 2 B otherwise), a 2-bit staged-slice
// code (only when the tuft's gather stages several slices) and the strength
// in, and the new state out: 3-3.25 B.
#include <algorithm>
#include <type_traits>
#include <cstring>

#include "bsm_internal.h"

namespace bsm {

// ------------------------------------------------------------------ F1

// Buffer resource over `bytes` bytes at base (raw, dword format): loads and
// stores through it take one 32-bit offset VGPR and return 0 beyond the range.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ inline Rsrc make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
// Non-temporal (streaming) cache policy, for the synapse streams' loads and
// stores (new states, learned strengths): each byte is touched once per cycle,
// so keeping it in L2 only evicts the axon windows and dendrite arrays that are
// reused.
// (Measured on the Config 2 pyramidal launch: the streaming store takes 6% off
// the launch, and 11% off the 64-area batched cycle.)
constexpr int BUF_NT = 2;

template <int W>
struct Chunk;
// 32 synapses per lane (check-free gather only): two 16-byte vectors
template <>
struct Chunk<32> {
    static constexpr int NW = 8;
    __device__ static inline void load_b(Rsrc r, uint32_t off, uint32_t* o) {
        u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, BUF_NT);
        u32x4 y = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off + 16, 0, BUF_NT);
        o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = x.w;
        o[4] = y.x; o[5] = y.y; o[6] = y.z; o[7] = y.w;
    }
    __device__ static inline void store_b(Rsrc r, uint32_t off, const uint32_t* o) {
        u32x4 x = {o[0], o[1], o[2], o[3]}, y = {o[4], o[5], o[6], o[7]};
        __builtin_amdgcn_raw_buffer_store_b128(x, r, (int)off, 0, BUF_NT);
        __builtin_amdgcn_raw_buffer_store_b128(y, r, (int)off + 16, 0, BUF_NT);
    }
};
template <>
struct Chunk<16> {
    static constexpr int NW = 4;
    __device__ static inline void load_b(Rsrc r, uint32_t off, uint32_t* o) {
        u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, BUF_NT);
        o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = x.w;
    }
    __device__ static inline void store_b(Rsrc r, uint32_t off, const uint32_t* o) {
        u32x4 v = {o[0], o[1], o[2], o[3]};
        __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, BUF_NT);
    }
    __device__ static inline void load(const void* p, uint32_t* o) {
        u32x4 x = __builtin_nontemporal_load(BSM_G(const u32x4, p));
        o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = x.w;
    }
    __device__ static inline void store(void* p, const uint32_t* o) {
        u32x4 v = {o[0], o[1], o[2], o[3]};
        __builtin_nontemporal_store(v, BSM_G(u32x4, p));
    }
};
template <>
struct Chunk<8> {
    static constexpr int NW = 2;
    __device__ static inline void load_b(Rsrc r, uint32_t off, uint32_t* o) {
        u32x2 x = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, BUF_NT);
        o[0] = x.x; o[1] = x.y;
    }
    __device__ static inline void store_b(Rsrc r, uint32_t off, const uint32_t* o) {
        u32x2 v = {o[0], o[1]};
        __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)off, 0, BUF_NT);
    }
    __device__ static inline void load(const void* p, uint32_t* o) {
        u32x2 x = __builtin_nontemporal_load(BSM_G(const u32x2, p));
        o[0] = x.x; o[1] = x.y;
    }
    __device__ static inline void store(void* p, const uint32_t* o) {
        u32x2 v = {o[0], o[1]};
        __builtin_nontemporal_store(v, BSM_G(u32x2, p));
    }
};
template <>
struct Chunk<4> {
    static constexpr int NW = 1;
    __device__ static inline void load_b(Rsrc r, uint32_t off, uint32_t* o) {
        o[0] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, BUF_NT);
    }
    __device__ static inline void store_b(Rsrc r, uint32_t off, const uint32_t* o) {
        __builtin_amdgcn_raw_buffer_store_b32(o[0], r, (int)off, 0, BUF_NT);
    }
    __device__ static inline void load(const void* p, uint32_t* o) {
        o[0] = __builtin_nontemporal_load(BSM_G(const uint32_t, p));
    }
    __device__ static inline void store(void* p, const uint32_t* o) {
        __builtin_nontemporal_store(o[0], BSM_G(uint32_t, p));
    }
};
template <>
struct Chunk<1> {
    static constexpr int NW = 1;
    __device__ static inline void load(const void* p, uint32_t* o) { o[0] = *BSM_G(const uint8_t, p); }
    __device__ static inline void store(void* p, const uint32_t* o) { *BSM_G(uint8_t, p) = (uint8_t)o[0]; }
};

// The 2-bit staged-slice codes of a W-synapse chunk: W / 4 bytes at byte `off`.
template <int W>
__device__ inline uint32_t load_codes(Rsrc r, uint32_t off) {
    if constexpr (W == 16) return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, BUF_NT);
    else if constexpr (W == 8) return __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, BUF_NT);
    else return __builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, BUF_NT);
}

// Bit 7 of each byte of x gathered into bits 0..3 (byte k -> bit k).
__device__ __forceinline__ uint32_t sign_bits4(uint32_t x) {
    return (((x >> 7) & 0x01010101u) * 0x01020408u) >> 24;
}
// Bits 0..3 of b spread to the low bit of bytes 0..3.
__device__ __forceinline__ uint32_t spread4(uint32_t b) { return (b * 0x00204081u) & 0x01010101u; }

// Strength-sign words: word g of a layer covers synapses 32g..32g+31; synapse i
// (w = (i % 32) / 4, k = i % 4) sits at bit 8k + w, so the four synapses of one
// 4-byte word of a synapse stream are bit w of each byte.
__device__ __forceinline__ uint32_t sg_pos(uint32_t i) { return 8u * (i & 3u) + ((i & 31u) >> 2); }

// The epilogue's global inputs of one tile, loaded when the tile starts so that
// their latency hides behind the synapse loop: the first TAIL_PD dendrites and
// TAIL_PC cell-tufts of each thread (bytes packed four to a register).  Every
// byte belongs to this tile alone (its dendrites, its cells), so nothing else
// writes it meanwhile.
constexpr uint32_t TAIL_PD = 4;
constexpr uint32_t TAIL_PC = 1;
// Quad path (4 | dendrites per cell): thread t takes the 4-dendrite groups
// t, t + nth, ... (4 consecutive dendrites of one cell, one 32-bit word of each
// dendrite array); the first group's words (TAIL_PQ) are loaded at tile start as
// whole words, so no byte has to be unpacked before the synapse loop.
constexpr uint32_t TAIL_PQ = 1;
struct TailPre {
    uint32_t act[TAIL_PQ];       // dendrite activity ratings (4 per word)
    uint32_t en[TAIL_PQ];        // SSC soma energies (dendrite index < cells)
    uint32_t tft[TAIL_PC];       // best_id, best_raw, best, state of a cell-tuft
};

// Whether a tile's epilogue takes the quad path: dendrites per cell a multiple
// of 4 (a group never spans two cells, and its words are 4-byte aligned), with
// SSC energies a multiple-of-4 cell count (a group is wholly below it or wholly
// above), and at least one group per thread (smaller tiles keep every thread
// busy on single dendrites: Config 2's spiny stellate tile, 512 dendrites, took
// 0.7 us longer with groups).
__device__ inline bool tail_quad(const TuftJob& J, uint32_t nd_t, uint32_t nth) {
    return (J.dens & 3u) == 0 && (J.den_idz & 3u) == 0 &&
           (J.ssc_energy == nullptr || (J.C & 3u) == 0) && nd_t >= 4u * nth;
}

__device__ inline uint32_t tile_cell(uint32_t ct, uint32_t tvn, uint32_t tun, uint32_t v0,
                                     uint32_t u0, uint32_t VU, uint32_t U) {
    const uint32_t rowq = ct / tun, cc = ct - rowq * tun;
    const uint32_t sl = rowq / tvn, r = rowq - sl * tvn;
    return sl * VU + (v0 + r) * U + (u0 + cc);
}

// n / d as a multiply-high with magic = ceil(2^32 / d) (d > 1), exact whenever
// n * d < 2^32 (the rounding error n * (magic * d - 2^32) / 2^32 stays below
// the gap to the next multiple of d); d == 1 passes n through.
__device__ __forceinline__ uint32_t magic_u32(uint32_t d) { return d > 1 ? 0xFFFFFFFFu / d + 1u : 0u; }
__device__ __forceinline__ uint32_t quot_u32(uint32_t n, uint32_t d, uint32_t magic) {
    return d > 1 ? __umulhi(n, magic) : n;
}
// The same quotient without a branch: d == 1 takes m = 2^32 - 1 and adds 1 to
// n first (floor((n + 1)(2^32 - 1) / 2^32) = n for n + 1 < 2^32).
struct Quot {
    uint32_t m, e;
    __device__ static Quot of(uint32_t d) { return d > 1 ? Quot{0xFFFFFFFFu / d + 1u, 0u} : Quot{0xFFFFFFFFu, 1u}; }
    __device__ uint32_t operator()(uint32_t n) const { return __umulhi(n + e, m); }
};

// Tile cell ct (slice-major rows of tun columns) -> layer cell, and tile
// dendrite x -> (cell, dendrite of the cell).  DivMap divides (checked path);
// RowMap reads the first cell of each tile row from LDS and divides by
// multiply-high (check-free path; every product stays below 2^32).
struct DivMap {
    uint32_t tvn, tun, v0, u0, VU, U, dens;
    __device__ uint32_t cell(uint32_t ct) const { return tile_cell(ct, tvn, tun, v0, u0, VU, U); }
    __device__ uint32_t den_cell(uint32_t x, uint32_t& dj) const {
        const uint32_t ct = x / dens;
        dj = x - ct * dens;
        return ct;
    }
};
struct RowMap {
    const uint32_t* rowc;
    uint32_t tun, dens;
    Quot q_tun, q_dens;
    __device__ uint32_t cell(uint32_t ct) const {
        const uint32_t r = q_tun(ct);
        return rowc[r] + (ct - r * tun);
    }
    __device__ uint32_t den_cell(uint32_t x, uint32_t& dj) const {
        const uint32_t ct = q_dens(x);
        dj = x - ct * dens;
        return ct;
    }
};

template <class Map>
__device__ inline void tail_prefetch(const TuftJob& J, const Map& map, uint32_t nd_t,
                                     uint32_t cells_t, uint32_t tid, uint32_t nth, TailPre& pre) {
    // every load first (indices clamped to valid bytes, results masked after),
    // so that they share one round trip
    const uint32_t dens = J.dens;
    auto den_act = BSM_G(const uint8_t, J.den_act);
    if (tail_quad(J, nd_t, nth)) {
        // whole words, kept as loaded (the epilogue picks their bytes); indices
        // past the tile are clamped to group 0 and the words go unused
        const bool has_en = J.ssc_energy != nullptr;
#pragma unroll
        for (uint32_t k = 0; k < TAIL_PQ; k++) {
            const uint32_t x0 = 4u * (tid + k * nth);
            uint32_t dj;
            const uint32_t ct = map.den_cell(x0 < nd_t ? x0 : 0u, dj);
            const uint32_t dl = map.cell(ct) * dens + dj;
            pre.act[k] = *BSM_G(const uint32_t, J.den_act + dl);
            const uint32_t dg = J.den_idz + dl;
            pre.en[k] = has_en ? *BSM_G(const uint32_t, J.ssc_energy + (dg < J.C ? dg : 0u)) : 0u;
        }
        pre.tft[0] = 0u;   // loaded at the epilogue's start
        return;
    }
    // no branch around a load: a layer without energies or tuft arrays reads
    // its dendrite activities in their place (masked out below)
    const bool has_en = J.ssc_energy != nullptr, has_tft = J.tft[0] != nullptr;
    auto energy = BSM_G(const uint8_t, has_en ? (const uint8_t*)J.ssc_energy : J.den_act);
    const uint32_t en_lim = has_en ? J.C : 1u;
    uint32_t av[TAIL_PD], ev[TAIL_PD], ok[TAIL_PD];
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PD; k++) {
        const uint32_t x = tid + k * nth;
        uint32_t dj;
        const uint32_t ct = map.den_cell(x < nd_t ? x : 0u, dj);
        const uint32_t dl = map.cell(ct) * dens + dj;
        const uint32_t dg = J.den_idz + dl;
        av[k] = den_act[dl];
        ev[k] = energy[dg < en_lim ? dg : 0u];
        ok[k] = (uint32_t)(x < nd_t) | ((uint32_t)(has_en && dg < J.C) << 1);
    }
    uint32_t tv[TAIL_PC][4];
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PC; k++) {
        const uint32_t ct = tid + k * nth;
        const uint32_t cel = map.cell(ct < cells_t ? ct : 0u);
#pragma unroll
        for (uint32_t b = 0; b < 4; b++)
            tv[k][b] = BSM_G(const uint8_t, has_tft ? (const uint8_t*)J.tft[b] : J.den_act)[has_tft ? cel : 0u];
    }
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PQ; k++) {
        pre.act[k] = 0;
        pre.en[k] = 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PD; k++) {
        if (ok[k] & 1u) pre.act[k / 4] |= av[k] << (8 * (k % 4));
        if (ok[k] == 3u) pre.en[k / 4] |= ev[k] << (8 * (k % 4));
    }
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PC; k++)
        pre.tft[k] = has_tft ? tv[k][0] | (tv[k][1] << 8) | (tv[k][2] << 16) | (tv[k][3] << 24) : 0u;
}

// The fired bits of each tile chunk of a distal tuft (LDS, W bits per chunk in
// a 16-bit slot, 32-bit for 32-synapse chunks), copied for each cell's best
// dendrite to the tuft's kept bits of this parity (dst; cells' words of wpd,
// bit i = synapse i of the dendrite).  With 16- or 32-synapse chunks the slots
// form one bit stream (a cell's synapses are contiguous), so a dendrite of any
// length is a bit range of it; narrower chunks are copied slot by slot, whole
// chunks only.
struct MaskOut {
    const uint16_t* lm;   // null: none kept (4-byte aligned, one spare word)
    uint32_t* dst;
    uint32_t cpt, spd, wpd, w;
};

// Dendrites (den_cycle_tft, ssc_cycle on the SSC soma) and tufts (tft_cycle) of
// one tile from the per-chunk (sum, raw) partials in LDS; ends with a barrier.
template <class Map>
__device__ inline void tuft_tail(const TuftJob& J, uint8_t* lds, const uint32_t* acc,
                                 uint32_t stage_cap, uint32_t nd_t, uint32_t cells_t,
                                 const Map& map, int32_t rnd, uint32_t tid, uint32_t nth,
                                 const TailPre& pre, const MaskOut& mo) {
    // ---- dendrites (den_cycle_tft), optional ssc_cycle on the SSC soma
    const uint32_t dens = J.dens;
    const bool do_tft = J.tft[0] != nullptr;
    uint8_t* denc = lds + stage_cap + nd_t * 4;
    auto den_act = BSM_G(uint8_t, J.den_act);
    auto den_raw = BSM_G(uint8_t, J.den_raw);
    auto den_st = BSM_G(uint8_t, J.den_st);
    auto energy = BSM_G(uint8_t, J.ssc_energy);
    auto den_one = [&](uint32_t x, uint32_t act0, uint32_t e0, bool have) {
        uint32_t dj;
        const uint32_t ct = map.den_cell(x, dj);
        const uint32_t pv = acc[x];
        uint32_t sum = pv & 0xFFFFu;
        const uint32_t raw = pv >> 16;
        const uint32_t cel = map.cell(ct);
        const uint32_t dl = cel * dens + dj;
        const uint32_t dg = J.den_idz + dl;
        const int32_t active = sum > J.thresh;
        sum = active ? sum : 0u;
        const uint32_t raw_o = min(raw >> 4, 255u);
        uint32_t st_o = min(sum >> 4, 255u);
        const int32_t seed = (int32_t)((raw + J.thresh) | dg);
        den_act[dl] = (uint8_t)activity(have ? act0 : den_act[dl], active, rnd, seed,
                                        DEN_ACT_DECAY);
        den_raw[dl] = (uint8_t)raw_o;
        if (J.ssc_energy != nullptr && dg < J.C) {
            // ssc_cycle (cl/bismit.cl:849-884) on soma index dg == dendrite dg
            const uint32_t e = have ? e0 : energy[dg];
            const uint32_t restless = (e > 191) & (st_o != 0);
            const uint32_t dark = e == 255;
            if (dark | restless) energy[dg] = (uint8_t)(e - 64);
            st_o = min(st_o + 255u * restless + 255u * dark, 255u);
        }
        den_st[dl] = (uint8_t)st_o;
        if (do_tft) {
            denc[x] = (uint8_t)raw_o;
            denc[nd_t + x] = (uint8_t)st_o;
        }
    };
    const bool quad = tail_quad(J, nd_t, nth);
    // quad path: the first cell's current tuft bytes, loaded now and used by
    // the tuft pass after the dendrites (their latency hides behind them)
    uint32_t tb[4] = {0, 0, 0, 0};
    if (quad && do_tft) {
        const uint32_t cel = map.cell(tid < cells_t ? tid : 0u);
#pragma unroll
        for (uint32_t b = 0; b < 4; b++) tb[b] = BSM_G(const uint8_t, J.tft[b])[cel];
    }
    // 4 dendrites of one cell at a time: words of every dendrite array
    auto den_quad = [&](uint32_t q, uint32_t actw, uint32_t enw, bool have) {
        const uint32_t x0 = 4u * q;
        uint32_t dj0;
        const uint32_t ct = map.den_cell(x0, dj0);
        const uint32_t dl0 = map.cell(ct) * dens + dj0;
        const uint32_t dg0 = J.den_idz + dl0;
        const bool en_on = J.ssc_energy != nullptr && dg0 < J.C;
        if (!have) {
            actw = *BSM_G(const uint32_t, J.den_act + dl0);
            enw = en_on ? *BSM_G(const uint32_t, J.ssc_energy + dg0) : 0u;
        }
        uint32_t ao = 0, ro = 0, so = 0, eo = enw;
#pragma unroll
        for (uint32_t b = 0; b < 4; b++) {
            const uint32_t pv = acc[x0 + b];
            uint32_t sum = pv & 0xFFFFu;
            const uint32_t raw = pv >> 16;
            const int32_t active = sum > J.thresh;
            sum = active ? sum : 0u;
            const uint32_t raw_o = min(raw >> 4, 255u);
            uint32_t st_o = min(sum >> 4, 255u);
            const int32_t seed = (int32_t)((raw + J.thresh) | (dg0 + b));
            ao |= activity((actw >> (8 * b)) & 0xFFu, active, rnd, seed, DEN_ACT_DECAY) << (8 * b);
            if (en_on) {
                // ssc_cycle (cl/bismit.cl:849-884) on soma index dg == dendrite dg
                const uint32_t e = (enw >> (8 * b)) & 0xFFu;
                const uint32_t restless = (e > 191) & (st_o != 0);
                const uint32_t dark = e == 255;
                if (dark | restless) eo = (eo & ~(0xFFu << (8 * b))) | ((e - 64) << (8 * b));
                st_o = min(st_o + 255u * restless + 255u * dark, 255u);
            }
            ro |= raw_o << (8 * b);
            so |= st_o << (8 * b);
        }
        *BSM_G(uint32_t, J.den_act + dl0) = ao;
        *BSM_G(uint32_t, J.den_raw + dl0) = ro;
        *BSM_G(uint32_t, J.den_st + dl0) = so;
        if (en_on) *BSM_G(uint32_t, J.ssc_energy + dg0) = eo;
        if (do_tft) {
            *(uint32_t*)(denc + x0) = ro;
            *(uint32_t*)(denc + nd_t + x0) = so;
        }
    };
    if (quad) {
        const uint32_t ng = nd_t / 4;
#pragma unroll
        for (uint32_t k = 0; k < TAIL_PQ; k++) {
            const uint32_t q = tid + k * nth;
            if (q < ng) den_quad(q, pre.act[k], pre.en[k], true);
        }
        for (uint32_t q = tid + TAIL_PQ * nth; q < ng; q += nth) den_quad(q, 0u, 0u, false);
    }
    // the inputs of the next TAIL_PE dendrites of each thread are loaded
    // together before any dendrite is finished (indices clamped, unused loads
    // discarded), so that they share one round trip
    constexpr uint32_t TAIL_PE = 2;
    const bool more = !quad && nd_t > TAIL_PD * nth;
    uint32_t a2[TAIL_PE], e2[TAIL_PE];
    if (more) {
        const bool has_en = J.ssc_energy != nullptr;
#pragma unroll
        for (uint32_t k = 0; k < TAIL_PE; k++) {
            const uint32_t x = tid + (TAIL_PD + k) * nth;
            uint32_t dj;
            const uint32_t ct = map.den_cell(x < nd_t ? x : 0u, dj);
            const uint32_t dl = map.cell(ct) * dens + dj;
            const uint32_t dg = J.den_idz + dl;
            a2[k] = den_act[dl];
            e2[k] = has_en ? (uint32_t)energy[dg < J.C ? dg : 0u] : 0u;
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < TAIL_PD; k++) {
        const uint32_t x = tid + k * nth;
        if (!quad && x < nd_t)
            den_one(x, (pre.act[k / 4] >> (8 * (k % 4))) & 0xFFu,
                    (pre.en[k / 4] >> (8 * (k % 4))) & 0xFFu, true);
    }
    if (more) {
#pragma unroll
        for (uint32_t k = 0; k < TAIL_PE; k++) {
            const uint32_t x = tid + (TAIL_PD + k) * nth;
            if (x < nd_t) den_one(x, a2[k], e2[k], true);
        }
    }
    if (!quad)
        for (uint32_t x = tid + (TAIL_PD + TAIL_PE) * nth; x < nd_t; x += nth) den_one(x, 0u, 0u, false);

    // ---- tufts (tft_cycle)
    if (do_tft) {
        __syncthreads();
        const uint32_t l2 = J.l2;
        for (uint32_t ct = tid, k = 0; ct < cells_t; ct += nth, k++) {
            uint32_t best_id = 0, best_raw = 0, best = 0, asum = 0;
            if (quad) {
                // the cell's dendrite bytes read four at a time
                for (uint32_t j4 = 0; j4 < dens; j4 += 4) {
                    const uint32_t wr = *(const uint32_t*)(denc + ct * dens + j4);
                    const uint32_t ws = *(const uint32_t*)(denc + nd_t + ct * dens + j4);
                    asum = __builtin_amdgcn_sad_u8(ws, 0u, asum);
#pragma unroll
                    for (uint32_t b = 0; b < 4; b++) {
                        const uint32_t sr = (wr >> (8 * b)) & 0xFFu;
                        if (sr > best_raw) { best_id = j4 + b; best_raw = sr; }
                        best = max(best, (ws >> (8 * b)) & 0xFFu);
                    }
                }
            } else {
                for (uint32_t j = 0; j < dens; j++) {
                    const uint32_t sr = denc[ct * dens + j], s = denc[nd_t + ct * dens + j];
                    if (sr > best_raw) { best_id = j; best_raw = sr; }
                    best = max(best, s);
                    asum += s;
                }
            }
            const uint32_t tstate = asum >= (1u << (l2 + 8)) ? 255u : (asum >> l2);
            const uint32_t cel = map.cell(ct);
            auto t0 = BSM_G(uint8_t, J.tft[0]), t1 = BSM_G(uint8_t, J.tft[1]);
            auto t2 = BSM_G(uint8_t, J.tft[2]), t3 = BSM_G(uint8_t, J.tft[3]);
            // the current values become prev_* (prefetched for the first TAIL_PC)
            uint32_t cur = 0;
#pragma unroll
            for (uint32_t q = 0; q < TAIL_PC; q++)
                if (k == q) cur = quad ? tb[0] | (tb[1] << 8) | (tb[2] << 16) | (tb[3] << 24) : pre.tft[q];
            if (k >= TAIL_PC)
                cur = (uint32_t)t0[cel] | ((uint32_t)t1[cel] << 8) | ((uint32_t)t2[cel] << 16) |
                      ((uint32_t)t3[cel] << 24);
            BSM_G(uint8_t, J.tft[4])[cel] = (uint8_t)cur;
            BSM_G(uint8_t, J.tft[5])[cel] = (uint8_t)(cur >> 8);
            BSM_G(uint8_t, J.tft[6])[cel] = (uint8_t)(cur >> 16);
            BSM_G(uint8_t, J.tft[7])[cel] = (uint8_t)(cur >> 24);
            t0[cel] = (uint8_t)best_id;
            t1[cel] = (uint8_t)best_raw;
            t2[cel] = (uint8_t)best;
            t3[cel] = (uint8_t)tstate;
            if (mo.lm) {
                auto dst = BSM_G(uint32_t, mo.dst + (size_t)cel * mo.wpd);
                if (mo.w >= 16) {
                    const uint32_t* l32 = (const uint32_t*)mo.lm;
                    const uint32_t start = ct * mo.cpt * mo.w + best_id * mo.spd;
                    for (uint32_t i = 0; i < mo.wpd; i++) {
                        const uint32_t bp = start + 32 * i, k = bp >> 5;
                        uint32_t v = __builtin_amdgcn_alignbit(l32[k + 1], l32[k], bp & 31u);
                        const uint32_t rem = mo.spd - 32 * i;
                        if (rem < 32) v &= (1u << rem) - 1u;
                        dst[i] = v;
                    }
                } else {
                    const uint32_t cpd = mo.spd / mo.w;
                    const uint16_t* src = mo.lm + ct * mo.cpt + best_id * cpd;
                    uint32_t word = 0, bit = 0, wi = 0;
                    for (uint32_t c = 0; c < cpd; c++) {
                        word |= (uint32_t)src[c] << bit;
                        bit += mo.w;
                        if (bit == 32) {
                            dst[wi++] = word;
                            word = 0;
                            bit = 0;
                        }
                    }
                    if (bit) dst[wi] = word;
                }
            }
        }
    }
    __syncthreads();
}

// Adds one chunk's per-word (sum, raw) partials to the per-dendrite packed
// accumulators (sum in bits 0-15, raw in bits 16-31; a dendrite's raw sum is at
// most 256*255 < 2^16, so the halves never carry into each other).  A chunk
// inside one dendrite takes one atomic; otherwise each 4-synapse word (never
// split across dendrites) takes its own.  dj = off / spd via a multiply-high
// with magic = ceil(2^32 / spd), exact since off * spd < 2^32.
template <int NW, int NB>
__device__ inline void acc_chunk(uint32_t* acc, uint32_t den0, uint32_t off0, const uint32_t* pw,
                                 bool whole, uint32_t magic) {
    if (whole) {
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) tot += pw[w];
        atomicAdd(&acc[den0 + __umulhi(off0, magic)], tot);
    } else {
#pragma unroll
        for (int w = 0; w < NW; w++)
            atomicAdd(&acc[den0 + __umulhi(off0 + (uint32_t)(w * NB), magic)], pw[w]);
    }
}

// A byte of LDS at byte address a.
#define LDS8(a) ((const __attribute__((address_space(3))) uint8_t*)(uintptr_t)(a))

// The check-free gather's loads of one chunk, kept as loaded: every value
// derived from them is computed in work(), so that a chunk's loads stay in
// flight while the previous chunk is gathered.
template <int NW>
struct ChunkMeta {
    uint32_t v[NW], u[NW];   // packed offsets (or v / u offsets)
    uint32_t c[(NW + 3) / 4];  // staged-slice code words of the chunk's 16-synapse groups
    uint32_t s;              // strength-sign word of the chunk's 32-synapse group
    uint32_t so, q;          // first synapse (tuft-relative), chunk index in the tile
    uint32_t dd;             // tile dendrite of the chunk's cell | synapse in cell << 16
    int32_t lb;              // window offset of the chunk's cell
};

// F1, check-free form: every synapse of each job sources a staged slice with
// |offsets| <= halo (proved on the host), and the staged slices are
// consecutive ids of one geometry (contiguous in the axon space).  Per tile:
// the window of each staged slice (tile + halo, zeros off the slice) is staged
// in LDS as syn_fire of each axon (cl/syns.cl:2-4), so the gather reads fired
// states directly; slot j of a several-slice window starts 32 rows after slot
// j - 1, so that with packed offsets the 2-bit staged-slice code merges into
// the v nibble (v + 8 + 32 * code, < 128) and one dot product per synapse forms
// its window address.  Synapse states are virtual: no state is written; the
// item's share of the snapshot of the staged slices is copied instead.  Per
// dendrite, (sum over non-negative strengths, raw sum) come from a dot product
// with the strength-sign bits and a sum of absolute differences.
// Five waves per SIMD (at most 96 VGPRs); launches of one staged slice, and of
// chunks under 32 synapses, fit in 80 VGPRs without spills and take six (on
// Config 3 the two-stream schedule, whose side launch runs beside the others,
// 0.342 -> 0.332 ms per batched cycle).
template <int W, int NST, bool PACKED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NST == 1 || W < 32) ? 6 : 5))) void kf_tuft_fast(const TuftJob* __restrict__ jobs,
                                                    const WorkItem* __restrict__ items,
                                                    uint32_t n_items, uint32_t par,
                                                    uint32_t rnd_ofs, uint32_t stage_cap,
                                                    const uint8_t* __restrict__ frame,
                                                    uint8_t* __restrict__ intake_dst,
                                                    uint32_t intake_job) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr uint32_t NTH = 256;   // launched with 256 threads
    // the LDS byte address of the dynamic LDS array
    const uint32_t lds0 = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint8_t*)lds;
    constexpr int NW = Chunk<W>::NW;
    const uint32_t tid = threadIdx.x;
    uint32_t* acc = (uint32_t*)(lds + stage_cap);

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem wi = items[it];
        const TuftJob& J = jobs[wi.job];
        const uint32_t V = J.V, U = J.U, depth = J.depth, spt = J.spt, dens = J.dens;
        const uint32_t TV = J.tile_v, TU = J.tile_u;
        const uint32_t v0 = (uint32_t)wi.tv * TV, u0 = (uint32_t)wi.tu * TU;
        const uint32_t tvn = min(TV, V - v0), tun = min(TU, U - u0);
        uint32_t p = par ^ J.par_xor;
        if (J.dyn) {  // dynamic batch: skip areas that do not cycle now (uniform per block)
            const uint32_t d = *J.dyn;
            if (!(d & 2u)) continue;
            p = d & 1u;
        }
        const uint32_t pu = __builtin_amdgcn_readfirstlane(p) & 1u;
        const uint32_t rows = depth * tvn, cells_t = rows * tun, nd_t = cells_t * dens;
        for (uint32_t x = tid; x < nd_t; x += NTH) acc[x] = 0u;
        const SliceRt s0 = J.st0;
        const uint32_t idx = (uint32_t)wi.tv * J.ntu + wi.tu;   // the item among the job's
        if (idx == 0 && tid == 0) {
            // bit pu: states virtual; bit 2 + pu: best-dendrite fired bits kept
            const uint32_t keep = J.bmask && (W >= 16 || (J.spd % W) == 0);
            auto vb = BSM_G(uint32_t, J.vbits);
            *vb = (*vb & ~(5u << pu)) | (J.virt ? (1u << pu) : 0u) | (keep ? (4u << pu) : 0u);
        }
        // uniform per item: the geometry in scalar registers
        const uint32_t vsc = __builtin_amdgcn_readfirstlane(s0.v_scale);
        const uint32_t usc = __builtin_amdgcn_readfirstlane(s0.u_scale);
        const int32_t vmid = (int32_t)__builtin_amdgcn_readfirstlane(s0.v_mid);
        const int32_t umid = (int32_t)__builtin_amdgcn_readfirstlane(s0.u_mid);
        const uint32_t svs = __builtin_amdgcn_readfirstlane(s0.v_size);
        const uint32_t sus = __builtin_amdgcn_readfirstlane(s0.u_size);
        const int32_t ih = (int32_t)J.halo;
        const int32_t a0 = (int32_t)((v0 * vsc) >> 4) + vmid - ih;
        const int32_t b0 = (int32_t)((u0 * usc) >> 4) + umid - ih;
        const uint32_t nr = (uint32_t)((int32_t)(((v0 + tvn - 1) * vsc) >> 4) -
                                       (int32_t)((v0 * vsc) >> 4) + 1 + 2 * ih);
        const uint32_t nc = (uint32_t)((int32_t)(((u0 + tun - 1) * usc) >> 4) -
                                       (int32_t)((u0 * usc) >> 4) + 1 + 2 * ih);
        const uint32_t WS = nr * nc;
        // slot stride of the staged windows (32 rows with packed offsets)
        const uint32_t WSS = (PACKED && NST > 1) ? 32u * nc : WS;
        const uint32_t n_st = J.n_staged, SS = svs * sus;
        const Quot q_nc = Quot::of(nc), q_ws = Quot::of(WS);
        const uint32_t VU = V * U;

        // ---- every global load of the prologue is issued before any is used:
        // the window bytes (syn_fire of each axon of the tile + halo of each
        // staged slice, zeros off the slice), the item's share of the snapshot
        // of the staged slices (virtual states), the epilogue's inputs
        // fused intake (one staged slice): the window and the snapshot read the
        // frame, which the item's snapshot share also writes into the slice
        const bool fin = frame != nullptr && wi.job == intake_job;
        auto axs = BSM_G(const uint8_t, fin ? frame : J.axn + s0.idz);
        const uint32_t wtot = n_st * WS;
        constexpr int WK = 6;   // window bytes per thread in the first pass
        uint32_t wv[WK], wok = 0;
#pragma unroll
        for (int k = 0; k < WK; k++) {
            // a byte off the slice or past the window reads byte 0 and is zeroed below
            const uint32_t x = tid + (uint32_t)k * NTH;
            const uint32_t j = q_ws(x), xx = x - j * WS;
            const uint32_t rr = q_nc(xx), cc = xx - rr * nc;
            const int32_t gv = a0 + (int32_t)rr, gu = b0 + (int32_t)cc;
            const bool in = x < wtot && (uint32_t)gv < svs && (uint32_t)gu < sus;
            wv[k] = axs[in ? j * SS + (uint32_t)gv * sus + (uint32_t)gu : 0u];
            wok |= (uint32_t)in << k;
        }
        const uint32_t snap_lo = idx * J.snap_per;
        const uint32_t snap_hi = J.virt ? min(snap_lo + J.snap_per, J.snap_len) : 0u;
        uint32_t sv = 0;
        if (snap_lo + tid < snap_hi) sv = axs[snap_lo + tid];
        // the epilogue's inputs, by tile cell index (no row table yet)
        const Quot q_tun = Quot::of(tun), q_tvn = Quot::of(tvn);
        struct DirectMap {
            uint32_t tun, tvn, v0, u0, VU, U, dens;
            Quot q_tun, q_tvn, q_dens;
            __device__ uint32_t cell(uint32_t ct) const {
                const uint32_t r = q_tun(ct), sl = q_tvn(r);
                return sl * VU + (v0 + r - sl * tvn) * U + u0 + (ct - r * tun);
            }
            __device__ uint32_t den_cell(uint32_t x, uint32_t& dj) const {
                const uint32_t ct = q_dens(x);
                dj = x - ct * dens;
                return ct;
            }
        };
        const DirectMap dmap{tun, tvn, v0, u0, VU, U, dens, q_tun, q_tvn, Quot::of(dens)};
        TailPre pre;
        tail_prefetch(J, dmap, nd_t, cells_t, tid, NTH, pre);
        // ---- then the stores
#pragma unroll
        for (int k = 0; k < WK; k++) {
            const uint32_t x = tid + (uint32_t)k * NTH;
            if (x < wtot) {
                const uint32_t j = q_ws(x);
                lds[j * WSS + (x - j * WS)] = (uint8_t)syn_fire(((wok >> k) & 1u) ? wv[k] : 0u);
            }
        }
        for (uint32_t x = tid + WK * NTH; x < wtot; x += NTH) {   // larger windows
            const uint32_t j = q_ws(x), xx = x - j * WS;
            const uint32_t rr = q_nc(xx), cc = xx - rr * nc;
            const int32_t gv = a0 + (int32_t)rr, gu = b0 + (int32_t)cc;
            uint32_t val = 0;
            if ((uint32_t)gv < svs && (uint32_t)gu < sus) val = axs[j * SS + (uint32_t)gv * sus + (uint32_t)gu];
            lds[j * WSS + xx] = (uint8_t)syn_fire(val);
        }
        {
            auto dst = BSM_G(uint8_t, J.snap + (size_t)pu * J.snap_len);
            auto ain = BSM_G(uint8_t, intake_dst);
            if (snap_lo + tid < snap_hi) {
                dst[snap_lo + tid] = (uint8_t)sv;
                if (fin) ain[snap_lo + tid] = (uint8_t)sv;
            }
            for (uint32_t x = snap_lo + tid + NTH; x < snap_hi; x += NTH) {
                const uint8_t b = axs[x];
                dst[x] = b;
                if (fin) ain[x] = b;
            }
        }
        // per tile row (slice, v): synapse offset and window offset of its first
        // cell, and the cell itself
        uint32_t* rowt = (uint32_t*)(lds + (((n_st - 1) * WSS + WS + 15) & ~15u));
        uint32_t* rowc = rowt + 2 * rows;
        for (uint32_t rq = tid; rq < rows; rq += NTH) {
            const uint32_t sl = q_tvn(rq), v = v0 + (rq - sl * tvn);
            const uint32_t c0 = sl * VU + v * U + u0;
            rowc[rq] = c0;
            rowt[2 * rq] = c0 * spt;
            rowt[2 * rq + 1] = (uint32_t)(((int32_t)((v * vsc) >> 4) + vmid - a0) * (int32_t)nc +
                                          umid - b0);
        }
        __syncthreads();
        const RowMap map{rowc, tun, dens, q_tun, Quot::of(dens)};
        const int32_t rnd = J.rnd[rnd_ofs + J.rnd_slot];

        // one buffer resource per synapse stream of the tuft (< 2^32 bytes)
        const uint32_t nrec = J.C * spt;
        const bool multi = NST > 1 && n_st > 1;
        // staged-slice codes: 32-bit words of 16 synapses (synapse (w, k) of a
        // group at bit 8k + 2w); a job with one staged slice reads a zero-length
        // stream, whose loads return 0 without touching memory
        const Rsrc r_slc = make_rsrc(J.sb, multi ? (nrec + 15) / 16 * 4 : 0u);
        const Rsrc r_v = make_rsrc(PACKED ? J.vu : (const uint8_t*)J.vofs, nrec);
        const Rsrc r_u = make_rsrc(J.uofs, nrec);
        const Rsrc r_sg = make_rsrc(J.sg, J.sg_len);
        const uint32_t sgb0 = J.sg_bit0;
        // a job whose states are not virtual writes them (its fired window bytes)
        const Rsrc r_out = make_rsrc(J.st[pu], J.virt ? 0u : nrec);
        const bool wr = !J.virt;
        // a virtual spiny stellate tuft stores each chunk's fired bits (W / 8
        // bytes at byte so / 8) for its layer's learning (launches of one staged
        // slice and chunks of 8 or more synapses; the planner keeps every other
        // spiny stellate tuft writing its states)
        constexpr bool FB = NST == 1 && W >= 8;
        const bool fbw = FB && J.fb != nullptr;
        const Rsrc r_fb = make_rsrc(J.fb, fbw ? (nrec + 7) / 8 : 0u);
        // packed offsets carry v + 8 and u + 8: their bias folds into the base
        const int32_t pbias = PACKED ? 8 * (int32_t)nc + 8 : 0;
        const uint32_t coefA = (nc & 0xFFu) | (1u << 8);
        const uint32_t coefB = ((nc & 0xFFu) << 16) | (1u << 24);
        const uint32_t CPT = spt / W, row_chunks = tun * CPT, n_chunks = rows * row_chunks;
        const Quot q_rc = Quot::of(row_chunks), q_cpt = Quot::of(CPT);
        const bool whole = (J.spd % W) == 0;
        const uint32_t m_spd = 0xFFFFFFFFu / J.spd + 1u;
        const uint32_t last = n_chunks - 1;
        // distal tufts: each chunk's fired bits, for the best dendrites' copy
        uint16_t* lmask = (J.bmask && (W >= 16 || whole))
                              ? (uint16_t*)(lds + ((stage_cap + nd_t * 6 + 3) & ~3u)) : nullptr;

        // MS: the job stages several slices (reads staged-slice codes); a job of
        // one staged slice in a several-slice launch runs the loop without them
        auto fetch = [&](auto MS_, uint32_t q, ChunkMeta<NW>& m) {
            constexpr bool MS = decltype(MS_)::value;
            m.q = q;
            const uint32_t rowq = q_rc(q);
            const uint32_t rem = q - __umul24(rowq, row_chunks);
            const uint32_t cc = q_cpt(rem);
            m.dd = __umul24(__umul24(rowq, tun) + cc, dens) |
                   (((rem - __umul24(cc, CPT)) * W) << 16);   // < 2^16 each (planner)
            const uint32_t so = rowt[2 * rowq] + rem * W;
            m.so = so;
            m.lb = (int32_t)rowt[2 * rowq + 1] + (int32_t)(__umul24(u0 + cc, usc) >> 4);
            if constexpr (MS && W == 32) {
                const u32x2 cc2 = __builtin_amdgcn_raw_buffer_load_b64(r_slc, (int)((so >> 4) << 2), 0, BUF_NT);
                m.c[0] = cc2.x;
                m.c[1] = cc2.y;
            } else if constexpr (MS) {
                m.c[0] = __builtin_amdgcn_raw_buffer_load_b32(r_slc, (int)((so >> 4) << 2), 0, BUF_NT);
            }
            m.s = __builtin_amdgcn_raw_buffer_load_b32(r_sg, (int)(((sgb0 + so) >> 5) << 2), 0, BUF_NT);
            Chunk<W>::load_b(r_v, so, m.v);   // packed: (v + 8) << 4 | (u + 8) per byte
            if constexpr (!PACKED) Chunk<W>::load_b(r_u, so, m.u);
        };
        auto work = [&](auto MS_, const ChunkMeta<NW>& m, bool valid) {
            constexpr bool MS = decltype(MS_)::value;
            // window addresses as LDS byte addresses (the dynamic LDS base folded in)
            const int32_t lbp = m.lb - pbias + (int32_t)lds0;
            // codes of word w (within its 16-synapse group: bits 8k + 2(w % 4))
            uint32_t cwv[(NW + 3) / 4] = {};
            if constexpr (MS) {
                if constexpr (W == 32) { cwv[0] = m.c[0]; cwv[1] = m.c[1]; }
                else cwv[0] = m.c[0] >> ((m.so & 12u) >> 1);
            }
            const uint32_t nsg = ~(m.s >> (((sgb0 + m.so) & 31u) >> 2));  // word w: bit 8k + w
                                                                          // set if strength >= 0
            const uint32_t den0 = m.dd & 0xFFFFu, off0 = m.dd >> 16;
            uint32_t x[NW], aa[NW][4];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                uint32_t* a = aa[w];
                if constexpr (PACKED) {
                    uint32_t vw = (m.v[w] >> 4) & 0x0F0F0F0Fu;
                    const uint32_t uw = m.v[w] & 0x0F0F0F0Fu;
                    if constexpr (MS) {
                        // code k of this word to bits 5-6 of byte k: v + 8 + 32 * code
                        const uint32_t cw = cwv[w >> 2];
                        const int w4 = w & 3;
                        const uint32_t cs = (2 * w4 <= 5) ? (cw << (5 - 2 * w4)) : (cw >> (2 * w4 - 5));
                        vw |= cs & 0x60606060u;
                    }
                    // [v_b, u_b, v_b+1, u_b+1] byte pairs -> dot4 gives base + v*cols + u
                    const uint32_t p01 = __builtin_amdgcn_perm(uw, vw, 0x05010400u);
                    const uint32_t p23 = __builtin_amdgcn_perm(uw, vw, 0x07030602u);
                    a[0] = __builtin_amdgcn_udot4(p01, coefA, (uint32_t)lbp, false);
                    a[1] = __builtin_amdgcn_udot4(p01, coefB, (uint32_t)lbp, false);
                    a[2] = __builtin_amdgcn_udot4(p23, coefA, (uint32_t)lbp, false);
                    a[3] = __builtin_amdgcn_udot4(p23, coefB, (uint32_t)lbp, false);
                } else {
                    const uint32_t p01 = __builtin_amdgcn_perm(m.u[w], m.v[w], 0x05010400u);
                    const uint32_t p23 = __builtin_amdgcn_perm(m.u[w], m.v[w], 0x07030602u);
                    int32_t l[4] = {lbp, lbp, lbp, lbp};
                    if constexpr (MS) {
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            l[k] = (int32_t)__umul24((cwv[w >> 2] >> (8 * k + 2 * (w & 3))) & 3u, WSS) + lbp;
                    }
                    a[0] = (uint32_t)__builtin_amdgcn_sdot4((int32_t)p01, (int32_t)coefA, l[0], false);
                    a[1] = (uint32_t)__builtin_amdgcn_sdot4((int32_t)p01, (int32_t)coefB, l[1], false);
                    a[2] = (uint32_t)__builtin_amdgcn_sdot4((int32_t)p23, (int32_t)coefA, l[2], false);
                    a[3] = (uint32_t)__builtin_amdgcn_sdot4((int32_t)p23, (int32_t)coefB, l[3], false);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            uint32_t bb[NW][4];
#pragma unroll
            for (int w = 0; w < NW; w++)
#pragma unroll
                for (int k = 0; k < 4; k++) bb[w][k] = *LDS8(aa[w][k]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int w = 0; w < NW; w++)
                x[w] = __builtin_amdgcn_perm(__builtin_amdgcn_perm(bb[w][3], bb[w][2], 0x0c0c0400u),
                                             __builtin_amdgcn_perm(bb[w][1], bb[w][0], 0x0c0c0400u),
                                             0x05040100u);
            if (wr && valid) Chunk<W>::store_b(r_out, m.so, x);
            if ((lmask || (FB && fbw)) && valid) {
                // fired = bit 7 of each state byte; bit 4w + k for byte k of word w
                uint32_t bt[(NW + 1) / 2];
#pragma unroll
                for (int h = 0; h < (NW + 1) / 2; h++) {
                    bt[h] = __builtin_amdgcn_udot4((x[2 * h] >> 7) & 0x01010101u, 0x08040201u, 0u, false);
                    if (2 * h + 1 < NW)
                        bt[h] = __builtin_amdgcn_udot4((x[2 * h + 1] >> 7) & 0x01010101u, 0x80402010u,
                                                       bt[h], false);
                }
                if constexpr (W == 32) {
                    const uint32_t fv = bt[0] | (bt[1] << 8) | (bt[2] << 16) | (bt[3] << 24);
                    if (lmask) ((uint32_t*)lmask)[m.q] = fv;
                    if constexpr (FB)
                        if (fbw) __builtin_amdgcn_raw_buffer_store_b32(fv, r_fb, (int)(m.so >> 3), 0, 0);
                } else if constexpr (NW > 2) {
                    const uint32_t fv = bt[0] | (bt[1] << 8);
                    if (lmask) lmask[m.q] = (uint16_t)fv;
                    if constexpr (FB)
                        if (fbw) __builtin_amdgcn_raw_buffer_store_b16((uint16_t)fv, r_fb, (int)(m.so >> 3), 0, 0);
                } else {
                    if (lmask) lmask[m.q] = (uint16_t)bt[0];
                    if constexpr (FB && W == 8)
                        if (fbw) __builtin_amdgcn_raw_buffer_store_b8((uint8_t)bt[0], r_fb, (int)(m.so >> 3), 0, 0);
                }
            }
            if (whole) {
                uint32_t ps = 0, rs = 0;
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    ps = __builtin_amdgcn_udot4(x[w], (nsg >> w) & 0x01010101u, ps, false);
                    rs = __builtin_amdgcn_sad_u8(x[w], 0u, rs);
                }
                if (valid) atomicAdd(&acc[den0 + __umulhi(off0, m_spd)], ps | (rs << 16));
            } else if (valid) {
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const uint32_t ps = __builtin_amdgcn_udot4(x[w], (nsg >> w) & 0x01010101u, 0u, false);
                    const uint32_t rs = __builtin_amdgcn_sad_u8(x[w], 0u, 0u);
                    atomicAdd(&acc[den0 + __umulhi(off0 + (uint32_t)(4 * w), m_spd)],
                              ps | (rs << 16));
                }
            }
        };
        // S chunk slots per lane (4 when the tile has more than two chunks per
        // lane, else 2), filled in order before the first is gathered; each slot
        // is refilled with its chunk of the next round as soon as it has been
        // gathered, so that a lane keeps S - 1 to S chunks' loads in flight
        // throughout.  The loads are unconditional, their indices clamped to the
        // last chunk: a chunk past the end re-gathers the last one and is not
        // used, and the last round's refills are never gathered (a branch around
        // them, or offsets past the streams' ends for them, measured slower).
        auto run = [&](auto S_, auto MS_) {
            constexpr int S = decltype(S_)::value;
            ChunkMeta<NW> M[S];
#pragma unroll
            for (int k = 0; k < S; k++) {
                fetch(MS_, min(tid + (uint32_t)k * NTH, last), M[k]);
                __builtin_amdgcn_sched_barrier(0);
            }
            for (uint32_t q0 = tid; q0 < n_chunks; q0 += S * NTH) {
#pragma unroll
                for (int k = 0; k < S; k++) {
                    work(MS_, M[k], q0 + (uint32_t)k * NTH < n_chunks);
                    fetch(MS_, min(q0 + (uint32_t)(S + k) * NTH, last), M[k]);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        };
        auto run_s = [&](auto MS_) {
            if (W < 32 && n_chunks > 2 * NTH) run(std::integral_constant<int, 4>{}, MS_);
            else run(std::integral_constant<int, 2>{}, MS_);
        };
        if constexpr (NST > 1 && W == 32) {
            if (__builtin_amdgcn_readfirstlane((uint32_t)multi)) run_s(std::true_type{});
            else run_s(std::false_type{});
        } else {
            // narrower chunks keep four chunks in flight: one loop body (a second
            // would raise the register count past five waves per SIMD)
            run_s(std::integral_constant<bool, (NST > 1)>{});
        }
        __syncthreads();
        const MaskOut mo{lmask, J.bmask + (size_t)pu * J.bm_len, CPT, J.spd, J.wpd, (uint32_t)W};
        tuft_tail(J, lds, acc, stage_cap, nd_t, cells_t, map, rnd, tid, NTH, pre, mo);
    }
}

// F1, checked form: any offsets, non-staged slices and the scalar char-wrapped
// mid quirk (Q10); staged windows with bounds checks and a global fallback.
template <int W, int NST>
__global__ __launch_bounds__(256) void kf_tuft(const TuftJob* __restrict__ jobs,
                                               const WorkItem* __restrict__ items, uint32_t n_items,
                                               uint32_t par, uint32_t rnd_ofs,
                                               uint32_t stage_cap) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int NB = W < 4 ? W : 4;               // bytes per word
    constexpr int NW = Chunk<W>::NW;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    uint32_t* acc = (uint32_t*)(lds + stage_cap);

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem wi = items[it];
        const TuftJob& J = jobs[wi.job];
        const uint32_t V = J.V, U = J.U, depth = J.depth, spt = J.spt, halo = J.halo;
        const uint32_t TV = J.tile_v, TU = J.tile_u;
        const uint32_t v0 = (uint32_t)wi.tv * TV, u0 = (uint32_t)wi.tu * TU;
        const uint32_t tvn = min(TV, V - v0), tun = min(TU, U - u0);
        uint32_t p = par ^ J.par_xor;
        if (J.dyn) {  // dynamic batch: skip areas that do not cycle now (uniform per block)
            const uint32_t d = *J.dyn;
            if (!(d & 2u)) continue;
            p = d & 1u;
        }
        const bool quirk = J.scalar_quirk != 0;
        const uint32_t cells_t = depth * tvn * tun, nd_t = cells_t * J.dens;
        const DivMap map{tvn, tun, v0, u0, V * U, U, J.dens};
        TailPre pre;
        tail_prefetch(J, map, nd_t, cells_t, tid, nth, pre);
        for (uint32_t x = tid; x < nd_t; x += nth) acc[x] = 0u;
        const bool whole = (J.spd % W) == 0;
        const uint32_t magic = 0xFFFFFFFFu / J.spd + 1u;

        // ---- stage the tile's hex neighbourhood of each staged source slice
        int32_t vlo[NST], ulo[NST], rows[NST], cols[NST], lbase[NST];
        uint32_t sid[NST], vsc[NST], usc[NST], vmd[NST], umd[NST];
        uint32_t off = 0;
#pragma unroll
        for (int j = 0; j < NST; j++) {
            rows[j] = 0; cols[j] = 0; lbase[j] = 0; vlo[j] = 0; ulo[j] = 0;
            sid[j] = 0xFFFFFFFFu; vsc[j] = usc[j] = vmd[j] = umd[j] = 0;
            if ((uint32_t)j < J.n_staged) {
                sid[j] = J.staged[j];
                SliceRt s = J.slices[sid[j]];
                vsc[j] = s.v_scale; usc[j] = s.u_scale; vmd[j] = s.v_mid; umd[j] = s.u_mid;
                int32_t a0 = (int32_t)((v0 * vsc[j]) >> 4) + (int32_t)vmd[j] - (int32_t)halo;
                int32_t a1 = (int32_t)(((v0 + tvn - 1) * vsc[j]) >> 4) + (int32_t)vmd[j] + (int32_t)halo;
                int32_t b0 = (int32_t)((u0 * usc[j]) >> 4) + (int32_t)umd[j] - (int32_t)halo;
                int32_t b1 = (int32_t)(((u0 + tun - 1) * usc[j]) >> 4) + (int32_t)umd[j] + (int32_t)halo;
                a0 = max(a0, 0); b0 = max(b0, 0);
                a1 = min(a1, (int32_t)s.v_size - 1); b1 = min(b1, (int32_t)s.u_size - 1);
                int32_t nr = a1 >= a0 ? a1 - a0 + 1 : 0, nc = b1 >= b0 ? b1 - b0 + 1 : 0;
                if (off + (uint32_t)(nr * nc) > stage_cap) { nr = 0; nc = 0; }
                vlo[j] = a0; ulo[j] = b0; rows[j] = nr; cols[j] = nc; lbase[j] = (int32_t)off;
                auto src = BSM_G(const uint8_t, J.axn + s.idz);
                const uint32_t n = (uint32_t)(nr * nc);
                for (uint32_t x = tid; x < n; x += nth) {
                    uint32_t rr = x / (uint32_t)nc, cc = x - rr * (uint32_t)nc;
                    lds[off + x] = src[(uint32_t)(a0 + (int32_t)rr) * s.u_size + (uint32_t)b0 + cc];
                }
                off += n;
            }
        }
        __syncthreads();

        // ---- synapse chunks: gather + fire + store + per-chunk dendrite partials
        const int32_t rnd = J.rnd[rnd_ofs + J.rnd_slot];
        const uint32_t CPT = spt / W;
        const uint32_t row_chunks = tun * CPT;
        const uint32_t n_chunks = depth * tvn * row_chunks;
        const uint32_t VU = V * U;
        uint8_t* __restrict__ st_out = J.st[p & 1];
        if (wi.tv == 0 && wi.tu == 0 && tid == 0) {  // parity p's states are written
            auto vb = BSM_G(uint32_t, J.vbits);
            *vb = *vb & ~(5u << (p & 1));
        }
        for (uint32_t q = tid; q < n_chunks; q += nth) {
            const uint32_t rowq = q / row_chunks, rem = q - rowq * row_chunks;
            const uint32_t sl = rowq / tvn, r = rowq - sl * tvn;
            const uint32_t cc = rem / CPT, k = rem - cc * CPT;
            const uint32_t v = v0 + r, u = u0 + cc;
            const uint32_t cel = sl * VU + v * U + u;
            const size_t so = (size_t)cel * spt + (size_t)k * W;
            int32_t av[NST], au[NST];
#pragma unroll
            for (int j = 0; j < NST; j++) {
                av[j] = (int32_t)((v * vsc[j]) >> 4) + (int32_t)vmd[j] - vlo[j];
                au[j] = (int32_t)((u * usc[j]) >> 4) + (int32_t)umd[j] - ulo[j];
            }
            uint32_t ms[NW], mv[NW], mu[NW], mt[NW], out[NW];
            Chunk<W>::load(J.slc + so, ms);
            Chunk<W>::load(J.vofs + so, mv);
            Chunk<W>::load(J.uofs + so, mu);
            Chunk<W>::load(J.str + so, mt);
            uint32_t pw[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                uint32_t ow = 0, psum = 0, praw = 0;
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const uint32_t sl_id = byte_of(ms[w], b);
                    const int32_t vo = sbyte_of(mv[w], b), uo = sbyte_of(mu[w], b);
                    int32_t lv = -1, lu = -1, nr = 0, nc = 0, lb = 0;
#pragma unroll
                    for (int j = 0; j < NST; j++) {
                        if (sl_id == sid[j]) {
                            lv = av[j] + vo; lu = au[j] + uo; nr = rows[j]; nc = cols[j]; lb = lbase[j];
                        }
                    }
                    uint32_t ax;
                    if ((uint32_t)lv < (uint32_t)nr && (uint32_t)lu < (uint32_t)nc) {
                        ax = lds[lb + lv * nc + lu];
                    } else {
                        ax = axn_gather(J.axn, J.slices, J.n_slices, sl_id, (int32_t)v, vo,
                                        (int32_t)u, uo, quirk);
                    }
                    const uint32_t stv = syn_fire(ax);
                    ow |= stv << (8 * b);
                    praw += stv;
                    psum += (sbyte_of(mt[w], b) >= 0) ? stv : 0u;
                }
                out[w] = ow;
                pw[w] = psum | (praw << 16);
            }
            Chunk<W>::store(st_out + so, out);
            acc_chunk<NW, NB>(acc, (rowq * tun + cc) * J.dens, k * W, pw, whole, magic);
        }
        __syncthreads();

        tuft_tail(J, lds, acc, stage_cap, nd_t, cells_t, map, rnd, tid, nth, pre, MaskOut{});
    }
}

// ------------------------------------------------------------------ F2
// One chunk of LTP over W synapses from index i0 (prx_syns__active__mtp_ltd,
// cl/bismit.cl:604-639, and dst_syns__active__mtpot_mtdep, :554-600): a synapse
// whose state byte is nonzero is incremented with mask 0x7F, any other one
// decremented with `dec_mask`; the seed is z + i with i the absolute synapse
// index (Q5).  load() and apply() are split so that a thread can have several
// chunks' loads in flight before it computes the first.
template <int W>
struct LtpChunk {
    static constexpr int NB = W < 4 ? W : 4;
    static constexpr int NW = Chunk<W>::NW;
    uint32_t ms[NW], mt[NW];
    __device__ inline void load(const uint8_t* __restrict__ st, const int8_t* __restrict__ str,
                                uint32_t i0) {
        Chunk<W>::load(st + i0, ms);
        Chunk<W>::load(str + i0, mt);
    }
    // Virtual states: the chunk's source axons in a snapshot stand in for the
    // state bytes (a state is nonzero exactly when its axon is).  i0 is the
    // layer index of the chunk, t0 the tuft's first synapse, cel its cell.
    __device__ inline void load_virt(const VirtTuft& T, const uint8_t* snap,
                                     const int8_t* __restrict__ str, uint32_t i0, uint32_t t0,
                                     uint32_t cel) {
        uint32_t sl[NW], vv[NW], uu[NW];
        Chunk<W>::load(T.slc + (i0 - t0), sl);
        Chunk<W>::load(T.vofs + (i0 - t0), vv);
        Chunk<W>::load(T.uofs + (i0 - t0), uu);
        Chunk<W>::load(str + i0, mt);
#pragma unroll
        for (int w = 0; w < NW; w++) {
            uint32_t o = 0;
#pragma unroll
            for (int b = 0; b < NB; b++)
                o |= virt_axon(T, snap, cel, byte_of(sl[w], b), sbyte_of(vv[w], b),
                               sbyte_of(uu[w], b)) << (8 * b);
            ms[w] = o;
        }
    }
    // Previous states from a distal tuft's kept best-dendrite fired bits (the
    // dendrite's words at bits; bytes 0 or 1): the chunk starts at synapse k0 of
    // the dendrite.
    __device__ inline void load_bits(const uint32_t* bits_d, uint32_t k0,
                                     const int8_t* __restrict__ str, uint32_t i0) {
        const uint32_t word = BSM_G(const uint32_t, bits_d)[k0 >> 5];
        const uint32_t bits = word >> (k0 & 31u);
        Chunk<W>::load(str + i0, mt);
#pragma unroll
        for (int w = 0; w < NW; w++) {
            if constexpr (W == 1) ms[w] = bits & 1u;
            else ms[w] = spread4((bits >> (4 * w)) & 0xFu);
        }
    }
    // New strengths of the chunk; the sign bits of those whose sign changed are
    // flipped in the layer's sign words (a chunk of W <= 16 synapses lies in one
    // word's 32 synapses).
    __device__ inline void apply(int8_t* __restrict__ str, uint32_t z, uint32_t i0, int32_t rnd,
                                 uint32_t dec_mask, uint32_t* __restrict__ sg) const {
        uint32_t out[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) {
            uint32_t ow = 0;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const uint32_t i = i0 + (uint32_t)(w * 4 + b);
                const int32_t seed = (int32_t)(z + i);
                int32_t s = sbyte_of(mt[w], b);
                if (byte_of(ms[w], b)) s += rnd_inc(rnd, seed, s, 0x7F);
                else s -= rnd_dec(rnd, seed, s, dec_mask);
                ow |= ((uint32_t)s & 0xFFu) << (8 * b);
            }
            out[w] = ow;
        }
        Chunk<W>::store(str + i0, out);
        uint32_t flip = 0;
        if constexpr (W == 1) {
            if ((mt[0] ^ out[0]) & 0x80u) flip = 1u << sg_pos(i0);
        } else {
#pragma unroll
            for (int w = 0; w < NW; w++)
                flip |= spread4(sign_bits4(mt[w] ^ out[w])) << (((i0 & 31u) >> 2) + (uint32_t)w);
        }
        if (flip) atomicXor(sg + (i0 >> 5), flip);
    }
};
// LTP chunks each thread loads before it applies the first: four of the
// narrow (4-byte) chunks batches of mixed dendrite lengths use, so that one
// memory latency covers sixteen synapses; wide chunks one at a time (more in
// flight cost registers and occupancy: measured slower).
template <int W>
constexpr int ltp_u() { return W <= 4 ? 4 : 1; }

// inhib_simple's test for radius R fully unrolled (cl/control.cl:28-135): a
// neighbour at hex distance d inhibits when n * (M - (d - c)^2) > cs * M, so
// only the largest neighbour state of each ring matters (the factor depends on
// d alone, and a non-positive factor never inhibits; the cell itself, d = 0,
// never does).  ctr is the cell's window byte, rows are WU bytes apart.
template <int R>
__device__ inline uint32_t hex_uninhibited(const uint8_t* win, uint32_t ctr, uint32_t WU) {
    constexpr int M = (R + INHIB_HORIZ_OFS) * (R + INHIB_HORIZ_OFS);
    uint32_t rmax[R + 1];
#pragma unroll
    for (int d = 0; d <= R; d++) rmax[d] = 0;
#pragma unroll
    for (int vo = -R; vo <= R; vo++) {
        // the row's first byte (u offset -R); the bytes read are u offsets within the hex
        const uint8_t* row = win + (ctr - (uint32_t)R + (uint32_t)(vo * (int)WU));
#pragma unroll
        for (int uo = -R; uo <= R; uo++) {
            const int w = -vo - uo;
            if (w < -R || w > R || (vo == 0 && uo == 0)) continue;
            const int d = ((vo < 0 ? -vo : vo) + (uo < 0 ? -uo : uo) + (w < 0 ? -w : w)) >> 1;
            rmax[d] = max(rmax[d], (uint32_t)row[uo + R]);
        }
    }
    const int32_t ci = (int32_t)win[ctr] * M;
    uint32_t unin = 1;
#pragma unroll
    for (int d = 1; d <= R; d++) {
        const int f = M - (d - INHIB_CENTER_OFS) * (d - INHIB_CENTER_OFS);
        if (f > 0) unin &= (uint32_t)(ci >= (int32_t)rmax[d] * f);
    }
    return unin;
}

constexpr uint32_t F2_MAXC = 4;  // tile cells per thread (tiles up to 4 x 256 cells)
constexpr uint32_t F2_WK = 7;    // soma window bytes per thread loaded ahead of their stores

template <int W>
__global__ __launch_bounds__(256) void kf_ssc_post(const SscJob* __restrict__ jobs,
                                                   const WorkItem* __restrict__ items,
                                                   uint32_t n_items, uint32_t par, uint32_t rnd_ofs,
                                                   uint32_t TV, uint32_t TU) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t n_act, n_hex;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem wi = items[it];
        const SscJob& J = jobs[wi.job];
        const uint32_t V = J.V, U = J.U, VU = V * U, slc = wi.slc;
        const uint32_t v0 = (uint32_t)wi.tv * TV, u0 = (uint32_t)wi.tu * TU;
        const uint32_t tvn = min(TV, V - v0), tun = min(TU, U - u0);
        const int32_t r = J.mode == 0 ? (int32_t)J.radius : 0;
        const uint32_t WV = tvn + 2 * r, WU = tun + 2 * r;
        uint32_t p = par ^ J.par_xor;
        if (J.dyn) {
            const uint32_t d = *J.dyn;
            if (!(d & 2u)) continue;
            p = d & 1u;
        }
        uint32_t* list = (uint32_t*)(lds + ((WV * WU + 15) & ~15u));  // learning cells
        uint16_t* hexl = (uint16_t*)(list + TV * TU);  // tile cells that need the hex test
        uint8_t* hexo = (uint8_t*)(hexl + TV * TU);    // their uninhibited flags
        const uint8_t* soma = J.soma + (size_t)slc * VU;
        // loads independent of the window, issued with it ahead of the barrier:
        // rnd values, the own slice entry, this thread's first cell's rating
        const int32_t rnd_c = J.mode == 0 ? J.rnd[rnd_ofs + J.slot_ctrl] : 0;
        const int32_t rnd_m = J.learn ? J.rnd[rnd_ofs + J.slot_mtp] : 0;
        // where this cycle's states of the learned tuft are (bit p: virtual)
        const uint32_t vb0 = J.learn ? *BSM_G(const uint32_t, J.vbits) : 0u;
        const SliceRt own = J.slices[J.base_slc + slc];
        uint32_t actp[F2_MAXC];  // activity ratings of this thread's cells
#pragma unroll
        for (uint32_t k = 0; k < F2_MAXC; k++) {
            const uint32_t x = tid + k * nth;
            actp[k] = 0;
            if (J.mode != 2 && x < tvn * tun) {
                const uint32_t tr = x / tun, tc = x - tr * tun;
                actp[k] = BSM_G(const uint8_t, J.cel_act)[slc * VU + (v0 + tr) * U + (u0 + tc)];
            }
        }
        // the soma window: every thread issues its first F2_WK bytes' loads
        // before it stores any (one round trip for windows up to F2_WK * 256
        // bytes, a 32 x 32 tile at radius 4), bytes off the layer read 0
        const uint32_t WN = WV * WU;
        uint32_t wv[F2_WK];
#pragma unroll
        for (uint32_t k = 0; k < F2_WK; k++) {
            const uint32_t x = tid + k * nth;
            const uint32_t wr = x / WU, wc = x - wr * WU;
            const int32_t nv = (int32_t)v0 - r + (int32_t)wr, nu = (int32_t)u0 - r + (int32_t)wc;
            wv[k] = 0;
            if (x < WN && nv >= 0 && nv < (int32_t)V && nu >= 0 && nu < (int32_t)U)
                wv[k] = soma[nv * (int32_t)U + nu];
        }
#pragma unroll
        for (uint32_t k = 0; k < F2_WK; k++) {
            const uint32_t x = tid + k * nth;
            if (x < WN) lds[x] = (uint8_t)wv[k];
        }
        for (uint32_t x = tid + F2_WK * nth; x < WN; x += nth) {
            const uint32_t wr = x / WU, wc = x - wr * WU;
            const int32_t nv = (int32_t)v0 - r + (int32_t)wr, nu = (int32_t)u0 - r + (int32_t)wc;
            uint8_t val = 0;
            if (nv >= 0 && nv < (int32_t)V && nu >= 0 && nu < (int32_t)U) val = soma[nv * U + nu];
            lds[x] = val;
        }
        if (tid == 0) {
            n_act = 0;
            n_hex = 0;
        }
        __syncthreads();

        const int32_t imax = (r + INHIB_HORIZ_OFS) * (r + INHIB_HORIZ_OFS);
        // The hex test decides only cells with a non-zero soma (a few percent of a
        // spiny stellate layer's), packed densely so that lanes of silent cells do
        // not run it: for radii other than the reference's (the generic loop), and
        // at radius 4 for tiles of more than one cell per thread (the unrolled
        // test, about once per tile instead of once per thread's cell).  Tiles of
        // one cell per thread run the unrolled test in the cell's own thread
        // below, with no packing barrier.
        const bool pack = J.mode == 0 && (r != 4 || tvn * tun > nth);
        if (pack) {
            for (uint32_t x = tid; x < tvn * tun; x += nth) {
                const uint32_t tr = x / tun, tc = x - tr * tun;
                int safe;
                (void)cel_axn_idx(own, v0 + tr, u0 + tc, &safe);
                if (safe && lds[(tr + r) * WU + tc + r]) hexl[atomicAdd(&n_hex, 1u)] = (uint16_t)x;
            }
            __syncthreads();
            const uint32_t nh = n_hex;
            for (uint32_t y = tid; y < nh; y += nth) {
                const uint32_t x = hexl[y];
                const uint32_t tr = x / tun, tc = x - tr * tun;
                uint32_t unin = 1;
                if (r == 4) {
                    unin = hex_uninhibited<4>(lds, (tr + r) * WU + tc + r, WU);
                } else {
                    const int32_t ci = (int32_t)lds[(tr + r) * WU + tc + r] * imax;
                    for (int32_t vo = -r; vo <= r && unin; vo++) {
                        const int32_t uz = max(-r, -vo - r), um = min(r, -vo + r);
                        const uint8_t* row = lds + (uint32_t)((int32_t)tr + r + vo) * WU + tc + r;
                        for (int32_t uo = uz; uo <= um; uo++) {
                            const int32_t w = -vo - uo;
                            const int32_t dist = (abs(vo) + abs(uo) + abs(w)) >> 1;
                            const int32_t dc = dist - INHIB_CENTER_OFS;
                            unin &= (uint32_t)(ci >= (int32_t)row[uo] * (imax - dc * dc));
                        }
                    }
                }
                hexo[x] = (uint8_t)unin;
            }
            __syncthreads();
        }
#pragma unroll
        for (uint32_t k = 0; k < F2_MAXC; k++) {
            const uint32_t x = tid + k * nth;
            if (x >= tvn * tun) break;
            const uint32_t tr = x / tun, tc = x - tr * tun;
            const uint32_t v = v0 + tr, u = u0 + tc;
            const uint32_t c = slc * VU + v * U + u;
            uint32_t out;
            if (J.mode == 2) {
                out = J.axn[own.idz + v * U + u];
            } else {
                int safe;
                const uint32_t axi = cel_axn_idx(own, v, u, &safe);
                uint32_t cs = lds[(tr + r) * WU + tc + r];
                uint32_t is_act;
                if (J.mode == 0) {
                    cs = safe ? cs : 0u;
                    // radius 4, the reference's (cl/control.cl:9): the unrolled test
                    const uint32_t unin =
                        cs ? (pack ? (uint32_t)hexo[x]
                                   : hex_uninhibited<4>(lds, (tr + r) * WU + tc + r, WU))
                           : 1u;
                    out = unin ? cs : 0u;
                    is_act = unin & (cs != 0);
                } else {
                    out = cs;
                    is_act = cs != 0;
                }
                J.axn[axi] = (uint8_t)out;
                J.cel_act[c] = (uint8_t)activity(actp[k], (int32_t)is_act, rnd_c,
                                                 (int32_t)(cs | c), CEL_ACT_DECAY);
            }
            if (J.learn && out) list[atomicAdd(&n_act, 1u)] = c;
        }
        __syncthreads();

        if (J.learn) {
            // ssc_mtp_simple on the listed cells: state nonzero -> increment (0x7F),
            // else decrement (0x1FF); ltp_u chunks per thread in flight at once (a
            // chunk past the list loads chunk 0 again and is not applied)
            const uint32_t na = n_act, spt = J.spt, CPC = spt / W, total = na * CPC;
            // this cycle's states: virtual -> the fired bits F1 stored, else the
            // state buffer
            const bool virt = (vb0 >> (p & 1)) & 1u;
            const uint8_t* __restrict__ st = J.st[p & 1];
            constexpr int LU = ltp_u<W>();
            for (uint32_t y0 = tid; y0 < total; y0 += LU * nth) {
                LtpChunk<W> ch[LU];
                uint32_t zz[LU], ii[LU];
#pragma unroll
                for (int j = 0; j < LU; j++) {
                    const uint32_t y = y0 + j * nth < total ? y0 + j * nth : 0u;
                    const uint32_t ci = y / CPC, k = y - ci * CPC;
                    zz[j] = list[ci] * spt;
                    ii[j] = zz[j] + k * W;
                    if (virt) ch[j].load_bits(J.fb, ii[j], J.str, ii[j]);   // tuft 0 starts the layer
                    else ch[j].load(st, J.str, ii[j]);
                }
#pragma unroll
                for (int j = 0; j < LU; j++)
                    if (y0 + j * nth < total) ch[j].apply(J.str, zz[j], ii[j], rnd_m, 0x1FF, J.sg);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ F3
constexpr uint32_t F3_COLS = 256;
constexpr uint32_t F3_G = 4;  // slices whose per-column inputs are loaded ahead of any store

// One work item = F3_COLS columns of one layer, one thread per column.  Pass 1:
// pyr_cycle of every slice and the intra-column running best; pass 2, F3_G
// slices at a time: column output (inhibition / passthrough), the distal
// learning decisions (queued in LDS), then the queued dendrites' LTP by the
// whole block.  Every load of a group is issued before its stores.
// PRE: the first group's pass-2 inputs are loaded with pass 1's (one memory
// round trip fewer per item, more registers: for launches of about one
// workgroup per CU, where occupancy does not matter).  Without PRE the kernel
// asks for at least five waves per SIMD (at most 96 VGPRs): many-item launches
// are latency-bound, and the compiler's own register choice left them at four
// to six waves (73-103 VGPRs), with the bound five to seven (70-89).
template <int W, bool PRE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PRE ? 1 : 5))) void kf_pyr_post(const PyrJob* __restrict__ jobs,
                                                   const WorkItem* __restrict__ items,
                                                   uint32_t n_items, uint32_t par, uint32_t rnd_ofs) {
    __shared__ uint32_t lz[F3_COLS * F3_G * BSM_MAX_TUFTS];  // queued dendrite: cell | best << 24
    __shared__ uint8_t lt[F3_COLS * F3_G * BSM_MAX_TUFTS];   // queued dendrite: its tuft
    __shared__ uint32_t n_l;
    const uint32_t tid = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem wi = items[it];
        const PyrJob& J = jobs[wi.job];
        const uint32_t U = J.U, VU = J.V * U, C = J.C, depth = J.depth;
        const uint32_t col = (uint32_t)wi.tv * F3_COLS + tid;
        const bool valid = col < VU;
        uint32_t p = par ^ J.par_xor;
        if (J.dyn) {
            const uint32_t d = *J.dyn;
            if (!(d & 2u)) continue;
            p = d & 1u;
        }
        auto tft_st = BSM_G(const uint8_t, J.tft_states);
        auto best_raw = BSM_G(const uint8_t, J.tft_best_raw);
        auto pb_raw = BSM_G(const uint8_t, J.prev_best_raw);
        auto pb_ids = BSM_G(const uint8_t, J.prev_best_ids);
        auto soma = BSM_G(uint8_t, J.soma);
        auto flags = BSM_G(uint8_t, J.cel_flags);
        auto axn = BSM_G(uint8_t, J.axn);
        const bool prx_on = (J.en & FLAG_BASAL_PROXIMAL) != 0;
        const bool dst_on = (J.en & FLAG_BASAL_DISTAL) != 0;
        const bool out_ctl = J.control == BSM_CTRL_INTRA_COLUMN || J.control == BSM_CTRL_PYR_OUTPUT;
        const bool passthru = J.control == BSM_CTRL_PYR_OUTPUT;
        const bool learn = J.learn && J.distal_mask;
        const uint32_t dmask = J.distal_mask, n_tufts = J.n_tufts;
        const uint32_t v = col / U, u = col - v * U;

        const uint32_t pq = (p & 1) ^ 1;   // the previous cycle's states
        // pass 2's per-cell inputs of a group of F3_G slices
        auto load2 = [&](uint32_t s0, uint32_t* fl, uint32_t* ain,
                         uint32_t (*pbr)[BSM_MAX_TUFTS], uint32_t (*pbi)[BSM_MAX_TUFTS]) {
#pragma unroll
            for (uint32_t g = 0; g < F3_G; g++) {
                const uint32_t c = (s0 + g) * VU + col;
                const bool in = valid && s0 + g < depth;
                fl[g] = (in && learn) ? flags[c] : 0u;
                ain[g] = (in && !out_ctl) ? axn[J.slices[J.base_slc].idz + c] : 0u;
#pragma unroll
                for (uint32_t t = 0; t < BSM_MAX_TUFTS; t++) {
                    const bool on = in && learn && t < n_tufts && ((dmask >> t) & 1u);
                    pbr[g][t] = on ? pb_raw[t * C + c] : 0u;
                    pbi[g][t] = on ? pb_ids[t * C + c] : 0u;
                }
            }
        };
        // loads that do not depend on pass 1, issued with its loads: the first
        // group's pass-2 inputs and where each distal tuft's previous states are
        // (uniform per item: bit 0 virtual, else the state buffer; bit 2 its best
        // dendrites' fired bits kept, else the snapshot)
        uint32_t fl0[F3_G], ain0[F3_G], pbr0[F3_G][BSM_MAX_TUFTS], pbi0[F3_G][BSM_MAX_TUFTS];
        if constexpr (PRE) load2(0, fl0, ain0, pbr0, pbi0);
        uint32_t pv[BSM_MAX_TUFTS];
#pragma unroll
        for (uint32_t t = 0; t < BSM_MAX_TUFTS; t++)
            pv[t] = (learn && t < n_tufts && ((dmask >> t) & 1u))
                        ? *BSM_G(const uint32_t, J.t_vbits[t]) >> pq : 0u;

        // pass 1: pyr_cycle, running intra-column best (Q3 reset quirk)
        uint32_t cs_r[F3_G] = {0, 0, 0, 0};
        uint32_t best = 0;
        if (valid) {
            for (uint32_t s0 = 0; s0 < depth; s0 += F3_G) {
                uint32_t pp[F3_G], dd[F3_G];
#pragma unroll
                for (uint32_t g = 0; g < F3_G; g++) {
                    const uint32_t c = (s0 + g) * VU + col;
                    const bool in = s0 + g < depth;
                    pp[g] = (in && prx_on) ? tft_st[J.prx * C + c] : 0u;
                    dd[g] = (in && dst_on) ? best_raw[J.bdst * C + c] : 0u;
                }
#pragma unroll
                for (uint32_t g = 0; g < F3_G; g++) {
                    if (s0 + g >= depth) break;
                    const uint32_t act = pp[g] != 0;
                    const uint32_t cs = (pp[g] >> 2) + ((pp[g] <= 3) && act) + act * (dd[g] >> 1);
                    soma[(s0 + g) * VU + col] = (uint8_t)cs;
                    best = (cs >= best) ? cs : 0u;
                    if (s0 == 0) cs_r[g] = cs;
                }
            }
        }

        // pass 2: outputs and learning, F3_G slices at a time
        for (uint32_t s0 = 0; s0 < depth; s0 += F3_G) {
            if (tid == 0) n_l = 0;
            __syncthreads();
            if (valid) {
                uint32_t cs[F3_G], fl[F3_G], ain[F3_G], pbr[F3_G][BSM_MAX_TUFTS],
                    pbi[F3_G][BSM_MAX_TUFTS];
                if (PRE && s0 == 0) {
#pragma unroll
                    for (uint32_t g = 0; g < F3_G; g++) {
                        cs[g] = cs_r[g];
                        fl[g] = fl0[g];
                        ain[g] = ain0[g];
#pragma unroll
                        for (uint32_t t = 0; t < BSM_MAX_TUFTS; t++) {
                            pbr[g][t] = pbr0[g][t];
                            pbi[g][t] = pbi0[g][t];
                        }
                    }
                } else {
                    load2(s0, fl, ain, pbr, pbi);
#pragma unroll
                    for (uint32_t g = 0; g < F3_G; g++)
                        cs[g] = s0 == 0 ? cs_r[g]
                                        : (s0 + g < depth ? (uint32_t)soma[(s0 + g) * VU + col] : 0u);
                }
#pragma unroll
                for (uint32_t g = 0; g < F3_G; g++) {
                    const uint32_t s = s0 + g;
                    if (s >= depth) break;
                    const uint32_t c = s * VU + col;
                    uint32_t act;
                    if (out_ctl) {
                        int safe;
                        const uint32_t axi = cel_axn_idx(J.slices[J.base_slc + s], v, u, &safe);
                        const uint32_t out = passthru ? cs[g] : ((cs[g] >= best) ? cs[g] : 0u);
                        axn[axi] = (uint8_t)out;
                        act = out != 0;
                    } else {
                        act = ain[g] != 0;
                    }
                    if (learn) {
                        // tft_dst_mtp per distal tuft in tuft order; the cell flag is
                        // shared by the tufts (Q9)
                        uint32_t f = fl[g];
#pragma unroll
                        for (uint32_t t = 0; t < BSM_MAX_TUFTS; t++) {
                            if (t >= n_tufts || !((dmask >> t) & 1u)) continue;
                            const uint32_t prev = (f & CEL_PREV_ACTIVE) != 0;
                            const uint32_t spd = J.t_spd[t];
                            if (!prev && act && pbr[g][t] != 0 && pbi[g][t] * spd < J.t_spt[t]) {
                                const uint32_t q = atomicAdd(&n_l, 1u);
                                lz[q] = c | (pbi[g][t] << 24);   // cell < 2^24 (planner)
                                lt[q] = (uint8_t)t;
                            }
                            f = (f & ~(uint32_t)CEL_PREV_ACTIVE) | (act ? CEL_PREV_ACTIVE : 0u);
                        }
                        flags[c] = (uint8_t)f;
                    }
                }
            }
            __syncthreads();
            if (learn) {
                const uint32_t nl = n_l;
                const uint8_t* __restrict__ prev_st = J.st[pq];
                const uint32_t maxch = J.max_chunks;
                const int32_t r0 = J.rnd[rnd_ofs + J.slot_mtp0];
                const int32_t r1 = n_tufts > 1 ? J.rnd[rnd_ofs + J.slot_mtp0 + 1] : 0;
                const int32_t r2 = n_tufts > 2 ? J.rnd[rnd_ofs + J.slot_mtp0 + 2] : 0;
                // tft_dst_mtp on the queued dendrites: previously active -> increment
                // (0x7F), else decrement (0xFFF); ltp_u chunks per thread in flight
                // at once (a chunk past a dendrite or the queue loads the queue's
                // first chunk and is not applied)
                const uint32_t total = nl * maxch;
                const Quot q_mc = Quot::of(maxch);
                // the distal tufts' geometry, uniform per item, picked per queued
                // dendrite by its tuft (no per-lane loads of job fields)
                uint32_t f_spd[BSM_MAX_TUFTS], f_spt[BSM_MAX_TUFTS], f_idz[BSM_MAX_TUFTS],
                    f_wpd[BSM_MAX_TUFTS];
                const uint32_t* f_bm[BSM_MAX_TUFTS];
#pragma unroll
                for (uint32_t t = 0; t < BSM_MAX_TUFTS; t++) {
                    f_spd[t] = J.t_spd[t];
                    f_spt[t] = J.t_spt[t];
                    f_idz[t] = J.t_syn_idz[t];
                    f_wpd[t] = J.t_wpd[t];
                    f_bm[t] = J.t_bmask[t] + (size_t)pq * J.t_bmlen[t];
                }
                auto pick = [](const auto* a, uint32_t t) { return t == 0 ? a[0] : (t == 1 ? a[1] : a[2]); };
                static_assert(BSM_MAX_TUFTS == 3, "pick() selects among three tufts");
                constexpr int LU = ltp_u<W>();
                for (uint32_t y0 = tid; y0 < total; y0 += LU * blockDim.x) {
                    LtpChunk<W> ch[LU];
                    uint32_t zz[LU], ii[LU], tt[LU];
                    bool ok[LU];
#pragma unroll
                    for (int j = 0; j < LU; j++) {
                        const uint32_t y = y0 + j * blockDim.x;
                        const uint32_t li = y < total ? q_mc(y) : 0u;
                        const uint32_t k = y < total ? y - li * maxch : 0u;
                        const uint32_t t = lt[li], e = lz[li];
                        const uint32_t cel = e & 0xFFFFFFu, best = e >> 24;
                        const uint32_t spd = pick(f_spd, t);
                        tt[j] = t;
                        ok[j] = y < total && k * W < spd;
                        zz[j] = cel * pick(f_spt, t) + pick(f_idz, t) + best * spd;
                        ii[j] = zz[j] + (ok[j] ? k * W : 0u);
                        const uint32_t vb = pick(pv, t);
                        if (vb & 1u) {
                            // virtual previous states: the kept fired bits of the
                            // best dendrite, else the snapshot
                            if (vb & 4u)
                                ch[j].load_bits(pick(f_bm, t) + (size_t)cel * pick(f_wpd, t),
                                                ii[j] - zz[j], J.str, ii[j]);
                            else
                                ch[j].load_virt(J.vt[t], J.vt[t].snap + (size_t)pq * J.vt[t].snap_len,
                                                J.str, ii[j], J.t_syn_idz[t], cel);
                        } else {
                            ch[j].load(prev_st, J.str, ii[j]);
                        }
                    }
#pragma unroll
                    for (int j = 0; j < LU; j++)
                        if (ok[j])
                            ch[j].apply(J.str, zz[j], ii[j],
                                        tt[j] == 0 ? r0 : (tt[j] == 1 ? r1 : r2), 0xFFF, J.sg);
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace bsm

// ====================================================================== host
namespace bsm {

static constexpr uint32_t LDS_LIMIT = 64 * 1024;
// LDS per F1 workgroup that lets five share a CU's 160 KiB (the kernel's
// launch bounds allow five per CU)
static constexpr uint32_t OCC_LDS = 160 * 1024 / 5;

static uint32_t pick_w(uint64_t a, uint64_t b, uint64_t c, uint32_t max_w) {
    const uint32_t ws[5] = {32, 16, 8, 4, 1};
    for (uint32_t w : ws)
        if (w <= max_w && a % w == 0 && b % w == 0 && c % w == 0) return w;
    return 1;
}

// Chunk width of a tuft: aligned to the tuft base and the cell stride; a
// 4-byte word must not straddle two dendrites.
uint32_t tuft_chunk(const bsm_area* a, uint32_t l, uint32_t t, uint32_t max_w) {
    const TuftRt& T = a->L[l].t[t];
    const uint32_t w = pick_w(T.syn_idz, T.spt, 0, max_w);
    return (T.d.syns_per_den % 4 == 0) ? w : 1u;
}

uint32_t ssc_chunk(const bsm_area* a, uint32_t l) { return pick_w(a->L[l].t[0].spt, 0, 0, 16); }

// Up to 4 distinct source slices, in the tuft's order, sorted; with the scalar
// quirk only slices without mids (where char and int offsets agree).
uint32_t stage_slices(const bsm_area* a, const TuftRt& T, uint32_t* staged) {
    const bool quirk = (T.spt % 4) != 0;
    uint32_t n = 0;
    for (uint32_t k = 0; k < T.d.n_src_slcs && n < 4; k++) {
        const uint32_t sid = T.d.src_slc_ids[k];
        bool dup = false;
        for (uint32_t m = 0; m < n; m++) dup |= staged[m] == sid;
        if (dup) continue;
        const SliceRt& s = a->slices_h[sid];
        if (quirk && (s.v_mid || s.u_mid)) continue;
        staged[n++] = sid;
    }
    std::sort(staged, staged + n);
    return n;
}

TuftJob make_tuft_job(bsm_area* a, uint32_t l, uint32_t t, const Tuning& tn, const int32_t* rnd) {
    LayerRt& L = a->L[l];
    TuftRt& T = L.t[t];
    TuftJob j{};
    j.axn = a->axn;
    j.slices = a->slices_d;
    j.slc = L.syn_slc + T.syn_idz;
    j.vofs = L.syn_v + T.syn_idz;
    j.uofs = L.syn_u + T.syn_idz;
    j.vu = L.syn_vu + T.syn_idz;
    j.sb = L.syn_sb + T.sb_idz;
    j.str = L.syn_str + T.syn_idz;
    j.st[0] = L.syn_state[0] + T.syn_idz;
    j.st[1] = L.syn_state[1] + T.syn_idz;
    j.den_raw = L.den_raw + T.den_idz;
    j.den_st = L.den_states + T.den_idz;
    j.den_act = L.den_act + T.den_idz;
    j.ssc_energy = (L.d.kind == BSM_LAYER_SSC) ? L.cel_energies : nullptr;
    if (L.d.kind == BSM_LAYER_PYR)
        for (int k = 0; k < 8; k++) j.tft[k] = L.tft[k] + (size_t)t * L.cells;
    j.rnd = rnd;
    j.rnd_slot = L.slot0 + t;
    j.n_slices = a->n_slices;
    j.V = a->V;
    j.U = a->U;
    j.depth = L.d.depth;
    j.C = L.cells;
    j.dens = T.d.dens_per_tft;
    j.spd = T.d.syns_per_den;
    j.spt = T.spt;
    j.thresh = T.d.threshold;
    j.den_idz = (uint32_t)T.den_idz;
    j.l2 = T.d.max_active_dens_l2;
    j.scalar_quirk = (T.spt % 4) != 0;
    j.halo = tn.halo;
    j.par_xor = 0;
    const uint32_t n = stage_slices(a, T, j.staged);
    j.n_staged = n;
    j.st0 = n ? a->slices_h[j.staged[0]] : SliceRt{};
    for (uint32_t k = 0; k < n; k++) {
        const SliceRt& s = a->slices_h[j.staged[k]];
        j.st_vsize[k] = s.v_size;
        j.st_usize[k] = s.u_size;
        j.st_vsc[k] = s.v_scale;
        j.st_usc[k] = s.u_scale;
    }
    // Check-free gather: every source slice present in the tuft is staged, the
    // staged slices are consecutive ids of one geometry, and the window of the
    // largest offset stays small enough for 8-bit dot-product coefficients.
    bool fast = n > 0 && T.reach_max <= 40;
    for (uint32_t w = 0; w < 4 && fast; w++) {
        uint64_t bits = T.slc_bits[w];
        while (bits && fast) {
            const uint32_t id = w * 64 + (uint32_t)__builtin_ctzll(bits);
            bits &= bits - 1;
            bool staged = false;
            for (uint32_t k = 0; k < n; k++) staged |= j.staged[k] == id;
            fast = staged;
        }
    }
    for (uint32_t k = 1; k < n && fast; k++) {
        const SliceRt& s0 = a->slices_h[j.staged[0]];
        const SliceRt& s = a->slices_h[j.staged[k]];
        fast = j.staged[k] == j.staged[0] + k && s.v_size == s0.v_size && s.u_size == s0.u_size &&
               s.v_scale == s0.v_scale && s.u_scale == s0.u_scale && s.v_mid == s0.v_mid &&
               s.u_mid == s0.u_mid;
    }
    if (fast) {
        const uint32_t cmax = (((tn.tile_u - 1) * j.st_usc[0]) >> 4) + 2 + 2 * T.reach_max;
        fast = cmax <= 127;
    }
    j.fast = fast ? 1u : 0u;
    if (fast) j.halo = T.reach_max;
    j.vbits = L.vbits + t;
    j.sg = (const uint8_t*)L.syn_sg;
    j.sg_bit0 = (uint32_t)T.syn_idz;
    j.sg_len = (uint32_t)L.n_sg;
    // a fast tuft's states are virtual: the staged slices (consecutive ids of one
    // geometry, so contiguous in the axon space) are its snapshot
    j.snap = L.snap[t];
    j.snap_len = fast ? n * (uint32_t)j.st_vsize[0] * j.st_usize[0] : 0u;
    // Every fast tuft's states are virtual.  Spiny stellate learning needs only
    // whether each synapse of an active cell fired: tuft 0 of a spiny stellate
    // layer stores those bits instead of the states (1 bit per synapse, not 8).
    j.virt = 1u;
    j.fb = (L.d.kind == BSM_LAYER_SSC && t == 0) ? L.fbits : nullptr;
    j.bmask = L.bm_wpd[t] ? L.bmask[t] : nullptr;
    j.wpd = L.bm_wpd[t];
    j.bm_len = L.cells * L.bm_wpd[t];
    if (j.snap_len > L.snap_cap[t]) {
        j.fast = 0;
        j.snap_len = 0;
    }
    // the packed offsets are exact when every |offset| fits in 3 bits + sign
    j.packed = (fast && T.reach_v <= PACK_REACH && T.reach_u <= PACK_REACH) ? 1u : 0u;
    return j;
}

SscJob make_ssc_job(bsm_area* a, uint32_t l, const int32_t* rnd) {
    LayerRt& L = a->L[l];
    SscJob j{};
    j.soma = L.den_states;
    j.axn = a->axn;
    j.slices = a->slices_d;
    j.cel_act = L.cel_act;
    j.st[0] = L.syn_state[0];
    j.st[1] = L.syn_state[1];
    j.str = L.syn_str;
    j.rnd = rnd;
    j.slot_ctrl = L.slot0 + L.d.n_tufts;
    j.slot_mtp = L.slot0 + L.d.n_tufts + 1;
    j.V = a->V;
    j.U = a->U;
    j.depth = L.d.depth;
    j.C = L.cells;
    j.spt = L.t[0].spt;
    j.base_slc = L.d.base_axn_slc;
    j.radius = (uint32_t)L.d.inhib_radius;
    j.mode = L.d.control == BSM_CTRL_INHIB_SIMPLE ? ((a->settings & BSM_SET_BYPASS_INHIB) ? 1u : 0u)
                                                   : 2u;
    j.learn = !(a->settings & BSM_SET_DISABLE_LEARNING);
    j.par_xor = 0;
    j.sg = L.syn_sg;
    j.vbits = L.vbits;
    j.fb = L.fbits;
    return j;
}

PyrJob make_pyr_job(bsm_area* a, uint32_t l, const int32_t* rnd, uint32_t* w3) {
    LayerRt& L = a->L[l];
    PyrJob j{};
    j.tft_states = L.tft[3];
    j.tft_best_raw = L.tft[1];
    j.prev_best_raw = L.tft[5];
    j.prev_best_ids = L.tft[4];
    j.soma = L.soma;
    j.cel_flags = L.cel_flags;
    j.axn = a->axn;
    j.slices = a->slices_d;
    j.st[0] = L.syn_state[0];
    j.st[1] = L.syn_state[1];
    j.str = L.syn_str;
    j.rnd = rnd;
    j.slot_mtp0 = L.slot0 + L.d.n_tufts + 1;
    j.V = a->V;
    j.U = a->U;
    j.depth = L.d.depth;
    j.C = L.cells;
    j.control = L.d.control;
    j.n_tufts = L.d.n_tufts;
    j.base_slc = L.d.base_axn_slc;
    j.learn = !(a->settings & BSM_SET_DISABLE_LEARNING);
    uint32_t w = 16;
    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
        const TuftRt& T = L.t[t];
        const bsm_tuft_desc& d = T.d;
        if (d.den_class == BSM_DEN_BASAL && d.den_kind == BSM_DEN_PROXIMAL) {
            j.en |= FLAG_BASAL_PROXIMAL;
            j.prx = t;
        } else if (d.den_class == BSM_DEN_BASAL && d.den_kind == BSM_DEN_DISTAL) {
            j.en |= FLAG_BASAL_DISTAL;
            j.bdst = t;
        } else {
            j.en |= FLAG_APICAL_DISTAL;
            j.adst = t;
        }
        if (d.den_kind == BSM_DEN_DISTAL) {
            j.distal_mask |= 1u << t;
            w = std::min(w, pick_w(d.syns_per_den, T.spt, T.syn_idz, 16));
        }
        j.t_spd[t] = d.syns_per_den;
        j.t_spt[t] = T.spt;
        j.t_syn_idz[t] = (uint32_t)T.syn_idz;
        j.t_vbits[t] = L.vbits + t;
        j.t_bmask[t] = L.bmask[t];
        j.t_bmlen[t] = L.cells * L.bm_wpd[t];
        j.t_wpd[t] = L.bm_wpd[t];
    }
    uint32_t mc = 1;
    for (uint32_t t = 0; t < L.d.n_tufts; t++)
        if ((j.distal_mask >> t) & 1u) mc = std::max(mc, j.t_spd[t] / w);
    j.max_chunks = mc;
    j.par_xor = 0;
    j.vt = L.vt_d;
    j.sg = L.syn_sg;
    *w3 = w;
    return j;
}

static uint32_t stage_bytes(const TuftJob& j, uint32_t TV, uint32_t TU, bool fast) {
    if (fast) {
        // windows (several staged slices with packed offsets: slots of 32 rows),
        // then 12 bytes per tile row (slice, v)
        const uint32_t r = (((TV - 1) * j.st_vsc[0]) >> 4) + 2 + 2 * j.halo;
        const uint32_t c = (((TU - 1) * j.st_usc[0]) >> 4) + 2 + 2 * j.halo;
        const uint32_t slot = (j.packed && j.n_staged > 1) ? 32 * c : r * c;
        return (((j.n_staged - 1) * slot + r * c + 15) & ~15u) + 12 * j.depth * TV;
    }
    uint32_t tot = 0;
    for (uint32_t k = 0; k < j.n_staged; k++) {
        uint32_t r = std::min<uint32_t>(j.st_vsize[k], (((TV - 1) * j.st_vsc[k]) >> 4) + 2 + 2 * j.halo);
        uint32_t c = std::min<uint32_t>(j.st_usize[k], (((TU - 1) * j.st_usc[k]) >> 4) + 2 + 2 * j.halo);
        tot += r * c;
    }
    return tot;
}

// Reorders items[from, end) by decreasing weight (weight(item)), keeping the
// order of equal weights: a counting sort over the distinct weights, linear in
// the items (a batch plans ~10^5 items each time its membership changes).
template <class Weight>
static void largest_first(std::vector<WorkItem>& items, size_t from, Weight&& weight) {
    const size_t n = items.size() - from;
    if (n < 2) return;
    std::vector<uint64_t> w(n);
    for (size_t i = 0; i < n; i++) w[i] = weight(items[from + i]);
    // the distinct weights, largest first; a job's items are adjacent and
    // mostly of one weight, so only a change from the previous item adds a key
    // and looks its rank up (a serving batch re-plans items of hundreds of areas)
    std::vector<uint64_t> keys;
    for (size_t i = 0; i < n; i++)
        if (i == 0 || w[i] != w[i - 1]) keys.push_back(w[i]);
    std::sort(keys.begin(), keys.end(), std::greater<uint64_t>());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (keys.size() == 1) return;
    std::vector<uint32_t> rank(n), start(keys.size() + 1, 0);
    uint32_t r = 0;
    for (size_t i = 0; i < n; i++) {
        if (i == 0 || w[i] != w[i - 1])
            r = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), w[i],
                                            std::greater<uint64_t>()) - keys.begin());
        rank[i] = r;
        start[r + 1]++;
    }
    for (size_t k = 0; k < keys.size(); k++) start[k + 1] += start[k];
    std::vector<WorkItem> out(n);
    for (size_t i = 0; i < n; i++) out[start[rank[i]]++] = items[from + i];
    std::copy(out.begin(), out.end(), items.begin() + from);
}

int plan_f1(std::vector<TuftJob>& jobs, const std::vector<uint32_t>& job_ids,
            const std::vector<uint32_t>& job_w, const Tuning& tn, std::vector<WorkItem>& items,
            Stage& st) {
    uint32_t W = 32, nst = 0;
    bool fast = true, packed = true;
    for (uint32_t k = 0; k < job_ids.size(); k++) {
        W = std::min(W, job_w[k]);
        nst = std::max(nst, jobs[job_ids[k]].n_staged);
        fast &= jobs[job_ids[k]].fast != 0;
        packed &= jobs[job_ids[k]].packed != 0;
    }
    fast &= W >= 4;
    if (!fast) W = std::min(W, 16u);   // 32-synapse chunks: check-free kernel only
    // a distal tuft's per-chunk fired bits (check-free path; whole chunks below
    // 16 synapses): 2 B per chunk of the tile, 4 B for 32-synapse chunks, plus
    // alignment and a spare word
    auto mask_bytes = [&](const TuftJob& j, uint32_t tiles) -> uint32_t {
        return (fast && j.bmask && (W >= 16 || j.spd % W == 0))
                   ? (W == 32 ? 4u : 2u) * j.depth * tiles * (j.spt / W) + 8 : 0u;
    };
    packed &= fast;
    // fired bits are stored whole bytes per chunk, in launches of one staged
    // slice: otherwise (or in the checked kernel) a spiny stellate tuft writes
    // its states instead
    for (uint32_t id : job_ids)
        if (jobs[id].fb && (!fast || W < 8 || nst > 1)) {
            jobs[id].fb = nullptr;
            jobs[id].virt = 0;
        }
    uint32_t TV = tn.tile_v, TU = tn.tile_u;
    // several staged slices with packed offsets: window slots of 32 rows, so a
    // window (tile rows + reach) must fit in 32 rows
    const bool slots32 = fast && packed && nst > 1;
    auto rows_fit = [&](uint32_t tv) {
        if (!slots32) return true;
        for (uint32_t id : job_ids) {
            const TuftJob& j = jobs[id];
            if ((((tv - 1) * j.st_vsc[0]) >> 4) + 2 + 2 * j.halo > 32) return false;
        }
        return true;
    };
    while (TV > 1 && !rows_fit(TV)) TV = (TV + 1) / 2;
    if (!rows_fit(TV)) {
        // reach too large for 32-row slots: every job takes the unpacked reads
        for (uint32_t id : job_ids) jobs[id].packed = 0;
        packed = false;
    }
    uint32_t cap = 0, lds = 0;
    for (;;) {
        cap = 0;
        uint32_t accb = 0, denc = 0;
        for (uint32_t id : job_ids) {
            const TuftJob& j = jobs[id];
            cap = std::max(cap, stage_bytes(j, TV, TU, fast));
            accb = std::max(accb, 4 * j.depth * TV * TU * j.dens);
            if (j.tft[0]) denc = std::max(denc, 2 * j.depth * TV * TU * j.dens + mask_bytes(j, TV * TU));
        }
        cap = (cap + 15) & ~15u;
        lds = cap + accb + denc;
        // the check-free path divides chunk indices by multiply-high (exact below 2^16 chunks)
        uint32_t chunks = 0;
        for (uint32_t id : job_ids)
            chunks = std::max(chunks, jobs[id].depth * TV * TU * (jobs[id].spt / std::max(W, 1u)));
        if (lds <= LDS_LIMIT && (!fast || chunks < 65536u)) break;
        if (TV == 1 && TU == 1) return fail("fused tuft kernel does not fit in LDS");
        if (TV >= TU && TV > 1) TV = (TV + 1) / 2;
        else TU = (TU + 1) / 2;
    }
    // Every job takes the tile (TV, TU) ...
    for (uint32_t id : job_ids) {
        jobs[id].tile_v = TV;
        jobs[id].tile_u = TU;
    }
    // ... except that a check-free item of under two chunks per lane spends
    // most of its time in the per-item steps (window staging, the dendrite/tuft
    // epilogue) rather than streaming synapses: such a job's tile grows (u
    // first, since a tile row is contiguous in the synapse streams) while it
    // stays within its area, the LDS limit and the 2^16 chunk bound.  Jobs of a
    // launch may differ by 12x in synapses per column (a batch of areas), so
    // each job grows on its own.  (Config 2's spiny stellate layer at a 2x16
    // tile: 15.2 us per launch; at 2x32: 12.6 us.)
    auto job_lds = [&](const TuftJob& j, uint32_t tv, uint32_t tu) {
        return ((stage_bytes(j, tv, tu, fast) + 15) & ~15u) + 4 * j.depth * tv * tu * j.dens +
               (j.tft[0] ? 2 * j.depth * tv * tu * j.dens + mask_bytes(j, tv * tu) : 0u);
    };
    if (fast) {
        for (uint32_t id : job_ids) {
            TuftJob& j = jobs[id];
            const uint32_t per_col = j.depth * (j.spt / std::max(W, 1u));
            for (;;) {
                if (per_col * j.tile_v * j.tile_u >= 2 * 256) break;
                uint32_t ntv = j.tile_v, ntu = j.tile_u;
                if (ntu < 32 && 2 * ntu <= j.U) ntu *= 2;
                else if (ntv < 64 && 2 * ntv <= j.V && rows_fit(2 * ntv)) ntv *= 2;
                else break;
                if (ntv * ntu > 1024 || per_col * ntv * ntu >= 65536u ||
                    job_lds(j, ntv, ntu) > LDS_LIMIT)
                    break;
                j.tile_v = ntv;
                j.tile_u = ntu;
            }
        }
    }
    // Occupancy: the launch's LDS is the largest window plus the largest sums
    // and tuft bytes over its jobs, so while that exceeds a fifth of a CU's LDS
    // (five workgroups per CU, the kernel's launch bounds) the job needing the
    // most shrinks its tile (rows first), as long as it keeps two chunks per lane.
    if (fast) {
        auto launch_lds = [&]() {
            uint32_t c = 0, a = 0, d = 0;
            for (uint32_t id : job_ids) {
                const TuftJob& j = jobs[id];
                const uint32_t tiles = j.tile_v * j.tile_u;
                c = std::max(c, stage_bytes(j, j.tile_v, j.tile_u, fast));
                a = std::max(a, 4 * j.depth * tiles * j.dens);
                if (j.tft[0]) d = std::max(d, 2 * j.depth * tiles * j.dens + mask_bytes(j, tiles));
            }
            return ((c + 15) & ~15u) + a + d;
        };
        std::vector<uint8_t> fixed(jobs.size(), 0);
        while (launch_lds() > OCC_LDS) {
            uint32_t pick = UINT32_MAX, most = 0;
            for (uint32_t id : job_ids) {
                const TuftJob& j = jobs[id];
                const uint32_t need = job_lds(j, j.tile_v, j.tile_u);
                if (!fixed[id] && need > most) { most = need; pick = id; }
            }
            if (pick == UINT32_MAX) break;
            TuftJob& j = jobs[pick];
            const uint32_t per_col = j.depth * (j.spt / std::max(W, 1u));
            uint32_t ntv = j.tile_v, ntu = j.tile_u;
            if (ntv > 1) ntv = (ntv + 1) / 2;
            else if (ntu > 1) ntu = (ntu + 1) / 2;
            if ((ntv == j.tile_v && ntu == j.tile_u) || per_col * ntv * ntu < 2 * 256) {
                fixed[pick] = 1;
                continue;
            }
            j.tile_v = ntv;
            j.tile_u = ntu;
        }
    }
    // the launch's LDS: the largest window (+ row table), sums and tuft bytes
    {
        uint32_t accb = 0, denc = 0;
        cap = 0;
        for (uint32_t id : job_ids) {
            const TuftJob& j = jobs[id];
            const uint32_t tiles = j.tile_v * j.tile_u;
            cap = std::max(cap, stage_bytes(j, j.tile_v, j.tile_u, fast));
            accb = std::max(accb, 4 * j.depth * tiles * j.dens);
            if (j.tft[0]) denc = std::max(denc, 2 * j.depth * tiles * j.dens + mask_bytes(j, tiles));
        }
        cap = (cap + 15) & ~15u;
        lds = cap + accb + denc;
        if (lds > LDS_LIMIT) return fail("fused tuft kernel does not fit in LDS");
    }
    st.kind = 1;
    st.item0 = (uint32_t)items.size();
    st.alg_bytes = 0;
    // per synapse, in eighth bytes: fast (virtual states): the 2-bit staged-slice
    // code when several slices are staged (nothing for one), the offsets (1 B
    // packed, 2 B otherwise) and the strength's sign bit, no state written (a
    // spiny stellate tuft's fired bit out; a fast tuft that is not virtual, its
    // state out); checked: the 1 B slice id, 2 B offsets, the strength in and
    // the state out
    for (uint32_t id : job_ids) {
        // per dendrite raw/state out + activity in/out; per cell-tuft 4 B in + 8 B
        // out (pyramidal); per SSC cell its energy; fast: the snapshot copy
        const TuftJob& j = jobs[id];
        const uint32_t syn_e = fast ? (j.n_staged <= 1 ? 0u : 2u) + (packed ? 8u : 16u) + 1u +
                                          (j.virt ? 0u : 8u) + (j.fb ? 1u : 0u)
                                    : 8u + 16u + 8u + 8u;
        st.alg_bytes += (uint64_t)j.C * j.spt * syn_e / 8 + (uint64_t)j.C * j.dens * 4 +
                        (j.tft[0] ? (uint64_t)j.C * 12 : 0) + (j.ssc_energy ? (uint64_t)j.C : 0) +
                        (fast && j.virt ? 2 * (uint64_t)j.snap_len : 0);
    }
    for (uint32_t id : job_ids) {
        TuftJob& j = jobs[id];
        uint32_t ntv = (j.V + j.tile_v - 1) / j.tile_v, ntu = (j.U + j.tile_u - 1) / j.tile_u;
        if (ntv > 65535 || ntu > 65535 || id > 65535) return fail("too many tiles");
        if (fast && j.virt && !j.snap_len) return fail("check-free tuft without a snapshot");
        j.ntu = ntu;
        j.n_tiles = ntv * ntu;
        j.snap_per = (j.snap_len + j.n_tiles - 1) / j.n_tiles;
        for (uint32_t a = 0; a < ntv; a++)
            for (uint32_t b = 0; b < ntu; b++)
                items.push_back(WorkItem{(uint16_t)id, 0, (uint16_t)a, (uint16_t)b});
    }
    // Largest items first: workgroups are dispatched in item order, so the
    // launch ends on the smallest items instead of on a late large one (a
    // batch mixes items of up to 12x different synapse counts; one item's
    // stream is latency-bound once few workgroups are left).
    auto item_syns = [&](const WorkItem& w) {
        const TuftJob& j = jobs[w.job];
        const uint64_t tvn = std::min<uint32_t>(j.tile_v, j.V - (uint32_t)w.tv * j.tile_v);
        const uint64_t tun = std::min<uint32_t>(j.tile_u, j.U - (uint32_t)w.tu * j.tile_u);
        return (uint64_t)j.depth * tvn * tun * j.spt;
    };
    largest_first(items, st.item0, item_syns);
    st.n_items = (uint32_t)items.size() - st.item0;
    st.lds_bytes = lds;
    st.W = W;
    st.nst = nst <= 1 ? 1 : (nst <= 2 ? 2 : 4);
    st.fast = fast ? 1u : 0u;
    st.packed = packed ? 1u : 0u;
    if (fast) st.nst = nst <= 1 ? 1 : 2;
    st.stage_cap = cap;
    st.tv = TV;
    st.tu = TU;
    st.grid = std::min<uint32_t>(st.n_items, 65535);
    return 0;
}

int plan_f2(const std::vector<SscJob>& jobs, const std::vector<uint32_t>& job_ids, const Tuning& tn,
            std::vector<WorkItem>& items, Stage& st) {
    uint32_t rmax = 0, W = 16;
    for (uint32_t id : job_ids) {
        const SscJob& j = jobs[id];
        if (j.mode == 0) rmax = std::max(rmax, j.radius);
        W = std::min(W, pick_w(j.spt, 0, 0, 16));
    }
    // Inhibition tile: the tuning's, or by default 16 x 16 unless that gives more
    // tiles than about one round of resident blocks, then 32 x 32 (each item is a
    // chain of dependent memory round trips; fewer, larger items overlap them).
    auto n_tiles = [&](uint32_t t) {
        uint64_t n = 0;
        for (uint32_t id : job_ids)
            n += (uint64_t)jobs[id].depth * ((jobs[id].V + t - 1) / t) * ((jobs[id].U + t - 1) / t);
        return n;
    };
    uint32_t side = tn.inhib_tile ? tn.inhib_tile : (n_tiles(16) > 2048 ? 32u : 16u);
    const uint32_t TV = side, TU = side;
    if (TV * TU > F2_MAXC * 256) return fail("inhibition tile larger than 4 x 256 cells");
    st.kind = 2;
    st.item0 = (uint32_t)items.size();
    for (uint32_t id : job_ids) {
        const SscJob& j = jobs[id];
        uint32_t ntv = (j.V + TV - 1) / TV, ntu = (j.U + TU - 1) / TU;
        for (uint32_t s = 0; s < j.depth; s++)
            for (uint32_t a = 0; a < ntv; a++)
                for (uint32_t b = 0; b < ntu; b++)
                    items.push_back(WorkItem{(uint16_t)id, (uint16_t)s, (uint16_t)a, (uint16_t)b});
    }
    // largest items first (cells x synapses per cell, which bounds the learning)
    auto item_w = [&](const WorkItem& w) {
        const SscJob& j = jobs[w.job];
        const uint64_t tvn = std::min<uint32_t>(TV, j.V - (uint32_t)w.tv * TV);
        const uint64_t tun = std::min<uint32_t>(TU, j.U - (uint32_t)w.tu * TU);
        return tvn * tun * (j.spt + 16);
    };
    largest_first(items, st.item0, item_w);
    st.n_items = (uint32_t)items.size() - st.item0;
    uint32_t win = (TV + 2 * rmax) * (TU + 2 * rmax);
    st.lds_bytes = ((win + 15) & ~15u) + 7 * TV * TU;  // window | learn list | hex list | flags
    if (st.lds_bytes > LDS_LIMIT) return fail("inhibition window does not fit in LDS");
    st.W = W;
    st.tv = TV;
    st.tu = TU;
    st.grid = std::min<uint32_t>(st.n_items, 65535);
    return 0;
}

int plan_f3(const std::vector<PyrJob>& jobs, const std::vector<uint32_t>& job_ids,
            std::vector<WorkItem>& items, Stage& st) {
    st.kind = 3;
    st.item0 = (uint32_t)items.size();
    for (uint32_t id : job_ids) {
        const PyrJob& j = jobs[id];
        if (j.C >= (1u << 24)) return fail("pyramidal layer of 2^24 cells or more");
        uint32_t nb = (j.V * j.U + F3_COLS - 1) / F3_COLS;
        if (nb > 65535) return fail("too many column blocks");
        for (uint32_t b = 0; b < nb; b++) items.push_back(WorkItem{(uint16_t)id, 0, (uint16_t)b, 0});
    }
    // largest items first (columns x slices x distal synapses per cell)
    auto item_w = [&](const WorkItem& w) {
        const PyrJob& j = jobs[w.job];
        const uint64_t cols = std::min<uint32_t>(F3_COLS, j.V * j.U - (uint32_t)w.tv * F3_COLS);
        uint64_t dst = 16;
        for (uint32_t t = 0; t < j.n_tufts; t++)
            if (j.distal_mask >> t & 1u) dst += j.t_spt[t];
        return cols * j.depth * dst;
    };
    largest_first(items, st.item0, item_w);
    st.n_items = (uint32_t)items.size() - st.item0;
    st.lds_bytes = 0;
    st.grid = std::min<uint32_t>(st.n_items, 65535);
    return 0;
}

int ensure_rnd(bsm_area* a, size_t n_ints) {
    if (n_ints <= a->rnd_cap) return 0;
    size_t cap = std::max<size_t>(n_ints, 4096);
    BSM_HIP(hipStreamSynchronize(a->stream));
    if (a->rnd_d) BSM_HIP(hipFree(a->rnd_d));
    a->rnd_d = nullptr;
    BSM_HIP(hipMalloc((void**)&a->rnd_d, cap * sizeof(int32_t)));
    BSM_HIP(hipMemsetAsync(a->rnd_d, 0, cap * sizeof(int32_t), a->stream));
    a->rnd_cap = cap;
    a->plan_dirty = true;
    return 0;
}

int build_plan(bsm_area* a) {
    if (ensure_rnd(a, 4096)) return 1;
    a->tjobs.clear();
    a->sjobs.clear();
    a->pjobs.clear();
    a->stages.clear();
    std::vector<WorkItem> items;
    a->flip_layers.clear();
    const bool sscs = !(a->settings & BSM_SET_DISABLE_SSCS);
    const bool pyrs = !(a->settings & BSM_SET_DISABLE_PYRS);
    // slices written in a cycle before the pyramidal tufts run
    SliceSet early;
    for (uint32_t l = 0; l < a->n_layers && sscs; l++)
        if (a->L[l].d.spatial) early.add_range(a->L[l].d.base_axn_slc, a->L[l].d.depth);
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t l = 0; l < a->n_layers; l++) {
            LayerRt& L = a->L[l];
            bool spatial = L.d.spatial != 0;
            if ((pass == 0) != spatial) continue;
            if (spatial && !sscs) continue;
            if (!spatial && !pyrs) continue;
            // pyramidal tufts that read nothing written earlier in the cycle may
            // run on the side stream from the cycle's start
            std::vector<uint32_t> ids, ws, side_ids, side_ws;
            for (uint32_t t = 0; t < L.d.n_tufts; t++) {
                const bool side = !spatial && a->tune.overlap && !early.meets(L.t[t].slc_bits);
                (side ? side_ids : ids).push_back((uint32_t)a->tjobs.size());
                (side ? side_ws : ws).push_back(tuft_chunk(a, l, t, a->tune.chunk));
                a->tjobs.push_back(make_tuft_job(a, l, t, a->tune, a->rnd_d));
            }
            if (ids.empty()) {  // keep one main-stream tuft stage per layer
                ids.swap(side_ids);
                ws.swap(side_ws);
            }
            int32_t join = -1;
            if (!side_ids.empty()) {
                Stage sd{};
                if (plan_f1(a->tjobs, side_ids, side_ws, a->tune, items, sd)) return 1;
                sd.layer = l;
                sd.stream = 1;
                join = (int32_t)a->stages.size();
                a->stages.push_back(sd);
            }
            Stage s1{};
            if (plan_f1(a->tjobs, ids, ws, a->tune, items, s1)) return 1;
            s1.layer = l;
            a->stages.push_back(s1);
            a->flip_layers.push_back(l);
            if (!spatial) early.add_range(L.d.base_axn_slc, L.d.depth);
            Stage s2{};
            if (spatial) {
                std::vector<uint32_t> sid{(uint32_t)a->sjobs.size()};
                a->sjobs.push_back(make_ssc_job(a, l, a->rnd_d));
                if (plan_f2(a->sjobs, sid, a->tune, items, s2)) return 1;
            } else {
                uint32_t w3;
                std::vector<uint32_t> pid{(uint32_t)a->pjobs.size()};
                a->pjobs.push_back(make_pyr_job(a, l, a->rnd_d, &w3));
                if (plan_f3(a->pjobs, pid, items, s2)) return 1;
                s2.W = w3;
                s2.join = join;
            }
            s2.layer = l;
            a->stages.push_back(s2);
            if (L.smooth_n) {
                // the layer's activity smoother (a control layer run after its
                // inhibition); it reads activities F2 writes and writes energies
                // only the next cycle's F1 reads
                Stage s4{};
                s4.kind = 4;
                s4.layer = l;
                s4.n_items = 1;
                a->stages.push_back(s4);
            }
        }
    }
    if (side_prepare(a->side, a->stages)) return 1;
    (void)hipFree(a->tjobs_d);
    (void)hipFree(a->sjobs_d);
    (void)hipFree(a->pjobs_d);
    (void)hipFree(a->items_d);
    a->tjobs_d = nullptr;
    a->sjobs_d = nullptr;
    a->pjobs_d = nullptr;
    a->items_d = nullptr;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        BSM_HIP(hipMalloc(dst, bytes ? bytes : 16));
        if (bytes) BSM_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if (up((void**)&a->tjobs_d, a->tjobs.data(), a->tjobs.size() * sizeof(TuftJob))) return 1;
    if (up((void**)&a->sjobs_d, a->sjobs.data(), a->sjobs.size() * sizeof(SscJob))) return 1;
    if (up((void**)&a->pjobs_d, a->pjobs.data(), a->pjobs.size() * sizeof(PyrJob))) return 1;
    if (up((void**)&a->items_d, items.data(), items.size() * sizeof(WorkItem))) return 1;
    // fused intake: the first stage is the check-free F1 of a layer of one tuft
    // (job 0, alone in its launch) staging one slice with virtual states, whose
    // snapshot covers the slice (so the launch's items copy every byte of it)
    a->fuse_job = -1;
    a->fuse_lo = a->fuse_len = 0;
    if (!a->stages.empty() && !a->tjobs.empty()) {
        const Stage& s0 = a->stages[0];
        const TuftJob& j = a->tjobs[0];
        if (s0.kind == 1 && s0.fast && s0.stream == 0 && a->L[s0.layer].d.n_tufts == 1 &&
            j.fast && j.virt && j.n_staged == 1) {
            const SliceRt& sl = a->slices_h[j.staged[0]];
            const uint32_t bytes = (uint32_t)sl.v_size * sl.u_size;
            if (bytes && j.snap_len == bytes) {
                a->fuse_job = 0;
                a->fuse_lo = (uint32_t)sl.idz;
                a->fuse_len = bytes;
            }
        }
    }
    a->plan_dirty = false;
    return 0;
}

#define F1_CASE(w, n)                                                                        \
    case (w * 8 + n):                                                                        \
        kf_tuft<w, n><<<st.grid, 256, st.lds_bytes, s>>>(tj, it, st.n_items, par, rnd_ofs, \
                                                         st.stage_cap);                      \
        break;
#define F1_FAST_CASE(w, n, pk)                                                               \
    case ((w * 8 + n) * 2 + pk):                                                             \
        kf_tuft_fast<w, n, pk != 0><<<st.grid, 256, st.lds_bytes, s>>>(                     \
            tj, it, st.n_items, par, rnd_ofs, st.stage_cap, frame, intake_dst, intake_job);  \
        break;

int stage_info(const std::vector<Stage>& stages, uint32_t max_stages, bsm_stage_info* out,
               uint32_t* n_stages) {
    *n_stages = (uint32_t)stages.size();
    if (stages.size() > max_stages) return fail("max_stages too small");
    for (size_t i = 0; i < stages.size() && out; i++) {
        const Stage& st = stages[i];
        bsm_stage_info& o = out[i];
        o.kind = st.kind;
        o.layer = st.layer;
        o.chunk = st.W;
        o.staged = st.kind == 1 ? st.nst : 0;
        o.fast = st.fast;
        o.tile_v = st.tv;
        o.tile_u = st.tu;
        o.n_items = st.n_items;
        o.lds_bytes = st.lds_bytes;
        o.grid = st.grid;
        o.stream = st.stream;
        o.alg_bytes = st.alg_bytes;
        o.packed = st.packed;
    }
    return 0;
}

int launch_stage(hipStream_t s, const Stage& st, const TuftJob* tj, const SscJob* sj,
                 const PyrJob* pj, const WorkItem* items, uint32_t par, uint32_t rnd_ofs,
                 const uint8_t* frame, uint8_t* intake_dst, uint32_t intake_job) {
    if (frame && !(st.kind == 1 && st.fast)) return fail("fused intake needs a check-free F1 stage");
    if (st.n_items == 0) return 0;
    const WorkItem* it = items + st.item0;
    if (st.kind == 1 && st.fast) {
        switch ((st.W * 8 + st.nst) * 2 + (st.packed ? 1 : 0)) {
            F1_FAST_CASE(32, 1, 0) F1_FAST_CASE(32, 2, 0)
            F1_FAST_CASE(32, 1, 1) F1_FAST_CASE(32, 2, 1)
            F1_FAST_CASE(16, 1, 0) F1_FAST_CASE(16, 2, 0)
            F1_FAST_CASE(8, 1, 0) F1_FAST_CASE(8, 2, 0)
            F1_FAST_CASE(4, 1, 0) F1_FAST_CASE(4, 2, 0)
            F1_FAST_CASE(16, 1, 1) F1_FAST_CASE(16, 2, 1)
            F1_FAST_CASE(8, 1, 1) F1_FAST_CASE(8, 2, 1)
            F1_FAST_CASE(4, 1, 1) F1_FAST_CASE(4, 2, 1)
        default: return fail("bad fast F1 variant");
        }
    } else if (st.kind == 1) {
        switch (st.W * 8 + st.nst) {
            F1_CASE(16, 1) F1_CASE(16, 2) F1_CASE(16, 4)
            F1_CASE(8, 1) F1_CASE(8, 2) F1_CASE(8, 4)
            F1_CASE(4, 1) F1_CASE(4, 2) F1_CASE(4, 4)
            F1_CASE(1, 1) F1_CASE(1, 2) F1_CASE(1, 4)
        default: return fail("bad F1 variant");
        }
    } else if (st.kind == 2) {
        switch (st.W) {
        case 16: kf_ssc_post<16><<<st.grid, 256, st.lds_bytes, s>>>(sj, it, st.n_items, par, rnd_ofs, st.tv, st.tu); break;
        case 8: kf_ssc_post<8><<<st.grid, 256, st.lds_bytes, s>>>(sj, it, st.n_items, par, rnd_ofs, st.tv, st.tu); break;
        case 4: kf_ssc_post<4><<<st.grid, 256, st.lds_bytes, s>>>(sj, it, st.n_items, par, rnd_ofs, st.tv, st.tu); break;
        default: kf_ssc_post<1><<<st.grid, 256, st.lds_bytes, s>>>(sj, it, st.n_items, par, rnd_ofs, st.tv, st.tu); break;
        }
    } else {
        // preload when the launch fits one round of four workgroups per CU (the
        // preloading form's occupancy) on MI355X's 256 CUs
        const bool pre = st.n_items <= 4 * 256;
        switch (st.W) {
#define F3_CASE(w)                                                                           \
    case w:                                                                                  \
        if (pre) kf_pyr_post<w, true><<<st.grid, F3_COLS, 0, s>>>(pj, it, st.n_items, par, rnd_ofs); \
        else kf_pyr_post<w, false><<<st.grid, F3_COLS, 0, s>>>(pj, it, st.n_items, par, rnd_ofs); \
        break;
        F3_CASE(16)
        F3_CASE(8)
        F3_CASE(4)
        default: F3_CASE(1)
#undef F3_CASE
        }
    }
    BSM_HIP(hipGetLastError());
    return 0;
}

// Device-to-device frame copy into the axon space (AxonSpace::intake,
// src/cortex/axon_space.rs:509-629): one block per INTAKE_SEG bytes (16 B per
// thread for a 64 KiB frame: 16 blocks, one load and one store each).
constexpr uint32_t INTAKE_SEG = 4096;
__global__ __launch_bounds__(256) void k_intake(uint8_t* __restrict__ dst,
                                                const uint8_t* __restrict__ src, uint32_t len) {
    const uint32_t lo = blockIdx.x * INTAKE_SEG;
    if (lo < len) block_copy(dst + lo, src + lo, min(INTAKE_SEG, len - lo), threadIdx.x, blockDim.x);
}

int launch_intake(hipStream_t s, uint8_t* dst, const uint8_t* src, uint32_t len) {
    if (!len) return 0;
    k_intake<<<(len + INTAKE_SEG - 1) / INTAKE_SEG, 256, 0, s>>>(dst, src, len);
    BSM_HIP(hipGetLastError());
    return 0;
}

// One plain cycle: every cycled layer's state parity flips once, then the
// stages run on the area stream (and its side stream).
int cycle_fused(bsm_area* a, uint32_t rnd_ofs, bool serial,
                const std::function<int(hipStream_t)>* intake, bool intake_first,
                const uint8_t* fused_frame) {
    for (uint32_t l : a->flip_layers) a->L[l].cur ^= 1;
    if (intake && intake_first && (*intake)(a->stream)) return 1;
    if (enqueue_stages(a->stream, a->side, a->stages, serial,
                       [&](const Stage& st, hipStream_t s) {
                           return launch_area_stage(a, s, st, rnd_ofs,
                                                    &st == &a->stages[0] ? fused_frame : nullptr);
                       },
                       [&](hipStream_t s) {
                           return intake && !intake_first ? (*intake)(s) : 0;
                       }))
        return 1;
    // readback of the area's samplers, after the cycle on its stream
    return sample_after_cycle(a->sample_plan, a->stream, &a, 1, nullptr, nullptr);
}

// Slices a tuft's gather may read: its declared sources and every source id
// its synapses hold.
static void tuft_reads(const TuftRt& T, SliceSet& s) {
    for (int k = 0; k < 4; k++) s.bits[k] |= T.slc_bits[k];
    for (uint32_t k = 0; k < T.d.n_src_slcs && k < BSM_MAX_TUFT_SRCS; k++)
        s.add_range(T.d.src_slc_ids[k], 1);
}

// Whether a cycle's side stream may start before the intake of the axon bytes
// [off, off + len): side stages run pyramidal tufts only, so it may when no
// pyramidal tuft reads a slice the intake writes.  The side tufts then stream
// from the cycle's start while the intake copy runs on the area stream.
static bool side_before_intake(const bsm_area* a, uint32_t off, uint32_t len) {
    SliceSet in, rd;
    for (uint32_t k = 0; k < a->n_slices; k++) {
        const uint64_t lo = a->slices_h[k].idz;
        const uint64_t hi = lo + (uint64_t)a->slices_h[k].v_size * a->slices_h[k].u_size;
        if (lo < (uint64_t)off + len && off < hi) in.add_range(k, 1);
    }
    for (uint32_t l = 0; l < a->n_layers; l++)
        if (!a->L[l].d.spatial)
            for (uint32_t t = 0; t < a->L[l].d.n_tufts; t++) tuft_reads(a->L[l].t[t], rd);
    return !in.meets(rd.bits);
}

// n_cycles plain cycles, each after its frame's intake when frames is set
// (frame c at frames + c * stride, len bytes into the axon space at axn_offset).
static int run_cycles(bsm_area* a, uint32_t axn_offset, const uint8_t* frames, uint64_t stride,
                      uint32_t len, uint32_t n_cycles) {
    const bool feed = frames && len;
    const bool first = !feed || !side_before_intake(a, axn_offset, len);
    // A frame of exactly the first stage's staged slice, which no pyramidal
    // tuft reads, goes in with that launch (no copy launch of its own): the
    // launch is the first to read the slice and the only one to write it.
    const bool fuse = feed && !first && a->fuse_job >= 0 && axn_offset == a->fuse_lo &&
                      len == a->fuse_len;
    for (uint32_t c = 0; c < n_cycles; c++) {
        const std::function<int(hipStream_t)> intake = [&](hipStream_t s) {
            return launch_intake(s, a->axn + axn_offset, frames + (size_t)c * stride, len);
        };
        if (cycle_fused(a, c * a->n_slots, false, feed && !fuse ? &intake : nullptr, first,
                        fuse ? frames + (size_t)c * stride : nullptr))
            return 1;
    }
    return 0;
}

// Upload n_cycles of rnd draws (user-given or drawn from the area stream) to
// the device through a pinned double buffer.
static int upload_rnd(bsm_area* a, const int32_t* rnd, uint32_t n_cycles) {
    size_t n = (size_t)n_cycles * a->n_slots;
    if (ensure_rnd(a, n)) return 1;
    if (a->plan_dirty && build_plan(a)) return 1;
    if (!n) return 0;
    int k = a->rnd_pin_next;
    a->rnd_pin_next ^= 1;
    BSM_HIP(hipEventSynchronize(a->rnd_ev[k]));
    if (a->rnd_pin_cap[k] < n) {
        if (a->rnd_pin[k]) BSM_HIP(hipHostFree(a->rnd_pin[k]));
        a->rnd_pin[k] = nullptr;
        size_t cap = std::max<size_t>(n, 4096);
        BSM_HIP(hipHostMalloc((void**)&a->rnd_pin[k], cap * sizeof(int32_t), hipHostMallocDefault));
        a->rnd_pin_cap[k] = cap;
    }
    int32_t* h = a->rnd_pin[k];
    if (rnd) {
        std::memcpy(h, rnd, n * sizeof(int32_t));
    } else {
        for (size_t i = 0; i < n; i++) h[i] = (int32_t)rng_next(a->rng);
    }
    BSM_HIP(hipMemcpyAsync(a->rnd_d, h, n * sizeof(int32_t), hipMemcpyHostToDevice, a->stream));
    BSM_HIP(hipEventRecord(a->rnd_ev[k], a->stream));
    return 0;
}

}  // namespace bsm

using namespace bsm;

extern "C" {

int bsm_area_cycle(bsm_area* a, const int32_t* rnd, uint32_t n_cycles) {
    if (!a) return fail("null area");
    if (a->batch_stream) return fail("area belongs to a batch; cycle it through the batch");
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (upload_rnd(a, rnd, n_cycles)) return 1;
    return run_cycles(a, 0, nullptr, 0, 0, n_cycles);
}

int bsm_cortex_cycle(bsm_area* const* areas, uint32_t n_areas, const uint8_t* const* dev_frames,
                     const uint64_t* stride, const uint32_t* len, const uint32_t* axn_offset,
                     uint32_t n_frames, uint32_t n_cycles) {
    if (!areas && n_areas) return fail("null areas");
    for (uint32_t i = 0; i < n_areas; i++) {
        bsm_area* a = areas[i];
        if (!a) return fail("null area");
        if (a->batch_stream) return fail("area belongs to a batch; cycle it through the batch");
        if (dev_frames && dev_frames[i] && len && len[i] &&
            ((axn_offset ? axn_offset[i] : 0) > a->axn_len ||
             len[i] > a->axn_len - (axn_offset ? axn_offset[i] : 0)))
            return fail("frame range out of the axon space of area " + std::to_string(i));
    }
    for (uint32_t c = 0; c < n_cycles; c++) {
        for (uint32_t i = 0; i < n_areas; i++) {
            bsm_area* a = areas[i];
            BSM_HIP(hipSetDevice(a->ctx->device));
            if (dev_frames && dev_frames[i] && len && len[i] && n_frames) {
                const uint64_t st = stride ? stride[i] : len[i];
                if (launch_intake(a->stream, a->axn + (axn_offset ? axn_offset[i] : 0),
                                  dev_frames[i] + (size_t)(c % n_frames) * st, len[i]))
                    return 1;
            }
            if (bsm_area_cycle(a, nullptr, 1)) return 1;
        }
    }
    return 0;
}

int bsm_area_run(bsm_area* a, uint32_t axn_offset, const uint8_t* dev_frames,
                 uint64_t frame_stride, uint32_t frame_len, uint32_t n_cycles) {
    if (!a) return fail("null area");
    if (a->batch_stream) return fail("area belongs to a batch; cycle it through the batch");
    if (dev_frames && (axn_offset > a->axn_len || frame_len > a->axn_len - axn_offset))
        return fail("frame range out of the axon space");
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (upload_rnd(a, nullptr, n_cycles)) return 1;
    return run_cycles(a, axn_offset, dev_frames, frame_stride, frame_len, n_cycles);
}

int bsm_area_time_cycles(bsm_area* a, uint32_t axn_offset, const uint8_t* dev_frames,
                         uint64_t frame_stride, uint32_t frame_len, uint32_t n_cycles,
                         float* ms_total) {
    if (!a || !ms_total) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch; cycle it through the batch");
    BSM_HIP(hipSetDevice(a->ctx->device));
    // draw + upload outside the timed span
    if (upload_rnd(a, nullptr, n_cycles)) return 1;
    BSM_HIP(hipStreamSynchronize(a->stream));
    BSM_HIP(hipEventRecord(a->ev0, a->stream));
    if (run_cycles(a, axn_offset, dev_frames, frame_stride, frame_len, n_cycles)) return 1;
    BSM_HIP(hipEventRecord(a->ev1, a->stream));
    BSM_HIP(hipEventSynchronize(a->ev1));
    BSM_HIP(hipEventElapsedTime(ms_total, a->ev0, a->ev1));
    return 0;
}

}  // extern "C"

namespace bsm {
__global__ __launch_bounds__(256) void k_calib_copy(const uint4* __restrict__ src,
                                                    uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// The synapse streams of F1 without the gather: NR 16-byte non-temporal read
// streams (n vectors each, back to back in src) and one 16-byte write stream,
// the load/store forms F1 uses.
template <int NR>
__global__ __launch_bounds__(256) void k_calib_stream(const u32x4* __restrict__ src,
                                                      u32x4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        u32x4 x = __builtin_nontemporal_load(BSM_G(const u32x4, src + i));
#pragma unroll
        for (int k = 1; k < NR; k++)
            x ^= __builtin_nontemporal_load(BSM_G(const u32x4, src + k * n + i));
        *BSM_G(u32x4, dst + i) = x;
    }
}

// The read streams alone (the check-free gather writes no synapse bytes): the
// loads feed an xor that is stored only if it takes one fixed value, so that
// they are not removed and almost nothing is written.
template <int NR>
__global__ __launch_bounds__(256) void k_calib_read(const u32x4* __restrict__ src,
                                                    u32x4* __restrict__ sink, size_t n) {
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < NR; k++)
            acc ^= __builtin_nontemporal_load(BSM_G(const u32x4, src + k * n + i));
    }
    if (acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u && acc.z == 0xF39CC060u &&
        acc.w == 0x5CEDC834u)
        *BSM_G(u32x4, sink) = acc;
}
}  // namespace bsm

extern "C" {

int bsm_area_profile_stages(bsm_area* a, uint32_t n_cycles, uint32_t max_stages, float* stage_ms,
                            int* stage_kind, uint32_t* stage_layer, uint32_t* n_stages) {
    if (!a || !n_stages) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch");
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (a->plan_dirty && build_plan(a)) return 1;
    if (upload_rnd(a, nullptr, n_cycles)) return 1;
    const uint32_t ns = (uint32_t)a->stages.size();
    *n_stages = ns;
    if (ns > max_stages) return fail("max_stages too small");
    std::vector<hipEvent_t> ev((size_t)2 * ns * n_cycles);
    for (auto& e : ev) BSM_HIP(hipEventCreate(&e));
    BSM_HIP(hipStreamSynchronize(a->stream));
    for (uint32_t c = 0; c < n_cycles; c++) {
        // one stream, stages in list order, so that each launch is timed alone
        for (uint32_t l : a->flip_layers) a->L[l].cur ^= 1;
        for (uint32_t i = 0; i < ns; i++) {
            const Stage& st = a->stages[i];
            BSM_HIP(hipEventRecord(ev[2 * ((size_t)c * ns + i)], a->stream));
            if (launch_area_stage(a, a->stream, st, c * a->n_slots)) return 1;
            BSM_HIP(hipEventRecord(ev[2 * ((size_t)c * ns + i) + 1], a->stream));
        }
        if (sample_after_cycle(a->sample_plan, a->stream, &a, 1, nullptr, nullptr)) return 1;
    }
    BSM_HIP(hipStreamSynchronize(a->stream));
    for (uint32_t i = 0; i < ns; i++) {
        float tot = 0;
        for (uint32_t c = 0; c < n_cycles; c++) {
            float ms = 0;
            BSM_HIP(hipEventElapsedTime(&ms, ev[2 * ((size_t)c * ns + i)],
                                        ev[2 * ((size_t)c * ns + i) + 1]));
            tot += ms;
        }
        if (stage_ms) stage_ms[i] = tot;
        if (stage_kind) stage_kind[i] = a->stages[i].kind;
        if (stage_layer) stage_layer[i] = a->stages[i].layer;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return 0;
}

int bsm_area_stage_plan(bsm_area* a, uint32_t max_stages, bsm_stage_info* out,
                        uint32_t* n_stages) {
    if (!a || !n_stages) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch");
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (a->plan_dirty && build_plan(a)) return 1;
    return stage_info(a->stages, max_stages, out, n_stages);
}

int bsm_calib_stream(bsm_ctx* ctx, void* dst, const void* src, size_t bytes, uint32_t n_reads,
                     float* ms) {
    if (!ctx || !src) return fail("null arg");
    if (bytes % 16) return fail("bytes must be a multiple of 16");
    if (n_reads < 1 || n_reads > 4) return fail("n_reads must be in [1, 4]");
    BSM_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0, e1;
    BSM_HIP(hipEventCreate(&e0));
    BSM_HIP(hipEventCreate(&e1));
    const size_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 16384);
    BSM_HIP(hipEventRecord(e0, 0));
    if (!dst) {
        // read-only: the sink is the first vector of src (written only if the xor
        // of all loads takes one fixed value)
        u32x4* sink = (u32x4*)const_cast<void*>(src);
        switch (n_reads) {
        case 1: k_calib_read<1><<<grid, 256, 0, 0>>>((const u32x4*)src, sink, n); break;
        case 2: k_calib_read<2><<<grid, 256, 0, 0>>>((const u32x4*)src, sink, n); break;
        case 3: k_calib_read<3><<<grid, 256, 0, 0>>>((const u32x4*)src, sink, n); break;
        default: k_calib_read<4><<<grid, 256, 0, 0>>>((const u32x4*)src, sink, n); break;
        }
    } else switch (n_reads) {
    case 1: k_calib_stream<1><<<grid, 256, 0, 0>>>((const u32x4*)src, (u32x4*)dst, n); break;
    case 2: k_calib_stream<2><<<grid, 256, 0, 0>>>((const u32x4*)src, (u32x4*)dst, n); break;
    case 3: k_calib_stream<3><<<grid, 256, 0, 0>>>((const u32x4*)src, (u32x4*)dst, n); break;
    default: k_calib_stream<4><<<grid, 256, 0, 0>>>((const u32x4*)src, (u32x4*)dst, n); break;
    }
    BSM_HIP(hipGetLastError());
    BSM_HIP(hipEventRecord(e1, 0));
    BSM_HIP(hipEventSynchronize(e1));
    float t = 0;
    BSM_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int bsm_calib_copy(bsm_ctx* ctx, void* dst, const void* src, size_t bytes, float* ms) {
    if (!ctx || !dst || !src) return fail("null arg");
    if (bytes % 16) return fail("bytes must be a multiple of 16");
    BSM_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0, e1;
    BSM_HIP(hipEventCreate(&e0));
    BSM_HIP(hipEventCreate(&e1));
    size_t n = bytes / 16;
    uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
    BSM_HIP(hipEventRecord(e0, 0));
    k_calib_copy<<<grid, 256, 0, 0>>>((const uint4*)src, (uint4*)dst, n);
    BSM_HIP(hipGetLastError());
    BSM_HIP(hipEventRecord(e1, 0));
    BSM_HIP(hipEventSynchronize(e1));
    float t = 0;
    BSM_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // extern "C"
