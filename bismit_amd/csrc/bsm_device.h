// bsm_device.h -- device-side integer helpers shared by every gfx950 kernel.
//
// All arithmetic follows OpenCL C integer semantics of the reference kernels:
// int32 wrap on left shifts, arithmetic right shift of signed ints, u8/i8
// truncating stores.  Nothing on the path is floating point.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsm {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// constants: src/cmn/mod.rs:323-336, cl/bismit.cl:16,22, cl/control.cl:6-13
constexpr uint8_t CEL_PREV_ACTIVE = 0x08;
constexpr uint8_t SYN_PREV_ACTIVE = 0x10;
constexpr uint8_t FLAG_BASAL_PROXIMAL = 0x01;
constexpr uint8_t FLAG_BASAL_DISTAL = 0x02;
constexpr uint8_t FLAG_APICAL_DISTAL = 0x04;
constexpr int DEN_ACT_DECAY = 1536;
constexpr int CEL_ACT_DECAY = 768;
constexpr int INHIB_CENTER_OFS = 1;
constexpr int INHIB_HORIZ_OFS = 3;

// Runtime slice table entry (the reference bakes these in as -D lists,
// cl/bismit.cl:108-114).  16 bytes, one per axon slice.
struct SliceRt {
    uint32_t idz;
    uint16_t v_size, u_size;
    uint8_t v_scale, u_scale, v_mid, u_mid;
    uint32_t _pad;
};

// Global-address-space views of device pointers.  Job tables carry plain
// pointers, which the compiler would otherwise access with flat instructions
// (these also count against the LDS wait counter and serialise LDS waits
// behind outstanding HBM loads).
#define BSM_G(T, p) ((__attribute__((address_space(1))) T*)(p))

// xorshift mix: cl/bismit.cl:426-431
__device__ __forceinline__ int32_t rnd_mix(int32_t rnd, int32_t seed) {
    uint32_t s = (uint32_t)seed;
    s ^= (s ^ (uint32_t)rnd) << 13;
    s ^= (uint32_t)(((int32_t)s) >> 17);
    s ^= s << 5;
    return (int32_t)s;
}

__device__ __forceinline__ int32_t abs_i8(int32_t v) { return v < 0 ? -v : v; }

// rnd_inc / rnd_dec: cl/bismit.cl:447-469 (val is a sign-extended i8)
__device__ __forceinline__ int32_t rnd_inc(int32_t rnd, int32_t seed, int32_t val, int32_t mask) {
    return (rnd_mix(rnd, seed) & mask) > (abs_i8(val) + (mask - 0x7f));
}
__device__ __forceinline__ int32_t rnd_dec(int32_t rnd, int32_t seed, int32_t val, int32_t mask) {
    int32_t is_max = val == 127;
    return (rnd_mix(rnd, seed) & mask) > ((abs_i8(val) - is_max) + (mask - 0x7f));
}

// update_activity_rating: cl/bismit.cl:531-540
__device__ __forceinline__ uint32_t activity(uint32_t a, int32_t is_active, int32_t rnd,
                                             int32_t seed, int32_t decay) {
    a = (a + (((rnd_mix(rnd, seed) & 0xFF) > (int32_t)a) & is_active)) & 0xFF;
    int32_t seed2 = (int32_t)((uint32_t)seed << 1);
    a = (a - (((rnd_mix(rnd, seed2) & 0xFFFF) < decay) & (a > 0))) & 0xFF;
    return a;
}

// syn_fire: cl/syns.cl:2-4
__device__ __forceinline__ uint32_t syn_fire(uint32_t ax) {
    return ax ? 128u + (ax >> 1) : 0u;
}

// Generic gather (slow path): axn_state_3d_safe_vec4, cl/bismit.cl:366-418.
// scalar_quirk selects the char-wrapped mid of axn_idx_3d_unsafe
// (cl/bismit.cl:331-332), used when syns_per_tft % 4 != 0.
__device__ __forceinline__ uint32_t axn_gather(const uint8_t* __restrict__ axn,
                                               const SliceRt* __restrict__ tab,
                                               uint32_t n_slices, uint32_t slc, int32_t v,
                                               int32_t vo, int32_t u, int32_t uo,
                                               bool scalar_quirk) {
    if (slc >= n_slices) return 0;
    const SliceRt s = tab[slc];
    int32_t vm = s.v_mid, um = s.u_mid;
    int32_t vof, uof;
    if (scalar_quirk) {
        vof = (int32_t)(int8_t)(uint8_t)(vo + vm);
        uof = (int32_t)(int8_t)(uint8_t)(uo + um);
    } else {
        vof = vo + vm;
        uof = uo + um;
    }
    int32_t vv = (int32_t)(((uint32_t)v * s.v_scale) >> 4) + vof;
    int32_t uu = (int32_t)(((uint32_t)u * s.u_scale) >> 4) + uof;
    if (vv < 0 || vv >= (int32_t)s.v_size || uu < 0 || uu >= (int32_t)s.u_size) return 0;
    return BSM_G(const uint8_t, axn)[s.idz + (uint32_t)(vv * (int32_t)s.u_size + uu)];
}

// The cell's own axon index (scalar axn_idx_3d_unsafe with zero offsets,
// as used by cl/control.cl:45-46,209,310-311,344-345).
__device__ __forceinline__ uint32_t cel_axn_idx(const SliceRt& s, uint32_t v, uint32_t u,
                                                int* safe) {
    int32_t vo = (int32_t)(int8_t)s.v_mid, uo = (int32_t)(int8_t)s.u_mid;
    int32_t vv = (int32_t)((v * s.v_scale) >> 4) + vo;
    int32_t uu = (int32_t)((u * s.u_scale) >> 4) + uo;
    *safe = (vv >= 0) & (vv < (int32_t)s.v_size) & (uu >= 0) & (uu < (int32_t)s.u_size);
    return s.idz + (uint32_t)(vv * (int32_t)s.u_size + uu);
}

// Block-cooperative copy of len bytes: 16-byte vectors when both ends are
// 16-byte aligned, four loads in flight per thread ahead of their stores.
__device__ inline void block_copy(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                                  uint32_t len, uint32_t tid, uint32_t nth) {
    auto d8 = BSM_G(uint8_t, dst);
    auto s8 = BSM_G(const uint8_t, src);
    uint32_t done = 0;
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0) {
        const uint32_t nv = len >> 4;
        auto d = BSM_G(u32x4, dst);
        auto s = BSM_G(const u32x4, src);
        uint32_t i = tid;
        for (; i + 3 * nth < nv; i += 4 * nth) {
            const u32x4 a = s[i], b = s[i + nth], c = s[i + 2 * nth], e = s[i + 3 * nth];
            d[i] = a;
            d[i + nth] = b;
            d[i + 2 * nth] = c;
            d[i + 3 * nth] = e;
        }
        for (; i < nv; i += nth) d[i] = s[i];
        done = nv << 4;
    }
    uint32_t i = done + tid;
    for (; i + 3 * nth < len; i += 4 * nth) {
        const uint8_t a = s8[i], b = s8[i + nth], c = s8[i + 2 * nth], e = s8[i + 3 * nth];
        d8[i] = a;
        d8[i + nth] = b;
        d8[i + 2 * nth] = c;
        d8[i + 3 * nth] = e;
    }
    for (; i < len; i += nth) d8[i] = s8[i];
}

// Block-cooperative zero fill of len bytes (16-byte stores over the aligned middle).
__device__ inline void block_zero(uint8_t* __restrict__ dst, uint32_t len, uint32_t tid,
                                  uint32_t nth) {
    auto d8 = BSM_G(uint8_t, dst);
    const uint32_t head = (uint32_t)((16u - (reinterpret_cast<uintptr_t>(dst) & 15u)) & 15u);
    const uint32_t h = head < len ? head : len;
    for (uint32_t i = tid; i < h; i += nth) d8[i] = 0;
    const uint32_t nv = (len - h) >> 4;
    auto d = BSM_G(u32x4, dst + h);
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (uint32_t i = tid; i < nv; i += nth) d[i] = z;
    for (uint32_t i = h + (nv << 4) + tid; i < len; i += nth) d8[i] = 0;
}

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int b) { return (w >> (8 * b)) & 0xFFu; }
__device__ __forceinline__ int32_t sbyte_of(uint32_t w, int b) {
    return (int32_t)(int8_t)(uint8_t)((w >> (8 * b)) & 0xFFu);
}

}  // namespace bsm
