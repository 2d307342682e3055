// bsm_api.hip -- context, area construction, buffer I/O, rnd schedule and
// synapse layout generation for the MI355X cortical-cycle engine.
//
// Replaces the reference's per-component ocl Buffer construction and host I/O
// (src/cortex/{synapses,dendrites,tufts,pyramidals,spiny_stellates,axon_space}.rs)
// with one device arena per area, and its entropy-seeded SmallRng draws with a
// documented deterministic xorshift128 stream.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bsm_internal.h"

namespace bsm {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(const std::string& msg) {
    g_err = msg;
    return 1;
}
int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return 2;
}

// splitmix64 (Steele, Lea, Flood 2014) and xorshift128 (Marsaglia 2003).
static inline uint64_t splitmix64_next(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void rng_seed(uint32_t st[4], uint64_t seed) {
    uint64_t x = seed;
    uint64_t a = splitmix64_next(&x), b = splitmix64_next(&x);
    st[0] = (uint32_t)a;
    st[1] = (uint32_t)(a >> 32);
    st[2] = (uint32_t)b;
    st[3] = (uint32_t)(b >> 32);
    if (!(st[0] | st[1] | st[2] | st[3])) st[0] = 1;
}

uint32_t rng_next(uint32_t st[4]) {
    uint32_t t = st[0] ^ (st[0] << 11);
    st[0] = st[1];
    st[1] = st[2];
    st[2] = st[3];
    st[3] = st[3] ^ (st[3] >> 19) ^ (t ^ (t >> 8));
    return st[3];
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------- validation
static int validate(const bsm_area_desc* d) {
    if (!d) return fail("null area descriptor");
    if (d->v_size == 0 || d->u_size == 0 || d->v_size > 65535 || d->u_size > 65535)
        return fail("area v_size/u_size must be in [1, 65535]");
    if (d->n_slices == 0 || d->n_slices > BSM_MAX_SLICES || !d->slices)
        return fail("n_slices must be in [1, 255]");
    if (d->n_layers > BSM_MAX_LAYERS || (d->n_layers && !d->layers))
        return fail("n_layers must be <= 8");
    uint64_t axn = 0;
    for (uint32_t s = 0; s < d->n_slices; s++) {
        const bsm_slice_desc& sl = d->slices[s];
        if (sl.v_size == 0 || sl.u_size == 0 || sl.v_size > 65535 || sl.u_size > 65535)
            return fail("slice " + std::to_string(s) + ": sizes must be in [1, 65535]");
        axn += (uint64_t)sl.v_size * sl.u_size;
    }
    if (axn >= (1ull << 31)) return fail("axon space too large");
    for (uint32_t l = 0; l < d->n_layers; l++) {
        const bsm_layer_desc& L = d->layers[l];
        std::string p = "layer " + std::to_string(l) + ": ";
        if (L.kind != BSM_LAYER_SSC && L.kind != BSM_LAYER_PYR) return fail(p + "bad kind");
        if (L.depth == 0) return fail(p + "depth must be >= 1");
        if ((uint32_t)L.base_axn_slc + L.depth > d->n_slices) return fail(p + "slices out of range");
        for (uint32_t s = L.base_axn_slc; s < (uint32_t)L.base_axn_slc + L.depth; s++) {
            const bsm_slice_desc& sl = d->slices[s];
            if (sl.v_size != d->v_size || sl.u_size != d->u_size || sl.v_scale != 16 ||
                sl.u_scale != 16 || sl.v_mid != 0 || sl.u_mid != 0)
                return fail(p + "own axon slices must be local spatial slices (area dims, "
                                "scale 16, mid 0)");
        }
        if (L.kind == BSM_LAYER_SSC) {
            // src/cortex/spiny_stellates.rs:54-56: exactly one tuft.
            if (L.n_tufts != 1) return fail(p + "spiny stellate layers take exactly 1 tuft");
            if (!L.spatial) return fail(p + "spiny stellate layers must carry the SPATIAL role");
            if (L.control != BSM_CTRL_INHIB_SIMPLE && L.control != BSM_CTRL_NONE)
                return fail(p + "spiny stellate control must be INHIB_SIMPLE or NONE");
        } else {
            if (L.n_tufts < 1 || L.n_tufts > 3) return fail(p + "pyramidal layers take 1..3 tufts");
            if (L.spatial) return fail(p + "pyramidal layers must not carry the SPATIAL role");
            if (L.control != BSM_CTRL_INTRA_COLUMN && L.control != BSM_CTRL_PYR_OUTPUT &&
                L.control != BSM_CTRL_NONE)
                return fail(p + "pyramidal control must be INTRA_COLUMN, PYR_OUTPUT or NONE");
            // src/cortex/pyramidals.rs:88-126: at most one tuft of each kind.
            int seen[3] = {0, 0, 0};
            for (uint32_t t = 0; t < L.n_tufts; t++) {
                const bsm_tuft_desc& T = L.tufts[t];
                int k = (T.den_class == BSM_DEN_BASAL && T.den_kind == BSM_DEN_PROXIMAL) ? 0
                        : (T.den_class == BSM_DEN_BASAL && T.den_kind == BSM_DEN_DISTAL)  ? 1
                        : (T.den_class == BSM_DEN_APICAL && T.den_kind == BSM_DEN_DISTAL) ? 2
                                                                                           : -1;
                if (k < 0) return fail(p + "unsupported tuft class/kind");
                if (seen[k]++) return fail(p + "duplicate tuft kind");
            }
        }
        if (L.control == BSM_CTRL_INHIB_SIMPLE && (L.inhib_radius < 0 || L.inhib_radius > 64))
            return fail(p + "inhib_radius must be in [0, 64]");
        uint64_t cells = (uint64_t)L.depth * d->v_size * d->u_size;
        uint64_t ns = 0;
        for (uint32_t t = 0; t < L.n_tufts; t++) {
            const bsm_tuft_desc& T = L.tufts[t];
            if (T.dens_per_tft == 0 || T.dens_per_tft > 255)
                return fail(p + "dens_per_tft must be in [1, 255]");
            if (T.syns_per_den == 0 || T.syns_per_den > 256)
                return fail(p + "syns_per_den must be in [1, 256]");
            if (T.max_active_dens_l2 > 23) return fail(p + "max_active_dens_l2 too large");
            if (T.n_src_slcs > BSM_MAX_TUFT_SRCS) return fail(p + "too many tuft sources");
            for (uint32_t k = 0; k < T.n_src_slcs; k++)
                if (T.src_slc_ids[k] >= d->n_slices) return fail(p + "tuft source slice out of range");
            ns += cells * T.dens_per_tft * T.syns_per_den;
        }
        if (ns >= (1ull << 32)) return fail(p + "more than 2^32 synapses in one layer");
    }
    return 0;
}

struct ArenaPlan {
    std::vector<std::pair<void**, size_t>> parts;
    size_t total = 0;
    void add(void* pp, size_t bytes) {
        parts.push_back({(void**)pp, bytes});
        total = align_up(total, 256) + align_up(bytes ? bytes : 1, 256);
    }
};

static void plan_area(bsm_area* a, ArenaPlan& ap) {
    ap.add(&a->axn, a->axn_len);
    ap.add(&a->slices_d, sizeof(SliceRt) * a->n_slices);
    for (uint32_t l = 0; l < a->n_layers; l++) {
        LayerRt& L = a->L[l];
        size_t ns = L.n_syn, nd = L.n_den, nct = L.n_celtft, C = L.cells;
        ap.add(&L.syn_state[0], ns);
        ap.add(&L.syn_state[1], ns);
        ap.add(&L.syn_flag_static, ns);
        ap.add(&L.syn_slc, ns);
        ap.add(&L.syn_v, ns);
        ap.add(&L.syn_u, ns);
        ap.add(&L.syn_vu, ns);
        ap.add(&L.syn_sb, L.n_sb);
        ap.add(&L.syn_str, ns);
        ap.add(&L.syn_sg, L.n_sg);
        for (uint32_t t = 0; t < L.d.n_tufts; t++) ap.add(&L.snap[t], 2 * (size_t)L.snap_cap[t]);
        for (uint32_t t = 0; t < L.d.n_tufts; t++)
            ap.add(&L.bmask[t], 2 * sizeof(uint32_t) * (size_t)L.cells * L.bm_wpd[t]);
        // spiny stellate layers: tuft 0's fired bits (one bit per synapse)
        ap.add(&L.fbits, (L.d.kind == BSM_LAYER_SSC && L.d.n_tufts)
                             ? ((size_t)L.cells * L.t[0].spt + 31) / 32 * 4 : 0);
        ap.add(&L.vbits, sizeof(uint32_t) * BSM_MAX_TUFTS);
        ap.add(&L.vt_d, sizeof(VirtTuft) * BSM_MAX_TUFTS);
        ap.add(&L.den_states, nd);
        ap.add(&L.den_raw, nd);
        ap.add(&L.den_energies, nd);
        ap.add(&L.den_act, nd);
        ap.add(&L.den_thr, nd);
        for (int k = 0; k < 8; k++) ap.add(&L.tft[k], nct);
        ap.add(&L.soma, C);
        ap.add(&L.cel_flags, C);
        ap.add(&L.cel_energies, C);
        ap.add(&L.cel_act, C);
    }
}

static void fill_layers(bsm_area* a, const bsm_area_desc* d) {
    uint32_t slot = 0;
    for (uint32_t l = 0; l < d->n_layers; l++) {
        LayerRt& L = a->L[l];
        L.d = d->layers[l];
        L.cells = (uint32_t)L.d.depth * d->v_size * d->u_size;
        uint64_t ns = 0, nd = 0, nsb = 0;
        for (uint32_t t = 0; t < L.d.n_tufts; t++) {
            TuftRt& T = L.t[t];
            T.d = L.d.tufts[t];
            T.spt = T.d.dens_per_tft * T.d.syns_per_den;
            T.syn_idz = ns;
            T.den_idz = nd;
            T.sb_idz = nsb;
            nsb += ((uint64_t)L.cells * T.spt + 15) / 16 * 4;
            // zero-filled buffers: slice 0, offsets 0
            T.reach_v = T.reach_u = T.reach_max = 0;
            T.slc_bits[0] = 1;
            T.slc_bits[1] = T.slc_bits[2] = T.slc_bits[3] = 0;
            ns += (uint64_t)L.cells * T.spt;
            nd += (uint64_t)L.cells * T.d.dens_per_tft;
            // snapshot room: every distinct source slice (a fast tuft's staged
            // slices are among them)
            uint64_t cap = 0;
            for (uint32_t k = 0; k < T.d.n_src_slcs && k < BSM_MAX_TUFT_SRCS; k++) {
                bool dup = false;
                for (uint32_t m = 0; m < k; m++) dup |= T.d.src_slc_ids[m] == T.d.src_slc_ids[k];
                const bsm_slice_desc& sd = d->slices[T.d.src_slc_ids[k]];
                if (!dup) cap += (uint64_t)sd.v_size * sd.u_size;
            }
            L.snap_cap[t] = (uint32_t)cap;
            L.bm_wpd[t] = (L.d.kind == BSM_LAYER_PYR && T.d.den_kind == BSM_DEN_DISTAL)
                              ? (T.d.syns_per_den + 31) / 32 : 0u;
        }
        L.n_syn = ns;
        L.n_sg = (ns + 31) / 32 * 4;
        L.n_den = nd;
        L.n_sb = nsb;
        L.n_celtft = (uint64_t)L.cells * L.d.n_tufts;
        L.cur = 0;
        L.slot0 = slot;
        slot += 2u * L.d.n_tufts + 1u;
    }
    a->n_slots = slot;
}

// Device form of bsm_gen_tuft_layout (same hash, same fields): one thread per synapse.
struct LayoutSrc {
    uint8_t ids[BSM_MAX_TUFT_SRCS];
};
__global__ void k_layout(uint64_t key, uint64_t n, uint32_t n_src, LayoutSrc src, int32_t reach,
                         uint8_t* __restrict__ slc, int8_t* __restrict__ v, int8_t* __restrict__ u) {
    const uint32_t span = 2u * (uint32_t)reach + 1u;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t z = key + i * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        const uint64_t h = z ^ (z >> 31);
        slc[i] = src.ids[(uint32_t)(h & 0xFFFF) % n_src];
        v[i] = (int8_t)((int32_t)((uint32_t)((h >> 16) & 0xFFFF) % span) - reach);
        u[i] = (int8_t)((int32_t)((uint32_t)((h >> 32) & 0xFFFF) % span) - reach);
    }
}

// Packed offsets for the gather: one byte (v + 8) << 4 | (u + 8) per synapse,
// exact when |v|, |u| <= 7 (the planner reads them only then).
__global__ __launch_bounds__(256) void k_pack_vu(const int8_t* __restrict__ v,
                                                 const int8_t* __restrict__ u,
                                                 uint8_t* __restrict__ vu, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        vu[i] = (uint8_t)((((v[i] + 8) & 15) << 4) | ((u[i] + 8) & 15));
}

// Staged-slice codes for the gather: 32-bit word g of a tuft holds
// (id - base) & 3 of its synapses 16g..16g+15, synapse 16g + 4w + k at bits
// 8k + 2w (tuft-relative), exact when every id lies in [base, base + 4) (the
// planner reads them only then).  Words [g0, g1) of one tuft of n synapses.
__global__ __launch_bounds__(256) void k_pack_sb(const uint8_t* __restrict__ slc,
                                                 uint32_t* __restrict__ sb, uint64_t n,
                                                 uint64_t g0, uint64_t g1, uint32_t base) {
    for (uint64_t g = g0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < g1;
         g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t code = 0;
        for (uint32_t j = 0; j < 16; j++) {
            const uint64_t i = 16 * g + j;
            if (i < n) code |= ((slc[i] - base) & 3u) << (8 * (j & 3u) + 2 * (j >> 2));
        }
        sb[g] = code;
    }
}

int launch_pack_vu(hipStream_t s, bsm_area* a, uint32_t layer, size_t offset, size_t len) {
    LayerRt& L = a->L[layer];
    if (offset >= L.n_syn || !len) return 0;
    len = std::min(len, (size_t)L.n_syn - offset);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((len + 255) / 256, 8192);
    k_pack_vu<<<grid, 256, 0, s>>>(L.syn_v + offset, L.syn_u + offset, L.syn_vu + offset, len);
    BSM_HIP(hipGetLastError());
    // the codes of every tuft the range touches, in whole four-synapse groups
    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
        const TuftRt& T = L.t[t];
        const uint64_t n = (uint64_t)L.cells * T.spt;
        const uint64_t lo = std::max<uint64_t>(offset, T.syn_idz);
        const uint64_t hi = std::min<uint64_t>(offset + len, T.syn_idz + n);
        if (lo >= hi) continue;
        uint32_t staged[4];
        const uint32_t base = stage_slices(a, T, staged) ? staged[0] : 0u;
        const uint64_t g0 = (lo - T.syn_idz) / 16, g1 = (hi - T.syn_idz + 15) / 16;
        const uint32_t g = (uint32_t)std::min<uint64_t>((g1 - g0 + 255) / 256, 8192);
        k_pack_sb<<<g, 256, 0, s>>>(L.syn_slc + T.syn_idz, (uint32_t*)(L.syn_sb + T.sb_idz), n, g0,
                                    g1, base);
        BSM_HIP(hipGetLastError());
    }
    return 0;
}

// Strength signs: word w of a layer holds (strength < 0) of synapses 32w..32w+31,
// synapse 32w + 4j + k at bit 8k + j.  Words [w0, w1) of a layer of n synapses.
__global__ __launch_bounds__(256) void k_pack_sg(const int8_t* __restrict__ str,
                                                 uint32_t* __restrict__ sg, uint64_t n,
                                                 uint64_t w0, uint64_t w1) {
    for (uint64_t w = w0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < w1;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = 0;
        for (uint32_t k = 0; k < 32; k++) {
            const uint64_t i = 32 * w + k;
            if (i < n && str[i] < 0) bits |= 1u << (8 * (k & 3u) + (k >> 2));
        }
        sg[w] = bits;
    }
}

int launch_pack_sg(hipStream_t s, bsm_area* a, uint32_t layer, size_t offset, size_t len) {
    LayerRt& L = a->L[layer];
    if (offset >= L.n_syn || !len) return 0;
    len = std::min(len, (size_t)L.n_syn - offset);
    const uint64_t w0 = offset / 32, w1 = (offset + len + 31) / 32;
    const uint32_t g = (uint32_t)std::min<uint64_t>((w1 - w0 + 255) / 256, 8192);
    k_pack_sg<<<g, 256, 0, s>>>(L.syn_str, L.syn_sg, L.n_syn, w0, w1);
    BSM_HIP(hipGetLastError());
    return 0;
}

// Parity q's states of tuft synapses [lo, hi) into st[q], for each parity whose
// states are virtual (cl/syns.cl:2-4 on the snapshot's source axon).
__global__ __launch_bounds__(256) void k_materialize(const VirtTuft* __restrict__ vt,
                                                     uint64_t lo, uint64_t hi) {
    const VirtTuft T = *vt;
    const uint32_t vb = *BSM_G(const uint32_t, T.vbits) & 3u;
    if (!vb) return;
    for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t cel = (uint32_t)(i / T.spt);
        const uint32_t sid = BSM_G(const uint8_t, T.slc)[i];
        const int32_t vo = BSM_G(const int8_t, T.vofs)[i], uo = BSM_G(const int8_t, T.uofs)[i];
        for (uint32_t q = 0; q < 2; q++)
            if ((vb >> q) & 1u)
                BSM_G(uint8_t, T.st[q])[i] =
                    (uint8_t)syn_fire(virt_axon(T, T.snap + (size_t)q * T.snap_len, cel, sid, vo, uo));
    }
}

__global__ void k_vbits_clear(uint32_t* vbits, uint32_t n) {
    if (threadIdx.x < n) BSM_G(uint32_t, vbits)[threadIdx.x] = 0u;
}

int settle_states(bsm_area* a, uint32_t layer, hipStream_t s, size_t offset, size_t len,
                  bool clear) {
    LayerRt& L = a->L[layer];
    if (clear) {
        offset = 0;
        len = L.n_syn;
    }
    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
        const TuftRt& T = L.t[t];
        const uint64_t n = (uint64_t)L.cells * T.spt;
        const uint64_t lo = std::max<uint64_t>(offset, T.syn_idz);
        const uint64_t hi = std::min<uint64_t>(offset + len, T.syn_idz + n);
        if (lo >= hi) continue;
        const uint32_t g = (uint32_t)std::min<uint64_t>((hi - lo + 255) / 256, 16384);
        k_materialize<<<g, 256, 0, s>>>(L.vt_d + t, lo - T.syn_idz, hi - T.syn_idz);
        BSM_HIP(hipGetLastError());
    }
    if (clear) {
        k_vbits_clear<<<1, 64, 0, s>>>(L.vbits, BSM_MAX_TUFTS);
        BSM_HIP(hipGetLastError());
    }
    return 0;
}

int settle_area(bsm_area* a, hipStream_t s) {
    for (uint32_t l = 0; l < a->n_layers; l++)
        if (settle_states(a, l, s, 0, 0, true)) return 1;
    return 0;
}

int upload_virt(bsm_area* a, hipStream_t s) {
    for (uint32_t l = 0; l < a->n_layers; l++) {
        LayerRt& L = a->L[l];
        std::vector<VirtTuft> h(BSM_MAX_TUFTS);
        for (uint32_t t = 0; t < L.d.n_tufts; t++) {
            const TuftRt& T = L.t[t];
            VirtTuft& v = h[t];
            uint32_t staged[4];
            const uint32_t ns = stage_slices(a, T, staged);
            v.snap = L.snap[t];
            v.st[0] = L.syn_state[0] + T.syn_idz;
            v.st[1] = L.syn_state[1] + T.syn_idz;
            v.slc = L.syn_slc + T.syn_idz;
            v.vofs = L.syn_v + T.syn_idz;
            v.uofs = L.syn_u + T.syn_idz;
            v.vbits = L.vbits + t;
            v.bmask = L.bm_wpd[t] ? L.bmask[t] : nullptr;
            v.wpd = L.bm_wpd[t];
            v.bm_len = L.cells * L.bm_wpd[t];
            v.n = (uint64_t)L.cells * T.spt;
            v.spt = T.spt;
            v.VU = a->V * a->U;
            v.U = a->U;
            if (ns) {
                const SliceRt& s0 = a->slices_h[staged[0]];
                v.base = staged[0];
                v.vs = s0.v_size;
                v.us = s0.u_size;
                v.vsc = s0.v_scale;
                v.usc = s0.u_scale;
                v.vmid = s0.v_mid;
                v.umid = s0.u_mid;
                v.slice_len = (uint32_t)s0.v_size * s0.u_size;
                v.snap_len = std::min(ns * v.slice_len, L.snap_cap[t]);
            }
        }
        BSM_HIP(hipStreamSynchronize(s));
        BSM_HIP(hipMemcpy(L.vt_d, h.data(), sizeof(VirtTuft) * BSM_MAX_TUFTS,
                          hipMemcpyHostToDevice));
    }
    return 0;
}

int launch_layout(hipStream_t s, uint64_t seed, uint32_t layer, uint32_t tuft, uint64_t n,
                  uint32_t n_src, const uint8_t* src_ids, int32_t reach, uint8_t* slc, int8_t* v,
                  int8_t* u) {
    if (!n) return 0;
    LayoutSrc src{};
    for (uint32_t k = 0; k < n_src && k < BSM_MAX_TUFT_SRCS; k++) src.ids[k] = src_ids[k];
    const uint64_t key = seed ^ ((uint64_t)(layer * 8u + tuft) << 56);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
    k_layout<<<grid, 256, 0, s>>>(key, n, n_src, src, reach, slc, v, u);
    BSM_HIP(hipGetLastError());
    return 0;
}

// ---- deterministic regrowth (CorticalArea::regrow, src/cortex/cortical_area.rs:792-806;
// Synapses::regrow_syn, src/cortex/synapses.rs:411-501; SynSrcSlices::gen_src,
// src/map/syn_src.rs:448-490) with counter-based draws instead of an entropy generator.

struct RegrowSrc {
    uint8_t ids[BSM_MAX_TUFT_SRCS];
    uint16_t v_size[BSM_MAX_TUFT_SRCS], u_size[BSM_MAX_TUFT_SRCS];
    uint8_t v_scale[BSM_MAX_TUFT_SRCS], u_scale[BSM_MAX_TUFT_SRCS];
    uint32_t tile[BSM_MAX_TUFT_SRCS];  // hex tile size (0: non-spatial source)
};

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Row r of the hex tile of radius rv (gen_syn_offs, src/map/syn_src.rs:31-107): the
// u range [max(-rv, -v - rv), min(rv, -v + rv)] rescaled by u_scale / v_scale.
__host__ __device__ inline void hex_row(int32_t v, int32_t rv, int32_t vsc, int32_t usc,
                                        int32_t* uz, int32_t* un) {
    const int32_t lo = v < 0 ? -v - rv : -rv, hi = v > 0 ? rv - v : rv;
    *uz = lo * usc / vsc;
    *un = hi * usc / vsc;
}

__host__ __device__ inline uint32_t hex_tile_size(int32_t rv, int32_t vsc, int32_t usc) {
    uint32_t n = 0;
    for (int32_t v = -rv; v <= rv; v++) {
        int32_t uz, un;
        hex_row(v, rv, vsc, usc, &uz, &un);
        n += (uint32_t)(un - uz + 1);
    }
    return n;
}

// One thread per dendrite of one tuft; synapses in order; a draw is kept when no
// other synapse of the dendrite has its (slice, v, u), the 32nd draw always.
__global__ __launch_bounds__(256) void k_regrow(uint64_t key, uint64_t n_den, uint32_t spd,
                                                uint32_t n_src, RegrowSrc src, int32_t reach,
                                                uint8_t* __restrict__ slc, int8_t* __restrict__ vo,
                                                int8_t* __restrict__ uo, int8_t* __restrict__ str) {
    const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_den) return;
    const uint64_t base = d * spd;
    for (uint32_t k = 0; k < spd; k++) {
        if (str[base + k] > -25) continue;
        const uint64_t ti = base + k;
        for (uint32_t att = 0; att < 32; att++) {
            const uint64_t h = mix64(key + (ti * 32 + att + 1) * 0x9E3779B97F4A7C15ull);
            const uint32_t slot = (uint32_t)(h & 0xFFFF) % n_src;
            int32_t v, u, s;
            if (src.tile[slot] == 0) {
                const int32_t rv = src.v_size[slot] / 2, ru = src.u_size[slot] / 2;
                v = (int32_t)(((h >> 16) & 0xFFFF) % (uint64_t)(2 * rv + 1)) - rv;
                u = (int32_t)(((h >> 32) & 0xFFFF) % (uint64_t)(2 * ru + 1)) - ru;
                s = 0;
            } else {
                const int32_t vsc = src.v_scale[slot], usc = src.u_scale[slot];
                const int32_t rv = (reach * vsc) >> 4, ru = (reach * usc) >> 4;
                uint32_t q = (uint32_t)(((h >> 16) & 0xFFFFFF) % src.tile[slot]);
                v = -rv;
                u = 0;
                for (int32_t row = -rv; row <= rv; row++) {
                    int32_t uz, un;
                    hex_row(row, rv, vsc, usc, &uz, &un);
                    const uint32_t w = (uint32_t)(un - uz + 1);
                    if (q < w) {
                        v = row;
                        u = uz + (int32_t)q;
                        break;
                    }
                    q -= w;
                }
                const int32_t inten = ((rv - (v < 0 ? -v : v)) + (ru - (u < 0 ? -u : u))) >> 3;
                s = inten * ((int32_t)(((h >> 40) & 0xFFFF) % 7) - 3);
            }
            const uint8_t sl = src.ids[slot];
            bool unique = true;
            for (uint32_t j = 0; j < spd && unique; j++)
                if (j != k && slc[base + j] == sl && vo[base + j] == (int8_t)v &&
                    uo[base + j] == (int8_t)u)
                    unique = false;
            if (unique || att == 31) {
                slc[base + k] = sl;
                vo[base + k] = (int8_t)v;
                uo[base + k] = (int8_t)u;
                str[base + k] = (int8_t)s;
                break;
            }
        }
    }
}

void layout_stats_reset(bsm_area* a, uint32_t l, uint32_t t, uint32_t reach, const uint8_t* ids,
                        uint32_t n_ids) {
    TuftRt& T = a->L[l].t[t];
    T.reach_v = T.reach_u = T.reach_max = reach;
    T.slc_bits[0] = T.slc_bits[1] = T.slc_bits[2] = T.slc_bits[3] = 0;
    for (uint32_t k = 0; k < n_ids; k++) T.slc_bits[ids[k] >> 6] |= 1ull << (ids[k] & 63);
    a->plan_dirty = true;
    a->layout_epoch++;
}

// Update of the per-tuft source summary after a write (data) or a fill
// (fill_val >= 0) of a synapse source buffer range: a range covering the whole
// tuft replaces that buffer's component, a partial range widens it.
void layout_stats_update(bsm_area* a, int buf, uint32_t layer, size_t offset, size_t len,
                         const uint8_t* data, int fill_val) {
    if (buf != BSM_SYN_SRC_SLC_IDS && buf != BSM_SYN_SRC_COL_V_OFFS &&
        buf != BSM_SYN_SRC_COL_U_OFFS)
        return;
    LayerRt& L = a->L[layer];
    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
        TuftRt& T = L.t[t];
        const size_t z = T.syn_idz, n = (size_t)L.cells * T.spt;
        const size_t lo = std::max(offset, z), hi = std::min(offset + len, z + n);
        if (lo >= hi) continue;
        const bool whole = lo == z && hi == z + n;
        uint32_t& reach = buf == BSM_SYN_SRC_COL_V_OFFS ? T.reach_v : T.reach_u;
        if (whole) {
            if (buf == BSM_SYN_SRC_SLC_IDS)
                T.slc_bits[0] = T.slc_bits[1] = T.slc_bits[2] = T.slc_bits[3] = 0;
            else
                reach = 0;
        }
        // a fill contributes its one value
        const size_t end = data ? hi : lo + 1;
        for (size_t i = lo; i < end; i++) {
            const uint8_t b = data ? data[i - offset] : (uint8_t)fill_val;
            if (buf == BSM_SYN_SRC_SLC_IDS) {
                T.slc_bits[b >> 6] |= 1ull << (b & 63);
            } else {
                const int32_t v = (int8_t)b;
                reach = std::max<uint32_t>(reach, (uint32_t)(v < 0 ? -v : v));
            }
        }
        T.reach_max = std::max(T.reach_v, T.reach_u);
        a->plan_dirty = true;
        a->layout_epoch++;
    }
}

// The side stream exists only while a plan has side stages: HIP spreads new
// streams over the hardware queues in creation order, so streams created
// eagerly beside every area stream would crowd the area streams onto fewer
// queues when many areas run concurrently.
int side_prepare(Side& s, const std::vector<Stage>& stages) {
    bool any = false;
    for (const Stage& st : stages) any |= st.stream == 1;
    if (!any) {
        side_destroy(s);
        return 0;
    }
    if (!s.stream) BSM_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    if (!s.fork) BSM_HIP(hipEventCreateWithFlags(&s.fork, hipEventDisableTiming));
    while (s.join.size() < stages.size()) {
        hipEvent_t e;
        BSM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s.join.push_back(e);
    }
    return 0;
}

void side_destroy(Side& s) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    for (hipEvent_t e : s.join) (void)hipEventDestroy(e);
    s.join.clear();
    if (s.fork) (void)hipEventDestroy(s.fork);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s.fork = nullptr;
    s.stream = nullptr;
}

// A tuning update: zero fields keep the current value; inhib_tile 1 returns
// the inhibition tile to the planner's choice.
int merge_tuning(const Tuning& cur, const bsm_tuning* t, Tuning* out) {
    Tuning nt = cur;
    if (t->tile_v) nt.tile_v = t->tile_v;
    if (t->tile_u) nt.tile_u = t->tile_u;
    if (t->halo) nt.halo = t->halo;
    if (t->chunk) nt.chunk = t->chunk;
    if (t->inhib_tile) nt.inhib_tile = t->inhib_tile == 1 ? 0u : t->inhib_tile;
    if (t->overlap) nt.overlap = t->overlap == 1 ? 1u : 0u;
    if (nt.tile_v > 64 || nt.tile_u > 64 || nt.tile_v * nt.tile_u > 1024)
        return fail("tile too large");
    if (nt.halo > 64) return fail("halo too large");
    if (nt.chunk != 1 && nt.chunk != 4 && nt.chunk != 8 && nt.chunk != 16 && nt.chunk != 32)
        return fail("chunk must be 1, 4, 8, 16 or 32");
    if (nt.inhib_tile != 0 && nt.inhib_tile != 8 && nt.inhib_tile != 16 && nt.inhib_tile != 32)
        return fail("inhib_tile must be 1 (planner's choice), 8, 16 or 32");
    if (t->overlap > 2) return fail("overlap must be 0 (keep), 1 (on) or 2 (off)");
    *out = nt;
    return 0;
}

int regrow_launch(bsm_area* a, hipStream_t s, uint64_t seed, uint64_t epoch, int32_t reach,
                  bool* widened) {
    *widened = false;
    for (uint32_t l = 0; l < a->n_layers; l++) {
        LayerRt& L = a->L[l];
        // the regrown synapses' states stay those of their old sources
        if (settle_states(a, l, s, 0, 0, true)) return 1;
        for (uint32_t t = 0; t < L.d.n_tufts; t++) {
            TuftRt& T = L.t[t];
            const uint32_t n_src = std::min<uint32_t>(T.d.n_src_slcs, BSM_MAX_TUFT_SRCS);
            if (!n_src) continue;
            RegrowSrc src{};
            uint32_t new_reach = 0;
            for (uint32_t k = 0; k < n_src; k++) {
                const SliceRt& sl = a->slices_h[T.d.src_slc_ids[k]];
                src.ids[k] = T.d.src_slc_ids[k];
                src.v_size[k] = sl.v_size;
                src.u_size[k] = sl.u_size;
                src.v_scale[k] = sl.v_scale;
                src.u_scale[k] = sl.u_scale;
                const bool spatial = sl.v_scale && sl.u_scale;
                const int32_t rv = (reach * sl.v_scale) >> 4;
                src.tile[k] = spatial ? hex_tile_size(rv, sl.v_scale, sl.u_scale) : 0u;
                // largest |offset| a regrown synapse of this source can take
                int32_t mo = spatial ? std::max(rv, (rv * sl.u_scale + sl.v_scale - 1) / sl.v_scale)
                                     : std::max<int32_t>(sl.v_size / 2, sl.u_size / 2);
                new_reach = std::max(new_reach, (uint32_t)mo);
            }
            const uint64_t key = mix64(seed ^ (epoch * 0xD1B54A32D192ED03ull)) ^
                                 ((uint64_t)(l * 8 + t) << 56);
            const uint64_t n_den = (uint64_t)L.cells * T.d.dens_per_tft;
            const uint32_t grid = (uint32_t)((n_den + 255) / 256);
            if (n_den) {
                k_regrow<<<grid, 256, 0, s>>>(key, n_den, T.d.syns_per_den, n_src, src, reach,
                                              L.syn_slc + T.syn_idz, (int8_t*)L.syn_v + T.syn_idz,
                                              (int8_t*)L.syn_u + T.syn_idz,
                                              (int8_t*)L.syn_str + T.syn_idz);
                BSM_HIP(hipGetLastError());
                if (launch_pack_vu(s, a, l, T.syn_idz, (size_t)L.cells * T.spt)) return 1;
                if (launch_pack_sg(s, a, l, T.syn_idz, (size_t)L.cells * T.spt)) return 1;
            }
            // the source summary widens to the regrown sources; the fused plan
            // stays exact while the summary is unchanged
            if (new_reach > T.reach_v || new_reach > T.reach_u) *widened = true;
            T.reach_v = std::max(T.reach_v, new_reach);
            T.reach_u = std::max(T.reach_u, new_reach);
            T.reach_max = std::max(T.reach_v, T.reach_u);
            for (uint32_t k = 0; k < n_src; k++) {
                const uint64_t bit = 1ull << (src.ids[k] & 63);
                if (!(T.slc_bits[src.ids[k] >> 6] & bit)) *widened = true;
                T.slc_bits[src.ids[k] >> 6] |= bit;
            }
        }
    }
    return 0;
}

// Buffers whose writes change synapse states as the cycle left them: the
// states themselves, the flags (their previous-active bit is the previous
// state) and the sources (a virtual state follows its synapse's source); and
// the best dendrite ids, which the distal tufts' kept fired bits follow.
static bool changes_states(int buf) {
    return buf == BSM_SYN_STATES || buf == BSM_SYN_FLAG_SETS || buf == BSM_SYN_SRC_SLC_IDS ||
           buf == BSM_SYN_SRC_COL_V_OFFS || buf == BSM_SYN_SRC_COL_U_OFFS ||
           // the best-dendrite fired bits follow the best dendrite ids the cycle set
           buf == BSM_TFT_BEST_DEN_IDS;
}

// Buffer I/O on an area inside a batch first waits for the batch's enqueued
// cycles.  bsm_batch_sync also copies a dynamic batch's per-layer state
// parities (kept on the device between masked cycles) into the areas' host
// copies, so the state buffer resolved afterwards is the current one.
static int batch_settle(bsm_area* a) {
    if (!a->batch) return 0;
    return bsm_batch_sync(a->batch);
}

}  // namespace bsm

using namespace bsm;

extern "C" {

const char* bsm_last_error(void) { return g_err.c_str(); }

uint64_t bsm_splitmix64(uint64_t* state) { return splitmix64_next(state); }

int bsm_xorshift128_fill(uint64_t seed, uint32_t* out, uint64_t n) {
    if (!out && n) return fail("null output");
    uint32_t st[4];
    rng_seed(st, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = rng_next(st);
    return 0;
}

int bsm_ctx_create(int device, bsm_ctx** out) {
    if (!out) return fail("null out");
    int n = 0;
    BSM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail("no such HIP device");
    BSM_HIP(hipSetDevice(device));
    bsm_ctx* c = new bsm_ctx();
    c->device = device;
    hipError_t e = hipGetDeviceProperties(&c->prop, device);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    if (filter_scratch_setup(device)) {
        delete c;
        return 1;
    }
    *out = c;
    return 0;
}

int bsm_ctx_destroy(bsm_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    frame_pool_destroy(ctx);
    delete ctx;
    return 0;
}

int bsm_sync(bsm_ctx* ctx) {
    if (!ctx) return fail("null ctx");
    BSM_HIP(hipSetDevice(ctx->device));
    BSM_HIP(hipDeviceSynchronize());
    return 0;
}

int bsm_device_info(bsm_ctx* ctx, char* name_buf, size_t name_len, int* cu_count,
                    size_t* total_mem) {
    if (!ctx) return fail("null ctx");
    if (name_buf && name_len) {
        std::snprintf(name_buf, name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
    }
    if (cu_count) *cu_count = ctx->prop.multiProcessorCount;
    if (total_mem) *total_mem = ctx->prop.totalGlobalMem;
    return 0;
}

int bsm_area_footprint(const bsm_area_desc* desc, size_t* bytes) {
    if (validate(desc)) return 1;
    bsm_area tmp{};
    tmp.n_slices = desc->n_slices;
    tmp.n_layers = desc->n_layers;
    uint64_t tot = 0;
    for (uint32_t s = 0; s < desc->n_slices; s++)
        tot += (uint64_t)desc->slices[s].v_size * desc->slices[s].u_size;
    tmp.axn_len = (uint32_t)tot;
    fill_layers(&tmp, desc);
    ArenaPlan ap;
    plan_area(&tmp, ap);
    *bytes = ap.total;
    return 0;
}

int bsm_area_create(bsm_ctx* ctx, const bsm_area_desc* desc, uint64_t rnd_seed, bsm_area** out) {
    if (!ctx || !out) return fail("null ctx/out");
    if (validate(desc)) return 1;
    BSM_HIP(hipSetDevice(ctx->device));
    bsm_area* a = new bsm_area();
    a->ctx = ctx;
    a->V = desc->v_size;
    a->U = desc->u_size;
    a->n_slices = desc->n_slices;
    a->n_layers = desc->n_layers;
    a->settings = desc->settings;
    a->slice_desc.assign(desc->slices, desc->slices + desc->n_slices);
    uint32_t tot = 0;
    for (uint32_t s = 0; s < desc->n_slices; s++) {
        const bsm_slice_desc& sl = desc->slices[s];
        SliceRt r{};
        r.idz = tot;
        r.v_size = (uint16_t)sl.v_size;
        r.u_size = (uint16_t)sl.u_size;
        r.v_scale = sl.v_scale;
        r.u_scale = sl.u_scale;
        r.v_mid = sl.v_mid;
        r.u_mid = sl.u_mid;
        a->slices_h.push_back(r);
        tot += sl.v_size * sl.u_size;
    }
    a->axn_len = tot;
    fill_layers(a, desc);
    rng_seed(a->rng, rnd_seed);
    ArenaPlan ap;
    plan_area(a, ap);
    hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete a;
        return hip_fail(e, "hipStreamCreate");
    }
    e = hipMalloc((void**)&a->arena, ap.total);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(a->stream);
        delete a;
        return hip_fail(e, "hipMalloc(area arena)");
    }
    a->arena_bytes = ap.total;
    size_t off = 0;
    for (auto& p : ap.parts) {
        off = align_up(off, 256);
        *p.first = a->arena + off;
        off += align_up(p.second ? p.second : 1, 256);
    }
    // Buffers start zero-filled like Buffer::builder().fill_val(0); the derived
    // gather streams follow from the zero offsets and slice ids.
    BSM_HIP(hipMemsetAsync(a->arena, 0, ap.total, a->stream));
    for (uint32_t l = 0; l < a->n_layers; l++)
        if (launch_pack_vu(a->stream, a, l, 0, a->L[l].n_syn)) return 1;
    BSM_HIP(hipMemcpyAsync(a->slices_d, a->slices_h.data(), sizeof(SliceRt) * a->n_slices,
                           hipMemcpyHostToDevice, a->stream));
    BSM_HIP(hipStreamSynchronize(a->stream));
    if (upload_virt(a, a->stream)) return 1;
    BSM_HIP(hipEventCreate(&a->ev0));
    BSM_HIP(hipEventCreate(&a->ev1));
    for (int i = 0; i < 2; i++) {
        BSM_HIP(hipEventCreateWithFlags(&a->rnd_ev[i], hipEventDisableTiming));
        a->rnd_pin[i] = nullptr;
        a->rnd_pin_cap[i] = 0;
    }
    a->rnd_d = nullptr;
    a->rnd_cap = 0;
    a->rnd_pin_next = 0;
    a->tjobs_d = nullptr;
    a->sjobs_d = nullptr;
    a->pjobs_d = nullptr;
    a->items_d = nullptr;
    a->plan_dirty = true;
    a->fuse_job = -1;
    a->fuse_lo = a->fuse_len = 0;
    a->batch_stream = nullptr;
    a->batch = nullptr;
    a->layout_epoch = 0;
    a->sampler_gen = 0;
    a->sample_plan = nullptr;
    a->cycles_done = 0;
    *out = a;
    return 0;
}

int bsm_area_destroy(bsm_area* a) {
    if (!a) return 0;
    if (a->batch_stream) return fail("area belongs to a batch; destroy the batch first");
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->stream);
    samplers_detach(a);
    for (uint32_t l = 0; l < a->n_layers; l++) smoother_free(a->L[l]);
    (void)hipFree(a->filt_stage);
    (void)hipFree(a->arena);
    (void)hipFree(a->rnd_d);
    (void)hipFree(a->tjobs_d);
    (void)hipFree(a->sjobs_d);
    (void)hipFree(a->pjobs_d);
    (void)hipFree(a->items_d);
    for (int i = 0; i < 2; i++) {
        if (a->rnd_pin[i]) hipHostFree(a->rnd_pin[i]);
        (void)hipEventDestroy(a->rnd_ev[i]);
    }
    (void)hipEventDestroy(a->ev0);
    (void)hipEventDestroy(a->ev1);
    side_destroy(a->side);
    (void)hipStreamDestroy(a->stream);
    delete a;
    return 0;
}

int bsm_area_sync(bsm_area* a) {
    if (!a) return fail("null area");
    BSM_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

int bsm_area_axn_len(bsm_area* a, uint32_t* len) {
    if (!a || !len) return fail("null arg");
    *len = a->axn_len;
    return 0;
}

int bsm_area_syn_count(bsm_area* a, uint64_t* n) {
    if (!a || !n) return fail("null arg");
    uint64_t k = 0;
    for (uint32_t l = 0; l < a->n_layers; l++) k += a->L[l].n_syn;
    *n = k;
    return 0;
}

int bsm_area_slice_idz(bsm_area* a, uint32_t slc, uint32_t* idz) {
    if (!a || !idz) return fail("null arg");
    if (slc >= a->n_slices) return fail("slice out of range");
    *idz = a->slices_h[slc].idz;
    return 0;
}

// Resolve (buf, layer) to a device range.  SYN_FLAG_SETS is virtual: static
// bits in syn_flag_static, SYN_PREV_ACTIVE from the previous state buffer.
}  // extern "C"

int bsm::resolve_buf(bsm_area* a, int buf, uint32_t layer, uint8_t** ptr, size_t* len) {
    if (buf == BSM_AXN_STATES) {
        *ptr = a->axn;
        *len = a->axn_len;
        return 0;
    }
    if (layer >= a->n_layers) return fail("layer out of range");
    LayerRt& L = a->L[layer];
    bool ssc = L.d.kind == BSM_LAYER_SSC;
    size_t nct = L.n_celtft;
    switch (buf) {
    case BSM_SOMA_STATES:
        if (ssc) { *ptr = L.den_states; *len = L.n_den; }
        else { *ptr = L.soma; *len = L.cells; }
        return 0;
    case BSM_SOMA_ENERGIES: *ptr = L.cel_energies; *len = L.cells; return 0;
    case BSM_SOMA_ACTIVITIES: *ptr = L.cel_act; *len = L.cells; return 0;
    case BSM_SOMA_FLAG_SETS:
        if (ssc) return fail("spiny stellate layers have no soma flag sets");
        *ptr = L.cel_flags; *len = L.cells; return 0;
    case BSM_TFT_STATES: case BSM_TFT_BEST_DEN_IDS: case BSM_TFT_BEST_DEN_STATES_RAW:
    case BSM_TFT_BEST_DEN_STATES: case BSM_TFT_PREV_STATES: case BSM_TFT_PREV_BEST_DEN_IDS:
    case BSM_TFT_PREV_BEST_DEN_STATES_RAW: case BSM_TFT_PREV_BEST_DEN_STATES: {
        if (ssc) return fail("spiny stellate layers have no tuft buffers");
        static const int map[8] = {3, 0, 1, 2, 7, 4, 5, 6};
        *ptr = L.tft[map[buf - BSM_TFT_STATES]];
        *len = nct;
        return 0;
    }
    case BSM_DEN_STATES: *ptr = L.den_states; *len = L.n_den; return 0;
    case BSM_DEN_STATES_RAW: *ptr = L.den_raw; *len = L.n_den; return 0;
    case BSM_DEN_ENERGIES: *ptr = L.den_energies; *len = L.n_den; return 0;
    case BSM_DEN_ACTIVITIES: *ptr = L.den_act; *len = L.n_den; return 0;
    case BSM_DEN_THRESHOLDS: *ptr = L.den_thr; *len = L.n_den; return 0;
    case BSM_SYN_STATES: *ptr = L.syn_state[L.cur]; *len = L.n_syn; return 0;
    case BSM_SYN_STRENGTHS: *ptr = (uint8_t*)L.syn_str; *len = L.n_syn; return 0;
    case BSM_SYN_SRC_SLC_IDS: *ptr = L.syn_slc; *len = L.n_syn; return 0;
    case BSM_SYN_SRC_COL_V_OFFS: *ptr = (uint8_t*)L.syn_v; *len = L.n_syn; return 0;
    case BSM_SYN_SRC_COL_U_OFFS: *ptr = (uint8_t*)L.syn_u; *len = L.n_syn; return 0;
    case BSM_SYN_FLAG_SETS: *ptr = L.syn_flag_static; *len = L.n_syn; return 0;
    default: return fail("unknown buffer kind");
    }
}

extern "C" {

int bsm_area_buf_len(bsm_area* a, int buf, uint32_t layer, size_t* len) {
    if (!a || !len) return fail("null arg");
    uint8_t* p;
    return resolve_buf(a, buf, layer, &p, len);
}

int bsm_area_read(bsm_area* a, int buf, uint32_t layer, size_t offset, size_t len, void* dst) {
    if (!a || (!dst && len)) return fail("null arg");
    if (batch_settle(a)) return 1;
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, buf, layer, &p, &n)) return 1;
    if (offset > n || len > n - offset) return fail("read range out of bounds");
    BSM_HIP(hipSetDevice(a->ctx->device));
    // virtual states of the range into the state buffers (the parity bits stay)
    if ((buf == BSM_SYN_STATES || buf == BSM_SYN_FLAG_SETS) && len &&
        settle_states(a, layer, a->stream, offset, len, false))
        return 1;
    BSM_HIP(hipStreamSynchronize(a->stream));
    if (len == 0) return 0;
    BSM_HIP(hipMemcpy(dst, p + offset, len, hipMemcpyDeviceToHost));
    if (buf == BSM_SYN_FLAG_SETS) {
        LayerRt& L = a->L[layer];
        std::vector<uint8_t> prev(len);
        BSM_HIP(hipMemcpy(prev.data(), L.syn_state[L.cur ^ 1] + offset, len,
                          hipMemcpyDeviceToHost));
        uint8_t* o = (uint8_t*)dst;
        for (size_t i = 0; i < len; i++)
            o[i] = (uint8_t)((o[i] & ~SYN_PREV_ACTIVE) | (prev[i] ? SYN_PREV_ACTIVE : 0));
    }
    return 0;
}

int bsm_area_write(bsm_area* a, int buf, uint32_t layer, size_t offset, size_t len,
                   const void* src) {
    if (!a || (!src && len)) return fail("null arg");
    if (batch_settle(a)) return 1;
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, buf, layer, &p, &n)) return 1;
    if (offset > n || len > n - offset) return fail("write range out of bounds");
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (len && changes_states(buf) && settle_states(a, layer, a->stream, 0, 0, true)) return 1;
    BSM_HIP(hipStreamSynchronize(a->stream));
    if (len == 0) return 0;
    if (buf == BSM_SYN_FLAG_SETS) {
        LayerRt& L = a->L[layer];
        const uint8_t* s = (const uint8_t*)src;
        std::vector<uint8_t> stat(len), prev(len);
        for (size_t i = 0; i < len; i++) {
            stat[i] = (uint8_t)(s[i] & ~SYN_PREV_ACTIVE);
            prev[i] = (s[i] & SYN_PREV_ACTIVE) ? 0x80 : 0;
        }
        BSM_HIP(hipMemcpy(p + offset, stat.data(), len, hipMemcpyHostToDevice));
        BSM_HIP(hipMemcpy(L.syn_state[L.cur ^ 1] + offset, prev.data(), len,
                          hipMemcpyHostToDevice));
        return 0;
    }
    BSM_HIP(hipMemcpy(p + offset, src, len, hipMemcpyHostToDevice));
    layout_stats_update(a, buf, layer, offset, len, (const uint8_t*)src, -1);
    if (buf == BSM_SYN_SRC_COL_V_OFFS || buf == BSM_SYN_SRC_COL_U_OFFS ||
        buf == BSM_SYN_SRC_SLC_IDS) {
        if (launch_pack_vu(a->stream, a, layer, offset, len)) return 1;
        BSM_HIP(hipStreamSynchronize(a->stream));
    }
    if (buf == BSM_SYN_STRENGTHS) {
        if (launch_pack_sg(a->stream, a, layer, offset, len)) return 1;
        BSM_HIP(hipStreamSynchronize(a->stream));
    }
    return 0;
}

int bsm_area_fill(bsm_area* a, int buf, uint32_t layer, uint8_t val) {
    if (!a) return fail("null area");
    if (batch_settle(a)) return 1;
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, buf, layer, &p, &n)) return 1;
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (changes_states(buf) && settle_states(a, layer, a->stream, 0, 0, true)) return 1;
    if (buf == BSM_SYN_FLAG_SETS) {
        LayerRt& L = a->L[layer];
        BSM_HIP(hipMemsetAsync(p, val & ~SYN_PREV_ACTIVE, n, a->stream));
        BSM_HIP(hipMemsetAsync(L.syn_state[L.cur ^ 1], (val & SYN_PREV_ACTIVE) ? 0x80 : 0, n,
                               a->stream));
    } else {
        BSM_HIP(hipMemsetAsync(p, val, n, a->stream));
        // a whole-buffer fill replaces the summary of every tuft of the layer
        if (buf != BSM_AXN_STATES) layout_stats_update(a, buf, layer, 0, n, nullptr, val);
        if (buf == BSM_SYN_SRC_COL_V_OFFS || buf == BSM_SYN_SRC_COL_U_OFFS ||
            buf == BSM_SYN_SRC_SLC_IDS)
            if (launch_pack_vu(a->stream, a, layer, 0, n)) return 1;
        if (buf == BSM_SYN_STRENGTHS && launch_pack_sg(a->stream, a, layer, 0, n)) return 1;
    }
    BSM_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

int bsm_area_intake(bsm_area* a, uint32_t axn_offset, const uint8_t* src, uint32_t len,
                    int src_is_device) {
    if (!a || (!src && len)) return fail("null arg");
    if (axn_offset > a->axn_len || len > a->axn_len - axn_offset)
        return fail("intake range out of the axon space");
    if (!len) return 0;
    BSM_HIP(hipSetDevice(a->ctx->device));
    if (src_is_device) {
        if (launch_intake(a->stream, a->axn + axn_offset, src, len)) return 1;
    } else {
        BSM_HIP(hipStreamSynchronize(a->stream));
        BSM_HIP(hipMemcpy(a->axn + axn_offset, src, len, hipMemcpyHostToDevice));
    }
    return 0;
}

int bsm_area_output(bsm_area* a, uint32_t axn_offset, uint8_t* dst, uint32_t len) {
    return bsm_area_read(a, BSM_AXN_STATES, 0, axn_offset, len, dst);
}

int bsm_area_rnd_slots(bsm_area* a, uint32_t* n) {
    if (!a || !n) return fail("null arg");
    *n = a->n_slots;
    return 0;
}

int bsm_area_rnd_slot_info(bsm_area* a, uint32_t slot, int* kind, uint32_t* layer,
                           uint32_t* tuft) {
    if (!a) return fail("null area");
    for (uint32_t l = 0; l < a->n_layers; l++) {
        const LayerRt& L = a->L[l];
        uint32_t T = L.d.n_tufts;
        if (slot >= L.slot0 && slot < L.slot0 + 2 * T + 1) {
            uint32_t k = slot - L.slot0;
            if (layer) *layer = l;
            if (k < T) { if (kind) *kind = BSM_RND_DEN; if (tuft) *tuft = k; }
            else if (k == T) { if (kind) *kind = BSM_RND_CTRL; if (tuft) *tuft = 0; }
            else { if (kind) *kind = BSM_RND_MTP; if (tuft) *tuft = k - T - 1; }
            return 0;
        }
    }
    return fail("slot out of range");
}

int bsm_area_draw_rnd(bsm_area* a, int32_t* out, uint32_t n_cycles) {
    if (!a || (!out && n_cycles)) return fail("null arg");
    uint64_t n = (uint64_t)n_cycles * a->n_slots;
    for (uint64_t i = 0; i < n; i++) out[i] = (int32_t)rng_next(a->rng);
    return 0;
}

int bsm_area_regrow(bsm_area* a, uint64_t seed, uint64_t epoch, int32_t reach) {
    if (!a) return fail("null area");
    if (a->batch_stream) return fail("area belongs to a batch (use bsm_batch_regrow)");
    if (reach < 0 || reach > 127) return fail("reach must be in [0, 127]");
    BSM_HIP(hipSetDevice(a->ctx->device));
    bool widened = false;
    if (regrow_launch(a, a->stream, seed, epoch, reach, &widened)) return 1;
    if (widened) {
        a->plan_dirty = true;
        a->layout_epoch++;
    }
    BSM_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

int bsm_area_set_tuning(bsm_area* a, const bsm_tuning* t) {
    if (!a || !t) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch");
    Tuning nt;
    if (merge_tuning(a->tune, t, &nt)) return 1;
    a->tune = nt;
    a->plan_dirty = true;
    return 0;
}

int bsm_area_get_tuning(bsm_area* a, bsm_tuning* t) {
    if (!a || !t) return fail("null arg");
    t->tile_v = a->tune.tile_v;
    t->tile_u = a->tune.tile_u;
    t->halo = a->tune.halo;
    t->chunk = a->tune.chunk;
    t->inhib_tile = a->tune.inhib_tile ? a->tune.inhib_tile : 1u;
    t->overlap = a->tune.overlap ? 1u : 2u;
    return 0;
}

int bsm_dev_alloc(bsm_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return fail("null arg");
    BSM_HIP(hipSetDevice(ctx->device));
    BSM_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    return 0;
}

int bsm_dev_free(bsm_ctx* ctx, void* dptr) {
    if (!ctx) return fail("null ctx");
    BSM_HIP(hipSetDevice(ctx->device));
    BSM_HIP(hipFree(dptr));
    return 0;
}

int bsm_dev_upload(bsm_ctx* ctx, void* dptr, const void* src, size_t bytes) {
    if (!ctx) return fail("null ctx");
    BSM_HIP(hipSetDevice(ctx->device));
    BSM_HIP(hipMemcpy(dptr, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

// Deterministic layout: h = mix64(key + (i+1)*golden) per synapse, with the
// splitmix64 finalizer; key = seed ^ ((layer*8 + tuft) << 56).
int bsm_gen_tuft_layout(uint64_t seed, uint32_t layer, uint32_t tuft, uint64_t n_syns,
                        uint32_t n_src, const uint8_t* src_ids, int32_t reach, uint8_t* out_slc,
                        int8_t* out_v, int8_t* out_u) {
    if (n_src == 0 || !src_ids) return fail("tuft has no source slices");
    if (reach < 0 || reach > 127) return fail("reach must be in [0, 127]");
    if (n_syns && (!out_slc || !out_v || !out_u)) return fail("null output");
    const uint64_t key = seed ^ ((uint64_t)(layer * 8u + tuft) << 56);
    const uint32_t span = 2u * (uint32_t)reach + 1u;
    for (uint64_t i = 0; i < n_syns; i++) {
        uint64_t x = key + i * 0x9E3779B97F4A7C15ull;
        uint64_t h = splitmix64_next(&x);
        out_slc[i] = src_ids[(uint32_t)(h & 0xFFFF) % n_src];
        out_v[i] = (int8_t)((int32_t)((uint32_t)((h >> 16) & 0xFFFF) % span) - reach);
        out_u[i] = (int8_t)((int32_t)((uint32_t)((h >> 32) & 0xFFFF) % span) - reach);
    }
    return 0;
}

int bsm_area_init_layout(bsm_area* a, uint64_t seed, int32_t reach) {
    if (!a) return fail("null area");
    if (reach < 0 || reach > 127) return fail("reach must be in [0, 127]");
    BSM_HIP(hipSetDevice(a->ctx->device));
    for (uint32_t l = 0; l < a->n_layers; l++) {
        LayerRt& L = a->L[l];
        if (settle_states(a, l, a->stream, 0, 0, true)) return 1;
        for (uint32_t t = 0; t < L.d.n_tufts; t++) {
            TuftRt& T = L.t[t];
            if (T.d.n_src_slcs == 0) return fail("tuft has no source slices");
            uint64_t n = (uint64_t)L.cells * T.spt;
            if (launch_layout(a->stream, seed, l, t, n, T.d.n_src_slcs, T.d.src_slc_ids, reach,
                              L.syn_slc + T.syn_idz, L.syn_v + T.syn_idz, L.syn_u + T.syn_idz))
                return 1;
            layout_stats_reset(a, l, t, (uint32_t)reach, T.d.src_slc_ids, T.d.n_src_slcs);
        }
        if (launch_pack_vu(a->stream, a, l, 0, L.n_syn)) return 1;
        BSM_HIP(hipMemsetAsync(L.syn_str, 0, L.n_syn, a->stream));
        if (launch_pack_sg(a->stream, a, l, 0, L.n_syn)) return 1;
    }
    BSM_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

int bsm_area_copy_axons(bsm_area* dst, uint32_t dst_offset, bsm_area* src, uint32_t src_offset,
                        uint32_t len) {
    if (!dst || !src) return fail("null area");
    if (dst->ctx != src->ctx) return fail("areas on different contexts");
    if (dst->batch_stream || src->batch_stream) return fail("area belongs to a batch");
    if (src_offset > src->axn_len || len > src->axn_len - src_offset || dst_offset > dst->axn_len ||
        len > dst->axn_len - dst_offset)
        return fail("copy range out of the axon space");
    if (!len) return 0;
    BSM_HIP(hipSetDevice(dst->ctx->device));
    if (src != dst) BSM_HIP(hipStreamSynchronize(src->stream));
    BSM_HIP(hipMemcpyAsync(dst->axn + dst_offset, src->axn + src_offset, len,
                           hipMemcpyDeviceToDevice, dst->stream));
    // complete on return, so later work on src cannot overwrite the range first
    BSM_HIP(hipStreamSynchronize(dst->stream));
    return 0;
}

}  // extern "C"
