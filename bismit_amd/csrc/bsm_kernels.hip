// bsm_kernels.hip -- per-kernel entry points with the reference's argument
// semantics (one launch per reference kernel, one thread per element).  These
// are the unit-parity surface of the ABI (the bsm_k_* functions); the whole
// cycle runs the fused kernels of bsm_fused.hip instead.
//
// Synapse flag bit SYN_PREV_ACTIVE is kept as "previous state buffer != 0":
// tft_set_syn_flags (cl/syns.cl:23-33) therefore becomes a copy of the current
// state buffer into the previous one, and tft_dst_mtp reads that buffer.
#include "bsm_internal.h"

namespace bsm {

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// layer_cycle_syns_wow_vec4 / _wow: cl/syns.cl:179-388 (direct per-synapse form)
__global__ void kr_syn_cycle(const uint8_t* __restrict__ axn, const SliceRt* __restrict__ tab,
                             uint32_t n_slices, const uint8_t* __restrict__ slc,
                             const int8_t* __restrict__ vo, const int8_t* __restrict__ uo,
                             uint8_t* __restrict__ st, uint64_t n, uint32_t spt, uint32_t V,
                             uint32_t U, uint32_t quirk) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = (uint32_t)(i / spt);
    uint32_t vu = c % (V * U);
    uint32_t v = vu / U, u = vu % U;
    uint32_t ax = axn_gather(axn, tab, n_slices, slc[i], (int32_t)v, vo[i], (int32_t)u, uo[i],
                             quirk != 0);
    st[i] = (uint8_t)syn_fire(ax);
}

// den_cycle_tft: cl/bismit.cl:695-747
__global__ void kr_den_cycle(const uint8_t* __restrict__ st, const int8_t* __restrict__ str,
                             uint8_t* __restrict__ den_raw, uint8_t* __restrict__ den_st,
                             uint8_t* __restrict__ den_act, uint32_t n_den, uint32_t spd,
                             uint32_t thresh, uint32_t den_idz, int32_t rnd) {
    uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_den) return;
    uint64_t z = (uint64_t)d * spd;
    int32_t sum = 0, raw = 0;
    for (uint32_t k = 0; k < spd; k++) {
        int32_t s = st[z + k];
        sum += (str[z + k] >= 0) ? s : 0;
        raw += s;
    }
    int32_t active = (uint32_t)sum > thresh;
    sum *= active;
    uint32_t dg = d + den_idz;
    int32_t seed = (int32_t)(((uint32_t)raw + thresh) | dg);
    den_act[d] = (uint8_t)activity(den_act[d], active, rnd, seed, DEN_ACT_DECAY);
    den_raw[d] = (uint8_t)min(raw >> 4, 255);
    den_st[d] = (uint8_t)min(sum >> 4, 255);
}

// tft_cycle: cl/bismit.cl:755-842
__global__ void kr_tft_cycle(const uint8_t* __restrict__ den_raw, const uint8_t* __restrict__ den_st,
                             uint8_t* bid, uint8_t* braw, uint8_t* bst, uint8_t* tst, uint8_t* pbid,
                             uint8_t* pbraw, uint8_t* pbst, uint8_t* ptst, uint32_t C, uint32_t dens,
                             uint32_t l2) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint64_t z = (uint64_t)c * dens;
    uint32_t best_id = 0, best_raw = 0, best = 0, asum = 0;
    for (uint32_t j = 0; j < dens; j++) {
        uint32_t sr = den_raw[z + j], s = den_st[z + j];
        if (sr > best_raw) { best_id = j; best_raw = sr; }
        best = max(best, s);
        asum += s;
    }
    uint32_t state = asum >= (1u << (l2 + 8)) ? 255u : (asum >> l2);
    pbid[c] = bid[c];
    pbraw[c] = braw[c];
    pbst[c] = bst[c];
    ptst[c] = tst[c];
    bid[c] = (uint8_t)best_id;
    braw[c] = (uint8_t)best_raw;
    bst[c] = (uint8_t)best;
    tst[c] = (uint8_t)state;
}

// ssc_cycle: cl/bismit.cl:849-884 (soma = dens.states[0..C))
__global__ void kr_ssc_cycle(uint8_t* __restrict__ energies, uint8_t* __restrict__ soma, uint32_t C) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t e = energies[c], s = soma[c];
    int restless = (e > 191) & (s != 0);
    int dark = e == 255;
    energies[c] = (uint8_t)((dark | restless) ? e - 64 : e);
    soma[c] = (uint8_t)min(s + 255u * restless + 255u * dark, 255u);
}

// inhib_simple (mode 0) / inhib_passthrough (mode 1): cl/control.cl:28-217
__global__ void kr_inhib(const uint8_t* __restrict__ soma, uint8_t* __restrict__ axn,
                         const SliceRt* __restrict__ tab, uint8_t* __restrict__ act, uint32_t C,
                         uint32_t V, uint32_t U, uint32_t base_slc, int32_t r, int32_t rnd,
                         int mode) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t slc = c / (V * U), vu = c % (V * U);
    int32_t v = (int32_t)(vu / U), u = (int32_t)(vu % U);
    int safe;
    uint32_t axi = cel_axn_idx(tab[slc + base_slc], (uint32_t)v, (uint32_t)u, &safe);
    uint32_t cs, is_act;
    if (mode == 0) {
        cs = safe ? soma[c] : 0u;
        int32_t imax = (r + INHIB_HORIZ_OFS) * (r + INHIB_HORIZ_OFS);
        int unin = 1;
        for (int32_t vo = -r; vo <= r; vo++) {
            int32_t uz = max(-r, -vo - r), um = min(r, -vo + r);
            for (int32_t uo = uz; uo <= um; uo++) {
                int32_t nv = v + vo, nu = u + uo;
                uint32_t ns = 0;
                if (nv >= 0 && nv < (int32_t)V && nu >= 0 && nu < (int32_t)U)
                    ns = soma[slc * V * U + (uint32_t)nv * U + (uint32_t)nu];
                int32_t w = -vo - uo;
                int32_t dist = (abs(vo) + abs(uo) + abs(w)) >> 1;
                int32_t dc = dist - INHIB_CENTER_OFS;
                unin &= ((int32_t)cs * imax) >= ((int32_t)ns * (imax - dc * dc));
            }
        }
        axn[axi] = (uint8_t)(unin * cs);
        is_act = unin & (cs != 0);
    } else {
        cs = soma[c];
        axn[axi] = (uint8_t)cs;
        is_act = cs != 0;
    }
    act[c] = (uint8_t)activity(act[c], (int32_t)is_act, rnd, (int32_t)(cs | c), CEL_ACT_DECAY);
}

// ssc_mtp_simple: cl/bismit.cl:888-915, prx_syns__active__mtp_ltd :604-639
__global__ void kr_ssc_mtp(const uint8_t* __restrict__ axn, uint32_t axn_idz,
                           const uint8_t* __restrict__ st, int8_t* __restrict__ str, uint64_t n,
                           uint32_t spt, int32_t rnd) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = (uint32_t)(i / spt);
    if (!axn[axn_idz + c]) return;
    uint32_t z = c * spt;
    int32_t seed = (int32_t)(z + (uint32_t)i);
    int32_t s = str[i];
    int32_t inc = rnd_inc(rnd, seed, s, 0x7F);
    int32_t dec = rnd_dec(rnd, seed, s, 0x1FF);
    s = st[i] ? s + inc : s - dec;
    str[i] = (int8_t)s;
}

// pyr_cycle: cl/bismit.cl:961-1022 (tft_prev_states bound to best_den_states_raw,
// src/cortex/pyramidals.rs:152-154; apical term dropped by the comma at :1021)
__global__ void kr_pyr_cycle(const uint8_t* __restrict__ tst, const uint8_t* __restrict__ braw,
                             uint8_t* __restrict__ soma, uint32_t C, uint32_t en, uint32_t prx,
                             uint32_t bdst) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t p = (en & FLAG_BASAL_PROXIMAL) ? tst[prx * C + c] : 0u;
    uint32_t d = (en & FLAG_BASAL_DISTAL) ? braw[bdst * C + c] : 0u;
    uint32_t act = p != 0;
    uint32_t pc = (p >> 2) + ((p <= 3) && act);
    uint32_t dc = act * (d >> 1);
    soma[c] = (uint8_t)(pc + dc);
}

// inhib_intra_column: cl/control.cl:319-351 (running best resets to 0)
__global__ void kr_intra_column(const uint8_t* __restrict__ soma, uint8_t* __restrict__ axn,
                                const SliceRt* __restrict__ tab, uint32_t V, uint32_t U,
                                uint32_t depth, uint32_t base_slc) {
    uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= V * U) return;
    uint32_t best = 0;
    for (uint32_t s = 0; s < depth; s++) {
        uint32_t cs = soma[s * V * U + col];
        best = (cs >= best) ? cs : 0u;
    }
    for (uint32_t s = 0; s < depth; s++) {
        uint32_t cs = soma[s * V * U + col];
        int safe;
        uint32_t axi = cel_axn_idx(tab[s + base_slc], col / U, col % U, &safe);
        axn[axi] = (uint8_t)((cs >= best) ? cs : 0u);
    }
}

// pyr_output: cl/control.cl:298-314
__global__ void kr_pyr_output(const uint8_t* __restrict__ soma, uint8_t* __restrict__ axn,
                              const SliceRt* __restrict__ tab, uint32_t C, uint32_t V, uint32_t U,
                              uint32_t base_slc) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t slc = c / (V * U), vu = c % (V * U);
    int safe;
    uint32_t axi = cel_axn_idx(tab[slc + base_slc], vu / U, vu % U, &safe);
    axn[axi] = soma[c];
}

// tft_dst_mtp: cl/bismit.cl:1195-1268, dst_syns__active__mtpot_mtdep :554-600
// (64 cell groups upstream; cells are independent, one thread per cell here).
__global__ void kr_tft_dst_mtp(const uint8_t* __restrict__ axn, uint32_t axn_idz,
                               const uint8_t* __restrict__ pbid, const uint8_t* __restrict__ pbraw,
                               const uint8_t* __restrict__ prev_st, int8_t* __restrict__ str,
                               uint8_t* __restrict__ cel_flags, uint32_t C, uint32_t tft_syn_idz,
                               uint32_t spt, uint32_t spd, int32_t rnd) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t f = cel_flags[c];
    uint32_t act = axn[axn_idz + c] != 0;
    uint32_t prev = (f & CEL_PREV_ACTIVE) == CEL_PREV_ACTIVE;
    // a best-dendrite id beyond the tuft (only possible from a user write) is
    // undefined upstream; it is skipped here
    if (!prev && act && pbraw[c] != 0 && (uint32_t)pbid[c] * spd < spt) {
        uint32_t z = c * spt + tft_syn_idz + (uint32_t)pbid[c] * spd;
        for (uint32_t i = z; i < z + spd; i++) {
            int32_t s = str[i];
            int32_t seed = (int32_t)(z + i);
            if (prev_st[i]) s += rnd_inc(rnd, seed, s, 0x7F);
            else s -= rnd_dec(rnd, seed, s, 0xFFF);
            str[i] = (int8_t)s;
        }
    }
    f = (f & ~(uint32_t)CEL_PREV_ACTIVE) | (act ? CEL_PREV_ACTIVE : 0u);
    cel_flags[c] = (uint8_t)f;
}

}  // namespace bsm

using namespace bsm;

#define GET_LAYER(a, l)                                    \
    if (!(a)) return fail("null area");                    \
    if ((l) >= (a)->n_layers) return fail("layer out of range"); \
    if ((a)->batch_stream) return fail("area belongs to a batch"); \
    LayerRt& L = (a)->L[(l)];                              \
    BSM_HIP(hipSetDevice((a)->ctx->device))

// Unit kernels read and write the state buffers directly: virtual states of
// the layer are settled into them first.
#define SETTLE(a, l) \
    if (settle_states((a), (l), (a)->stream, 0, 0, true)) return 1

#define CHECK_TUFT(t) \
    if ((t) >= L.d.n_tufts) return fail("tuft out of range")

#define LAUNCH_CHECK() BSM_HIP(hipGetLastError())

extern "C" {

int bsm_k_syn_flags(bsm_area* a, uint32_t layer) {
    GET_LAYER(a, layer);
    SETTLE(a, layer);
    BSM_HIP(hipMemcpyAsync(L.syn_state[L.cur ^ 1], L.syn_state[L.cur], L.n_syn,
                           hipMemcpyDeviceToDevice, a->stream));
    return 0;
}

int bsm_k_syn_cycle(bsm_area* a, uint32_t layer, uint32_t tuft) {
    GET_LAYER(a, layer);
    CHECK_TUFT(tuft);
    SETTLE(a, layer);
    TuftRt& T = L.t[tuft];
    uint64_t n = (uint64_t)L.cells * T.spt;
    if (!n) return 0;
    kr_syn_cycle<<<nblk(n, 256), 256, 0, a->stream>>>(
        a->axn, a->slices_d, a->n_slices, L.syn_slc + T.syn_idz, L.syn_v + T.syn_idz,
        L.syn_u + T.syn_idz, L.syn_state[L.cur] + T.syn_idz, n, T.spt, a->V, a->U,
        (T.spt % 4) != 0);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_den_cycle(bsm_area* a, uint32_t layer, uint32_t tuft, int32_t rnd) {
    GET_LAYER(a, layer);
    CHECK_TUFT(tuft);
    SETTLE(a, layer);
    TuftRt& T = L.t[tuft];
    uint32_t nd = L.cells * T.d.dens_per_tft;
    kr_den_cycle<<<nblk(nd, 256), 256, 0, a->stream>>>(
        L.syn_state[L.cur] + T.syn_idz, L.syn_str + T.syn_idz, L.den_raw + T.den_idz,
        L.den_states + T.den_idz, L.den_act + T.den_idz, nd, T.d.syns_per_den, T.d.threshold,
        (uint32_t)T.den_idz, rnd);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_tft_cycle(bsm_area* a, uint32_t layer, uint32_t tuft) {
    GET_LAYER(a, layer);
    CHECK_TUFT(tuft);
    if (L.d.kind != BSM_LAYER_PYR) return fail("tft_cycle needs a pyramidal layer");
    SETTLE(a, layer);   // it sets the best dendrite ids
    TuftRt& T = L.t[tuft];
    uint32_t o = tuft * L.cells;
    kr_tft_cycle<<<nblk(L.cells, 256), 256, 0, a->stream>>>(
        L.den_raw + T.den_idz, L.den_states + T.den_idz, L.tft[0] + o, L.tft[1] + o, L.tft[2] + o,
        L.tft[3] + o, L.tft[4] + o, L.tft[5] + o, L.tft[6] + o, L.tft[7] + o, L.cells,
        T.d.dens_per_tft, T.d.max_active_dens_l2);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_ssc_cycle(bsm_area* a, uint32_t layer) {
    GET_LAYER(a, layer);
    if (L.d.kind != BSM_LAYER_SSC) return fail("ssc_cycle needs a spiny stellate layer");
    kr_ssc_cycle<<<nblk(L.cells, 256), 256, 0, a->stream>>>(L.cel_energies, L.den_states, L.cells);
    LAUNCH_CHECK();
    return 0;
}

static int inhib_common(bsm_area* a, uint32_t layer, int32_t rnd, int mode) {
    GET_LAYER(a, layer);
    const uint8_t* soma = L.d.kind == BSM_LAYER_SSC ? L.den_states : L.soma;
    kr_inhib<<<nblk(L.cells, 256), 256, 0, a->stream>>>(soma, a->axn, a->slices_d, L.cel_act,
                                                         L.cells, a->V, a->U, L.d.base_axn_slc,
                                                         L.d.inhib_radius, rnd, mode);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_inhib_simple(bsm_area* a, uint32_t layer, int32_t rnd) {
    return inhib_common(a, layer, rnd, 0);
}

int bsm_k_inhib_passthrough(bsm_area* a, uint32_t layer, int32_t rnd) {
    return inhib_common(a, layer, rnd, 1);
}

int bsm_k_ssc_mtp(bsm_area* a, uint32_t layer, uint32_t rnd) {
    GET_LAYER(a, layer);
    if (L.d.kind != BSM_LAYER_SSC) return fail("ssc_mtp needs a spiny stellate layer");
    SETTLE(a, layer);
    TuftRt& T = L.t[0];
    uint64_t n = (uint64_t)L.cells * T.spt;
    if (!n) return 0;
    kr_ssc_mtp<<<nblk(n, 256), 256, 0, a->stream>>>(a->axn, a->slices_h[L.d.base_axn_slc].idz,
                                                     L.syn_state[L.cur], L.syn_str, n, T.spt,
                                                     (int32_t)rnd);
    LAUNCH_CHECK();
    return launch_pack_sg(a->stream, a, layer, 0, L.n_syn);
}

static void pyr_tuft_ids(const LayerRt& L, uint32_t* en, uint32_t* prx, uint32_t* bdst,
                         uint32_t* adst) {
    *en = 0;
    *prx = *bdst = *adst = 0;
    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
        const bsm_tuft_desc& d = L.t[t].d;
        if (d.den_class == BSM_DEN_BASAL && d.den_kind == BSM_DEN_PROXIMAL) {
            *en |= FLAG_BASAL_PROXIMAL;
            *prx = t;
        } else if (d.den_class == BSM_DEN_BASAL && d.den_kind == BSM_DEN_DISTAL) {
            *en |= FLAG_BASAL_DISTAL;
            *bdst = t;
        } else {
            *en |= FLAG_APICAL_DISTAL;
            *adst = t;
        }
    }
}

int bsm_k_pyr_cycle(bsm_area* a, uint32_t layer) {
    GET_LAYER(a, layer);
    if (L.d.kind != BSM_LAYER_PYR) return fail("pyr_cycle needs a pyramidal layer");
    uint32_t en, prx, bdst, adst;
    pyr_tuft_ids(L, &en, &prx, &bdst, &adst);
    kr_pyr_cycle<<<nblk(L.cells, 256), 256, 0, a->stream>>>(L.tft[3], L.tft[1], L.soma, L.cells,
                                                             en, prx, bdst);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_inhib_intra_column(bsm_area* a, uint32_t layer) {
    GET_LAYER(a, layer);
    const uint8_t* soma = L.d.kind == BSM_LAYER_SSC ? L.den_states : L.soma;
    kr_intra_column<<<nblk(a->V * a->U, 256), 256, 0, a->stream>>>(
        soma, a->axn, a->slices_d, a->V, a->U, L.d.depth, L.d.base_axn_slc);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_pyr_output(bsm_area* a, uint32_t layer) {
    GET_LAYER(a, layer);
    const uint8_t* soma = L.d.kind == BSM_LAYER_SSC ? L.den_states : L.soma;
    kr_pyr_output<<<nblk(L.cells, 256), 256, 0, a->stream>>>(soma, a->axn, a->slices_d, L.cells,
                                                              a->V, a->U, L.d.base_axn_slc);
    LAUNCH_CHECK();
    return 0;
}

int bsm_k_tft_dst_mtp(bsm_area* a, uint32_t layer, uint32_t tuft, int32_t rnd) {
    GET_LAYER(a, layer);
    CHECK_TUFT(tuft);
    if (L.d.kind != BSM_LAYER_PYR) return fail("tft_dst_mtp needs a pyramidal layer");
    SETTLE(a, layer);
    TuftRt& T = L.t[tuft];
    uint32_t o = tuft * L.cells;
    kr_tft_dst_mtp<<<nblk(L.cells, 256), 256, 0, a->stream>>>(
        a->axn, a->slices_h[L.d.base_axn_slc].idz, L.tft[4] + o, L.tft[5] + o,
        L.syn_state[L.cur ^ 1], L.syn_str, L.cel_flags, L.cells, (uint32_t)T.syn_idz, T.spt,
        T.d.syns_per_den, rnd);
    LAUNCH_CHECK();
    return launch_pack_sg(a->stream, a, layer, T.syn_idz, (size_t)L.cells * T.spt);
}

}  // extern "C"
