/*
 * bsm_oracle.c -- CPU restatement of bismit's per-cycle cortical pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; nothing in
 * bismit_amd/ links or calls it.
 *
 * Each function restates the reference kernel/helper named in its comment,
 * element by element, in plain scalar C.  OpenCL integer semantics are kept:
 * int32 two's-complement wrap on left shifts, arithmetic right shift of signed
 * ints (gcc implements >> on negative ints as arithmetic), uchar/char stores
 * truncate.  orc_set_threads(n > 1) parallelises the element loops with OpenMP
 * (every kernel is element-parallel with no cross-element dependence), which is
 * the CPU baseline; parity runs use one thread.
 *
 * Parity status of each function and the checks that pin it are listed in the
 * tests (tests/test_oracle_pins.py); there is no reference golden vector for
 * this path, and the reference's OpenCL/Rust build cannot run in this image.
 */
#include "bsm_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants: src/cmn/mod.rs:323-336, cl/bismit.cl:16,22, cl/control.cl:6-13 */
#define CEL_PREV_ACTIVE_FLAG 0x08
#define SYN_PREV_ACTIVE_FLAG 0x10
#define DEN_BASAL_PROXIMAL_FLAG 0x01
#define DEN_BASAL_DISTAL_FLAG 0x02
#define DEN_APICAL_DISTAL_FLAG 0x04
#define SYNAPSE_AXON_BIAS_LOG2 1
#define DENDRITE_ACTIVITY_DECAY_FACTOR 1536
#define CELL_ACTIVITY_DECAY_FACTOR 768
#define INHIB_INFL_CENTER_OFFSET 1
#define INHIB_INFL_HORIZ_OFFSET 3
#define SLC_SCL_COEFF_L2 4

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/* ------------------------------------------------------------------ helpers */

/* rnd_mix: cl/bismit.cl:426-431.  seed ^= (seed ^ rnd) << 13; seed ^= seed >> 17
 * (arithmetic, int); seed ^= seed << 5. */
int32_t orc_rnd_mix(int32_t rnd, int32_t seed) {
    uint32_t s = (uint32_t)seed;
    s ^= (s ^ (uint32_t)rnd) << 13;
    s ^= (uint32_t)(((int32_t)s) >> 17);
    s ^= s << 5;
    return (int32_t)s;
}

/* abs() of an OpenCL char returns uchar: |-128| = 128. */
static inline int32_t abs_char(int8_t v) { return v < 0 ? -(int32_t)v : (int32_t)v; }

/* rnd_inc: cl/bismit.cl:447-454 */
int32_t orc_rnd_inc(int32_t rnd, int32_t seed, int8_t val, int32_t mask) {
    return (orc_rnd_mix(rnd, seed) & mask) > (abs_char(val) + (mask - 0x7f));
}

/* rnd_dec: cl/bismit.cl:457-469 */
int32_t orc_rnd_dec(int32_t rnd, int32_t seed, int8_t val, int32_t mask) {
    int32_t is_max = val == 127;
    return (orc_rnd_mix(rnd, seed) & mask) > ((abs_char(val) - is_max) + (mask - 0x7f));
}

/* lshft_mask applied to 0x7F << l2i: cl/bismit.cl:473-475, 572-577, 616-623 */
int32_t orc_lshft_mask(int32_t l2i) {
    int32_t mask = 0x7F << l2i;
    for (int32_t i = 0; i < l2i; i++) mask |= (1 << i);
    return mask;
}

/* update_activity_rating: cl/bismit.cl:531-540 with rnd_inc_u (:482-484) and
 * rnd_0xFFFF (:525-527).  The decrement tests the already-incremented rating. */
uint8_t orc_update_activity_rating(uint8_t a, int32_t is_active, int32_t rnd, int32_t seed,
                                   uint16_t decay) {
    a = (uint8_t)(a + (((orc_rnd_mix(rnd, seed) & 0xFF) > a) & is_active));
    int32_t seed2 = (int32_t)((uint32_t)seed << 1);
    a = (uint8_t)(a - (((orc_rnd_mix(rnd, seed2) & 0xFFFF) < (int32_t)decay) & (a > 0)));
    return a;
}

/* syn_fire: cl/syns.cl:2-4 */
uint8_t orc_syn_fire(uint8_t axn) {
    return (uint8_t)(((axn != 0) << 7) + (axn >> SYNAPSE_AXON_BIAS_LOG2));
}

/* ------------------------------------------------------------------- area */

typedef struct orc_tuft {
    bsm_tuft_desc d;
    uint32_t spt;          /* synapses per cell-tuft */
    uint64_t syn_idz;      /* tft_syn_idz (src/cortex/synapses.rs:171-199) */
    uint64_t den_idz;      /* tft_den_idz (src/cortex/dendrites.rs:57-68) */
} orc_tuft;

typedef struct orc_layer {
    bsm_layer_desc d;
    uint32_t cells;        /* depth * v * u */
    uint64_t n_syn, n_den, n_celtft;
    orc_tuft t[BSM_MAX_TUFTS];
    /* synapses (src/cortex/synapses.rs:206-211) */
    uint8_t *syn_states, *syn_flags, *syn_slc;
    int8_t *syn_str, *syn_v, *syn_u;
    /* dendrites (src/cortex/dendrites.rs:77-81) */
    uint8_t *den_states, *den_states_raw, *den_energies, *den_activities, *den_thresholds;
    /* tufts (src/cortex/tufts.rs:63-70) */
    uint8_t *tft_best_ids, *tft_best_raw, *tft_best, *tft_states;
    uint8_t *tft_prev_best_ids, *tft_prev_best_raw, *tft_prev_best, *tft_prev_states;
    /* cells (src/cortex/pyramidals.rs:66-70, src/cortex/spiny_stellates.rs:61-62) */
    uint8_t *soma, *cel_flags, *cel_energies, *cel_activities;
    /* activity smoother hosted by the layer: radius, centers (v, u), per group */
    int32_t smooth_r;
    uint32_t smooth_n, smooth_counts[6];
    int32_t* smooth_c;
} orc_layer;

struct orc_area {
    uint32_t v_size, u_size, n_slices, n_layers, settings;
    bsm_slice_desc slc[BSM_MAX_SLICES];
    uint32_t idz[BSM_MAX_SLICES];
    uint32_t axn_len;
    uint8_t* axn;
    orc_layer L[BSM_MAX_LAYERS];
};

static void* zalloc(uint64_t n) { return calloc(n ? n : 1, 1); }

int orc_area_create(const bsm_area_desc* d, orc_area** out) {
    if (!d || !out || d->n_slices == 0 || d->n_slices > BSM_MAX_SLICES ||
        d->n_layers > BSM_MAX_LAYERS)
        return 1;
    orc_area* a = (orc_area*)calloc(1, sizeof(orc_area));
    a->v_size = d->v_size;
    a->u_size = d->u_size;
    a->n_slices = d->n_slices;
    a->n_layers = d->n_layers;
    a->settings = d->settings;
    /* axon_idzs prefix sum: src/map/slice_map.rs:43-65 */
    uint32_t tot = 0;
    for (uint32_t s = 0; s < d->n_slices; s++) {
        a->slc[s] = d->slices[s];
        a->idz[s] = tot;
        tot += d->slices[s].v_size * d->slices[s].u_size;
    }
    a->axn_len = tot;
    a->axn = (uint8_t*)zalloc(tot);
    for (uint32_t l = 0; l < d->n_layers; l++) {
        orc_layer* L = &a->L[l];
        L->d = d->layers[l];
        L->cells = (uint32_t)L->d.depth * d->v_size * d->u_size;
        uint64_t ns = 0, nd = 0;
        for (uint32_t t = 0; t < L->d.n_tufts; t++) {
            orc_tuft* T = &L->t[t];
            T->d = L->d.tufts[t];
            T->spt = T->d.dens_per_tft * T->d.syns_per_den;
            T->syn_idz = ns;
            T->den_idz = nd;
            ns += (uint64_t)L->cells * T->spt;
            nd += (uint64_t)L->cells * T->d.dens_per_tft;
        }
        L->n_syn = ns;
        L->n_den = nd;
        L->n_celtft = (uint64_t)L->cells * L->d.n_tufts;
        L->syn_states = zalloc(ns); L->syn_flags = zalloc(ns); L->syn_slc = zalloc(ns);
        L->syn_str = zalloc(ns); L->syn_v = zalloc(ns); L->syn_u = zalloc(ns);
        L->den_states = zalloc(nd); L->den_states_raw = zalloc(nd); L->den_energies = zalloc(nd);
        L->den_activities = zalloc(nd); L->den_thresholds = zalloc(nd);
        uint64_t nct = L->n_celtft;
        L->tft_best_ids = zalloc(nct); L->tft_best_raw = zalloc(nct); L->tft_best = zalloc(nct);
        L->tft_states = zalloc(nct); L->tft_prev_best_ids = zalloc(nct);
        L->tft_prev_best_raw = zalloc(nct); L->tft_prev_best = zalloc(nct);
        L->tft_prev_states = zalloc(nct);
        L->soma = zalloc(L->cells); L->cel_flags = zalloc(L->cells);
        L->cel_energies = zalloc(L->cells); L->cel_activities = zalloc(L->cells);
    }
    *out = a;
    return 0;
}

void orc_area_destroy(orc_area* a) {
    if (!a) return;
    free(a->axn);
    for (uint32_t l = 0; l < a->n_layers; l++) {
        orc_layer* L = &a->L[l];
        void* ps[] = {L->syn_states, L->syn_flags, L->syn_slc, L->syn_str, L->syn_v, L->syn_u,
                      L->den_states, L->den_states_raw, L->den_energies, L->den_activities,
                      L->den_thresholds, L->tft_best_ids, L->tft_best_raw, L->tft_best,
                      L->tft_states, L->tft_prev_best_ids, L->tft_prev_best_raw,
                      L->tft_prev_best, L->tft_prev_states, L->soma, L->cel_flags,
                      L->cel_energies, L->cel_activities};
        for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
        free(L->smooth_c);
    }
    free(a);
}

int orc_slice_idz(orc_area* a, uint32_t slc, uint32_t* idz) {
    if (slc >= a->n_slices) return 1;
    *idz = a->idz[slc];
    return 0;
}

int orc_rnd_slots(orc_area* a, uint32_t* n) {
    uint32_t k = 0;
    for (uint32_t l = 0; l < a->n_layers; l++) k += 2u * a->L[l].d.n_tufts + 1u;
    *n = k;
    return 0;
}

int orc_buf_ptr(orc_area* a, int buf, uint32_t layer, void** ptr, size_t* len) {
    if (buf == BSM_AXN_STATES) { *ptr = a->axn; *len = a->axn_len; return 0; }
    if (layer >= a->n_layers) return 1;
    orc_layer* L = &a->L[layer];
    int ssc = L->d.kind == BSM_LAYER_SSC;
    switch (buf) {
    case BSM_SOMA_STATES:
        if (ssc) { *ptr = L->den_states; *len = L->n_den; }   /* soma == dens.states */
        else { *ptr = L->soma; *len = L->cells; }
        return 0;
    case BSM_SOMA_ENERGIES: *ptr = L->cel_energies; *len = L->cells; return 0;
    case BSM_SOMA_ACTIVITIES: *ptr = L->cel_activities; *len = L->cells; return 0;
    case BSM_SOMA_FLAG_SETS: if (ssc) return 1; *ptr = L->cel_flags; *len = L->cells; return 0;
    case BSM_TFT_STATES: *ptr = L->tft_states; *len = L->n_celtft; return ssc;
    case BSM_TFT_BEST_DEN_IDS: *ptr = L->tft_best_ids; *len = L->n_celtft; return ssc;
    case BSM_TFT_BEST_DEN_STATES_RAW: *ptr = L->tft_best_raw; *len = L->n_celtft; return ssc;
    case BSM_TFT_BEST_DEN_STATES: *ptr = L->tft_best; *len = L->n_celtft; return ssc;
    case BSM_TFT_PREV_STATES: *ptr = L->tft_prev_states; *len = L->n_celtft; return ssc;
    case BSM_TFT_PREV_BEST_DEN_IDS: *ptr = L->tft_prev_best_ids; *len = L->n_celtft; return ssc;
    case BSM_TFT_PREV_BEST_DEN_STATES_RAW: *ptr = L->tft_prev_best_raw; *len = L->n_celtft; return ssc;
    case BSM_TFT_PREV_BEST_DEN_STATES: *ptr = L->tft_prev_best; *len = L->n_celtft; return ssc;
    case BSM_DEN_STATES: *ptr = L->den_states; *len = L->n_den; return 0;
    case BSM_DEN_STATES_RAW: *ptr = L->den_states_raw; *len = L->n_den; return 0;
    case BSM_DEN_ENERGIES: *ptr = L->den_energies; *len = L->n_den; return 0;
    case BSM_DEN_ACTIVITIES: *ptr = L->den_activities; *len = L->n_den; return 0;
    case BSM_DEN_THRESHOLDS: *ptr = L->den_thresholds; *len = L->n_den; return 0;
    case BSM_SYN_STATES: *ptr = L->syn_states; *len = L->n_syn; return 0;
    case BSM_SYN_STRENGTHS: *ptr = L->syn_str; *len = L->n_syn; return 0;
    case BSM_SYN_SRC_SLC_IDS: *ptr = L->syn_slc; *len = L->n_syn; return 0;
    case BSM_SYN_SRC_COL_V_OFFS: *ptr = L->syn_v; *len = L->n_syn; return 0;
    case BSM_SYN_SRC_COL_U_OFFS: *ptr = L->syn_u; *len = L->n_syn; return 0;
    case BSM_SYN_FLAG_SETS: *ptr = L->syn_flags; *len = L->n_syn; return 0;
    default: return 1;
    }
}

/* ----------------------------------------------------------- axon indexing */

/* axn_state_3d_safe_vec4 / axn_idx_3d_unsafe_vec4: cl/bismit.cl:366-391,406-418.
 * Offsets and mids are added in int.  scalar_quirk != 0 selects the scalar
 * axn_idx_3d_unsafe (cl/bismit.cl:323-363): the mid is added into the char
 * offset, which wraps.  The host uses the scalar kernel only when
 * syns_per_tft % 4 != 0 (src/cortex/synapses.rs:261-268).  A slice id beyond
 * the table is undefined upstream; here it reads as 0. */
static inline uint32_t axn_index(const orc_area* a, uint32_t slc, int32_t v_id, int8_t v_ofs,
                                 int32_t u_id, int8_t u_ofs, int scalar_quirk, int* safe) {
    const bsm_slice_desc* s = &a->slc[slc];
    int32_t vsz = (int32_t)s->v_size, usz = (int32_t)s->u_size;
    int32_t vo, uo;
    if (scalar_quirk) {
        vo = (int8_t)(uint8_t)((uint8_t)v_ofs + s->v_mid);
        uo = (int8_t)(uint8_t)((uint8_t)u_ofs + s->u_mid);
    } else {
        vo = (int32_t)v_ofs + s->v_mid;
        uo = (int32_t)u_ofs + s->u_mid;
    }
    int32_t vs = (int32_t)(((uint32_t)v_id * s->v_scale) >> SLC_SCL_COEFF_L2);
    int32_t us = (int32_t)(((uint32_t)u_id * s->u_scale) >> SLC_SCL_COEFF_L2);
    int32_t vv = vs + vo, uu = us + uo;
    *safe = (vv >= 0) & (vv < vsz) & (uu >= 0) & (uu < usz);
    return a->idz[slc] + (uint32_t)(vv * usz + uu);
}

static inline uint8_t axn_state(const orc_area* a, uint32_t slc, int32_t v_id, int8_t v_ofs,
                                int32_t u_id, int8_t u_ofs, int scalar_quirk) {
    if (slc >= a->n_slices) return 0;
    int safe;
    uint32_t idx = axn_index(a, slc, v_id, v_ofs, u_id, u_ofs, scalar_quirk, &safe);
    return safe ? a->axn[idx] : 0;
}

uint8_t orc_axn_state(orc_area* a, uint32_t slc, uint32_t v_id, int8_t v_ofs, uint32_t u_id,
                      int8_t u_ofs, int32_t scalar_quirk, uint32_t* idx, int32_t* safe) {
    int sf = 0;
    uint32_t i = 0;
    if (slc < a->n_slices)
        i = axn_index(a, slc, (int32_t)v_id, v_ofs, (int32_t)u_id, u_ofs, scalar_quirk, &sf);
    if (idx) *idx = i;
    if (safe) *safe = sf;
    return sf ? a->axn[i] : 0;
}

/* axn_idx_3d_unsafe with zero offsets (the cell's own axon), scalar form:
 * cl/bismit.cl:323-363 as called from cl/control.cl:45-46, 209, 310-311, 344-345. */
static inline uint32_t cel_axn_idx(const orc_area* a, uint32_t slc, uint32_t v_id, uint32_t u_id,
                                   int* safe) {
    const bsm_slice_desc* s = &a->slc[slc];
    int32_t vo = (int8_t)(uint8_t)s->v_mid, uo = (int8_t)(uint8_t)s->u_mid;
    int32_t vs = (int32_t)((v_id * s->v_scale) >> SLC_SCL_COEFF_L2);
    int32_t us = (int32_t)((u_id * s->u_scale) >> SLC_SCL_COEFF_L2);
    int32_t vv = vs + vo, uu = us + uo;
    *safe = (vv >= 0) & (vv < (int32_t)s->v_size) & (uu >= 0) & (uu < (int32_t)s->u_size);
    return a->idz[slc] + (uint32_t)(vv * (int32_t)s->u_size + uu);
}

#define CHECK_LAYER(a, l) do { if ((l) >= (a)->n_layers) return 1; } while (0)
#define CHECK_TUFT(L, t) do { if ((t) >= (L)->d.n_tufts) return 1; } while (0)

/* ------------------------------------------------------------------ kernels */

/* tft_set_syn_flags: cl/syns.cl:23-33, over every synapse of the layer
 * (src/cortex/synapses.rs:222-227). */
int orc_k_syn_flags(orc_area* a, uint32_t layer) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    int64_t n = (int64_t)L->n_syn;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t i = 0; i < n; i++) {
        uint8_t f = L->syn_flags[i] & (uint8_t)~SYN_PREV_ACTIVE_FLAG;
        L->syn_flags[i] = f | ((L->syn_states[i] != 0) ? SYN_PREV_ACTIVE_FLAG : 0);
    }
    return 0;
}

/* layer_cycle_syns_wow_vec4 / layer_cycle_syns_wow: cl/syns.cl:179-388, bound
 * at src/cortex/synapses.rs:271-286.  The workgroup's tile-strided walk visits
 * every (cell, synapse) of the tile exactly once, so each synapse's state is
 * syn_fire(axn_state(src_slc, v, v_ofs, u, u_ofs)) for its own cell. */
int orc_k_syn_cycle(orc_area* a, uint32_t layer, uint32_t tuft) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    CHECK_TUFT(L, tuft);
    orc_tuft* T = &L->t[tuft];
    const uint32_t V = a->v_size, U = a->u_size, spt = T->spt;
    const int scalar_quirk = (spt % 4) != 0;
    const int64_t cells = L->cells;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < cells; c++) {
        uint32_t vu = (uint32_t)(c % ((int64_t)V * U));
        uint32_t v = vu / U, u = vu % U;
        uint64_t base = T->syn_idz + (uint64_t)c * spt;
        for (uint32_t k = 0; k < spt; k++) {
            uint64_t i = base + k;
            uint8_t ax = axn_state(a, L->syn_slc[i], (int32_t)v, L->syn_v[i], (int32_t)u,
                                   L->syn_u[i], scalar_quirk);
            L->syn_states[i] = orc_syn_fire(ax);
        }
    }
    return 0;
}

/* den_cycle_tft: cl/bismit.cl:695-747, bound at src/cortex/dendrites.rs:117-134. */
int orc_k_den_cycle(orc_area* a, uint32_t layer, uint32_t tuft, int32_t rnd) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    CHECK_TUFT(L, tuft);
    orc_tuft* T = &L->t[tuft];
    const uint32_t spd = T->d.syns_per_den, thresh = T->d.threshold;
    const int64_t nd = (int64_t)L->cells * T->d.dens_per_tft;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t d = 0; d < nd; d++) {
        uint64_t z = T->syn_idz + (uint64_t)d * spd;
        int32_t sum = 0, raw = 0;
        for (uint32_t k = 0; k < spd; k++) {
            int8_t st = L->syn_str[z + k];
            uint8_t s = L->syn_states[z + k];
            sum += (st >= 0) * (int32_t)s;
            raw += s;
        }
        int32_t active = (uint32_t)sum > thresh;   /* int vs uint compare -> unsigned */
        sum = active * sum;
        uint32_t den_idx = (uint32_t)(d + T->den_idz);
        int32_t seed = (int32_t)(((uint32_t)raw + thresh) | den_idx);
        L->den_activities[den_idx] = orc_update_activity_rating(
            L->den_activities[den_idx], active, rnd, seed, DENDRITE_ACTIVITY_DECAY_FACTOR);
        int32_t r4 = raw >> 4, s4 = sum >> 4;
        L->den_states_raw[den_idx] = (uint8_t)(r4 < 0 ? 0 : (r4 > 255 ? 255 : r4));
        L->den_states[den_idx] = (uint8_t)(s4 < 0 ? 0 : (s4 > 255 ? 255 : s4));
    }
    return 0;
}

/* tft_cycle: cl/bismit.cl:755-842, bound at src/cortex/tufts.rs:107-127. */
int orc_k_tft_cycle(orc_area* a, uint32_t layer, uint32_t tuft) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    CHECK_TUFT(L, tuft);
    orc_tuft* T = &L->t[tuft];
    const uint32_t dens = T->d.dens_per_tft, l2 = T->d.max_active_dens_l2;
    const int64_t C = L->cells;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        uint64_t dz = T->den_idz + (uint64_t)c * dens;
        uint32_t best_id = 0, best_raw = 0, best = 0, asum = 0;
        for (uint32_t j = 0; j < dens; j++) {
            uint32_t sr = L->den_states_raw[dz + j], s = L->den_states[dz + j];
            if (sr > best_raw) { best_id = j; best_raw = sr; }
            if (s > best) best = s;
            asum += s;
        }
        int is_max = asum >= (1u << (l2 + 8));
        uint8_t st = (uint8_t)((uint8_t)(is_max * 255u) + (uint8_t)(!is_max * (asum >> l2)));
        uint64_t ct = (uint64_t)tuft * C + c;
        L->tft_prev_best_ids[ct] = L->tft_best_ids[ct];
        L->tft_prev_best_raw[ct] = L->tft_best_raw[ct];
        L->tft_prev_best[ct] = L->tft_best[ct];
        L->tft_prev_states[ct] = L->tft_states[ct];
        L->tft_best_ids[ct] = (uint8_t)best_id;
        L->tft_best_raw[ct] = (uint8_t)best_raw;
        L->tft_best[ct] = (uint8_t)best;
        L->tft_states[ct] = st;
    }
    return 0;
}

/* ssc_cycle: cl/bismit.cl:849-884 (src/cortex/spiny_stellates.rs:78-84); the
 * SSC soma is dens.states[0..C) (src/cortex/spiny_stellates.rs:83,283). */
int orc_k_ssc_cycle(orc_area* a, uint32_t layer) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const int64_t C = L->cells;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        uint32_t energy = L->cel_energies[c], state = L->den_states[c];
        int is_active = state != 0;
        int restless = (energy > 191) & is_active;
        int dark = energy == 255;
        L->cel_energies[c] = (uint8_t)((dark | restless) ? energy - 64 : energy);
        uint32_t st = state + (uint32_t)restless * 255u + (uint32_t)dark * 255u;
        L->den_states[c] = (uint8_t)(st > 255 ? 255 : st);
    }
    return 0;
}

static inline uint8_t* layer_soma(orc_layer* L) {
    return L->d.kind == BSM_LAYER_SSC ? L->den_states : L->soma;
}

/* inhib_simple: cl/control.cl:28-191, launched (depth, v, u) by
 * src/cortex/iinn.rs:39-54 with the host layer's soma and activities. */
int orc_k_inhib_simple(orc_area* a, uint32_t layer, int32_t rnd) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const int32_t V = (int32_t)a->v_size, U = (int32_t)a->u_size, r = L->d.inhib_radius;
    const uint8_t* soma = layer_soma(L);
    const int32_t horizon = r + INHIB_INFL_HORIZ_OFFSET;
    const int32_t imax = horizon * horizon;
    const int64_t C = L->cells;
    uint8_t* out = a->axn;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        int32_t slc = (int32_t)(c / ((int64_t)V * U));
        int32_t vu = (int32_t)(c % ((int64_t)V * U));
        int32_t v = vu / U, u = vu % U;
        int safe;
        uint32_t axi = cel_axn_idx(a, (uint32_t)(slc + L->d.base_axn_slc), (uint32_t)v,
                                   (uint32_t)u, &safe);
        uint8_t cs = (uint8_t)(safe * soma[c]);
        int unin = 1;
        for (int32_t vo = -r; vo <= r; vo++) {
            int32_t uz = (-r > -vo - r) ? -r : -vo - r;
            int32_t um = (r < -vo + r) ? r : -vo + r;
            for (int32_t uo = uz; uo <= um; uo++) {
                int32_t nv = v + vo, nu = u + uo;
                uint8_t ns = 0;
                if (nv >= 0 && nv < V && nu >= 0 && nu < U)
                    ns = soma[(int64_t)slc * V * U + (int64_t)nv * U + nu];
                int32_t w = -vo - uo;
                int32_t dist = ((vo < 0 ? -vo : vo) + (uo < 0 ? -uo : uo) + (w < 0 ? -w : w)) >> 1;
                int32_t dc = dist - INHIB_INFL_CENTER_OFFSET;
                int32_t nei_f = imax - dc * dc;
                unin &= ((int32_t)cs * imax) >= ((int32_t)ns * nei_f);
            }
        }
        out[axi] = (uint8_t)(unin * cs);
        int act = unin & (cs != 0);
        int32_t seed = (int32_t)((uint32_t)cs | (uint32_t)c);
        L->cel_activities[c] = orc_update_activity_rating(L->cel_activities[c], act, rnd, seed,
                                                          CELL_ACTIVITY_DECAY_FACTOR);
    }
    return 0;
}

/* inhib_passthrough: cl/control.cl:194-217. */
int orc_k_inhib_passthrough(orc_area* a, uint32_t layer, int32_t rnd) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const int32_t V = (int32_t)a->v_size, U = (int32_t)a->u_size;
    const uint8_t* soma = layer_soma(L);
    const int64_t C = L->cells;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        int32_t slc = (int32_t)(c / ((int64_t)V * U));
        int32_t vu = (int32_t)(c % ((int64_t)V * U));
        int safe;
        uint32_t axi = cel_axn_idx(a, (uint32_t)(slc + L->d.base_axn_slc), (uint32_t)(vu / U),
                                   (uint32_t)(vu % U), &safe);
        uint8_t cs = soma[c];
        a->axn[axi] = cs;
        int32_t seed = (int32_t)((uint32_t)cs | (uint32_t)c);
        L->cel_activities[c] = orc_update_activity_rating(L->cel_activities[c], cs != 0, rnd,
                                                          seed, CELL_ACTIVITY_DECAY_FACTOR);
    }
    return 0;
}

/* ssc_mtp_simple: cl/bismit.cl:888-915 with prx_syns__active__mtp_ltd
 * (:604-639); host src/cortex/spiny_stellates.rs:98-112,256-268.  Seeds are
 * syn_idz + i with absolute i, i.e. 2*syn_idz + k. */
int orc_k_ssc_mtp(orc_area* a, uint32_t layer, uint32_t rnd_u) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const int32_t rnd = (int32_t)rnd_u;
    const uint32_t spt = L->t[0].spt;
    const uint32_t axn_idz = a->idz[L->d.base_axn_slc];
    const int32_t pr_mask = orc_lshft_mask(0), dr_mask = orc_lshft_mask(2);
    const int64_t C = L->cells;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        if (!a->axn[axn_idz + (uint32_t)c]) continue;
        uint32_t z = (uint32_t)c * spt;
        for (uint32_t i = z; i < z + spt; i++) {
            int32_t seed = (int32_t)(z + i);
            int8_t s = L->syn_str[i];
            int32_t inc = orc_rnd_inc(rnd, seed, s, pr_mask);
            int32_t dec = orc_rnd_dec(rnd, seed, s, dr_mask);
            int32_t act = L->syn_states[i] != 0;
            s = (int8_t)(s + act * inc);
            s = (int8_t)(s - (!act) * dec);
            L->syn_str[i] = s;
        }
    }
    return 0;
}

static void tuft_ids(const orc_layer* L, uint32_t* en, uint32_t* prx, uint32_t* bdst,
                     uint32_t* adst) {
    *en = 0; *prx = *bdst = *adst = 0;
    for (uint32_t t = 0; t < L->d.n_tufts; t++) {
        const bsm_tuft_desc* d = &L->t[t].d;
        if (d->den_class == BSM_DEN_BASAL && d->den_kind == BSM_DEN_PROXIMAL) {
            *en |= DEN_BASAL_PROXIMAL_FLAG; *prx = t;
        } else if (d->den_class == BSM_DEN_BASAL && d->den_kind == BSM_DEN_DISTAL) {
            *en |= DEN_BASAL_DISTAL_FLAG; *bdst = t;
        } else if (d->den_class == BSM_DEN_APICAL && d->den_kind == BSM_DEN_DISTAL) {
            *en |= DEN_APICAL_DISTAL_FLAG; *adst = t;
        }
    }
}

/* pyr_cycle: cl/bismit.cl:961-1022, bound at src/cortex/pyramidals.rs:148-165:
 * the parameter named tft_prev_states receives best_den_states_raw, and the
 * comma operator at :1021 discards the apical term. */
int orc_k_pyr_cycle(orc_area* a, uint32_t layer) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    uint32_t en, prx, bdst, adst;
    tuft_ids(L, &en, &prx, &bdst, &adst);
    const int64_t C = L->cells;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        uint8_t p = (en & DEN_BASAL_PROXIMAL_FLAG) ? L->tft_states[(int64_t)prx * C + c] : 0;
        uint8_t d = (en & DEN_BASAL_DISTAL_FLAG) ? L->tft_best_raw[(int64_t)bdst * C + c] : 0;
        int act = p != 0;
        int is_min = (p <= 3) && act;
        uint8_t prx_c = (uint8_t)((p >> 2) + is_min);
        uint8_t dst_c = (uint8_t)(act * (d >> 1));
        L->soma[c] = (uint8_t)(prx_c + dst_c);
    }
    return 0;
}

/* inhib_intra_column: cl/control.cl:319-351 (src/cortex/intra_column_inhib.rs:36-46).
 * The running "best" resets to 0 on a smaller state. */
int orc_k_inhib_intra_column(orc_area* a, uint32_t layer) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const uint32_t V = a->v_size, U = a->u_size, depth = L->d.depth;
    const uint8_t* soma = layer_soma(L);
    const int64_t cols = (int64_t)V * U;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t col = 0; col < cols; col++) {
        uint32_t v = (uint32_t)(col / U), u = (uint32_t)(col % U);
        uint8_t best = 0;
        for (uint32_t s = 0; s < depth; s++) {
            uint8_t cs = soma[(int64_t)s * cols + col];
            best = (uint8_t)((cs >= best) * cs);
        }
        for (uint32_t s = 0; s < depth; s++) {
            uint8_t cs = soma[(int64_t)s * cols + col];
            int safe;
            uint32_t axi = cel_axn_idx(a, s + L->d.base_axn_slc, v, u, &safe);
            a->axn[axi] = (uint8_t)((cs >= best) * cs);
        }
    }
    return 0;
}

/* pyr_output: cl/control.cl:298-314. */
int orc_k_pyr_output(orc_area* a, uint32_t layer) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    const uint32_t V = a->v_size, U = a->u_size;
    const uint8_t* soma = layer_soma(L);
    const int64_t C = L->cells;
    for (int64_t c = 0; c < C; c++) {
        uint32_t slc = (uint32_t)(c / ((int64_t)V * U)), vu = (uint32_t)(c % ((int64_t)V * U));
        int safe;
        uint32_t axi = cel_axn_idx(a, slc + L->d.base_axn_slc, vu / U, vu % U, &safe);
        a->axn[axi] = soma[c];
    }
    return 0;
}

/* tft_dst_mtp: cl/bismit.cl:1195-1268 with dst_syns__active__mtpot_mtdep
 * (:554-600); host src/cortex/tufts.rs:151-213,304-333 (pr_l2i 0, dr_l2i 5).
 * 64 cell groups in the reference, each walking its cells in order; cells are
 * independent, so a flat loop over cells is the same computation. */
int orc_k_tft_dst_mtp(orc_area* a, uint32_t layer, uint32_t tuft, int32_t rnd) {
    CHECK_LAYER(a, layer);
    orc_layer* L = &a->L[layer];
    CHECK_TUFT(L, tuft);
    orc_tuft* T = &L->t[tuft];
    const uint32_t spd = T->d.syns_per_den, spt = T->spt;
    const uint32_t axn_idz = a->idz[L->d.base_axn_slc];
    const int32_t pr_mask = orc_lshft_mask(0), dr_mask = orc_lshft_mask(5);
    const int64_t C = L->cells;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads) if (g_threads > 1)
    for (int64_t c = 0; c < C; c++) {
        uint8_t f = L->cel_flags[c];
        int act = a->axn[axn_idz + (uint32_t)c] != 0;
        int prev = (f & CEL_PREV_ACTIVE_FLAG) == CEL_PREV_ACTIVE_FLAG;
        int newly = !prev & act;
        uint64_t ct = (uint64_t)tuft * C + c;
        int tpa = L->tft_prev_best_raw[ct] != 0;
        /* an out-of-range best-dendrite id (user-written) is undefined upstream; skipped */
        if (newly & tpa && (uint32_t)L->tft_prev_best_ids[ct] * spd < spt) {
            uint32_t z = (uint32_t)((uint32_t)c * spt + (uint32_t)T->syn_idz) +
                         (uint32_t)L->tft_prev_best_ids[ct] * spd;
            for (uint32_t i = z; i < z + spd; i++) {
                int8_t s = L->syn_str[i];
                int32_t seed = (int32_t)(z + i);
                if ((L->syn_flags[i] & SYN_PREV_ACTIVE_FLAG) == SYN_PREV_ACTIVE_FLAG)
                    s = (int8_t)(s + orc_rnd_inc(rnd, seed, s, pr_mask));
                else
                    s = (int8_t)(s - orc_rnd_dec(rnd, seed, s, dr_mask));
                L->syn_str[i] = s;
            }
        }
        f &= (uint8_t)~CEL_PREV_ACTIVE_FLAG;
        f |= (uint8_t)(act * CEL_PREV_ACTIVE_FLAG);
        L->cel_flags[c] = f;
    }
    return 0;
}

/* ------------------------------------------------------------- area cycle */

/* CorticalArea::cycle enqueue order (src/cortex/cortical_area.rs:716-789),
 * executed in order:
 *   SPATIAL layers: flags, gather(t), den(t), ssc_cycle, inhib
 *     (src/cortex/spiny_stellates.rs:220-252, dendrites.rs:191-209, synapses.rs:363-387)
 *   SPATIAL layers: ssc_mtp (src/cortex/spiny_stellates.rs:256-268)
 *   other layers: flags, gather(t), den(t), tft_cycle(t), pyr_cycle, control,
 *     tft_dst_mtp(t) for distal tufts (src/cortex/pyramidals.rs:297-335, tufts.rs:304-358)
 * rnd slots per layer (descriptor order): den[t], ctrl, mtp[t].  A bypassed
 * inhibition runs inhib_passthrough with rnd 0 (its rnd argument is never set,
 * src/cortex/iinn.rs:57-66,105-111). */
static int layer_slot0(orc_area* a, uint32_t layer) {
    uint32_t k = 0;
    for (uint32_t l = 0; l < layer; l++) k += 2u * a->L[l].d.n_tufts + 1u;
    return (int)k;
}

static int cycle_dens(orc_area* a, uint32_t l, const int32_t* r) {
    orc_layer* L = &a->L[l];
    int s0 = layer_slot0(a, l);
    int e = orc_k_syn_flags(a, l);
    for (uint32_t t = 0; t < L->d.n_tufts; t++) e |= orc_k_syn_cycle(a, l, t);
    for (uint32_t t = 0; t < L->d.n_tufts; t++) e |= orc_k_den_cycle(a, l, t, r[s0 + t]);
    return e;
}

static int run_control(orc_area* a, uint32_t l, const int32_t* r) {
    orc_layer* L = &a->L[l];
    int32_t rc = r[layer_slot0(a, l) + L->d.n_tufts];
    switch (L->d.control) {
    case BSM_CTRL_INHIB_SIMPLE:
        if (a->settings & BSM_SET_BYPASS_INHIB) return orc_k_inhib_passthrough(a, l, 0);
        return orc_k_inhib_simple(a, l, rc);
    case BSM_CTRL_INTRA_COLUMN: return orc_k_inhib_intra_column(a, l);
    case BSM_CTRL_PYR_OUTPUT: return orc_k_pyr_output(a, l);
    default: return 0;
    }
}

int orc_area_cycle(orc_area* a, const int32_t* rnd, uint32_t n_cycles) {
    uint32_t ns;
    orc_rnd_slots(a, &ns);
    const int learn = !(a->settings & BSM_SET_DISABLE_LEARNING);
    for (uint32_t cyc = 0; cyc < n_cycles; cyc++) {
        const int32_t* r = rnd + (size_t)cyc * ns;
        int e = 0;
        if (!(a->settings & BSM_SET_DISABLE_SSCS)) {
            for (uint32_t l = 0; l < a->n_layers; l++) {
                if (!a->L[l].d.spatial) continue;
                e |= cycle_dens(a, l, r);
                e |= orc_k_ssc_cycle(a, l);
                e |= run_control(a, l, r);
                if (a->L[l].smooth_n) e |= orc_k_smooth(a, l);
            }
            for (uint32_t l = 0; l < a->n_layers && learn; l++) {
                if (!a->L[l].d.spatial) continue;
                e |= orc_k_ssc_mtp(a, l, (uint32_t)r[layer_slot0(a, l) + a->L[l].d.n_tufts + 1]);
            }
        }
        if (!(a->settings & BSM_SET_DISABLE_PYRS)) {
            for (uint32_t l = 0; l < a->n_layers; l++) {
                orc_layer* L = &a->L[l];
                if (L->d.spatial) continue;
                e |= cycle_dens(a, l, r);
                for (uint32_t t = 0; t < L->d.n_tufts; t++) e |= orc_k_tft_cycle(a, l, t);
                e |= orc_k_pyr_cycle(a, l);
                e |= run_control(a, l, r);
                if (L->smooth_n) e |= orc_k_smooth(a, l);
                if (learn) {
                    int s0 = layer_slot0(a, l) + L->d.n_tufts + 1;
                    for (uint32_t t = 0; t < L->d.n_tufts; t++)
                        if (L->t[t].d.den_kind == BSM_DEN_DISTAL)
                            e |= orc_k_tft_dst_mtp(a, l, t, r[s0 + t]);
                }
            }
        }
        if (e) return e;
    }
    return 0;
}

/* ------------------------------------------------------------------ regrowth */

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Hex tile of gen_syn_offs (src/map/syn_src.rs:31-107) for a spatial slice:
 * position q (row-major) -> (v, u); returns the tile size when q is past it. */
static uint32_t hex_tile_at(int32_t rv, int32_t vsc, int32_t usc, uint32_t q, int32_t* v,
                            int32_t* u) {
    uint32_t n = 0;
    for (int32_t vo = -rv; vo <= rv; vo++) {
        const int32_t lo = (vo < 0 ? -vo - rv : -rv);   /* max(-rv, -vo - rv) */
        const int32_t hi = (vo > 0 ? rv - vo : rv);     /* min(rv, -vo + rv) */
        const int32_t uz = lo * usc / vsc, un = hi * usc / vsc;
        const uint32_t w = (uint32_t)(un - uz + 1);
        if (q < n + w) {
            *v = vo;
            *u = uz + (int32_t)(q - n);
            return 0;
        }
        n += w;
    }
    return n;
}

int orc_area_regrow(orc_area* a, uint64_t seed, uint64_t epoch, int32_t reach) {
    if (reach < 0 || reach > 127) return 1;
    for (uint32_t l = 0; l < a->n_layers; l++) {
        orc_layer* L = &a->L[l];
        for (uint32_t t = 0; t < L->d.n_tufts; t++) {
            const orc_tuft* T = &L->t[t];
            const uint32_t n_src = T->d.n_src_slcs, spd = T->d.syns_per_den;
            if (n_src == 0) continue;
            const uint64_t key = mix64(seed ^ (epoch * 0xD1B54A32D192ED03ull)) ^
                                 ((uint64_t)(l * 8 + t) << 56);
            /* tile sizes per source */
            uint32_t tile[8];
            for (uint32_t k = 0; k < n_src; k++) {
                const bsm_slice_desc* s = &a->slc[T->d.src_slc_ids[k]];
                if (s->v_scale == 0 || s->u_scale == 0) {
                    tile[k] = 0;
                } else {
                    int32_t dv, du;
                    tile[k] = hex_tile_at((reach * s->v_scale) >> 4, s->v_scale, s->u_scale,
                                          0xFFFFFFFFu, &dv, &du);
                }
            }
            const uint64_t n_den = (uint64_t)L->cells * T->d.dens_per_tft;
            for (uint64_t d = 0; d < n_den; d++) {
                const uint64_t base = T->syn_idz + d * spd;
                for (uint32_t k = 0; k < spd; k++) {
                    const uint64_t i = base + k;
                    if (L->syn_str[i] > -25) continue;
                    const uint64_t ti = d * spd + k;   /* tuft-relative synapse index */
                    for (uint32_t att = 0; att < 32; att++) {
                        const uint64_t h = mix64(key + (ti * 32 + att + 1) * 0x9E3779B97F4A7C15ull);
                        const uint32_t slot = (uint32_t)(h & 0xFFFF) % n_src;
                        const uint8_t slc = T->d.src_slc_ids[slot];
                        const bsm_slice_desc* s = &a->slc[slc];
                        int32_t v, u, str;
                        if (tile[slot] == 0) {
                            const int32_t rv = (int32_t)(s->v_size / 2), ru = (int32_t)(s->u_size / 2);
                            v = (int32_t)(((h >> 16) & 0xFFFF) % (uint64_t)(2 * rv + 1)) - rv;
                            u = (int32_t)(((h >> 32) & 0xFFFF) % (uint64_t)(2 * ru + 1)) - ru;
                            str = 0;
                        } else {
                            const int32_t rv = (reach * s->v_scale) >> 4, ru = (reach * s->u_scale) >> 4;
                            hex_tile_at(rv, s->v_scale, s->u_scale,
                                        (uint32_t)(((h >> 16) & 0xFFFFFF) % tile[slot]), &v, &u);
                            const int32_t inten = ((rv - abs(v)) + (ru - abs(u))) >> 3;
                            str = inten * ((int32_t)(((h >> 40) & 0xFFFF) % 7) - 3);
                        }
                        int unique = 1;
                        for (uint32_t j = 0; j < spd && unique; j++) {
                            if (j == k) continue;
                            unique = !(L->syn_slc[base + j] == slc && L->syn_v[base + j] == (int8_t)v &&
                                       L->syn_u[base + j] == (int8_t)u);
                        }
                        if (unique || att == 31) {
                            L->syn_slc[i] = slc;
                            L->syn_v[i] = (int8_t)v;
                            L->syn_u[i] = (int8_t)u;
                            L->syn_str[i] = (int8_t)str;
                            break;
                        }
                    }
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------ sensory filter */

/* retina (cl/filters.cl:34-100) with get_neighbors_avg (:8-31): per cell, bit 0
 * of the cell index selects off-center, bit 1 xor bit 0 the large field (radius
 * 5, else 3); the mean runs over the whole hex field, off-grid cells adding 0
 * (cel_state_3d_safe, cl/bismit.cl:285-297) but counting; RETNAL_THRESHOLD = 48
 * (cl/bismit.cl:19). */
int orc_retina(orc_area* a, uint32_t slc, uint32_t depth, const uint8_t* frame) {
    if (!depth || slc >= a->n_slices || depth > a->n_slices - slc) return 1;
    const int32_t V = a->slc[slc].v_size, U = a->slc[slc].u_size;
    for (uint32_t k = 1; k < depth; k++)
        if ((int32_t)a->slc[slc + k].v_size != V || (int32_t)a->slc[slc + k].u_size != U) return 1;
    for (uint32_t s = 0; s < depth; s++) {
        const uint8_t* cs = frame + (size_t)s * V * U;
        uint8_t* out = a->axn + a->idz[slc + s];
        for (int32_t v = 0; v < V; v++) {
            for (int32_t u = 0; u < U; u++) {
                const uint32_t cel_idx = s * (uint32_t)(V * U) + (uint32_t)(v * U + u);
                const int32_t one_of_two = (int32_t)(cel_idx & 1);
                const int32_t two_of_four = (int32_t)((cel_idx & 3) >> 1) ^ one_of_two;
                const int32_t radius = two_of_four ? 5 : 3;
                int32_t sum = 0, count = 0;
                for (int32_t vo = -radius; vo <= radius; vo++) {
                    const int32_t uz = -vo - radius > -radius ? -vo - radius : -radius;
                    const int32_t um = -vo + radius < radius ? -vo + radius : radius;
                    for (int32_t uo = uz; uo <= um; uo++) {
                        const int32_t vv = v + vo, uu = u + uo;
                        if (vv >= 0 && vv < V && uu >= 0 && uu < U) sum += cs[vv * U + uu];
                        count++;
                    }
                }
                const int32_t avg = sum / count;
                const int32_t center = cs[v * U + u];
                int32_t st;
                if (one_of_two) {
                    const int32_t cst = center + 48;
                    st = avg > cst ? avg - cst : 0;
                } else {
                    const int32_t nat = avg + 48;
                    st = center > nat ? center - nat : 0;
                }
                out[v * U + u] = (uint8_t)st;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ smoother */

/* gen_grp_centers (src/cortex/smoother.rs:33-58): six groups seeded at
 * (0, d), (-d, d), (-d, 0), (0, -d), (d, -d), (d, 0) with d = (r + 1) / 2; each
 * group is the set HexGroupCenters::populate (src/cmn/mod.rs:892-921) reaches
 * from its seed through add_surrounds' six steps (:867-888, clockwise bias:
 * l = r + 1, s = r) inside [-r, dims + r) (add_center, :852-864).  The set is
 * what is fixed; the reference's HashSet order is not, so groups list their
 * centers row-major. */
int orc_smoother_centers(int32_t r, int32_t V, int32_t U, int32_t* ov, int32_t* ou,
                         uint32_t* counts, uint32_t cap, uint32_t* n) {
    if (r < 1 || V < 1 || U < 1) return 1;
    const int32_t l = r + 1, s = r, ls = l + s, d = (r + 1) / 2;
    const int32_t seeds[6][2] = {{0, d}, {-d, d}, {-d, 0}, {0, -d}, {d, -d}, {d, 0}};
    const int32_t sur[6][2] = {{-s, ls}, {l, s}, {ls, -l}, {s, -ls}, {-l, -s}, {-ls, l}};
    const int32_t W = V + 2 * r, H = U + 2 * r;
    uint8_t* in_set = (uint8_t*)malloc((size_t)W * H);
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)W * H);
    uint32_t tot = 0;
    for (int g = 0; g < 6; g++) {
        memset(in_set, 0, (size_t)W * H);
        size_t qn = 0, qi = 0;
        int32_t sv = seeds[g][0], su = seeds[g][1];
        if (sv >= -r && su >= -r && sv < V + r && su < U + r) {
            in_set[(size_t)(sv + r) * H + (su + r)] = 1;
            q[2 * qn] = sv;
            q[2 * qn + 1] = su;
            qn++;
        }
        while (qi < qn) {   /* breadth-first over new centers */
            const int32_t cv = q[2 * qi], cu = q[2 * qi + 1];
            qi++;
            for (int k = 0; k < 6; k++) {
                const int32_t nv = cv + sur[k][0], nu = cu + sur[k][1];
                if (nv < -r || nu < -r || nv >= V + r || nu >= U + r) continue;
                uint8_t* m = &in_set[(size_t)(nv + r) * H + (nu + r)];
                if (*m) continue;
                *m = 1;
                q[2 * qn] = nv;
                q[2 * qn + 1] = nu;
                qn++;
            }
        }
        counts[g] = (uint32_t)qn;
        for (int32_t i = 0; i < W; i++)
            for (int32_t j = 0; j < H; j++)
                if (in_set[(size_t)i * H + j]) {
                    if (tot < cap) {
                        if (ov) ov[tot] = i - r;
                        if (ou) ou[tot] = j - r;
                    }
                    tot++;
                }
    }
    free(in_set);
    free(q);
    *n = tot;
    return 0;
}

int orc_set_smoother(orc_area* a, uint32_t layer, int32_t radius) {
    if (layer >= a->n_layers || radius < 0) return 1;
    orc_layer* L = &a->L[layer];
    free(L->smooth_c);
    L->smooth_c = NULL;
    L->smooth_n = 0;
    L->smooth_r = 0;
    if (!radius) return 0;
    uint32_t n = 0;
    if (orc_smoother_centers(radius, a->v_size, a->u_size, NULL, NULL, L->smooth_counts, 0, &n))
        return 1;
    int32_t* v = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* u = (int32_t*)malloc(sizeof(int32_t) * n);
    orc_smoother_centers(radius, a->v_size, a->u_size, v, u, L->smooth_counts, n, &n);
    L->smooth_c = (int32_t*)malloc(sizeof(int32_t) * 2 * n);
    for (uint32_t i = 0; i < n; i++) {
        L->smooth_c[2 * i] = v[i];
        L->smooth_c[2 * i + 1] = u[i];
    }
    free(v);
    free(u);
    L->smooth_n = n;
    L->smooth_r = radius;
    return 0;
}

/* smooth_activity (cl/control.cl:227-295), bound as in ActivitySmoother::new
 * (src/cortex/smoother.rs:97-109: the host layer's dims, depth, activities and
 * energies).  One work-item's body per center, centers of a group in any order
 * (their fields are disjoint), groups in seed order. */
int orc_k_smooth(orc_area* a, uint32_t layer) {
    if (layer >= a->n_layers) return 1;
    orc_layer* L = &a->L[layer];
    const int32_t V = a->v_size, U = a->u_size, r = L->smooth_r;
    for (uint32_t c = 0; c < L->smooth_n; c++) {
        const int32_t cv = L->smooth_c[2 * c], cu = L->smooth_c[2 * c + 1];
        uint32_t least_idx = 0, most_idx = 0;
        uint8_t least_actv = 255, most_actv = 0;
        for (uint32_t slc = 0; slc < L->d.depth; slc++) {
            for (int32_t vo = -r; vo <= r; vo++) {
                const int32_t uz = -vo - r > -r ? -vo - r : -r;
                const int32_t um = -vo + r < r ? -vo + r : r;
                for (int32_t uo = uz; uo <= um; uo++) {
                    const int32_t vv = cv + vo, uu = cu + uo;
                    const int safe = vv >= 0 && vv < V && uu >= 0 && uu < U;
                    if (!safe) continue;
                    const uint32_t idx = slc * (uint32_t)(V * U) + (uint32_t)(vv * U + uu);
                    const uint8_t x = L->cel_activities[idx];
                    if (x <= least_actv) { least_idx = idx; least_actv = x; }
                    if (x >= most_actv) { most_idx = idx; most_actv = x; }
                }
            }
        }
        const int differ = least_idx != most_idx;
        const uint8_t le = L->cel_energies[least_idx];
        L->cel_energies[least_idx] = (uint8_t)(le + ((le < 255) & (most_actv < 255) & differ));
        const uint8_t me = L->cel_energies[most_idx];
        L->cel_energies[most_idx] = (uint8_t)(me - ((me > 0) & (most_actv > 0) & differ));
    }
    return 0;
}
