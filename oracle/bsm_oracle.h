/*
 * bsm_oracle.h -- CPU restatement of bismit's per-cycle cortical pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP engine is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product path (bismit_amd) never links, loads or calls it.
 *
 * Every function restates one reference kernel or helper and cites the
 * reference file:line it follows.  Buffers use the reference layouts exactly
 * (see include/bismit_hip.h) and the same area descriptor type.
 */
#ifndef BSM_ORACLE_H
#define BSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/bismit_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_area orc_area;

/* helpers (pinned individually by tests) */
int32_t orc_rnd_mix(int32_t rnd, int32_t seed);
int32_t orc_rnd_inc(int32_t rnd, int32_t seed, int8_t val, int32_t mask);
int32_t orc_rnd_dec(int32_t rnd, int32_t seed, int8_t val, int32_t mask);
uint8_t orc_update_activity_rating(uint8_t a, int32_t is_active, int32_t rnd, int32_t seed,
                                   uint16_t decay);
uint8_t orc_syn_fire(uint8_t axn);
int32_t orc_lshft_mask(int32_t l2i);

/* area */
int orc_area_create(const bsm_area_desc* desc, orc_area** out);
void orc_area_destroy(orc_area* a);
int orc_buf_ptr(orc_area* a, int buf, uint32_t layer, void** ptr, size_t* len);
int orc_slice_idz(orc_area* a, uint32_t slc, uint32_t* idz);
void orc_set_threads(int n);
int orc_rnd_slots(orc_area* a, uint32_t* n);

/* axn_state_3d_safe (cl/bismit.cl:397-418) of one synapse source: the state
 * the gather reads for a cell at (v_id, u_id) with offsets (v_ofs, u_ofs) into
 * slice slc; scalar_quirk selects the scalar index form (char-wrapped mids,
 * cl/bismit.cl:323-363) over the vec4 form (int mids, :366-391).  *idx gets the
 * axon index (also when unsafe) and *safe the bounds test.  Test probe. */
uint8_t orc_axn_state(orc_area* a, uint32_t slc, uint32_t v_id, int8_t v_ofs, uint32_t u_id,
                      int8_t u_ofs, int32_t scalar_quirk, uint32_t* idx, int32_t* safe);

/* kernels (same argument meaning as the bsm_k_* entry points) */
int orc_k_syn_flags(orc_area* a, uint32_t layer);
int orc_k_syn_cycle(orc_area* a, uint32_t layer, uint32_t tuft);
int orc_k_den_cycle(orc_area* a, uint32_t layer, uint32_t tuft, int32_t rnd);
int orc_k_tft_cycle(orc_area* a, uint32_t layer, uint32_t tuft);
int orc_k_ssc_cycle(orc_area* a, uint32_t layer);
int orc_k_inhib_simple(orc_area* a, uint32_t layer, int32_t rnd);
int orc_k_inhib_passthrough(orc_area* a, uint32_t layer, int32_t rnd);
int orc_k_ssc_mtp(orc_area* a, uint32_t layer, uint32_t rnd);
int orc_k_pyr_cycle(orc_area* a, uint32_t layer);
int orc_k_inhib_intra_column(orc_area* a, uint32_t layer);
int orc_k_pyr_output(orc_area* a, uint32_t layer);
int orc_k_tft_dst_mtp(orc_area* a, uint32_t layer, uint32_t tuft, int32_t rnd);

/* whole cycle, reference enqueue order; rnd = n_cycles * n_slots draws */
int orc_area_cycle(orc_area* a, const int32_t* rnd, uint32_t n_cycles);

/* Deterministic regrowth (CorticalArea::regrow, src/cortex/cortical_area.rs:792-806;
 * Synapses::grow / regrow_syn, src/cortex/synapses.rs:411-501; SynSrcSlices::gen_src,
 * src/map/syn_src.rs:448-490).  Every synapse with strength <= -25
 * (SYNAPSE_STRENGTH_FLOOR) gets a new source drawn from a counter-based hash instead of
 * an entropy-seeded generator:
 *   key = mix64(seed ^ epoch * 0xD1B54A32D192ED03) ^ ((layer * 8 + tuft) << 56)
 *   draw a of tuft-relative synapse i: h = mix64(key + (i * 32 + a + 1) * 0x9E3779B97F4A7C15)
 *   slice = src_slc_ids[(h & 0xFFFF) % n_src]
 *   spatial slice: rv = (reach * v_scale) >> 4, ru = (reach * u_scale) >> 4; offsets
 *     enumerate the hex tile of gen_syn_offs (src/map/syn_src.rs:31-107): rows v = -rv..rv,
 *     u = trunc(max(-rv, -v - rv) * u_scale / v_scale) .. trunc(min(rv, -v + rv) * u_scale /
 *     v_scale); position ((h >> 16) & 0xFFFFFF) % tile size in row-major order;
 *     strength = (((rv - |v|) + (ru - |u|)) >> 3) * ((((h >> 40) & 0xFFFF) % 7) - 3)
 *   non-spatial slice (a scale of 0): v = ((h >> 16) & 0xFFFF) % (2 rv + 1) - rv with
 *     rv = v_size / 2, u likewise with u_size, strength 0
 *   the draw is kept if no other synapse of the dendrite has the same (slice, v, u)
 *   (SynSrcIdxCache::insert), else the next draw is tried; the 32nd draw is always kept
 *   (the reference's "source saturated" case).
 * Dendrites are independent; within a dendrite synapses are visited in order.
 * mix64 is the splitmix64 output function (Steele, Lea, Flood 2014). */
int orc_area_regrow(orc_area* a, uint64_t seed, uint64_t epoch, int32_t reach);

/* retina filter (cl/filters.cl:34-100): frame of `depth` slices of the dims of
 * axon slice slc, slice-major, into axon slices [slc, slc + depth). */
int orc_retina(orc_area* a, uint32_t slc, uint32_t depth, const uint8_t* frame);
/* gen_grp_centers (src/cortex/smoother.rs:33-58) with HexGroupCenters
 * (src/cmn/mod.rs:807-943): the six groups' centers, group-major, each group in
 * row-major order; *n = total (written only when it fits cap). */
int orc_smoother_centers(int32_t radius, int32_t v_size, int32_t u_size, int32_t* v, int32_t* u,
                         uint32_t* group_counts, uint32_t cap, uint32_t* n);
/* Attaches an ActivitySmoother of field radius `radius` to `layer` (0 removes);
 * orc_area_cycle then runs it after the layer's control kernel every cycle. */
int orc_set_smoother(orc_area* a, uint32_t layer, int32_t radius);
/* smooth_activity (cl/control.cl:227-295), one center at a time, group by group
 * in start order. */
int orc_k_smooth(orc_area* a, uint32_t layer);

#ifdef __cplusplus
}
#endif
#endif
