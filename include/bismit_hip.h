/*
 * bismit_hip.h -- C ABI of the MI355X cortical-cycle engine.
 *
 * This is the drop-in boundary that replaces bismit's OpenCL work queue (the
 * `ocl` crate's ProQue/Buffer/Kernel/Event use inside src/cortex, src/map and
 * src/flywheel.rs).  A Rust FFI crate would declare every function below as
 * `extern "C"`; this repository's C++ orchestrator, batch/serve loops and the
 * Python test harness (ctypes) call it directly.
 *
 * Conventions
 *   - Every function returns an int status: 0 = OK, non-zero = error.  The
 *     message of the last error on the calling thread is bsm_last_error().
 *     No exception or panic crosses the ABI.
 *   - The library owns all device memory.  Host pointers stay owned by the
 *     caller; a read/write is complete when the call returns.
 *   - Buffer layouts are the reference's (SoA, tuft-major synapse space):
 *       axon   (slice s, v, u)          -> axn_idz[s] + v*u_size[s] + u
 *       synapse(tuft t, cell c, den d, k) -> tft_syn_idz[t] + c*dens*spd + d*spd + k
 *       dendrite(tuft t, cell c, den d)   -> tft_den_idz[t] + c*dens + d
 *       cell-tuft(tuft t, cell c)         -> t*C + c,   c = slc*v*u + v_id*u_size + u_id
 *     (src/map/slice_map.rs:43-65, src/cortex/synapses.rs:816-844,
 *      src/cortex/dendrites.rs:485-508, src/cortex/tufts.rs:91).
 *   - One HIP stream per area (or per batch).  Calls on one area are not
 *     re-entrant; different areas may be driven from different host threads.
 */
#ifndef BISMIT_HIP_H
#define BISMIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSM_MAX_SLICES 255
#define BSM_MAX_LAYERS 8
#define BSM_MAX_TUFTS 3
#define BSM_MAX_TUFT_SRCS 8

/* Layer kinds (src/cortex/spiny_stellates.rs, src/cortex/pyramidals.rs). */
enum bsm_layer_kind { BSM_LAYER_SSC = 1, BSM_LAYER_PYR = 2 };
/* Dendrite class/kind of a tuft (src/map/scheme/cell.rs TuftScheme). */
enum bsm_den_class { BSM_DEN_BASAL = 0, BSM_DEN_APICAL = 1 };
enum bsm_den_kind { BSM_DEN_PROXIMAL = 0, BSM_DEN_DISTAL = 1 };
/* Control layer attached to a host layer (src/cortex/iinn.rs,
 * src/cortex/intra_column_inhib.rs, src/cortex/pyr_outputter.rs). */
enum bsm_control_kind {
    BSM_CTRL_NONE = 0,
    BSM_CTRL_INHIB_SIMPLE = 1,   /* basket-cell hex WTA (passthrough when bypassed) */
    BSM_CTRL_INTRA_COLUMN = 2,
    BSM_CTRL_PYR_OUTPUT = 3
};
/* CorticalAreaSettings switches (src/cortex/cortical_area.rs:253-338).
 * Regrowth is not a cycle setting here: the caller schedules it with
 * bsm_area_regrow / bsm_batch_regrow (every 201 cycles upstream). */
enum bsm_settings {
    BSM_SET_BYPASS_INHIB = 1u << 0,
    BSM_SET_DISABLE_PYRS = 1u << 1,
    BSM_SET_DISABLE_SSCS = 1u << 2,
    BSM_SET_DISABLE_LEARNING = 1u << 3
};

/* One axon slice: the runtime form of the -D slice tables baked into every
 * reference program (cl/bismit.cl:108-114, src/map/area_map.rs:51-77).
 * axn_idz is the prefix sum of v_size*u_size in slice order. */
typedef struct bsm_slice_desc {
    uint32_t v_size;
    uint32_t u_size;
    uint8_t v_scale;  /* 16 = 1:1, 0 = non-spatial (src/cmn/slice_dims.rs:200-223) */
    uint8_t u_scale;
    uint8_t v_mid;    /* non-spatial slices: size/2 */
    uint8_t u_mid;
} bsm_slice_desc;

typedef struct bsm_tuft_desc {
    uint8_t den_class;          /* enum bsm_den_class */
    uint8_t den_kind;           /* enum bsm_den_kind */
    uint8_t max_active_dens_l2; /* tft_cycle scaling (cl/bismit.cl:825) */
    uint8_t n_src_slcs;         /* source slices (layout generation, LDS staging) */
    uint8_t src_slc_ids[BSM_MAX_TUFT_SRCS];
    uint32_t dens_per_tft;
    uint32_t syns_per_den;
    uint32_t threshold;         /* den_cycle_tft den_threshold (src/cortex/dendrites.rs:102) */
} bsm_tuft_desc;

typedef struct bsm_layer_desc {
    uint8_t kind;          /* enum bsm_layer_kind */
    uint8_t spatial;       /* LayerTags::SPATIAL role (cycled+learned before the others) */
    uint8_t depth;         /* slices of this layer */
    uint8_t base_axn_slc;  /* first axon slice of this layer */
    uint8_t n_tufts;       /* SSC: exactly 1; PYR: 1..3 */
    uint8_t control;       /* enum bsm_control_kind */
    uint8_t _pad[2];
    int32_t inhib_radius;  /* BSM_CTRL_INHIB_SIMPLE field radius */
    bsm_tuft_desc tufts[BSM_MAX_TUFTS];
} bsm_layer_desc;

typedef struct bsm_area_desc {
    uint32_t v_size;       /* area (cell layer) dimensions */
    uint32_t u_size;
    uint32_t n_slices;
    uint32_t n_layers;
    uint32_t settings;     /* enum bsm_settings */
    uint32_t _pad;
    const bsm_slice_desc* slices;
    const bsm_layer_desc* layers;   /* in the area's cycle order */
} bsm_area_desc;

/* Buffers, mirroring SamplerKind (src/cortex/cortical_area.rs:61-90). */
enum bsm_buf {
    BSM_AXN_STATES = 0,             /* whole axon space (layer ignored) */
    BSM_SOMA_STATES,                /* PYR: soma u8[C]; SSC: dens.states (aliased) */
    BSM_SOMA_ENERGIES,
    BSM_SOMA_ACTIVITIES,
    BSM_SOMA_FLAG_SETS,             /* PYR only */
    BSM_TFT_STATES,
    BSM_TFT_BEST_DEN_IDS,
    BSM_TFT_BEST_DEN_STATES_RAW,
    BSM_TFT_BEST_DEN_STATES,
    BSM_TFT_PREV_STATES,
    BSM_TFT_PREV_BEST_DEN_IDS,
    BSM_TFT_PREV_BEST_DEN_STATES_RAW,
    BSM_TFT_PREV_BEST_DEN_STATES,
    BSM_DEN_STATES,
    BSM_DEN_STATES_RAW,
    BSM_DEN_ENERGIES,
    BSM_DEN_ACTIVITIES,
    BSM_DEN_THRESHOLDS,
    BSM_SYN_STATES,
    BSM_SYN_STRENGTHS,              /* i8 */
    BSM_SYN_SRC_SLC_IDS,
    BSM_SYN_SRC_COL_V_OFFS,         /* i8 */
    BSM_SYN_SRC_COL_U_OFFS,         /* i8 */
    BSM_SYN_FLAG_SETS,
    BSM_BUF_COUNT
};

/* rnd slot kinds: one int32 per launch per cycle (src/cortex/dendrites.rs:198,
 * src/cortex/iinn.rs:113, src/cortex/spiny_stellates.rs:259, src/cortex/tufts.rs:312).
 * Slot order: for each layer in descriptor order: DEN[t] for each tuft, CTRL,
 * MTP[t] for each tuft (always allocated, used or not). */
enum bsm_rnd_kind { BSM_RND_DEN = 0, BSM_RND_CTRL = 1, BSM_RND_MTP = 2 };

typedef struct bsm_ctx bsm_ctx;
typedef struct bsm_area bsm_area;
typedef struct bsm_batch bsm_batch;
typedef struct bsm_sampler bsm_sampler;

/* ---- context (replaces ProQue/Context construction, src/cortex/cortex.rs:51-90) ---- */
const char* bsm_last_error(void);
int bsm_ctx_create(int device, bsm_ctx** out);
int bsm_ctx_destroy(bsm_ctx* ctx);
int bsm_sync(bsm_ctx* ctx);
int bsm_device_info(bsm_ctx* ctx, char* name_buf, size_t name_len, int* cu_count,
                    size_t* total_mem);

/* ---- area (replaces CorticalArea::new buffer/kernel setup,
 *      src/cortex/cortical_area.rs:376-642) ---- */
int bsm_area_create(bsm_ctx* ctx, const bsm_area_desc* desc, uint64_t rnd_seed,
                    bsm_area** out);
int bsm_area_destroy(bsm_area* area);
int bsm_area_sync(bsm_area* area);
/* Bytes of device memory the area holds (admission by HBM footprint). */
int bsm_area_footprint(const bsm_area_desc* desc, size_t* bytes);
int bsm_area_axn_len(bsm_area* area, uint32_t* len);
int bsm_area_syn_count(bsm_area* area, uint64_t* n);
int bsm_area_slice_idz(bsm_area* area, uint32_t slc, uint32_t* idz);

/* ---- buffer I/O (replaces Buffer read/write/fill and the test traits,
 *      src/cortex/synapses.rs:536-549, src/cortex/axon_space.rs:650-668).
 *      On an area inside a batch these first wait for the batch's enqueued
 *      cycles (bsm_batch_sync), so they see the state after the last cycle
 *      enqueued, static or masked. ---- */
int bsm_area_buf_len(bsm_area* area, int buf, uint32_t layer, size_t* len);
int bsm_area_read(bsm_area* area, int buf, uint32_t layer, size_t offset, size_t len,
                  void* host_dst);
int bsm_area_write(bsm_area* area, int buf, uint32_t layer, size_t offset, size_t len,
                   const void* host_src);
int bsm_area_fill(bsm_area* area, int buf, uint32_t layer, uint8_t val);

/* ---- axon intake/output (replaces AxonSpace::intake/output map+memcpy,
 *      src/cortex/axon_space.rs:509-629).  src_is_device != 0: src is a device
 *      pointer and the copy is enqueued on the area stream (async). ---- */
int bsm_area_intake(bsm_area* area, uint32_t axn_offset, const uint8_t* src, uint32_t len,
                    int src_is_device);
int bsm_area_output(bsm_area* area, uint32_t axn_offset, uint8_t* host_dst, uint32_t len);

/* ---- sensory filter and activity smoother ---- */
enum bsm_filter_kind { BSM_FILTER_RETINA = 1 };
/* Intake through a sensory filter, replacing SensoryFilter::write + cycle with
 * the `retina` filter (src/cortex/sensory_filter.rs:24-174, cl/filters.cl:34-100)
 * on a filter chain of one: the frame (depth slices of the dims of axon slice
 * slc, slice-major) is filtered into axon slices [slc, slc + depth), which must
 * share their dimensions.  Enqueued on the area stream (a host frame is first
 * copied to a device staging buffer). */
int bsm_area_intake_filtered(bsm_area* area, int filter, uint32_t slc, uint32_t depth,
                             const uint8_t* src, int src_is_device);
/* bsm_area_intake_filtered from a host frame without allocating: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops).  The frame is staged in a
 * static device array of BSM_FILTER_STATIC_MAX bytes per device (held until the
 * process ends) instead of the area's lazily allocated staging buffer; larger
 * frames are refused with an error.  Completes on return.  Calls that overlap
 * on one device (from several host threads) race on the shared staging array. */
#define BSM_FILTER_STATIC_MAX (4u << 20)
int bsm_area_intake_filtered_noalloc(bsm_area* area, int filter, uint32_t slc, uint32_t depth,
                                     const uint8_t* src);
/* The retina kernel alone on a device frame (cl/filters.cl:34-100). */
int bsm_k_retina(bsm_area* area, uint32_t slc, uint32_t depth, const uint8_t* dev_src);
/* Attaches an ActivitySmoother control layer with field radius `radius` to
 * `layer` (0 removes it), replacing src/cortex/smoother.rs:62-180 and
 * cl/control.cl:227-295.  It runs after the layer's inhibition/learning stage in
 * every cycle of the plain form -- as the reference does, whose cycle counter
 * never advances (src/cortex/smoother.rs:162) -- reading the layer's soma
 * activities and adjusting its soma energies.  The six center groups apply in
 * start order (the reference launches them together and races where their
 * fields meet).  Areas with a smoother cycle in the plain form only. */
int bsm_area_set_smoother(bsm_area* area, uint32_t layer, int32_t radius);
/* The smoother kernel alone (cl/control.cl:227-295, groups in order). */
int bsm_k_smooth(bsm_area* area, uint32_t layer);
/* Centers of the smoother's six hex groups for a v_size x u_size layer
 * (gen_grp_centers + HexGroupCenters::populate, src/cortex/smoother.rs:33-58,
 * src/cmn/mod.rs:807-943), group-major, row-major within a group; *n is the
 * total, written only when it fits cap.  Host-only. */
int bsm_smoother_centers(int32_t radius, uint32_t v_size, uint32_t u_size, int32_t* v,
                         int32_t* u, uint32_t* group_counts, uint32_t cap, uint32_t* n);

/* ---- samplers: per-cycle readback (replaces CorticalArea::sampler and
 *      cycle_samplers, src/cortex/cortical_area.rs:808-1284, with the single-buffer
 *      tract channel of src/subcortex/tract_channel.rs:369-480) ----
 * A sampler reads bytes [offset, offset + len) of one buffer kind of a layer
 * (layer ignored for BSM_AXN_STATES) at the end of every cycle the area steps,
 * plain or batched.  Channel semantics (one held frame, the default): a cycle
 * sends a frame when the previous one has been received; otherwise it waits for the receiver when backpressure
 * is on and skips the frame when it is off, so a receiver always gets the frame
 * of the first cycle after its last receive.  Frames are gathered on the device
 * into pinned host slots (one kernel and one copy per cycle for all samplers of
 * an area or batch), ordered after the cycle on its stream.
 * Create and destroy on the thread that drives the area; recv and backpressure
 * calls may come from any thread.  Destroying an area closes its samplers (recv
 * then reports no frame; the handle still needs bsm_sampler_destroy). */
int bsm_sampler_create(bsm_area* area, int buf, uint32_t layer, uint64_t offset, uint64_t len,
                       int backpressure, bsm_sampler** out);
/* Takes the pending frame into dst (dst_len >= the sampler length): *got = 1 and
 * *cycle = the area's cycle count at that frame (1 = the first cycle after the
 * area was created).  No frame pending: waits for the next when wait_for_frame
 * is nonzero (returns *got = 0 if the sampler closes meanwhile), else *got = 0. */
int bsm_sampler_recv(bsm_sampler* s, int wait_for_frame, void* dst, uint64_t dst_len,
                     uint64_t* cycle, int* got);
/* bsm_sampler_recv over n samplers in order (one call per cycle for a server
 * that hands every session its output); cycles may be null. */
int bsm_sampler_recv_many(bsm_sampler* const* s, uint32_t n, int wait_for_frame,
                          void* const* dst, const uint64_t* dst_len, uint64_t* cycles, int* got);
/* Sets backpressure and returns the prior state in *prior (may be null). */
int bsm_sampler_set_backpressure(bsm_sampler* s, int backpressure, int* prior);
/* Frames the sampler holds at most, 1 to 3: SamplerBufferKind Single, Double and
 * Triple (src/cortex/cortical_area.rs:52-58; the reference implements Single
 * only, :1041-1048, and its channel buffer kinds Double and Triple are
 * unimplemented, src/subcortex/tract_channel.rs:259-290).  A cycle sends a
 * frame while fewer are held, else skips it or, with backpressure, waits; recv
 * takes the oldest held frame, so frames arrive in cycle order.  Default 1. */
int bsm_sampler_set_buffers(bsm_sampler* s, uint32_t n_frames);
int bsm_sampler_info(bsm_sampler* s, uint64_t* len, int* backpressure, int* pending);
int bsm_sampler_destroy(bsm_sampler* s);

/* ---- per-kernel entry points, reference argument semantics (parity units) ---- */
int bsm_k_syn_flags(bsm_area* a, uint32_t layer);                        /* cl/syns.cl:23-33 */
int bsm_k_syn_cycle(bsm_area* a, uint32_t layer, uint32_t tuft);         /* cl/syns.cl:285-388 */
int bsm_k_den_cycle(bsm_area* a, uint32_t layer, uint32_t tuft, int32_t rnd); /* cl/bismit.cl:695-747 */
int bsm_k_tft_cycle(bsm_area* a, uint32_t layer, uint32_t tuft);         /* cl/bismit.cl:755-842 */
int bsm_k_ssc_cycle(bsm_area* a, uint32_t layer);                        /* cl/bismit.cl:849-884 */
int bsm_k_inhib_simple(bsm_area* a, uint32_t layer, int32_t rnd);        /* cl/control.cl:28-191 */
int bsm_k_inhib_passthrough(bsm_area* a, uint32_t layer, int32_t rnd);   /* cl/control.cl:194-217 */
int bsm_k_ssc_mtp(bsm_area* a, uint32_t layer, uint32_t rnd);            /* cl/bismit.cl:888-915 */
int bsm_k_pyr_cycle(bsm_area* a, uint32_t layer);                        /* cl/bismit.cl:961-1022 */
int bsm_k_inhib_intra_column(bsm_area* a, uint32_t layer);               /* cl/control.cl:319-351 */
int bsm_k_pyr_output(bsm_area* a, uint32_t layer);                       /* cl/control.cl:298-314 */
int bsm_k_tft_dst_mtp(bsm_area* a, uint32_t layer, uint32_t tuft, int32_t rnd); /* cl/bismit.cl:1195-1268 */

/* ---- whole cycle (replaces CorticalArea::cycle's enqueue sequence,
 *      src/cortex/cortical_area.rs:716-789) ---- */
int bsm_area_rnd_slots(bsm_area* area, uint32_t* n_slots);
int bsm_area_rnd_slot_info(bsm_area* area, uint32_t slot, int* kind, uint32_t* layer,
                           uint32_t* tuft);
/* rnd: n_cycles*n_slots int32 draws (cycle-major) or NULL for the area's own
 * xorshift128 stream (seeded by splitmix64(rnd_seed) at bsm_area_create).
 * The intake for each cycle must be enqueued before the call (or use
 * bsm_area_run). */
int bsm_area_cycle(bsm_area* area, const int32_t* rnd, uint32_t n_cycles);
/* n_cycles of intake + cycle: frame i (len bytes at dev_frames + i*frame_stride,
 * a device pointer) is copied to axn_offset before cycle i. */
int bsm_area_run(bsm_area* area, uint32_t axn_offset, const uint8_t* dev_frames,
                 uint64_t frame_stride, uint32_t frame_len, uint32_t n_cycles);
/* Cortex::cycle (src/cortex/cortex.rs:100-120) over plain-form areas, driven from
 * the host: for each of n_cycles cycles, each area in turn takes its frame
 * (len[i] bytes at dev_frames[i] + (c % n_frames) * stride[i], device memory,
 * copied to axn_offset[i]; no intake where dev_frames or dev_frames[i] is NULL)
 * and one bsm_area_cycle on its own stream.  Returns once enqueued. */
int bsm_cortex_cycle(bsm_area* const* areas, uint32_t n_areas, const uint8_t* const* dev_frames,
                     const uint64_t* stride, const uint32_t* len, const uint32_t* axn_offset,
                     uint32_t n_frames, uint32_t n_cycles);
/* Draw n_cycles*n_slots values from the area's own stream (host side) without
 * running anything; advances the stream exactly as bsm_area_cycle(NULL) would. */
int bsm_area_draw_rnd(bsm_area* area, int32_t* out, uint32_t n_cycles);

/* Tile parameters of the fused kernels (autotuner hook). 0 = default. */
typedef struct bsm_tuning {
    uint32_t tile_v;     /* columns per workgroup tile, v */
    uint32_t tile_u;     /* columns per workgroup tile, u */
    uint32_t halo;       /* LDS halo (columns of source reach staged) */
    uint32_t chunk;      /* synapses per lane chunk: 4, 8 or 16 */
    uint32_t inhib_tile; /* side of the square inhibition tile: 8, 16 or 32;
                            1 = the planner's choice (the default) */
    uint32_t overlap;    /* 1: pyramidal tufts whose sources no earlier stage of the
                            cycle writes run on a second stream beside the spatial
                            layers' stages (the default); 2: one stream */
} bsm_tuning;
int bsm_area_set_tuning(bsm_area* area, const bsm_tuning* t);
int bsm_area_get_tuning(bsm_area* area, bsm_tuning* t);

/* ---- device memory helpers for resident inputs ---- */
int bsm_dev_alloc(bsm_ctx* ctx, size_t bytes, void** dptr);
int bsm_dev_free(bsm_ctx* ctx, void* dptr);
int bsm_dev_upload(bsm_ctx* ctx, void* dptr, const void* host_src, size_t bytes);

/* ---- timing (HIP events on the area/batch stream) ---- */
int bsm_area_time_cycles(bsm_area* area, uint32_t axn_offset, const uint8_t* dev_frames,
                         uint64_t frame_stride, uint32_t frame_len, uint32_t n_cycles,
                         float* ms_total);

/* Per-stage device time of n_cycles plain cycles (HIP events around every fused
 * launch on the area stream; no intake).  stage_ms[i] is the total over the
 * cycles; stage_kind[i] is 1 (gather+dendrites+tufts), 2 (SSC inhibition+LTP)
 * or 3 (pyramidal soma+column inhibition+LTP); stage_layer[i] its layer. */
int bsm_area_profile_stages(bsm_area* area, uint32_t n_cycles, uint32_t max_stages,
                            float* stage_ms, int* stage_kind, uint32_t* stage_layer,
                            uint32_t* n_stages);
/* Per-stage device time of n_cycles cycles of a static batch, enqueued without
 * graph capture (HIP events around every persistent launch on the batch
 * stream).  stage_ms[i] is the total over the cycles, stage_kind[i] as for
 * bsm_area_profile_stages.  The cycles advance the areas like bsm_batch_cycle. */
int bsm_batch_profile_stages(bsm_batch* batch, uint32_t n_cycles, uint32_t max_stages,
                             float* stage_ms, int* stage_kind, uint32_t* n_stages);
/* The fused launch plan (planned now if stale): one entry per persistent launch
 * of a cycle, in launch order.  Lets callers and tests see which kernel
 * variant runs (chunk width, check-free gather) and its tiling. */
typedef struct bsm_stage_info {
    int32_t kind;        /* 1 gather+dendrites+tufts, 2 SSC post, 3 pyramidal post,
                            4 activity smoother */
    uint32_t layer;      /* plain form: the layer; batches: 0 */
    uint32_t chunk;      /* synapses per lane chunk (W) */
    uint32_t staged;     /* F1: staged-slice template */
    uint32_t fast;       /* F1: check-free LDS gather */
    uint32_t tile_v, tile_u;
    uint32_t n_items;    /* work items (tiles) of the launch */
    uint32_t lds_bytes;
    uint32_t grid;       /* workgroups launched */
    uint32_t stream;     /* 0: the area/batch stream; 1: the side stream (runs beside
                            the spatial layers' stages, joined before the layer's
                            pyramidal post stage) */
    uint64_t alg_bytes;  /* gather+dendrites+tufts stages: HBM bytes this launch must
                            move: per synapse the source slice (a 2-bit staged-slice code
                            for tufts whose gather stages several slices, the 1 B id in
                            a checked gather), the offsets
                            (1 B packed, else 2 B), the strength and the new state; 4 B
                            per dendrite, 12 B per pyramidal cell-tuft, 1 B per SSC
                            cell; 0 for other stages */
    uint32_t packed;     /* F1: the gather reads the packed offsets */
    uint32_t _reserved;
} bsm_stage_info;
int bsm_area_stage_plan(bsm_area* area, uint32_t max_stages, bsm_stage_info* out,
                        uint32_t* n_stages);
int bsm_batch_stage_plan(bsm_batch* batch, uint32_t max_stages, bsm_stage_info* out,
                         uint32_t* n_stages);
/* Plain device-to-device copy kernel of `bytes` (multiple of 16): the known-size
 * reference for scaling HBM traffic counters. */
int bsm_calib_copy(bsm_ctx* ctx, void* dst, const void* src, size_t bytes, float* ms);
/* The fused synapse stage's HBM pattern without its gather: src holds n_reads
 * (1-4) streams of `bytes` (multiple of 16) back to back, read with
 * non-temporal 16-byte loads; dst receives one stream of `bytes`.  The practical
 * bandwidth ceiling for that stage's read/write mix.  dst null: the read
 * streams alone, nothing written (the check-free gather's pattern). */
int bsm_calib_stream(bsm_ctx* ctx, void* dst, const void* src, size_t bytes, uint32_t n_reads,
                     float* ms);

/* ---- host-side synapse layout generator (replaces Synapses::grow(init=true),
 *      src/cortex/synapses.rs:454-501, entropy-seeded upstream): src slice
 *      uniform over the tuft's sources, v/u offsets uniform in [-reach, reach],
 *      from splitmix64(seed, layer, tuft, index). ---- */
int bsm_gen_tuft_layout(uint64_t seed, uint32_t layer, uint32_t tuft, uint64_t n_syns,
                        uint32_t n_src, const uint8_t* src_ids, int32_t reach,
                        uint8_t* out_slc, int8_t* out_v, int8_t* out_u);
/* Fill every tuft of every layer of the area with bsm_gen_tuft_layout and
 * zero strengths (the same generator, run on the device). */
int bsm_area_init_layout(bsm_area* area, uint64_t seed, int32_t reach);
/* Deterministic regrowth, replacing CorticalArea::regrow (src/cortex/cortical_area.rs:
 * 792-806) -> Synapses::grow / regrow_syn (src/cortex/synapses.rs:411-501) and
 * SynSrcSlices::gen_src (src/map/syn_src.rs:448-490): every synapse with strength
 * <= -25 (SYNAPSE_STRENGTH_FLOOR) takes a new source and strength.  The reference draws
 * them from an entropy-seeded generator on the host; here they come from a
 * counter-based hash of (seed, epoch, layer, tuft, synapse, attempt), drawn on the
 * device: a uniform source slice of the tuft, a uniform offset in the hex tile of
 * radius (reach * scale) >> 4 (non-spatial slices: half the slice size), strength
 * ((rv - |v|) + (ru - |u|)) >> 3 times a uniform draw in -3..3 (0 when non-spatial),
 * redrawn up to 31 times while another synapse of the dendrite has the same source.
 * The reference calls it every SYNAPSE_REGROWTH_INTERVAL + 1 = 201 cycles unless
 * regrowth is disabled; the caller decides when (epoch numbers the calls).
 * Completes on return.  Areas inside a batch regrow through bsm_batch_regrow. */
int bsm_area_regrow(bsm_area* area, uint64_t seed, uint64_t epoch, int32_t reach);

/* ---- synthetic input streams of the workloads (host only; the reference's
 *      encoders and input generators, src/encode/, src/subcortex/
 *      input_generator.rs, are not part of this engine).  An SDR stream over
 *      n_axons keeps exactly k = round(density * n_axons) axons at 255; each
 *      next frame keeps round(overlap * k) of the previous active axons (a
 *      uniform subset) and draws the rest uniformly from outside the previous
 *      set, from xorshift128 seeded by splitmix64(seed). ---- */
typedef struct bsm_sdr bsm_sdr;
int bsm_sdr_create(uint32_t n_axons, double density, double overlap, uint64_t seed,
                   bsm_sdr** out);
int bsm_sdr_destroy(bsm_sdr* s);
int bsm_sdr_info(bsm_sdr* s, uint32_t* n_axons, uint32_t* k);
/* Writes the current frame (n_axons bytes) and advances the stream. */
int bsm_sdr_next(bsm_sdr* s, uint8_t* frame);
int bsm_sdr_frames(bsm_sdr* s, uint32_t n_frames, uint8_t* frames);
/* The next frame of each of n streams, stream i's at buf + ofs[i]. */
int bsm_sdr_next_packed(bsm_sdr* const* s, uint32_t n, uint8_t* buf, const uint64_t* ofs);
/* The current frame as its k active axon indices (in draw order), then advance. */
int bsm_sdr_next_indices(bsm_sdr* s, uint32_t* idx);
/* The next frame of each of n streams as an SDR entry of bsm_batch_cycle_masked
 * (uint32 k, then k indices) at buf + ofs[i] (4-byte aligned). */
int bsm_sdr_next_packed_sdr(bsm_sdr* const* s, uint32_t n, uint8_t* buf, const uint64_t* ofs);

/* ---- deterministic rnd schedule helpers ---- */
uint64_t bsm_splitmix64(uint64_t* state);
/* xorshift128 (Marsaglia 2003) stream seeded from splitmix64(seed). */
int bsm_xorshift128_fill(uint64_t seed, uint32_t* out, uint64_t n);

/* ---- batched form: many areas per stage in one persistent launch ---- */
int bsm_batch_create(bsm_ctx* ctx, bsm_area* const* areas, uint32_t n_areas, bsm_batch** out);
/* flags: BSM_BATCH_DYNAMIC -- each cycle steps only the areas flagged active in
 * bsm_batch_cycle_masked (continuous batching for serving); every area keeps its
 * own state parity, rnd stream and input frame. */
enum bsm_batch_flags { BSM_BATCH_DYNAMIC = 1u };
int bsm_batch_create_ex(bsm_ctx* ctx, bsm_area* const* areas, uint32_t n_areas, uint32_t flags,
                        bsm_batch** out);
/* Dynamic batches: areas join and leave between cycles (continuous batching).
 * Both wait for the batch's enqueued cycles; the next cycle re-plans.  Joining
 * areas bring their input ranges; leaving areas continue in the plain form with
 * their state as the batch left it.  A dynamic batch may be created empty. */
int bsm_batch_add_areas(bsm_batch* b, bsm_area* const* areas, uint32_t n,
                        const uint32_t* axn_offset, const uint32_t* len);
int bsm_batch_remove_areas(bsm_batch* b, bsm_area* const* areas, uint32_t n);
/* Dynamic batches: axon range each area's frame is copied to. */
int bsm_batch_set_input_ranges(bsm_batch* b, const uint32_t* axn_offset, const uint32_t* len);
/* Dynamic batches: one shared cycle of the areas with active[i] != 0.  frames is a
 * host buffer of frame_bytes; area i's frame starts at frame_ofs[i] (UINT64_MAX:
 * no intake).  An offset with BSM_FRAME_SDR set names an SDR entry instead of a
 * dense frame: at (offset & ~BSM_FRAME_SDR), 4-byte aligned, a uint32 count n and
 * n uint32 axon indices relative to the area's input range (each below its
 * length); the input range becomes 0 except those axons, which become 255.
 * Returns after enqueueing (asynchronous); bsm_batch_query polls. */
#define BSM_FRAME_SDR (1ull << 63)
int bsm_batch_cycle_masked(bsm_batch* b, const uint8_t* active, const uint8_t* frames,
                           const uint64_t* frame_ofs, uint64_t frame_bytes, int use_graph);
int bsm_batch_query(bsm_batch* b, int* done);
/* Dynamic batches: device time (HIP events on the batch stream) of every masked
 * cycle so far, from its frame upload to its samplers' gather; waits for the
 * last two. */
int bsm_batch_busy_ms(bsm_batch* b, double* ms);
/* Regrowth of batched areas (the batch's form of bsm_area_regrow, same draws):
 * area i regrows with seeds[i] and epochs[i] when which is null or which[i] is
 * nonzero, so each area of a dynamic batch keeps its own 201-cycle schedule
 * (src/cortex/cortical_area.rs:792-806).  Enqueued on the batch stream after
 * its enqueued cycles and before the next (buffer I/O on the areas waits for
 * it); a widened source summary re-plans the next cycle. */
int bsm_batch_regrow(bsm_batch* b, const uint8_t* which, const uint64_t* seeds,
                     const uint64_t* epochs, int32_t reach);
/* Inter-area tracts (thalamic routing, src/subcortex/thalamus/thalamus.rs:71-285,
 * src/cortex/axon_space.rs:509-629): at the start of every batch cycle, before
 * any area cycles, each tract copies len axon states of area src_area (as the
 * previous cycle left them) to area dst_area.  Indices refer to the batch's
 * area order.  Static batches only. */
typedef struct bsm_tract {
    uint32_t src_area, src_offset, dst_area, dst_offset, len;
} bsm_tract;
int bsm_batch_set_tracts(bsm_batch* b, const bsm_tract* tracts, uint32_t n);
/* Plain form of one tract: copy src's axon range into dst's, after everything
 * enqueued on src; complete when the call returns. */
int bsm_area_copy_axons(bsm_area* dst, uint32_t dst_offset, bsm_area* src, uint32_t src_offset,
                        uint32_t len);
int bsm_batch_destroy(bsm_batch* b);
/* use_graph != 0: each cycle replays a captured hipGraph. */
int bsm_batch_cycle(bsm_batch* b, uint32_t n_cycles, int use_graph);
/* Per-area input frames: frame i of area a comes from dev_frames[a] + i*stride[a]
 * (len[a] bytes, to axn_offset[a]).  NULL dev_frames: no intake. */
int bsm_batch_set_inputs(bsm_batch* b, const uint8_t* const* dev_frames,
                         const uint64_t* stride, const uint32_t* len,
                         const uint32_t* axn_offset, uint32_t n_frames);
int bsm_batch_sync(bsm_batch* b);
int bsm_batch_time_cycles(bsm_batch* b, uint32_t n_cycles, int use_graph, float* ms_total);
int bsm_batch_set_tuning(bsm_batch* b, const bsm_tuning* t);

#ifdef __cplusplus
}
#endif
#endif /* BISMIT_HIP_H */
