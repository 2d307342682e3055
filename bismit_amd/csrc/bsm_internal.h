// bsm_internal.h -- host-side engine structures shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/bismit_hip.h"
#include "bsm_device.h"

namespace bsm {

void set_error(const std::string& msg);
int fail(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define BSM_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::bsm::hip_fail(e_, #call); \
    } while (0)

// ---------------------------------------------------------------- virtual states
// Synapse states of a check-free tuft are virtual: a state is syn_fire of the
// source axon the gather read, so the gather keeps a snapshot of the staged
// slices (contiguous in the axon space: consecutive ids of one geometry)
// instead of writing one state byte per synapse.  Parity q's snapshot sits at
// snap + q * snap_len; bit q of *vbits says that st[q] is stale and parity q's
// states come from that snapshot.  Every reader recomputes them (the learning
// stages, buffer reads, samplers); every writer of states or sources first
// materialises both parities and clears the bits (settle_states).
struct VirtTuft {
    const uint8_t* snap;
    uint8_t* st[2];           // tuft base in the two state buffers
    const uint8_t* slc;       // tuft base of the source arrays (authoritative)
    const int8_t* vofs;
    const int8_t* uofs;
    uint32_t* vbits;
    uint64_t n;               // synapses of the tuft
    uint32_t snap_len;        // bytes per parity: n_staged * slice_len
    uint32_t slice_len;       // vs * us
    uint32_t base;            // first staged slice id
    uint32_t spt, VU, U;      // synapses per cell; layer columns
    uint16_t vs, us;          // staged slices' geometry
    uint8_t vsc, usc, vmid, umid;
    // distal tufts: per parity and cell, the fired bits of the cell's best
    // dendrite as the gather found them (bit i = synapse i of the dendrite),
    // valid for parity q while bit 2 + q of *vbits is set
    const uint32_t* bmask;    // parity q at bmask + q * bm_len
    uint32_t bm_len, wpd;     // words per parity, words per dendrite
};

// The source axon of one synapse of cell `cel` (layer index) in a snapshot:
// the staged window of the gather, zero off the slice (coord_is_safe).
__device__ __forceinline__ uint32_t virt_axon(const VirtTuft& T, const uint8_t* snap,
                                              uint32_t cel, uint32_t sid, int32_t vo,
                                              int32_t uo) {
    const uint32_t vu = cel % T.VU, v = vu / T.U, u = vu - v * T.U;
    const int32_t gv = (int32_t)((v * T.vsc) >> 4) + (int32_t)T.vmid + vo;
    const int32_t gu = (int32_t)((u * T.usc) >> 4) + (int32_t)T.umid + uo;
    const uint32_t code = sid - T.base;
    if ((uint32_t)gv >= T.vs || (uint32_t)gu >= T.us || code * T.slice_len >= T.snap_len)
        return 0u;
    return BSM_G(const uint8_t, snap)[code * T.slice_len + (uint32_t)gv * T.us + (uint32_t)gu];
}

// ---------------------------------------------------------------- jobs
// F1 work unit: one tuft of one layer of one area (gather + dendrites [+ tufts]).
struct TuftJob {
    const uint8_t* axn;
    const SliceRt* slices;
    const uint8_t* slc;
    const int8_t* vofs;
    const int8_t* uofs;
    const uint8_t* vu;     // packed offsets: (v + 8) << 4 | (u + 8), valid when |v|, |u| <= 7
    const uint8_t* sb;     // staged-slice codes: (slice id - staged[0]) & 3, 2 bits per
                           // synapse in 32-bit words of 16 synapses (synapse 4w + k of a
                           // word's group at bit 8k + 2w); valid when fast
    const int8_t* str;
    uint8_t* st[2];        // tuft base in the two state buffers; st[par] is written
    uint8_t* den_raw;      // tuft dendrite base
    uint8_t* den_st;
    uint8_t* den_act;
    uint8_t* ssc_energy;   // SSC layers: cell energies (ssc_cycle fused on dens < C)
    uint8_t* tft[8];       // cell-tuft arrays (offset t*C) or null: best_ids, best_raw,
                           // best, states, prev_best_ids, prev_best_raw, prev_best, prev_states
    const int32_t* rnd;    // per-area rnd array
    uint32_t rnd_slot;
    uint32_t n_slices;
    uint32_t V, U, depth, C;
    uint32_t dens, spd, spt, thresh;
    uint32_t den_idz;
    uint32_t l2;
    uint32_t n_staged;
    uint32_t staged[4];
    SliceRt st0;           // geometry of staged[0] (the check-free gather's windows),
                           // copied here so that a tile reads it with the job
    uint32_t scalar_quirk; // syns_per_tft % 4 != 0: char-wrapped mids (scalar kernel)
    uint32_t halo;         // staged reach around the tile, in source-slice columns
    uint32_t par_xor;      // state buffer written = st[(par ^ par_xor) & 1]
    uint32_t fast;         // every synapse sources a staged slice within `halo`
    uint32_t packed;       // fast, and every |offset| <= 7: the gather reads `vu`
    const uint32_t* dyn;   // dynamic batches: per (area, layer) word, bit0 parity, bit1 active
    // virtual states (fast jobs with virt set): snapshot of the staged slices,
    // parity bits; a fast job without virt writes its states
    uint32_t virt;
    uint32_t* bmask;       // distal tufts: best-dendrite fired bits (see VirtTuft), or null
    uint32_t bm_len, wpd;
    // spiny stellate tufts with virtual states: the fired bit of every synapse of
    // the cycle (bit i = tuft synapse i), which the layer's learning reads; null
    // for other tufts
    uint32_t* fb;
    uint8_t* snap;         // parity q at snap + q * snap_len
    uint32_t snap_len;
    uint32_t* vbits;
    // strength signs, one bit per synapse of the layer (bit sg_bit0 + i for
    // synapse i of the tuft), read by the fast gather instead of the strengths
    const uint8_t* sg;
    uint32_t sg_bit0, sg_len;
    uint32_t ntu, n_tiles; // the job's tiles per row and in all (set by the F1 planner)
    uint32_t snap_per;     // snapshot bytes each tile copies (the last tile: the rest)
    // host-side planning copies of the staged slices' geometry
    uint16_t st_vsize[4], st_usize[4];
    uint8_t st_vsc[4], st_usc[4];
    uint32_t tile_v, tile_u;  // the job's column tile (set by the F1 planner)
};

// F2 work unit: SSC layer post-processing (inhibition + SSC LTP).
struct SscJob {
    const uint8_t* soma;   // dens.states of the layer (soma alias)
    uint8_t* axn;
    const SliceRt* slices;
    uint8_t* cel_act;
    uint8_t* st[2];        // layer synapse state buffers
    int8_t* str;
    const int32_t* rnd;
    uint32_t slot_ctrl, slot_mtp;
    uint32_t V, U, depth, C;
    uint32_t spt, base_slc, radius;
    // where this cycle's states of tuft 0 are: bit p of *vbits set -> virtual,
    // fired bits in fb (see TuftJob::fb); else the state buffer st[p]
    const uint32_t* vbits;
    const uint32_t* fb;
    uint32_t mode;         // 0 inhib_simple, 1 passthrough(rnd 0), 2 none
    uint32_t learn;
    uint32_t par_xor;
    const uint32_t* dyn;
    uint32_t* sg;          // the layer's strength signs (32 per word)
};

// F3 work unit: pyramidal layer post-processing (soma + column control + distal LTP).
struct PyrJob {
    const uint8_t* tft_states;
    const uint8_t* tft_best_raw;
    const uint8_t* prev_best_raw;
    const uint8_t* prev_best_ids;
    uint8_t* soma;
    uint8_t* cel_flags;
    uint8_t* axn;
    const SliceRt* slices;
    uint8_t* st[2];
    int8_t* str;
    const int32_t* rnd;
    uint32_t slot_mtp0;
    uint32_t V, U, depth, C;
    uint32_t en, prx, bdst, adst;
    uint32_t control;      // BSM_CTRL_INTRA_COLUMN / BSM_CTRL_PYR_OUTPUT / NONE
    uint32_t n_tufts;
    uint32_t base_slc;
    uint32_t learn;
    uint32_t distal_mask;
    uint32_t par_xor;
    uint32_t max_chunks;   // max over distal tufts of syns_per_den / chunk
    const uint32_t* dyn;
    const VirtTuft* vt;    // the layer's tufts (device array)
    uint32_t* sg;          // the layer's strength signs (32 per word)
    uint32_t t_spd[BSM_MAX_TUFTS];
    uint32_t t_spt[BSM_MAX_TUFTS];
    uint32_t t_syn_idz[BSM_MAX_TUFTS];
    const uint32_t* t_vbits[BSM_MAX_TUFTS];   // each tuft's parity bits (see VirtTuft)
    const uint32_t* t_bmask[BSM_MAX_TUFTS];   // its kept best-dendrite fired bits
    uint32_t t_bmlen[BSM_MAX_TUFTS], t_wpd[BSM_MAX_TUFTS];
};

struct WorkItem {
    uint16_t job, slc, tv, tu;
};

// ---------------------------------------------------------------- area
struct TuftRt {
    bsm_tuft_desc d;
    uint32_t spt;
    uint64_t syn_idz, den_idz;
    uint64_t sb_idz;            // first byte of the tuft's staged-slice codes (LayerRt::syn_sb)
    // Host-side summary of the tuft's synapse sources, kept conservative on
    // every write (never shrinks): the largest |v/u offset| and the set of
    // source slice ids present.  Lets the planner prove that a staged LDS
    // window covers every synapse of the tuft.
    // Source summary of the tuft's synapses (bounds that only widen on partial
    // writes, exact after a write or fill of the whole tuft range):
    uint32_t reach_v, reach_u;  // max |v offset|, max |u offset|
    uint32_t reach_max;         // max(reach_v, reach_u)
    uint64_t slc_bits[4];       // set of source slice ids
};

struct LayerRt {
    bsm_layer_desc d;
    uint32_t cells;
    uint64_t n_syn, n_den, n_celtft;
    uint64_t n_sb;           // bytes of syn_sb: 4 * ceil(synapses / 16) per tuft
    TuftRt t[BSM_MAX_TUFTS];
    uint8_t* syn_state[2];   // double-buffered synapse states
    int cur;                 // index of the current state buffer
    uint8_t* syn_flag_static;  // flag bits other than SYN_PREV_ACTIVE (never touched by kernels)
    uint8_t* syn_slc;
    int8_t* syn_v;
    int8_t* syn_u;
    uint8_t* syn_vu;         // packed copy of the offsets for the gather (see TuftJob::vu)
    uint8_t* syn_sb;         // staged-slice codes of each tuft for the gather (see TuftJob::sb)
    int8_t* syn_str;
    uint32_t* syn_sg;        // strength signs: bit i of the layer = (strength i < 0)
    uint64_t n_sg;           // bytes of syn_sg (whole 32-bit words)
    // virtual states (see VirtTuft): per tuft the two snapshots, the parity
    // bits (one word per tuft) and the device descriptors
    uint8_t* snap[BSM_MAX_TUFTS];
    uint32_t snap_cap[BSM_MAX_TUFTS];  // bytes per parity
    uint32_t* bmask[BSM_MAX_TUFTS];    // distal pyramidal tufts: best-dendrite fired bits
    uint32_t bm_wpd[BSM_MAX_TUFTS];    // words per dendrite (0: none kept)
    uint32_t* fbits;                   // spiny stellate layers: tuft 0's fired bits
    uint32_t* vbits;
    VirtTuft* vt_d;
    uint8_t *den_states, *den_raw, *den_energies, *den_act, *den_thr;
    uint8_t* tft[8];
    uint8_t *soma, *cel_flags, *cel_energies, *cel_act;
    uint32_t slot0;          // first rnd slot of this layer
    // activity smoother hosted by this layer: hex-group centers (6 groups), the
    // center covering each column in each group, per-center selections
    int32_t smooth_r;
    uint32_t smooth_n;
    int32_t* smooth_centers;   // (v, u) per center, group-major
    uint32_t* smooth_cover;    // SMOOTH_GROUPS x (v_size * u_size) center indices
    uint32_t* smooth_sel;      // (least active, most active, its activity) per center
};

struct Tuning {
    uint32_t tile_v = 8, tile_u = 8, halo = 8, chunk = 32;
    uint32_t inhib_tile = 0;  // side of the square inhibition tile; 0 = chosen by the planner
    uint32_t overlap = 1;     // run independent pyramidal tufts on the side stream
};

// One launch of a fused stage over a contiguous range of work items.
struct Stage {
    int kind;            // 1 = F1 tufts, 2 = F2 ssc post, 3 = F3 pyr post, 4 = smoother
    uint32_t item0, n_items;
    uint32_t lds_bytes;
    uint32_t layer;      // plain form: layer whose state parity flips at F1
    uint32_t W;          // synapse chunk width (bytes per lane)
    uint32_t nst;        // F1: staged-slice template (1, 2 or 4)
    uint32_t fast;       // F1: all jobs take the check-free gather path
    uint32_t packed;     // F1: all jobs read packed offsets
    uint32_t stage_cap;  // F1: LDS bytes reserved for staged windows
    uint32_t tv, tu;     // tile extents in columns
    uint32_t grid;
    uint64_t alg_bytes = 0;  // F1: algorithmic HBM bytes of one launch
    uint32_t stream = 0; // 0: main stream; 1: side stream, forked after the cycle's intake
    int32_t join = -1;   // main-stream stages: index of the side stage to wait for, or -1
};

// A second stream for stages that may run beside the main sequence of a cycle
// (pyramidal tufts whose sources nothing earlier in the cycle writes).
struct Side {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;  // one per stage index (side stages record theirs)
};
int side_prepare(Side& s, const std::vector<Stage>& stages);
void side_destroy(Side& s);

// Enqueues one cycle's stages.  Side stages fork from `main` at this point and
// record their join events; then pre_main(main) runs (work the main stages
// need and the side stages do not touch, such as the cycle's intake), then the
// main stages, each waiting for the side stage it names; any side stage no main
// stage named is joined at the end, so work on `main` after this call always
// follows every stage.  serial: pre_main and all stages on `main` in order.
template <class Launch, class Pre>
int enqueue_stages(hipStream_t main, Side& side, const std::vector<Stage>& stages, bool serial,
                   Launch&& launch, Pre&& pre_main) {
    bool any_side = false;
    for (const Stage& st : stages) any_side |= st.stream == 1;
    if (serial || !any_side) {
        if (pre_main(main)) return 1;
        for (const Stage& st : stages)
            if (launch(st, main)) return 1;
        return 0;
    }
    if (side.join.size() < stages.size()) return fail("side-stream events not prepared");
    BSM_HIP(hipEventRecord(side.fork, main));
    BSM_HIP(hipStreamWaitEvent(side.stream, side.fork, 0));
    std::vector<char> joined(stages.size(), 1);
    for (size_t i = 0; i < stages.size(); i++) {
        if (stages[i].stream != 1) continue;
        if (launch(stages[i], side.stream)) return 1;
        BSM_HIP(hipEventRecord(side.join[i], side.stream));
        joined[i] = 0;
    }
    if (pre_main(main)) return 1;
    for (const Stage& st : stages) {
        if (st.stream == 1) continue;
        if (st.join >= 0 && (size_t)st.join < stages.size() && !joined[st.join]) {
            BSM_HIP(hipStreamWaitEvent(main, side.join[st.join], 0));
            joined[st.join] = 1;
        }
        if (launch(st, main)) return 1;
    }
    for (size_t i = 0; i < stages.size(); i++)
        if (!joined[i]) BSM_HIP(hipStreamWaitEvent(main, side.join[i], 0));
    return 0;
}
template <class Launch>
int enqueue_stages(hipStream_t main, Side& side, const std::vector<Stage>& stages, bool serial,
                   Launch&& launch) {
    return enqueue_stages(main, side, stages, serial, launch, [](hipStream_t) { return 0; });
}

// Slices written during a cycle before a pyramidal layer's tufts run: the
// spatial layers' axon slices and those of pyramidal layers planned earlier.
struct SliceSet {
    uint64_t bits[4] = {0, 0, 0, 0};
    void add_range(uint32_t first, uint32_t n) {
        for (uint32_t s = first; s < first + n && s < 256; s++) bits[s >> 6] |= 1ull << (s & 63);
    }
    bool meets(const uint64_t* other) const {
        return (bits[0] & other[0]) | (bits[1] & other[1]) | (bits[2] & other[2]) |
               (bits[3] & other[3]);
    }
};

}  // namespace bsm

namespace bsm {
struct FramePool;     // pooled pinned host blocks and events for sampler frames
struct SamplePlan;    // device descriptors of the samplers one launch serves
}  // namespace bsm

struct bsm_ctx {
    int device;
    hipDeviceProp_t prop;
    bsm::FramePool* frames = nullptr;
    std::atomic<uint64_t> sampler_gen{0};   // bumped by every sampler create/destroy
};

struct bsm_area {
    bsm_ctx* ctx;
    hipStream_t stream;
    uint32_t V, U, n_slices, n_layers, settings;
    std::vector<bsm_slice_desc> slice_desc;
    std::vector<bsm::SliceRt> slices_h;
    bsm::SliceRt* slices_d;
    uint32_t axn_len;
    uint8_t* axn;
    bsm::LayerRt L[BSM_MAX_LAYERS];
    uint32_t n_slots;
    uint32_t rng[4];         // xorshift128 state (host side)
    uint8_t* arena;          // single device allocation backing all buffers
    size_t arena_bytes;
    int32_t* rnd_d;          // device rnd draws (capacity rnd_cap ints)
    size_t rnd_cap;
    int32_t* rnd_pin[2];     // pinned staging for per-call rnd uploads
    size_t rnd_pin_cap[2];
    hipEvent_t rnd_ev[2];
    int rnd_pin_next;
    bsm::Tuning tune;
    // fused-path job tables (device) and launch geometry
    bsm::TuftJob* tjobs_d;
    bsm::SscJob* sjobs_d;
    bsm::PyrJob* pjobs_d;
    bsm::WorkItem* items_d;
    std::vector<bsm::TuftJob> tjobs;
    std::vector<bsm::SscJob> sjobs;
    std::vector<bsm::PyrJob> pjobs;
    std::vector<bsm::Stage> stages;
    std::vector<uint32_t> flip_layers;  // layers whose state parity flips each cycle
    bsm::Side side;
    bool plan_dirty;
    // Fused intake: the plan's first stage is the F1 of one tuft (job 0) that
    // stages one slice, virtual states, whose snapshot covers the whole slice;
    // a frame of exactly that slice's bytes [fuse_lo, fuse_lo + fuse_len) is
    // then taken by that launch (fuse_job 0), else fuse_job is -1.
    int32_t fuse_job;
    uint32_t fuse_lo, fuse_len;
    hipEvent_t ev0, ev1;
    hipStream_t batch_stream;  // non-null while the area belongs to a batch
    bsm_batch* batch;          // that batch (buffer I/O syncs it first, which also brings
                               // a dynamic batch's device-side state parities to the host)
    uint32_t layout_epoch;     // bumped whenever the synapse source summary changes
    // readback (samplers): the area's open samplers, the generation of that
    // set, the plan serving the plain form, and the number of cycles done
    std::vector<bsm_sampler*> samplers;
    uint64_t sampler_gen;
    bsm::SamplePlan* sample_plan;
    uint64_t cycles_done;
    uint8_t* filt_stage;       // filtered intake: device copy of a host frame
    size_t filt_cap;
};

namespace bsm {
// planning and launchers implemented in bsm_fused.hip
TuftJob make_tuft_job(bsm_area* a, uint32_t l, uint32_t t, const Tuning& tn, const int32_t* rnd);
SscJob make_ssc_job(bsm_area* a, uint32_t l, const int32_t* rnd);
PyrJob make_pyr_job(bsm_area* a, uint32_t l, const int32_t* rnd, uint32_t* w3);
uint32_t tuft_chunk(const bsm_area* a, uint32_t l, uint32_t t, uint32_t max_w);
uint32_t ssc_chunk(const bsm_area* a, uint32_t l);
// Appends the items of one stage and fills its geometry; jobs[j] indexes the
// job arrays.  Returns 0 or an error.
int plan_f1(std::vector<TuftJob>& jobs, const std::vector<uint32_t>& job_ids,
            const std::vector<uint32_t>& job_w, const Tuning& tn, std::vector<WorkItem>& items,
            Stage& st);
int plan_f2(const std::vector<SscJob>& jobs, const std::vector<uint32_t>& job_ids, const Tuning& tn,
            std::vector<WorkItem>& items, Stage& st);
int plan_f3(const std::vector<PyrJob>& jobs, const std::vector<uint32_t>& job_ids,
            std::vector<WorkItem>& items, Stage& st);
int build_plan(bsm_area* a);
int merge_tuning(const Tuning& cur, const bsm_tuning* t, Tuning* out);
int launch_intake(hipStream_t s, uint8_t* dst, const uint8_t* src, uint32_t len);
// Refreshes the gather's derived streams of synapses [offset, offset + len) of a
// layer after any write to their offsets or slice ids: the packed offsets, and
// the staged-slice codes of every four-synapse group the range touches.
int launch_pack_vu(hipStream_t s, bsm_area* a, uint32_t layer, size_t offset, size_t len);
// Refreshes the strength-sign bits of synapses [offset, offset + len) of a
// layer (whole 32-bit words) after any write to their strengths.
int launch_pack_sg(hipStream_t s, bsm_area* a, uint32_t layer, size_t offset, size_t len);
// Writes parity q's states of the layer's synapses [offset, offset + len) into
// st[q] wherever they are virtual (both parities).  clear: the whole layer,
// and its parity bits are cleared afterwards, so every state lives in the
// state buffers again -- required before states, flags or synapse sources
// change by any route other than the cycle.
int settle_states(bsm_area* a, uint32_t layer, hipStream_t s, size_t offset, size_t len,
                  bool clear);
// settle_states(clear) of every layer of the area.
int settle_area(bsm_area* a, hipStream_t s);
// The device descriptors of the area's virtual-state tufts (at creation).
int upload_virt(bsm_area* a, hipStream_t s);
// The source slices the fused gather stages for a tuft (at most 4, sorted; with
// the scalar quirk only slices without mids).  Depends on the area's scheme only.
uint32_t stage_slices(const bsm_area* a, const TuftRt& T, uint32_t* staged);
constexpr uint32_t PACK_REACH = 7;   // |offset| bound of the packed form
// Enqueues the deterministic regrowth of every tuft of the area on stream s and
// widens the tufts' source summaries; *widened reports whether a summary grew
// (the fused plan is then stale).
int regrow_launch(bsm_area* a, hipStream_t s, uint64_t seed, uint64_t epoch, int32_t reach,
                  bool* widened);
int stage_info(const std::vector<Stage>& stages, uint32_t max_stages, bsm_stage_info* out,
               uint32_t* n_stages);
// An F1 launch may take a frame for one of its jobs (intake_job): that job's one
// staged slice is read from the frame, and the launch writes the frame into
// the slice's axon bytes (intake_dst) beside the job's snapshot copy, in place
// of a separate intake copy before it.
int launch_stage(hipStream_t s, const Stage& st, const TuftJob* tj, const SscJob* sj,
                 const PyrJob* pj, const WorkItem* items, uint32_t par, uint32_t rnd_ofs,
                 const uint8_t* frame = nullptr, uint8_t* intake_dst = nullptr,
                 uint32_t intake_job = 0xFFFFFFFFu);
// One plain cycle on the area stream (and its side stream).  intake, when
// given, is launched on the area stream before the main stages; the side
// stages start before it when intake_first is false.
// fused_frame: this cycle's frame, taken by the area's first stage (its
// fused intake, bsm_area::fuse_job) instead of an intake copy.
int cycle_fused(bsm_area* a, uint32_t rnd_ofs, bool serial,
                const std::function<int(hipStream_t)>* intake = nullptr,
                bool intake_first = true, const uint8_t* fused_frame = nullptr);
// Device pointer and length of one buffer kind of the area (the current state
// buffer for BSM_SYN_STATES; the static flag bits for BSM_SYN_FLAG_SETS).
int resolve_buf(bsm_area* a, int buf, uint32_t layer, uint8_t** ptr, size_t* len);
// Readback after a cycle (bsm_sample.hip).  Counts the cycle of every area that
// stepped (active == null: all) and, where those areas have samplers, enqueues
// their frames on stream s.  Parities come from the host (a->L[l].cur) unless
// dyn (BSM_MAX_LAYERS device words per area, bit 0 the parity) is given.
int sample_after_cycle(SamplePlan*& plan, hipStream_t s, bsm_area* const* areas, size_t n,
                       const uint8_t* active, const uint32_t* dyn);
void sample_plan_free(SamplePlan*& plan);
// Closes the area's samplers (their frames end) before the area goes away.
void samplers_detach(bsm_area* a);
void frame_pool_destroy(bsm_ctx* c);
// Activity smoother and sensory filter (bsm_filters.hip).
constexpr uint32_t SMOOTH_GROUPS = 6;
// Static staging of bsm_area_intake_filtered_noalloc (bytes per device) and the
// devices whose staging address is kept.
constexpr size_t FILTER_SCRATCH_MAX = BSM_FILTER_STATIC_MAX;
constexpr int MAX_DEVICES = 64;
// Takes the device address of the static staging array on `device`, which must
// be current; once per device, thread-safe.
int filter_scratch_setup(int device);
int launch_smoother(bsm_area* a, uint32_t layer, hipStream_t s);
void smoother_free(LayerRt& L);
// A stage of the area's plain plan on stream s (smoother stages included).
// frame: the cycle's frame for the area's fused intake (first stage only).
int launch_area_stage(bsm_area* a, hipStream_t s, const Stage& st, uint32_t rnd_ofs,
                      const uint8_t* frame = nullptr);
int ensure_rnd(bsm_area* a, size_t n_ints);
void layout_stats_reset(bsm_area* a, uint32_t l, uint32_t t, uint32_t reach, const uint8_t* ids,
                        uint32_t n_ids);
void layout_stats_update(bsm_area* a, int buf, uint32_t layer, size_t offset, size_t len,
                         const uint8_t* data, int fill_val);
int launch_layout(hipStream_t s, uint64_t seed, uint32_t layer, uint32_t tuft, uint64_t n,
                  uint32_t n_src, const uint8_t* src_ids, int32_t reach, uint8_t* slc, int8_t* v,
                  int8_t* u);
void rng_seed(uint32_t st[4], uint64_t seed);
uint32_t rng_next(uint32_t st[4]);
}  // namespace bsm
