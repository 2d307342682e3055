// bsm_batch.hip -- batched form: many independent areas of uneven sizes step
// through each fused stage in one persistent launch over a flattened work-item
// list, with the whole cycle captured as a hipGraph.
//
// Per batch cycle (stream order == reference enqueue order per area,
// src/cortex/cortical_area.rs:716-789):
//   prologue : per area, draw its rnd slots from its xorshift128 stream and copy
//              its next input frame into its axon space (AxonSpace::intake); the
//              frame counter advances
//   for each spatial layer rank k : F1 over all areas' SSC tufts, F2 over all SSC layers
//   for each pyramidal layer rank k: F1 over all areas' pyramidal tufts, F3
// Pyramidal tufts whose sources no earlier stage writes run their F1 on a side
// stream from the prologue on, joined before their layer's F3.  Each area's
// layers keep their own state-buffer parity through a per-job xor.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bsm_internal.h"

namespace bsm {

struct AreaIo {
    uint32_t* rng;          // device xorshift128 state (4 words)
    int32_t* rnd;           // the area's per-cycle rnd slots
    uint32_t n_slots;
    uint32_t len;           // frame bytes (0: no intake)
    const uint8_t* frames;  // device frame ring
    uint64_t stride;
    uint8_t* dst;           // axon space + offset
    uint32_t n_frames;
    uint32_t layer_mask;    // dynamic batches: layers whose parity flips when the area cycles
    uint32_t* dyn;          // dynamic batches: per-layer words (bit0 parity, bit1 active)
    const uint8_t* act;     // dynamic batches: this cycle's active flag of the area
    const uint64_t* frame_ofs;  // dynamic batches: frame offset in the staging buffer (~0: none)
    const uint8_t* stage;   // dynamic batches: staged frames of this cycle
};

// Frame bytes copied by one prologue / tract block.
constexpr uint32_t COPY_SEG = 16384;

// Blocks (area, part): block (a, 0) draws area a's rnd slots (thread 0,
// xorshift128, Marsaglia 2003) and updates its dynamic words; every block
// copies part `part` (COPY_SEG bytes) of the area's frame.
__global__ __launch_bounds__(256) void kb_prologue(const AreaIo* __restrict__ io,
                                                   uint32_t* __restrict__ counter, uint32_t par) {
    const AreaIo A = io[blockIdx.x];
    const uint32_t part = blockIdx.y;
    // frame counter: this cycle reads slot par, the next cycle (other parity) slot par ^ 1
    if (blockIdx.x == 0 && part == 0 && threadIdx.x == 0) counter[(par & 1) ^ 1] = counter[par & 1] + 1;
    if (A.dyn) {
        // dynamic batch: only areas with a frame this cycle step; their layers
        // flip parity and are marked active, the others are marked idle
        const uint32_t active = *A.act;
        if (part == 0 && threadIdx.x < BSM_MAX_LAYERS && ((A.layer_mask >> threadIdx.x) & 1u)) {
            const uint32_t d = A.dyn[threadIdx.x];
            A.dyn[threadIdx.x] = active ? (((d & 1u) ^ 1u) | 2u) : (d & 1u);
        }
        if (!active) return;
    }
    if (part == 0 && threadIdx.x == 0) {
        uint32_t x = A.rng[0], y = A.rng[1], z = A.rng[2], w = A.rng[3];
        for (uint32_t k = 0; k < A.n_slots; k++) {
            uint32_t t = x ^ (x << 11);
            x = y;
            y = z;
            z = w;
            w = w ^ (w >> 19) ^ (t ^ (t >> 8));
            A.rnd[k] = (int32_t)w;
        }
        A.rng[0] = x;
        A.rng[1] = y;
        A.rng[2] = z;
        A.rng[3] = w;
    }
    const uint8_t* src = nullptr;
    if (A.dyn) {
        const uint64_t fo = *A.frame_ofs;
        if (A.len && fo != ~0ull && (fo & BSM_FRAME_SDR)) {
            // SDR entry: a count, then that many axon indices of the input range
            // (checked on the host); this block clears its segment, then sets
            // the active axons that fall in it
            const uint32_t* e = (const uint32_t*)(A.stage + (fo & ~BSM_FRAME_SDR));
            const uint32_t lo = part * COPY_SEG;
            if (lo >= A.len) return;
            const uint32_t seg = min(COPY_SEG, A.len - lo);
            block_zero(A.dst + lo, seg, threadIdx.x, blockDim.x);
            __syncthreads();
            const uint32_t n = e[0];
            auto d8 = BSM_G(uint8_t, A.dst);
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t x = e[1 + i] - lo;
                if (x < seg) d8[lo + x] = 255;
            }
            return;
        }
        if (A.len && fo != ~0ull) src = A.stage + fo;
    } else if (A.len && A.n_frames) {
        src = A.frames + (size_t)(counter[par & 1] % A.n_frames) * A.stride;
    }
    const uint32_t lo = part * COPY_SEG;
    if (src && lo < A.len)
        block_copy(A.dst + lo, src + lo, min(COPY_SEG, A.len - lo), threadIdx.x, blockDim.x);
}


struct TractRt {
    const uint8_t* src;
    uint8_t* dst;
    uint32_t len;
    uint32_t _pad;
};

// Blocks (tract, part): copy part `part` (COPY_SEG bytes) of the source
// area's axon range into the target's.
__global__ __launch_bounds__(256) void kb_tracts(const TractRt* __restrict__ tr) {
    const TractRt T = tr[blockIdx.x];
    const uint32_t lo = blockIdx.y * COPY_SEG;
    if (lo < T.len)
        block_copy(T.dst + lo, T.src + lo, min(COPY_SEG, T.len - lo), threadIdx.x, blockDim.x);
}

}  // namespace bsm

struct bsm_batch {
    bsm_ctx* ctx;
    hipStream_t stream;
    std::vector<bsm_area*> areas;
    std::vector<bsm::TuftJob> tjobs;
    std::vector<bsm::SscJob> sjobs;
    std::vector<bsm::PyrJob> pjobs;
    std::vector<bsm::Stage> stages;
    std::vector<bsm::AreaIo> io;
    bsm::TuftJob* tjobs_d = nullptr;
    bsm::SscJob* sjobs_d = nullptr;
    bsm::PyrJob* pjobs_d = nullptr;
    bsm::WorkItem* items_d = nullptr;
    // host copies for masked cycles: every stage's work items and, per job of
    // each kind (F1 tufts, F2, F3), its area
    std::vector<bsm::WorkItem> items_h;
    std::vector<uint32_t> job_area[3];
    // per stage, its items grouped by area: the areas in the order their first
    // (largest) item comes in the stage, each area's items in stage order
    struct StageAreas {
        std::vector<uint32_t> order;          // areas with items in the stage
        std::vector<uint32_t> first, count;   // per area: its run in `idx`
        std::vector<uint32_t> idx;            // item indices, grouped by area
    };
    std::vector<StageAreas> by_area;
    bsm::AreaIo* io_d = nullptr;
    uint32_t* rng_d = nullptr;      // 4 words per area
    uint32_t* counter_d = nullptr;  // frame counters, one per cycle parity
    uint32_t bpar = 0;
    bsm::Tuning tune;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool dirty = true;
    hipEvent_t ev0, ev1;
    // dynamic mode (serving): per-cycle active mask and staged frames
    bool dynamic = false;
    uint32_t* dyn_d = nullptr;      // BSM_MAX_LAYERS words per area
    uint8_t* xfer_d = nullptr;      // [act n | pad][frame_ofs n x u64][frames]
    size_t xfer_cap = 0;
    uint8_t* xfer_h[2] = {nullptr, nullptr};
    size_t xfer_h_cap[2] = {0, 0};
    hipEvent_t xfer_ev[2];
    int xfer_next = 0;
    bsm::Side side;             // second stream of the cycle (see enqueue_stages)
    bsm::TractRt* tracts_d = nullptr;
    uint32_t n_tracts = 0;
    uint32_t io_parts = 1;      // prologue blocks per area (largest axon space / COPY_SEG)
    uint32_t tract_parts = 1;   // tract blocks per tract (longest tract / COPY_SEG)
    std::vector<uint32_t> epochs;   // areas' layout epochs the plan was made for
    bsm::SamplePlan* sample_plan = nullptr;   // samplers of the batched areas
    // dynamic batches: each area's input range (kept across membership changes)
    std::vector<uint32_t> in_ofs, in_len;
    size_t area_cap = 0;        // areas the per-area device arrays hold
    size_t cap_t = 0, cap_s = 0, cap_p = 0, cap_i = 0;   // plan table capacities (bytes)
    // device time of masked cycles: events around each cycle, one pair per
    // transfer slot, added up once the slot comes round again
    hipEvent_t busy_ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    bool busy_rec[2] = {false, false};
    double busy_ms = 0.0;
};

namespace bsm {

static void free_graphs(bsm_batch* b) {
    for (int i = 0; i < 2; i++) {
        if (b->exec[i]) hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) hipGraphDestroy(b->graph[i]);
        b->exec[i] = nullptr;
        b->graph[i] = nullptr;
    }
}

static int batch_plan(bsm_batch* b) {
    free_graphs(b);
    // jobs are planned so that the next cycle (par = 1) writes each layer's cur ^ 1
    b->bpar = 0;
    b->tjobs.clear();
    b->sjobs.clear();
    b->pjobs.clear();
    b->stages.clear();
    b->epochs.clear();
    for (auto& v : b->job_area) v.clear();
    for (bsm_area* a : b->areas) b->epochs.push_back(a->layout_epoch);
    std::vector<WorkItem> items;
    const uint32_t na = (uint32_t)b->areas.size();
    // rank of each layer among the area's spatial / pyramidal layers
    uint32_t max_rank[2] = {0, 0};
    for (bsm_area* a : b->areas) {
        uint32_t r[2] = {0, 0};
        for (uint32_t l = 0; l < a->n_layers; l++) r[a->L[l].d.spatial ? 0 : 1]++;
        max_rank[0] = std::max(max_rank[0], r[0]);
        max_rank[1] = std::max(max_rank[1], r[1]);
    }
    // per area: slices written in a cycle before its pyramidal tufts of the
    // current rank run (spatial outputs, then lower-rank pyramidal outputs)
    std::vector<SliceSet> early(na);
    for (uint32_t ai = 0; ai < na; ai++) {
        bsm_area* a = b->areas[ai];
        if (a->settings & BSM_SET_DISABLE_SSCS) continue;
        for (uint32_t l = 0; l < a->n_layers; l++)
            if (a->L[l].d.spatial) early[ai].add_range(a->L[l].d.base_axn_slc, a->L[l].d.depth);
    }
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t k = 0; k < max_rank[pass]; k++) {
            std::vector<uint32_t> t_ids, t_w, s_ids, p_ids, side_ids, side_w;
            std::vector<std::pair<uint32_t, uint32_t>> ranked;  // (area, layer) of this rank
            uint32_t w3 = 16;
            for (uint32_t ai = 0; ai < na; ai++) {
                bsm_area* a = b->areas[ai];
                const bool sscs = !(a->settings & BSM_SET_DISABLE_SSCS);
                const bool pyrs = !(a->settings & BSM_SET_DISABLE_PYRS);
                uint32_t rank = 0;
                for (uint32_t l = 0; l < a->n_layers; l++) {
                    LayerRt& L = a->L[l];
                    const bool spatial = L.d.spatial != 0;
                    if ((pass == 0) != spatial) continue;
                    if (rank++ != k) continue;
                    if ((spatial && !sscs) || (!spatial && !pyrs)) continue;
                    // first batch cycle runs with par = 1 and must write buffer cur ^ 1
                    const uint32_t px = (uint32_t)L.cur;
                    const uint32_t* dw = b->dynamic ? b->dyn_d + ai * BSM_MAX_LAYERS + l : nullptr;
                    ranked.push_back({ai, l});
                    for (uint32_t t = 0; t < L.d.n_tufts; t++) {
                        TuftJob j = make_tuft_job(a, l, t, b->tune, a->rnd_d);
                        j.par_xor = px;
                        j.dyn = dw;
                        const bool side = !spatial && b->tune.overlap &&
                                          !early[ai].meets(L.t[t].slc_bits);
                        (side ? side_ids : t_ids).push_back((uint32_t)b->tjobs.size());
                        (side ? side_w : t_w).push_back(tuft_chunk(a, l, t, b->tune.chunk));
                        b->tjobs.push_back(j);
                        b->job_area[0].push_back(ai);
                    }
                    if (spatial) {
                        SscJob j = make_ssc_job(a, l, a->rnd_d);
                        j.par_xor = px;
                        j.dyn = dw;
                        s_ids.push_back((uint32_t)b->sjobs.size());
                        b->sjobs.push_back(j);
                        b->job_area[1].push_back(ai);
                    } else {
                        uint32_t w;
                        PyrJob j = make_pyr_job(a, l, a->rnd_d, &w);
                        j.par_xor = px;
                        j.dyn = dw;
                        w3 = std::min(w3, w);
                        p_ids.push_back((uint32_t)b->pjobs.size());
                        b->pjobs.push_back(j);
                        b->job_area[2].push_back(ai);
                    }
                }
            }
            if (t_ids.empty()) {
                t_ids.swap(side_ids);
                t_w.swap(side_w);
            }
            if (t_ids.empty()) continue;
            int32_t join = -1;
            if (!side_ids.empty()) {
                Stage sd{};
                if (plan_f1(b->tjobs, side_ids, side_w, b->tune, items, sd)) return 1;
                sd.stream = 1;
                join = (int32_t)b->stages.size();
                b->stages.push_back(sd);
            }
            Stage s1{};
            if (plan_f1(b->tjobs, t_ids, t_w, b->tune, items, s1)) return 1;
            b->stages.push_back(s1);
            if (pass == 1)
                for (auto& al : ranked)
                    early[al.first].add_range(b->areas[al.first]->L[al.second].d.base_axn_slc,
                                              b->areas[al.first]->L[al.second].d.depth);
            Stage s2{};
            if (pass == 0) {
                if (plan_f2(b->sjobs, s_ids, b->tune, items, s2)) return 1;
            } else {
                // a common chunk width for every job of the launch
                for (uint32_t id : p_ids) {
                    PyrJob& j = b->pjobs[id];
                    uint32_t mc = 1;
                    for (uint32_t t = 0; t < j.n_tufts; t++)
                        if ((j.distal_mask >> t) & 1u) mc = std::max(mc, j.t_spd[t] / w3);
                    j.max_chunks = mc;
                }
                if (plan_f3(b->pjobs, p_ids, items, s2)) return 1;
                s2.W = w3;
                s2.join = join;
            }
            b->stages.push_back(s2);
        }
    }
    if (side_prepare(b->side, b->stages)) return 1;
    // the tables only grow (membership changes re-plan often); the batch stream
    // is idle whenever a plan is made, so they are rewritten in place
    BSM_HIP(hipStreamSynchronize(b->stream));
    auto up = [&](void** dst, size_t& cap, const void* src, size_t bytes) -> int {
        if (bytes > cap || !*dst) {
            (void)hipFree(*dst);
            *dst = nullptr;
            const size_t c = std::max<size_t>({bytes, 2 * cap, 256});
            BSM_HIP(hipMalloc(dst, c));
            cap = c;
        }
        if (bytes) BSM_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if (up((void**)&b->tjobs_d, b->cap_t, b->tjobs.data(), b->tjobs.size() * sizeof(TuftJob)) ||
        up((void**)&b->sjobs_d, b->cap_s, b->sjobs.data(), b->sjobs.size() * sizeof(SscJob)) ||
        up((void**)&b->pjobs_d, b->cap_p, b->pjobs.data(), b->pjobs.size() * sizeof(PyrJob)) ||
        up((void**)&b->items_d, b->cap_i, items.data(), items.size() * sizeof(WorkItem)))
        return 1;
    b->items_h.swap(items);
    // masked cycles gather the active areas' items per stage from these groups
    b->by_area.assign(b->stages.size(), {});
    for (size_t si = 0; si < b->stages.size(); si++) {
        const Stage& st = b->stages[si];
        const std::vector<uint32_t>& ja = b->job_area[st.kind == 1 ? 0 : (st.kind == 2 ? 1 : 2)];
        auto& g = b->by_area[si];
        g.first.assign(na, 0);
        g.count.assign(na, 0);
        for (uint32_t i = st.item0; i < st.item0 + st.n_items; i++) {
            const uint32_t ai = ja[b->items_h[i].job];
            if (g.count[ai]++ == 0) g.order.push_back(ai);
        }
        uint32_t at = 0;
        for (uint32_t ai : g.order) {
            g.first[ai] = at;
            at += g.count[ai];
        }
        g.idx.resize(at);
        std::vector<uint32_t> fill(na, 0);
        for (uint32_t i = st.item0; i < st.item0 + st.n_items; i++) {
            const uint32_t ai = ja[b->items_h[i].job];
            g.idx[g.first[ai] + fill[ai]++] = i;
        }
    }
    b->dirty = false;
    return 0;
}

// One batched cycle on the batch stream.  With ev, events ev[2i], ev[2i+1]
// bracket stage i (never while capturing).
static int enqueue_cycle(bsm_batch* b, uint32_t par, hipEvent_t* ev = nullptr,
                         const std::vector<Stage>* stg = nullptr,
                         const WorkItem* items_d = nullptr) {
    const std::vector<Stage>& stages = stg ? *stg : b->stages;
    const WorkItem* items = items_d ? items_d : b->items_d;
    const uint32_t na = (uint32_t)b->areas.size();
    if (b->n_tracts) {
        kb_tracts<<<dim3(b->n_tracts, b->tract_parts), 256, 0, b->stream>>>(b->tracts_d);
        BSM_HIP(hipGetLastError());
    }
    kb_prologue<<<dim3(na, b->io_parts), 256, 0, b->stream>>>(b->io_d, b->counter_d, par);
    BSM_HIP(hipGetLastError());
    if (ev) {
        // profiling: one stream, stages in list order, each launch timed alone
        for (size_t i = 0; i < stages.size(); i++) {
            BSM_HIP(hipEventRecord(ev[2 * i], b->stream));
            if (launch_stage(b->stream, stages[i], b->tjobs_d, b->sjobs_d, b->pjobs_d,
                             items, par, 0))
                return 1;
            BSM_HIP(hipEventRecord(ev[2 * i + 1], b->stream));
        }
    } else if (enqueue_stages(b->stream, b->side, stages, false,
                              [&](const Stage& st, hipStream_t s) {
                                  return launch_stage(s, st, b->tjobs_d, b->sjobs_d, b->pjobs_d,
                                                      items, par, 0);
                              })) {
        return 1;
    }
    return 0;
}

static int capture(bsm_batch* b) {
    free_graphs(b);
    for (int par = 0; par < 2; par++) {
        BSM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int e = enqueue_cycle(b, (uint32_t)par);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (e) {
            if (g) hipGraphDestroy(g);
            return 1;
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        b->graph[par] = g;
        BSM_HIP(hipGraphInstantiate(&b->exec[par], g, nullptr, nullptr, 0));
    }
    return 0;
}

static bool layouts_changed(const bsm_batch* b) {
    for (size_t i = 0; i < b->areas.size(); i++)
        if (i >= b->epochs.size() || b->areas[i]->layout_epoch != b->epochs[i]) return true;
    return false;
}

static int run_cycles(bsm_batch* b, uint32_t n, int use_graph, hipEvent_t* ev = nullptr,
                      const std::vector<Stage>* stg = nullptr, const WorkItem* items_d = nullptr) {
    if (b->areas.empty()) return 0;   // a dynamic batch between its last and next member
    if (!b->dirty && layouts_changed(b)) {
        // a layout write may have changed which gather path is exact: re-plan
        // against the areas' current parities
        BSM_HIP(hipStreamSynchronize(b->stream));
        if (bsm_batch_sync(b)) return 1;
        b->dirty = true;
    }
    if (b->dirty && batch_plan(b)) return 1;
    if (use_graph && !b->exec[0] && capture(b)) return 1;
    for (uint32_t c = 0; c < n; c++) {
        b->bpar ^= 1u;
        if (use_graph) {
            BSM_HIP(hipGraphLaunch(b->exec[b->bpar], b->stream));
        } else if (enqueue_cycle(b, b->bpar, ev ? ev + (size_t)2 * c * b->stages.size() : nullptr,
                                 stg, items_d)) {
            return 1;
        }
        // dynamic batches keep parities on the device (read back at sync) and
        // sample after the masked cycle
        if (b->dynamic) continue;
        // host bookkeeping: every batched layer's current buffer flips each cycle
        for (bsm_area* a : b->areas) {
            const bool sscs = !(a->settings & BSM_SET_DISABLE_SSCS);
            const bool pyrs = !(a->settings & BSM_SET_DISABLE_PYRS);
            for (uint32_t l = 0; l < a->n_layers; l++) {
                LayerRt& L = a->L[l];
                const bool spatial = L.d.spatial != 0;
                if ((spatial && sscs) || (!spatial && pyrs)) L.cur ^= 1;
            }
        }
        if (sample_after_cycle(b->sample_plan, b->stream, b->areas.data(), b->areas.size(),
                               nullptr, nullptr))
            return 1;
    }
    return 0;
}

// Device transfer buffer of a dynamic batch: [act n | pad to 16][frame_ofs n][frames].
static size_t xfer_hdr(size_t n) { return ((n + 15) & ~(size_t)15) + 8 * n; }

static int grow_xfer(bsm_batch* b, size_t frame_bytes, bool repoint = false) {
    const size_t n = b->areas.size();
    const size_t need = xfer_hdr(n) + frame_bytes;
    if (need <= b->xfer_cap && !repoint) return 0;
    if (need > b->xfer_cap) {
        const size_t cap = std::max(need, b->xfer_cap * 2);
        BSM_HIP(hipStreamSynchronize(b->stream));
        if (b->xfer_d) BSM_HIP(hipFree(b->xfer_d));
        b->xfer_d = nullptr;
        BSM_HIP(hipMalloc((void**)&b->xfer_d, cap));
        BSM_HIP(hipMemset(b->xfer_d, 0, cap));
        b->xfer_cap = cap;
    }
    const size_t a0 = (n + 15) & ~(size_t)15;
    for (size_t i = 0; i < n; i++) {
        b->io[i].act = b->xfer_d + i;
        b->io[i].frame_ofs = (const uint64_t*)(b->xfer_d + a0) + i;
        b->io[i].stage = b->xfer_d + xfer_hdr(n);
    }
    if (b->io_d)
        BSM_HIP(hipMemcpy(b->io_d, b->io.data(), sizeof(AreaIo) * n, hipMemcpyHostToDevice));
    return 0;
}

// Adds the device time of the masked cycle last recorded in transfer slot k.
static int busy_harvest(bsm_batch* b, int k) {
    if (!b->busy_rec[k]) return 0;
    BSM_HIP(hipEventSynchronize(b->busy_ev[k][1]));
    float ms = 0.f;
    BSM_HIP(hipEventElapsedTime(&ms, b->busy_ev[k][0], b->busy_ev[k][1]));
    b->busy_ms += ms;
    b->busy_rec[k] = false;
    return 0;
}

static int check_joinable(bsm_batch* b, bsm_area* a, const std::vector<bsm_area*>& others) {
    if (!a || a->ctx != b->ctx) return fail("area not on this context");
    if (a->batch_stream) return fail("area already belongs to a batch");
    for (bsm_area* o : others)
        if (o == a) return fail("duplicate area in batch");
    if ((a->axn_len + COPY_SEG - 1) / COPY_SEG > 65535)
        return fail("axon space too large for the batch prologue");
    for (uint32_t l = 0; l < a->n_layers; l++)
        if (a->L[l].smooth_n) return fail("areas with an activity smoother cycle in the plain form");
    return 0;
}

// Loads the per-area device state of the batch from its areas' host state: the
// xorshift128 words, the dynamic parity words, the AreaIo entries (which keep the
// static inputs or the dynamic input ranges) and the transfer layout.  Device
// arrays only grow.  The batch stream must be idle; the plan becomes stale.
static int load_areas(bsm_batch* b, const std::vector<AreaIo>& keep_io) {
    const size_t n = b->areas.size();
    if (n > b->area_cap) {
        const size_t cap = std::max(n, b->area_cap * 2);
        (void)hipFree(b->rng_d);
        (void)hipFree(b->dyn_d);
        (void)hipFree(b->io_d);
        b->rng_d = nullptr;
        b->dyn_d = nullptr;
        b->io_d = nullptr;
        BSM_HIP(hipMalloc((void**)&b->rng_d, sizeof(uint32_t) * 4 * cap));
        if (b->dynamic) BSM_HIP(hipMalloc((void**)&b->dyn_d, sizeof(uint32_t) * BSM_MAX_LAYERS * cap));
        BSM_HIP(hipMalloc((void**)&b->io_d, sizeof(AreaIo) * cap));
        b->area_cap = cap;
    }
    std::vector<uint32_t> rng(4 * n);
    std::vector<uint32_t> dw(b->dynamic ? (size_t)BSM_MAX_LAYERS * n : 0, 0);
    b->io.assign(n, AreaIo{});
    b->io_parts = 1;
    for (size_t i = 0; i < n; i++) {
        bsm_area* a = b->areas[i];
        if (ensure_rnd(a, std::max<uint32_t>(a->n_slots, 1))) return 1;
        std::memcpy(&rng[4 * i], a->rng, sizeof(a->rng));
        AreaIo& io = b->io[i];
        if (i < keep_io.size()) io = keep_io[i];
        io.rng = b->rng_d + 4 * i;
        io.rnd = a->rnd_d;
        io.n_slots = a->n_slots;
        b->io_parts = std::max(b->io_parts, (uint32_t)((a->axn_len + COPY_SEG - 1) / COPY_SEG));
        if (b->dynamic) {
            uint32_t mask = 0;
            for (uint32_t l = 0; l < a->n_layers; l++) {
                const bool spatial = a->L[l].d.spatial != 0;
                const bool on = spatial ? !(a->settings & BSM_SET_DISABLE_SSCS)
                                        : !(a->settings & BSM_SET_DISABLE_PYRS);
                dw[i * BSM_MAX_LAYERS + l] = (uint32_t)a->L[l].cur;
                if (on) mask |= 1u << l;
            }
            io.layer_mask = mask;
            io.dyn = b->dyn_d + i * BSM_MAX_LAYERS;
            io.dst = a->axn + b->in_ofs[i];
            io.len = b->in_len[i];
        }
    }
    if (n) {
        BSM_HIP(hipMemcpy(b->rng_d, rng.data(), rng.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
        if (b->dynamic)
            BSM_HIP(hipMemcpy(b->dyn_d, dw.data(), dw.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice));
    }
    // dynamic batches: point the entries at the transfer buffer (uploads io_d)
    if (b->dynamic) {
        if (grow_xfer(b, 1 << 16, true)) return 1;
    } else if (n) {
        BSM_HIP(hipMemcpy(b->io_d, b->io.data(), sizeof(AreaIo) * n, hipMemcpyHostToDevice));
    }
    free_graphs(b);
    b->dirty = true;
    return 0;
}

}  // namespace bsm

using namespace bsm;

extern "C" {

int bsm_batch_create(bsm_ctx* ctx, bsm_area* const* areas, uint32_t n_areas, bsm_batch** out) {
    return bsm_batch_create_ex(ctx, areas, n_areas, 0, out);
}

int bsm_batch_create_ex(bsm_ctx* ctx, bsm_area* const* areas, uint32_t n_areas, uint32_t flags,
                        bsm_batch** out) {
    if (!ctx || !out || (!areas && n_areas)) return fail("null arg");
    const bool dynamic = (flags & BSM_BATCH_DYNAMIC) != 0;
    if ((n_areas == 0 && !dynamic) || n_areas > 65535)
        return fail("n_areas must be in [1, 65535] ([0, 65535] for dynamic batches)");
    BSM_HIP(hipSetDevice(ctx->device));
    bsm_batch* b = new bsm_batch();
    b->ctx = ctx;
    b->dynamic = dynamic;
    for (uint32_t i = 0; i < n_areas; i++) {
        if (check_joinable(b, areas[i], b->areas)) {
            delete b;
            return 1;
        }
        b->areas.push_back(areas[i]);
    }
    BSM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BSM_HIP(hipEventCreate(&b->ev0));
    BSM_HIP(hipEventCreate(&b->ev1));
    BSM_HIP(hipMalloc((void**)&b->counter_d, 2 * sizeof(uint32_t)));
    BSM_HIP(hipMemset(b->counter_d, 0, 2 * sizeof(uint32_t)));
    for (int i = 0; i < 2; i++) {
        BSM_HIP(hipEventCreateWithFlags(&b->xfer_ev[i], hipEventDisableTiming));
        BSM_HIP(hipEventCreate(&b->busy_ev[i][0]));
        BSM_HIP(hipEventCreate(&b->busy_ev[i][1]));
    }
    for (bsm_area* a : b->areas) {
        BSM_HIP(hipStreamSynchronize(a->stream));
        a->batch_stream = b->stream;
        a->batch = b;
    }
    b->in_ofs.assign(n_areas, 0);
    b->in_len.assign(n_areas, 0);
    if (load_areas(b, {})) return 1;
    *out = b;
    return 0;
}

int bsm_batch_add_areas(bsm_batch* b, bsm_area* const* areas, uint32_t n,
                        const uint32_t* axn_offset, const uint32_t* len) {
    if (!b || (n && (!areas || !axn_offset || !len))) return fail("null arg");
    if (!b->dynamic) return fail("areas join dynamic batches only");
    if (b->areas.size() + n > 65535) return fail("a batch holds at most 65535 areas");
    for (uint32_t i = 0; i < n; i++) {
        std::vector<bsm_area*> others(b->areas);
        others.insert(others.end(), areas, areas + i);
        if (check_joinable(b, areas[i], others)) return 1;
        if (axn_offset[i] > areas[i]->axn_len || len[i] > areas[i]->axn_len - axn_offset[i])
            return fail("input range out of the axon space of joining area " + std::to_string(i));
    }
    if (!n) return 0;
    BSM_HIP(hipSetDevice(b->ctx->device));
    // the enqueued cycles finish and the members' device state comes back to
    // their host copies, which load_areas uploads again with the new members
    if (bsm_batch_sync(b)) return 1;
    for (uint32_t i = 0; i < n; i++) {
        bsm_area* a = areas[i];
        BSM_HIP(hipStreamSynchronize(a->stream));
        b->areas.push_back(a);
        b->in_ofs.push_back(axn_offset[i]);
        b->in_len.push_back(len[i]);
        a->batch_stream = b->stream;
        a->batch = b;
    }
    return load_areas(b, {});
}

int bsm_batch_remove_areas(bsm_batch* b, bsm_area* const* areas, uint32_t n) {
    if (!b || (n && !areas)) return fail("null arg");
    if (!b->dynamic) return fail("areas leave dynamic batches only");
    std::vector<char> drop(b->areas.size(), 0);
    for (uint32_t i = 0; i < n; i++) {
        size_t k = 0;
        while (k < b->areas.size() && b->areas[k] != areas[i]) k++;
        if (k == b->areas.size()) return fail("area to remove is not in the batch");
        drop[k] = 1;
    }
    if (!n) return 0;
    BSM_HIP(hipSetDevice(b->ctx->device));
    if (bsm_batch_sync(b)) return 1;
    std::vector<bsm_area*> keep;
    std::vector<uint32_t> ko, kl;
    for (size_t k = 0; k < b->areas.size(); k++) {
        bsm_area* a = b->areas[k];
        if (drop[k]) {
            a->plan_dirty = true;
            a->batch_stream = nullptr;
            a->batch = nullptr;
            continue;
        }
        keep.push_back(a);
        ko.push_back(b->in_ofs[k]);
        kl.push_back(b->in_len[k]);
    }
    b->areas.swap(keep);
    b->in_ofs.swap(ko);
    b->in_len.swap(kl);
    return load_areas(b, {});
}

int bsm_batch_sync(bsm_batch* b) {
    if (!b) return fail("null batch");
    BSM_HIP(hipSetDevice(b->ctx->device));
    BSM_HIP(hipStreamSynchronize(b->stream));
    if (b->areas.empty()) return 0;
    // the areas' host rnd streams continue where the device streams stopped
    std::vector<uint32_t> rng(4 * b->areas.size());
    BSM_HIP(hipMemcpy(rng.data(), b->rng_d, rng.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < b->areas.size(); i++)
        std::memcpy(b->areas[i]->rng, &rng[4 * i], sizeof(b->areas[i]->rng));
    if (b->dynamic) {
        std::vector<uint32_t> dw((size_t)BSM_MAX_LAYERS * b->areas.size());
        BSM_HIP(hipMemcpy(dw.data(), b->dyn_d, dw.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < b->areas.size(); i++)
            for (uint32_t l = 0; l < b->areas[i]->n_layers; l++)
                b->areas[i]->L[l].cur = (int)(dw[i * BSM_MAX_LAYERS + l] & 1u);
    }
    return 0;
}

int bsm_batch_destroy(bsm_batch* b) {
    if (!b) return 0;
    bsm_batch_sync(b);
    (void)hipSetDevice(b->ctx->device);
    free_graphs(b);
    if (side_prepare(b->side, b->stages)) return 1;
    (void)hipFree(b->tjobs_d);
    (void)hipFree(b->sjobs_d);
    (void)hipFree(b->pjobs_d);
    (void)hipFree(b->items_d);
    (void)hipFree(b->io_d);
    (void)hipFree(b->dyn_d);
    (void)hipFree(b->tracts_d);
    (void)hipFree(b->xfer_d);
    for (int i = 0; i < 2; i++) {
        if (b->xfer_h[i]) (void)hipHostFree(b->xfer_h[i]);
        (void)hipEventDestroy(b->xfer_ev[i]);
        (void)hipEventDestroy(b->busy_ev[i][0]);
        (void)hipEventDestroy(b->busy_ev[i][1]);
    }
    (void)hipFree(b->rng_d);
    (void)hipFree(b->counter_d);
    sample_plan_free(b->sample_plan);
    (void)hipEventDestroy(b->ev0);
    (void)hipEventDestroy(b->ev1);
    side_destroy(b->side);
    (void)hipStreamDestroy(b->stream);
    for (bsm_area* a : b->areas) {
        a->plan_dirty = true;
        a->batch_stream = nullptr;
        a->batch = nullptr;
    }
    delete b;
    return 0;
}

int bsm_batch_set_inputs(bsm_batch* b, const uint8_t* const* dev_frames, const uint64_t* stride,
                         const uint32_t* len, const uint32_t* axn_offset, uint32_t n_frames) {
    if (!b) return fail("null batch");
    BSM_HIP(hipSetDevice(b->ctx->device));
    BSM_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->areas.size(); i++) {
        AreaIo& io = b->io[i];
        bsm_area* a = b->areas[i];
        if (!dev_frames || !dev_frames[i] || !len || !len[i]) {
            io.len = 0;
            io.frames = nullptr;
            continue;
        }
        uint32_t ofs = axn_offset ? axn_offset[i] : 0;
        if (ofs > a->axn_len || len[i] > a->axn_len - ofs)
            return fail("frame range out of the axon space of area " + std::to_string(i));
        io.len = len[i];
        io.frames = dev_frames[i];
        io.stride = stride ? stride[i] : len[i];
        io.dst = a->axn + ofs;
        io.n_frames = n_frames;
    }
    BSM_HIP(hipMemcpy(b->io_d, b->io.data(), sizeof(AreaIo) * b->io.size(), hipMemcpyHostToDevice));
    BSM_HIP(hipMemset(b->counter_d, 0, 2 * sizeof(uint32_t)));
    return 0;
}

int bsm_batch_cycle(bsm_batch* b, uint32_t n_cycles, int use_graph) {
    if (!b) return fail("null batch");
    BSM_HIP(hipSetDevice(b->ctx->device));
    return run_cycles(b, n_cycles, use_graph);
}

int bsm_batch_time_cycles(bsm_batch* b, uint32_t n_cycles, int use_graph, float* ms_total) {
    if (!b || !ms_total) return fail("null arg");
    BSM_HIP(hipSetDevice(b->ctx->device));
    if (b->dirty && batch_plan(b)) return 1;
    if (use_graph && !b->exec[0] && capture(b)) return 1;
    BSM_HIP(hipStreamSynchronize(b->stream));
    BSM_HIP(hipEventRecord(b->ev0, b->stream));
    if (run_cycles(b, n_cycles, use_graph)) return 1;
    BSM_HIP(hipEventRecord(b->ev1, b->stream));
    BSM_HIP(hipEventSynchronize(b->ev1));
    BSM_HIP(hipEventElapsedTime(ms_total, b->ev0, b->ev1));
    return 0;
}

int bsm_batch_profile_stages(bsm_batch* b, uint32_t n_cycles, uint32_t max_stages,
                             float* stage_ms, int* stage_kind, uint32_t* n_stages) {
    if (!b || !n_stages) return fail("null arg");
    BSM_HIP(hipSetDevice(b->ctx->device));
    if (b->dynamic) return fail("stage profiling needs a static batch");
    if (!b->dirty && layouts_changed(b)) {
        BSM_HIP(hipStreamSynchronize(b->stream));
        if (bsm_batch_sync(b)) return 1;
        b->dirty = true;
    }
    if (b->dirty && batch_plan(b)) return 1;
    const uint32_t ns = (uint32_t)b->stages.size();
    *n_stages = ns;
    if (ns > max_stages) return fail("max_stages too small");
    std::vector<hipEvent_t> ev((size_t)2 * ns * n_cycles);
    for (auto& e : ev) BSM_HIP(hipEventCreate(&e));
    BSM_HIP(hipStreamSynchronize(b->stream));
    int rc = run_cycles(b, n_cycles, 0, ev.data());
    if (rc == 0) BSM_HIP(hipStreamSynchronize(b->stream));
    for (uint32_t i = 0; i < ns && rc == 0; i++) {
        float tot = 0;
        for (uint32_t c = 0; c < n_cycles; c++) {
            float ms = 0;
            BSM_HIP(hipEventElapsedTime(&ms, ev[2 * ((size_t)c * ns + i)],
                                        ev[2 * ((size_t)c * ns + i) + 1]));
            tot += ms;
        }
        if (stage_ms) stage_ms[i] = tot;
        if (stage_kind) stage_kind[i] = b->stages[i].kind;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

int bsm_batch_stage_plan(bsm_batch* b, uint32_t max_stages, bsm_stage_info* out,
                         uint32_t* n_stages) {
    if (!b || !n_stages) return fail("null arg");
    BSM_HIP(hipSetDevice(b->ctx->device));
    if (!b->dirty && layouts_changed(b)) {
        BSM_HIP(hipStreamSynchronize(b->stream));
        if (bsm_batch_sync(b)) return 1;
        b->dirty = true;
    }
    if (b->dirty && batch_plan(b)) return 1;
    return stage_info(b->stages, max_stages, out, n_stages);
}

int bsm_batch_set_tuning(bsm_batch* b, const bsm_tuning* t) {
    if (!b || !t) return fail("null arg");
    Tuning nt;
    if (merge_tuning(b->tune, t, &nt)) return 1;
    BSM_HIP(hipStreamSynchronize(b->stream));
    b->tune = nt;
    b->dirty = true;
    return 0;
}


int bsm_batch_set_input_ranges(bsm_batch* b, const uint32_t* axn_offset, const uint32_t* len) {
    if (!b || !axn_offset || !len) return fail("null arg");
    if (!b->dynamic) return fail("input ranges apply to dynamic batches");
    BSM_HIP(hipSetDevice(b->ctx->device));
    BSM_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->areas.size(); i++) {
        bsm_area* a = b->areas[i];
        if (axn_offset[i] > a->axn_len || len[i] > a->axn_len - axn_offset[i])
            return fail("input range out of the axon space of area " + std::to_string(i));
    }
    for (size_t i = 0; i < b->areas.size(); i++) {
        bsm_area* a = b->areas[i];
        b->in_ofs[i] = axn_offset[i];
        b->in_len[i] = len[i];
        b->io[i].dst = a->axn + axn_offset[i];
        b->io[i].len = len[i];
    }
    BSM_HIP(hipMemcpy(b->io_d, b->io.data(), sizeof(AreaIo) * b->io.size(), hipMemcpyHostToDevice));
    return 0;
}

int bsm_batch_cycle_masked(bsm_batch* b, const uint8_t* active, const uint8_t* frames,
                           const uint64_t* frame_ofs, uint64_t frame_bytes, int use_graph) {
    if (!b || !active) return fail("null arg");
    if (!b->dynamic) return fail("masked cycles need a dynamic batch");
    BSM_HIP(hipSetDevice(b->ctx->device));
    const size_t n = b->areas.size();
    for (size_t i = 0; frame_ofs && i < n; i++) {
        const uint64_t fo = frame_ofs[i];
        if (fo == ~0ull) continue;
        if (fo & BSM_FRAME_SDR) {
            // every index is checked here: the prologue writes where they point
            const uint64_t at = fo & ~BSM_FRAME_SDR;
            if ((at & 3) || at > frame_bytes || frame_bytes - at < 4)
                return fail("SDR entry of area " + std::to_string(i) +
                            " misaligned or out of the packed buffer");
            const uint32_t* e = (const uint32_t*)(frames + at);
            const uint32_t cnt = e[0];
            if ((frame_bytes - at - 4) / 4 < cnt)
                return fail("SDR entry of area " + std::to_string(i) + " runs past the buffer");
            const uint32_t len = b->io[i].len;
            for (uint32_t k = 0; k < cnt; k++)
                if (e[1 + k] >= len)
                    return fail("SDR index of area " + std::to_string(i) +
                                " beyond its input range");
        } else if (fo > frame_bytes || b->io[i].len > frame_bytes - fo) {
            return fail("frame of area " + std::to_string(i) + " out of the packed buffer");
        }
    }
    if (n == 0) return 0;
    if (!b->dirty && layouts_changed(b)) {
        // as in run_cycles: a layout write may have changed which gather path is exact
        BSM_HIP(hipStreamSynchronize(b->stream));
        if (bsm_batch_sync(b)) return 1;
        b->dirty = true;
    }
    if (b->dirty && batch_plan(b)) return 1;
    if (use_graph && !b->exec[0] && capture(b)) return 1;
    // Without a graph the cycle launches only the active areas' work items: the
    // transfer carries each stage's items of active areas (in the plan's order,
    // largest first), so a cycle's launches scale with the areas it steps.
    const bool compact = !use_graph;
    std::vector<Stage> cst;
    size_t n_items = 0;
    if (compact) {
        cst = b->stages;
        for (size_t si = 0; si < cst.size(); si++) {
            const auto& g = b->by_area[si];
            uint32_t m = 0;
            for (uint32_t ai : g.order) m += active[ai] ? g.count[ai] : 0u;
            Stage& st = cst[si];
            st.item0 = (uint32_t)n_items;
            st.n_items = m;
            st.grid = std::min<uint32_t>(m, 65535);
            n_items += m;
        }
    }
    const size_t hdr = xfer_hdr(n);
    const size_t items_at = (hdr + frame_bytes + 15) & ~(size_t)15;
    const size_t total = compact ? items_at + n_items * sizeof(WorkItem) : hdr + frame_bytes;
    if (grow_xfer(b, total - hdr)) return 1;
    const int k = b->xfer_next;
    b->xfer_next ^= 1;
    BSM_HIP(hipEventSynchronize(b->xfer_ev[k]));
    if (busy_harvest(b, k)) return 1;
    if (b->xfer_h_cap[k] < total) {
        if (b->xfer_h[k]) BSM_HIP(hipHostFree(b->xfer_h[k]));
        b->xfer_h[k] = nullptr;
        const size_t cap = std::max(total, (size_t)1 << 16);
        BSM_HIP(hipHostMalloc((void**)&b->xfer_h[k], cap, hipHostMallocDefault));
        b->xfer_h_cap[k] = cap;
    }
    uint8_t* h = b->xfer_h[k];
    std::memcpy(h, active, n);
    uint64_t* fo = (uint64_t*)(h + ((n + 15) & ~(size_t)15));
    for (size_t i = 0; i < n; i++) fo[i] = frame_ofs ? frame_ofs[i] : ~0ull;
    if (frame_bytes) std::memcpy(h + hdr, frames, frame_bytes);
    if (compact) {
        WorkItem* out = (WorkItem*)(h + items_at);
        size_t q = 0;
        for (size_t si = 0; si < cst.size(); si++) {
            const auto& g = b->by_area[si];
            for (uint32_t ai : g.order) {
                if (!active[ai]) continue;
                const uint32_t* ix = g.idx.data() + g.first[ai];
                for (uint32_t j = 0; j < g.count[ai]; j++) out[q++] = b->items_h[ix[j]];
            }
        }
        if (q != n_items) return fail("masked cycle: work item count changed");
    }
    BSM_HIP(hipEventRecord(b->busy_ev[k][0], b->stream));
    BSM_HIP(hipMemcpyAsync(b->xfer_d, h, total, hipMemcpyHostToDevice, b->stream));
    BSM_HIP(hipEventRecord(b->xfer_ev[k], b->stream));
    if (run_cycles(b, 1, use_graph, nullptr, compact ? &cst : nullptr,
                   compact ? (const WorkItem*)(b->xfer_d + items_at) : nullptr))
        return 1;
    if (sample_after_cycle(b->sample_plan, b->stream, b->areas.data(), n, active, b->dyn_d))
        return 1;
    BSM_HIP(hipEventRecord(b->busy_ev[k][1], b->stream));
    b->busy_rec[k] = true;
    return 0;
}

int bsm_batch_busy_ms(bsm_batch* b, double* ms) {
    if (!b || !ms) return fail("null arg");
    BSM_HIP(hipSetDevice(b->ctx->device));
    for (int k = 0; k < 2; k++)
        if (busy_harvest(b, k)) return 1;
    *ms = b->busy_ms;
    return 0;
}

int bsm_batch_regrow(bsm_batch* b, const uint8_t* which, const uint64_t* seeds,
                     const uint64_t* epochs, int32_t reach) {
    if (!b || !seeds || !epochs) return fail("null arg");
    if (reach < 0 || reach > 127) return fail("reach must be in [0, 127]");
    BSM_HIP(hipSetDevice(b->ctx->device));
    // enqueued on the batch stream: after every cycle enqueued so far (the side
    // stream joins the batch stream within each cycle), before the next one
    for (size_t i = 0; i < b->areas.size(); i++) {
        if (which && !which[i]) continue;
        bsm_area* a = b->areas[i];
        bool widened = false;
        if (regrow_launch(a, b->stream, seeds[i], epochs[i], reach, &widened)) return 1;
        if (widened) {
            // the next batched cycle re-plans (layout epochs differ from the plan's)
            a->plan_dirty = true;
            a->layout_epoch++;
        }
    }
    return 0;
}

int bsm_batch_query(bsm_batch* b, int* done) {
    if (!b || !done) return fail("null arg");
    hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) { *done = 1; return 0; }
    if (e == hipErrorNotReady) { *done = 0; return 0; }
    return hip_fail(e, "hipStreamQuery");
}

int bsm_batch_set_tracts(bsm_batch* b, const bsm_tract* tracts, uint32_t n) {
    if (!b || (!tracts && n)) return fail("null arg");
    if (b->dynamic && n) return fail("tracts apply to static batches");
    BSM_HIP(hipSetDevice(b->ctx->device));
    BSM_HIP(hipStreamSynchronize(b->stream));
    std::vector<TractRt> rt(n);
    for (uint32_t i = 0; i < n; i++) {
        const bsm_tract& t = tracts[i];
        if (t.src_area >= b->areas.size() || t.dst_area >= b->areas.size())
            return fail("tract " + std::to_string(i) + ": area index out of range");
        bsm_area* sa = b->areas[t.src_area];
        bsm_area* da = b->areas[t.dst_area];
        if (t.src_offset > sa->axn_len || t.len > sa->axn_len - t.src_offset ||
            t.dst_offset > da->axn_len || t.len > da->axn_len - t.dst_offset)
            return fail("tract " + std::to_string(i) + ": range out of the axon space");
        if (t.src_area == t.dst_area && t.src_offset < t.dst_offset + t.len &&
            t.dst_offset < t.src_offset + t.len)
            return fail("tract " + std::to_string(i) + ": overlapping ranges");
        rt[i] = TractRt{sa->axn + t.src_offset, da->axn + t.dst_offset, t.len, 0};
        b->tract_parts = std::max(b->tract_parts, (t.len + COPY_SEG - 1) / COPY_SEG);
    }
    (void)hipFree(b->tracts_d);
    b->tracts_d = nullptr;
    if (n) {
        BSM_HIP(hipMalloc((void**)&b->tracts_d, sizeof(TractRt) * n));
        BSM_HIP(hipMemcpy(b->tracts_d, rt.data(), sizeof(TractRt) * n, hipMemcpyHostToDevice));
    }
    b->n_tracts = n;
    if (!n) b->tract_parts = 1;
    free_graphs(b);
    return 0;
}

}  // extern "C"
