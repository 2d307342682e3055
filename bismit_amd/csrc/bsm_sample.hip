// bsm_sample.hip -- samplers: per-cycle readback of area buffers to the host.
//
// The reference enqueues one async buffer read per sampler per cycle, ordered
// after the kernels that write the buffer, into a single-buffer tract channel
// (CorticalArea::cycle_samplers, src/cortex/cortical_area.rs:808-991;
// TractInner::send/recv, src/subcortex/tract_channel.rs:381-480).  Here, after a
// cycle, one gather kernel packs the ranges of every sampler due this cycle into
// a staging buffer -- resolving the state parity of double-buffered buffers and
// composing the previous-active flag bit on the device -- and one copy moves the
// packed frames into a pinned host block.  The block, with the event that marks
// it complete, is shared by the frames it holds and returns to a per-context
// pool when the last of them has been received or dropped.  A sampler holds up
// to 1, 2 or 3 frames in order (the reference's SamplerBufferKind Single,
// Double, Triple, src/cortex/cortical_area.rs:52-58, of which it implements
// Single): a frame is sent when a slot is free; otherwise it is skipped, or,
// with backpressure, the cycle waits for the receiver.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bsm_internal.h"

namespace bsm {

constexpr uint32_t SAMPLE_MAX = 256;        // samplers per gather launch
constexpr uint32_t SAMPLE_SEG = 65536;      // bytes per gather block
constexpr size_t SAMPLE_ALIGN = 256;        // frame alignment in the staging buffer
constexpr size_t BLOCK_MIN = 64u << 10;     // smallest pooled host block

struct SampleDesc {
    const uint8_t* src[2];      // by layer parity (equal when not double-buffered)
    const uint8_t* prev[2];     // BSM_SYN_FLAG_SETS: previous states by parity, else null
    const uint32_t* par_word;   // dynamic batches: parity word on the device, else null
    uint64_t len;
};

struct SampleArgs {
    uint32_t n;
    uint32_t par_bits[SAMPLE_MAX / 32];   // host parity of launch entry j
    uint16_t idx[SAMPLE_MAX];             // descriptor of launch entry j
    uint32_t dst[SAMPLE_MAX];             // staging offset of entry j, in SAMPLE_ALIGN units
};

// Blocks (entry j, part): copy part `part` of entry j's range into the staging
// buffer at the entry's offset.
__global__ __launch_bounds__(256) void k_sample(const SampleDesc* __restrict__ desc,
                                                uint8_t* __restrict__ stage, SampleArgs args) {
    const uint32_t j = blockIdx.x;
    const SampleDesc d = desc[args.idx[j]];
    const uint64_t lo = (uint64_t)blockIdx.y * SAMPLE_SEG;
    if (lo >= d.len) return;
    const uint32_t par = d.par_word ? (*d.par_word & 1u)
                                    : ((args.par_bits[j >> 5] >> (j & 31u)) & 1u);
    const uint32_t n = (uint32_t)min((uint64_t)SAMPLE_SEG, d.len - lo);
    const uint8_t* src = d.src[par] + lo;
    uint8_t* dst = stage + (uint64_t)args.dst[j] * SAMPLE_ALIGN + lo;
    if (!d.prev[0]) {
        block_copy(dst, src, n, threadIdx.x, blockDim.x);
        return;
    }
    auto s8 = BSM_G(const uint8_t, src);
    auto p8 = BSM_G(const uint8_t, d.prev[par] + lo);
    auto d8 = BSM_G(uint8_t, dst);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        d8[i] = (uint8_t)((s8[i] & ~SYN_PREV_ACTIVE) | (p8[i] ? SYN_PREV_ACTIVE : 0));
}

struct FramePool {
    std::mutex m;
    std::map<size_t, std::vector<uint8_t*>> free_;   // capacity -> pinned host blocks
    std::vector<hipEvent_t> events;
};

// One cycle's packed frames in pinned host memory, and their completion.
struct FrameBlock {
    FramePool* pool;
    uint8_t* host;
    size_t cap;
    hipEvent_t ev;
    ~FrameBlock() {
        // the copy into the block must have ended before another cycle reuses it
        (void)hipEventSynchronize(ev);
        std::lock_guard<std::mutex> g(pool->m);
        pool->free_[cap].push_back(host);
        pool->events.push_back(ev);
    }
};

struct SamplePlan {
    std::vector<bsm_sampler*> smp;
    std::vector<uint32_t> area_of;        // index into the areas of the launch
    std::vector<std::pair<bsm_area*, uint64_t>> key;   // areas and their sampler generations
    const uint32_t* dyn = nullptr;
    SampleDesc* desc_d = nullptr;
    size_t desc_cap = 0;
    uint8_t* stage_d = nullptr;           // packed frames of one cycle
    size_t stage_cap = 0;
};

enum SampleState { SMP_EMPTY = 0, SMP_FILLING = 1, SMP_PENDING = 2 };
constexpr uint32_t SMP_SLOTS = 3;           // frames a sampler holds at most

// One held frame: being filled (its gather and copy enqueued) or pending.
struct SampleSlot {
    int state = SMP_EMPTY;
    uint64_t cycle = 0;
    std::shared_ptr<FrameBlock> blk;        // the block holding the frame
    size_t off = 0;
};

}  // namespace bsm

struct bsm_sampler {
    bsm_area* area;
    int buf;
    uint32_t layer;
    uint64_t offset, len, bytes;          // bytes: len rounded up to the staging alignment
    std::mutex m;
    std::condition_variable cv;
    bool backpressure = false;
    bool closed = false;
    uint32_t depth = 1;                     // frames held at most (1..SMP_SLOTS)
    uint32_t head = 0, count = 0;           // oldest held frame, frames held
    bsm::SampleSlot slot[bsm::SMP_SLOTS];
    bsm::SampleSlot& oldest() { return slot[head]; }
    bool has_frame() { return count > 0 && slot[head].state == bsm::SMP_PENDING; }
    void pop() {
        slot[head] = bsm::SampleSlot();
        head = (head + 1) % bsm::SMP_SLOTS;
        count--;
    }
    void clear() {
        for (auto& x : slot) x = bsm::SampleSlot();
        head = count = 0;
    }
};

namespace bsm {

static FramePool* pool_of(bsm_ctx* c) {
    static std::mutex init;
    std::lock_guard<std::mutex> g(init);
    if (!c->frames) c->frames = new FramePool();
    return c->frames;
}

static int frame_block(bsm_ctx* c, size_t bytes, std::shared_ptr<FrameBlock>* out) {
    FramePool* P = pool_of(c);
    size_t cap = BLOCK_MIN;
    while (cap < bytes) cap <<= 1;
    uint8_t* host = nullptr;
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> g(P->m);
        auto& v = P->free_[cap];
        if (!v.empty()) {
            host = v.back();
            v.pop_back();
        }
        if (!P->events.empty()) {
            ev = P->events.back();
            P->events.pop_back();
        }
    }
    if (!host) {
        hipError_t e = hipHostMalloc((void**)&host, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            if (ev) {
                std::lock_guard<std::mutex> g(P->m);
                P->events.push_back(ev);
            }
            return hip_fail(e, "hipHostMalloc (sampler frames)");
        }
    }
    if (!ev) {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> g(P->m);
            P->free_[cap].push_back(host);
            return hip_fail(e, "hipEventCreateWithFlags");
        }
    }
    auto b = std::make_shared<FrameBlock>();
    b->pool = P;
    b->host = host;
    b->cap = cap;
    b->ev = ev;
    *out = std::move(b);
    return 0;
}

static int make_desc(bsm_sampler* sp, const uint32_t* dyn_area, SampleDesc* d) {
    bsm_area* a = sp->area;
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, sp->buf, sp->layer, &p, &n)) return 1;
    *d = SampleDesc{};
    d->len = sp->len;
    d->par_word = dyn_area ? dyn_area + sp->layer : nullptr;
    if (sp->buf == BSM_SYN_STATES) {
        const LayerRt& L = a->L[sp->layer];
        d->src[0] = L.syn_state[0] + sp->offset;
        d->src[1] = L.syn_state[1] + sp->offset;
    } else if (sp->buf == BSM_SYN_FLAG_SETS) {
        const LayerRt& L = a->L[sp->layer];
        d->src[0] = d->src[1] = p + sp->offset;
        d->prev[0] = L.syn_state[1] + sp->offset;
        d->prev[1] = L.syn_state[0] + sp->offset;
    } else {
        d->src[0] = d->src[1] = p + sp->offset;
        d->par_word = nullptr;   // single-buffered: parity does not matter
    }
    return 0;
}

static bool plan_stale(const SamplePlan* P, bsm_area* const* areas, size_t n,
                       const uint32_t* dyn) {
    if (!P || P->dyn != dyn || P->key.size() != n) return true;
    for (size_t i = 0; i < n; i++)
        if (P->key[i].first != areas[i] || P->key[i].second != areas[i]->sampler_gen) return true;
    return false;
}

static int plan_build(SamplePlan*& P, bsm_area* const* areas, size_t n, const uint32_t* dyn) {
    if (!P) P = new SamplePlan();
    P->smp.clear();
    P->area_of.clear();
    P->key.clear();
    P->dyn = dyn;
    size_t total = 0;
    for (size_t i = 0; i < n; i++) {
        P->key.emplace_back(areas[i], areas[i]->sampler_gen);
        for (bsm_sampler* sp : areas[i]->samplers) {
            P->smp.push_back(sp);
            P->area_of.push_back((uint32_t)i);
            total += sp->bytes;
        }
    }
    if (P->smp.size() > 65535) return fail("too many samplers in one launch set");
    if (total / SAMPLE_ALIGN > 0xFFFFFFFFull) return fail("sampled bytes exceed the staging range");
    if (P->smp.empty()) return 0;
    if (P->desc_cap < P->smp.size()) {
        if (P->desc_d) BSM_HIP(hipFree(P->desc_d));
        P->desc_d = nullptr;
        P->desc_cap = std::max<size_t>(P->smp.size(), 64);
        BSM_HIP(hipMalloc((void**)&P->desc_d, P->desc_cap * sizeof(SampleDesc)));
    }
    if (P->stage_cap < total) {
        if (P->stage_d) BSM_HIP(hipFree(P->stage_d));
        P->stage_d = nullptr;
        P->stage_cap = std::max(total, BLOCK_MIN);
        BSM_HIP(hipMalloc((void**)&P->stage_d, P->stage_cap));
    }
    std::vector<SampleDesc> h(P->smp.size());
    for (size_t k = 0; k < h.size(); k++) {
        const uint32_t* da = dyn ? dyn + (size_t)P->area_of[k] * BSM_MAX_LAYERS : nullptr;
        if (make_desc(P->smp[k], da, &h[k])) return 1;
    }
    BSM_HIP(hipMemcpy(P->desc_d, h.data(), h.size() * sizeof(SampleDesc), hipMemcpyHostToDevice));
    return 0;
}

int sample_after_cycle(SamplePlan*& P, hipStream_t s, bsm_area* const* areas, size_t n,
                       const uint8_t* active, const uint32_t* dyn) {
    bool any = false;
    for (size_t i = 0; i < n; i++) {
        if (active && !active[i]) continue;
        areas[i]->cycles_done++;
        any |= !areas[i]->samplers.empty();
    }
    if (!any) return 0;
    if (plan_stale(P, areas, n, dyn) && plan_build(P, areas, n, dyn)) return 1;
    // which samplers send this cycle (TractInner::send)
    std::vector<uint32_t> inc;
    std::vector<size_t> at;   // staging offset of each included frame
    size_t total = 0;
    for (uint32_t k = 0; k < P->smp.size(); k++) {
        const uint32_t i = P->area_of[k];
        if (active && !active[i]) continue;
        bsm_sampler* sp = P->smp[k];
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) continue;
        if (sp->count >= sp->depth) {
            if (!sp->backpressure) continue;   // frame skipped: the receiver keeps the older ones
            sp->cv.wait(lk, [&] { return sp->count < sp->depth || sp->closed || !sp->backpressure; });
            if (sp->closed || sp->count >= sp->depth) continue;
        }
        SampleSlot& x = sp->slot[(sp->head + sp->count) % SMP_SLOTS];
        sp->count++;
        x.state = SMP_FILLING;
        x.cycle = areas[i]->cycles_done;
        inc.push_back(k);
        at.push_back(total);
        total += sp->bytes;
    }
    if (inc.empty()) return 0;
    auto reset = [&] {
        for (uint32_t k : inc) {
            bsm_sampler* sp = P->smp[k];
            std::lock_guard<std::mutex> g(sp->m);
            if (!sp->closed && sp->count) {
                // the frame reserved last is the one that will not come
                sp->slot[(sp->head + sp->count - 1) % SMP_SLOTS] = SampleSlot();
                sp->count--;
            }
            sp->cv.notify_all();
        }
    };
    // virtual synapse states of sampled ranges into the state buffers (both
    // parities: states read the current one, flags the previous one)
    for (uint32_t k : inc) {
        const bsm_sampler* sp = P->smp[k];
        if ((sp->buf == BSM_SYN_STATES || sp->buf == BSM_SYN_FLAG_SETS) &&
            settle_states(areas[P->area_of[k]], sp->layer, s, sp->offset, sp->len, false)) {
            reset();
            return 1;
        }
    }
    for (size_t g0 = 0; g0 < inc.size(); g0 += SAMPLE_MAX) {
        SampleArgs args{};
        args.n = (uint32_t)std::min<size_t>(SAMPLE_MAX, inc.size() - g0);
        uint64_t parts = 1;
        for (uint32_t j = 0; j < args.n; j++) {
            const uint32_t k = inc[g0 + j];
            const bsm_sampler* sp = P->smp[k];
            args.idx[j] = (uint16_t)k;
            args.dst[j] = (uint32_t)(at[g0 + j] / SAMPLE_ALIGN);
            const uint32_t par = sp->buf == BSM_AXN_STATES
                                     ? 0u : (uint32_t)(areas[P->area_of[k]]->L[sp->layer].cur & 1);
            args.par_bits[j >> 5] |= par << (j & 31u);
            parts = std::max<uint64_t>(parts, (sp->len + SAMPLE_SEG - 1) / SAMPLE_SEG);
        }
        k_sample<<<dim3(args.n, (uint32_t)parts), 256, 0, s>>>(P->desc_d, P->stage_d, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            reset();
            return hip_fail(e, "k_sample");
        }
    }
    std::shared_ptr<FrameBlock> blk;
    if (frame_block(areas[0]->ctx, total, &blk)) {
        reset();
        return 1;
    }
    hipError_t e = hipMemcpyAsync(blk->host, P->stage_d, total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(blk->ev, s);
    if (e != hipSuccess) {
        reset();
        return hip_fail(e, "sampler frame copy");
    }
    for (size_t q = 0; q < inc.size(); q++) {
        bsm_sampler* sp = P->smp[inc[q]];
        std::lock_guard<std::mutex> g(sp->m);
        if (sp->closed || !sp->count) continue;
        SampleSlot& x = sp->slot[(sp->head + sp->count - 1) % SMP_SLOTS];
        x.blk = blk;
        x.off = at[q];
        x.state = SMP_PENDING;
        sp->cv.notify_all();
    }
    return 0;
}

void sample_plan_free(SamplePlan*& P) {
    if (!P) return;
    if (P->desc_d) (void)hipFree(P->desc_d);
    if (P->stage_d) (void)hipFree(P->stage_d);
    delete P;
    P = nullptr;
}

// Ends the sampler's frames; a receiver copying one keeps its block alive.
static void sampler_close(bsm_sampler* sp) {
    {
        std::lock_guard<std::mutex> g(sp->m);
        if (sp->closed) return;
        sp->closed = true;
        sp->clear();             // pending frames end unreceived
        sp->cv.notify_all();
    }
    bsm_area* a = sp->area;
    auto& v = a->samplers;
    v.erase(std::remove(v.begin(), v.end(), sp), v.end());
    a->sampler_gen = ++a->ctx->sampler_gen;
}

void samplers_detach(bsm_area* a) {
    while (!a->samplers.empty()) sampler_close(a->samplers.back());
    sample_plan_free(a->sample_plan);
}

void frame_pool_destroy(bsm_ctx* c) {
    if (!c->frames) return;
    for (auto& kv : c->frames->free_)
        for (uint8_t* h : kv.second) (void)hipHostFree(h);
    for (hipEvent_t e : c->frames->events) (void)hipEventDestroy(e);
    delete c->frames;
    c->frames = nullptr;
}

}  // namespace bsm

using namespace bsm;

extern "C" {

int bsm_sampler_create(bsm_area* a, int buf, uint32_t layer, uint64_t offset, uint64_t len,
                       int backpressure, bsm_sampler** out) {
    if (!a || !out) return fail("null arg");
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, buf, layer, &p, &n)) return 1;
    if (len == 0) return fail("empty sampler range");
    if (offset > n || len > n - offset) return fail("sampler range out of the buffer");
    if ((len + SAMPLE_SEG - 1) / SAMPLE_SEG > 65535) return fail("sampler range too long");
    bsm_sampler* sp = new bsm_sampler();
    sp->area = a;
    sp->buf = buf;
    sp->layer = buf == BSM_AXN_STATES ? 0u : layer;
    sp->offset = offset;
    sp->len = len;
    sp->bytes = (len + SAMPLE_ALIGN - 1) & ~(uint64_t)(SAMPLE_ALIGN - 1);
    sp->backpressure = backpressure != 0;
    a->samplers.push_back(sp);
    a->sampler_gen = ++a->ctx->sampler_gen;
    *out = sp;
    return 0;
}

int bsm_sampler_recv(bsm_sampler* sp, int wait_for_frame, void* dst, uint64_t dst_len,
                     uint64_t* cycle, int* got) {
    if (!sp || !got) return fail("null arg");
    *got = 0;
    if (dst_len < sp->len || !dst) return fail("destination shorter than the sampler range");
    std::shared_ptr<FrameBlock> blk;
    size_t off;
    uint64_t fc;
    {
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) return 0;
        if (!sp->has_frame()) {
            if (!wait_for_frame) return 0;
            sp->cv.wait(lk, [&] { return sp->has_frame() || sp->closed; });
            if (sp->closed || !sp->has_frame()) return 0;
        }
        blk = sp->oldest().blk;
        off = sp->oldest().off;
        fc = sp->oldest().cycle;
    }
    hipError_t e = hipEventSynchronize(blk->ev);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize (sampler frame)");
    std::memcpy(dst, blk->host + off, sp->len);
    {
        std::lock_guard<std::mutex> g(sp->m);
        if (!sp->closed && sp->count && sp->oldest().blk == blk && sp->oldest().off == off)
            sp->pop();
        sp->cv.notify_all();
    }
    if (cycle) *cycle = fc;
    *got = 1;
    return 0;
}

int bsm_sampler_recv_many(bsm_sampler* const* s, uint32_t n, int wait_for_frame,
                          void* const* dst, const uint64_t* dst_len, uint64_t* cycles, int* got) {
    if ((!s || !dst || !dst_len || !got) && n) return fail("null arg");
    // take every pending frame (waiting where asked), then copy them out --
    // split over threads when there is much to copy -- then release them
    struct Take {
        uint32_t i;
        std::shared_ptr<FrameBlock> blk;
        size_t off;
        uint64_t fc;
    };
    std::vector<Take> takes;
    size_t bytes = 0;
    for (uint32_t i = 0; i < n; i++) {
        got[i] = 0;
        bsm_sampler* sp = s[i];
        if (!sp) return fail("null sampler");
        if (dst_len[i] < sp->len || !dst[i]) return fail("destination shorter than the sampler range");
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) continue;
        if (!sp->has_frame()) {
            if (!wait_for_frame) continue;
            sp->cv.wait(lk, [&] { return sp->has_frame() || sp->closed; });
            if (sp->closed || !sp->has_frame()) continue;
        }
        takes.push_back(Take{i, sp->oldest().blk, sp->oldest().off, sp->oldest().cycle});
        bytes += sp->len;
    }
    const FrameBlock* last = nullptr;
    for (const Take& t : takes) {
        if (t.blk.get() == last) continue;
        last = t.blk.get();
        hipError_t e = hipEventSynchronize(t.blk->ev);
        if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize (sampler frame)");
    }
    auto copy = [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; k++)
            std::memcpy(dst[takes[k].i], takes[k].blk->host + takes[k].off, s[takes[k].i]->len);
    };
    const size_t nt = bytes >= ((size_t)4 << 20) ? std::min<size_t>(8, takes.size()) : 1;
    if (nt > 1) {
        std::vector<std::thread> pool;
        for (size_t w = 0; w < nt; w++)
            pool.emplace_back(copy, takes.size() * w / nt, takes.size() * (w + 1) / nt);
        for (std::thread& th : pool) th.join();
    } else {
        copy(0, takes.size());
    }
    for (const Take& t : takes) {
        bsm_sampler* sp = s[t.i];
        {
            std::lock_guard<std::mutex> g(sp->m);
            if (!sp->closed && sp->count && sp->oldest().blk == t.blk && sp->oldest().off == t.off)
                sp->pop();
            sp->cv.notify_all();
        }
        if (cycles) cycles[t.i] = t.fc;
        got[t.i] = 1;
    }
    return 0;
}

int bsm_sampler_set_backpressure(bsm_sampler* sp, int backpressure, int* prior) {
    if (!sp) return fail("null sampler");
    std::lock_guard<std::mutex> g(sp->m);
    if (prior) *prior = sp->backpressure ? 1 : 0;
    sp->backpressure = backpressure != 0;
    sp->cv.notify_all();
    return 0;
}

int bsm_sampler_set_buffers(bsm_sampler* sp, uint32_t n_frames) {
    if (!sp) return fail("null sampler");
    if (n_frames < 1 || n_frames > SMP_SLOTS) return fail("a sampler holds 1 to 3 frames");
    std::lock_guard<std::mutex> g(sp->m);
    sp->depth = n_frames;
    sp->cv.notify_all();
    return 0;
}

int bsm_sampler_info(bsm_sampler* sp, uint64_t* len, int* backpressure, int* pending) {
    if (!sp) return fail("null sampler");
    std::lock_guard<std::mutex> g(sp->m);
    if (len) *len = sp->len;
    if (backpressure) *backpressure = sp->backpressure ? 1 : 0;
    if (pending) *pending = sp->has_frame() ? 1 : 0;
    return 0;
}

int bsm_sampler_destroy(bsm_sampler* sp) {
    if (!sp) return 0;
    if (!sp->closed) sampler_close(sp);
    delete sp;
    return 0;
}

}  // extern "C"
