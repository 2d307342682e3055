// bsm_sample.hip -- samplers: per-cycle readback of area buffers to the host.
//
// The reference enqueues one async buffer read per sampler per cycle, ordered
// after the kernels that write the buffer, into a single-buffer tract channel
// (CorticalArea::cycle_samplers, src/cortex/cortical_area.rs:808-991;
// TractInner::send/recv, src/subcortex/tract_channel.rs:381-480).  Here every
// sampler owns a slot in a per-context arena of matched device and pinned host
// chunks.  After a cycle, one gather kernel copies the ranges of all samplers
// due this cycle into their device slots -- resolving the state parity of
// double-buffered buffers and composing the previous-active flag bit on the
// device -- and one copy per contiguous run of slots moves them to the host.
// A shared event marks the frames complete; a receiver waits for it only.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "bsm_internal.h"

namespace bsm {

constexpr uint32_t SAMPLE_MAX = 256;        // samplers per gather launch
constexpr uint32_t SAMPLE_SEG = 65536;      // bytes per gather block
constexpr size_t SAMPLE_ALIGN = 256;
constexpr size_t SAMPLE_CHUNK = 8u << 20;   // default arena chunk

struct SampleDesc {
    const uint8_t* src[2];      // by layer parity (equal when not double-buffered)
    const uint8_t* prev[2];     // BSM_SYN_FLAG_SETS: previous states by parity, else null
    uint8_t* dst;               // device slot
    const uint32_t* par_word;   // dynamic batches: parity word on the device, else null
    uint64_t len;
};

struct SampleArgs {
    uint32_t n;
    uint32_t par_bits[SAMPLE_MAX / 32];   // host parity of launch entry j
    uint16_t idx[SAMPLE_MAX];             // descriptor of launch entry j
};

// Blocks (entry j, part): copy part `part` of entry j's range into its slot.
__global__ __launch_bounds__(256) void k_sample(const SampleDesc* __restrict__ desc,
                                                SampleArgs args) {
    const uint32_t j = blockIdx.x;
    const SampleDesc d = desc[args.idx[j]];
    const uint64_t lo = (uint64_t)blockIdx.y * SAMPLE_SEG;
    if (lo >= d.len) return;
    const uint32_t par = d.par_word ? (*d.par_word & 1u)
                                    : ((args.par_bits[j >> 5] >> (j & 31u)) & 1u);
    const uint32_t n = (uint32_t)min((uint64_t)SAMPLE_SEG, d.len - lo);
    const uint8_t* src = d.src[par] + lo;
    uint8_t* dst = d.dst + lo;
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

struct SampleChunk {
    uint8_t* dev = nullptr;
    uint8_t* host = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes
};

struct SampleArena {
    std::mutex m;
    std::vector<SampleChunk> chunks;
    std::vector<hipEvent_t> ev_pool;
};

// Completion of one cycle's readback, shared by the frames it carries.
struct FrameEvent {
    SampleArena* arena;
    hipEvent_t ev;
    ~FrameEvent() {
        std::lock_guard<std::mutex> g(arena->m);
        arena->ev_pool.push_back(ev);
    }
};

struct SamplePlan {
    std::vector<bsm_sampler*> smp;        // arena order
    std::vector<uint32_t> area_of;        // index into the areas of the launch
    std::vector<std::pair<bsm_area*, uint64_t>> key;   // areas and their sampler generations
    const uint32_t* dyn = nullptr;
    SampleDesc* desc_d = nullptr;
    size_t desc_cap = 0;
};

enum SampleState { SMP_EMPTY = 0, SMP_FILLING = 1, SMP_PENDING = 2 };

}  // namespace bsm

struct bsm_sampler {
    bsm_area* area;
    int buf;
    uint32_t layer;
    uint64_t offset, len;
    uint32_t chunk;
    size_t off, bytes;
    uint8_t* dev_slot;    // this sampler's device and pinned host slots
    uint8_t* host_slot;
    std::mutex m;
    std::condition_variable cv;
    int state = bsm::SMP_EMPTY;
    bool backpressure = false;
    bool closed = false;
    bool reading = false;   // a receiver is copying the frame out of the host slot
    uint64_t frame_cycle = 0;
    std::shared_ptr<bsm::FrameEvent> ev;
};

namespace bsm {

static int arena_alloc(bsm_ctx* c, size_t bytes, uint32_t* chunk, size_t* off, uint8_t** dev,
                       uint8_t** host) {
    if (!c->samples) c->samples = new SampleArena();
    SampleArena& A = *c->samples;
    std::lock_guard<std::mutex> g(A.m);
    for (uint32_t k = 0; k < A.chunks.size(); k++) {
        auto& fl = A.chunks[k].free_;
        for (auto it = fl.begin(); it != fl.end(); ++it) {
            if (it->second < bytes) continue;
            *chunk = k;
            *off = it->first;
            const size_t o = it->first + bytes, rest = it->second - bytes;
            fl.erase(it);
            if (rest) fl[o] = rest;
            *dev = A.chunks[k].dev + *off;
            *host = A.chunks[k].host + *off;
            return 0;
        }
    }
    SampleChunk ch;
    ch.size = std::max(SAMPLE_CHUNK, (bytes + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1));
    BSM_HIP(hipMalloc((void**)&ch.dev, ch.size));
    hipError_t e = hipHostMalloc((void**)&ch.host, ch.size, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(ch.dev);
        return hip_fail(e, "hipHostMalloc");
    }
    if (ch.size > bytes) ch.free_[bytes] = ch.size - bytes;
    *dev = ch.dev;
    *host = ch.host;
    A.chunks.push_back(std::move(ch));
    *chunk = (uint32_t)(A.chunks.size() - 1);
    *off = 0;
    return 0;
}

static void arena_free(bsm_ctx* c, uint32_t chunk, size_t off, size_t bytes) {
    SampleArena& A = *c->samples;
    std::lock_guard<std::mutex> g(A.m);
    auto& fl = A.chunks[chunk].free_;
    auto it = fl.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != fl.end() && it->first + it->second == nx->first) {   // merge with the next hole
        it->second += nx->second;
        fl.erase(nx);
    }
    if (it != fl.begin()) {                                        // and with the previous one
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            fl.erase(it);
        }
    }
}

static int frame_event(bsm_ctx* c, hipStream_t s, std::shared_ptr<FrameEvent>* out) {
    SampleArena& A = *c->samples;
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> g(A.m);
        if (!A.ev_pool.empty()) {
            ev = A.ev_pool.back();
            A.ev_pool.pop_back();
        }
    }
    if (!ev) BSM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    auto fe = std::make_shared<FrameEvent>();
    fe->arena = &A;
    fe->ev = ev;
    BSM_HIP(hipEventRecord(ev, s));
    *out = std::move(fe);
    return 0;
}

static int make_desc(bsm_sampler* sp, const uint32_t* dyn_area, uint8_t* slot, SampleDesc* d) {
    bsm_area* a = sp->area;
    uint8_t* p;
    size_t n;
    if (resolve_buf(a, sp->buf, sp->layer, &p, &n)) return 1;
    *d = SampleDesc{};
    d->dst = slot;
    d->len = sp->len;
    d->par_word = (dyn_area && sp->buf != BSM_AXN_STATES) ? dyn_area + sp->layer : nullptr;
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
    for (size_t i = 0; i < n; i++) {
        P->key.emplace_back(areas[i], areas[i]->sampler_gen);
        for (bsm_sampler* sp : areas[i]->samplers) {
            P->smp.push_back(sp);
            P->area_of.push_back((uint32_t)i);
        }
    }
    // arena order, so that the frames of consecutive samplers form one copy
    std::vector<uint32_t> ord(P->smp.size());
    for (uint32_t k = 0; k < ord.size(); k++) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
        const bsm_sampler *a = P->smp[x], *b = P->smp[y];
        return a->chunk != b->chunk ? a->chunk < b->chunk : a->off < b->off;
    });
    std::vector<bsm_sampler*> smp(ord.size());
    std::vector<uint32_t> area_of(ord.size());
    for (size_t k = 0; k < ord.size(); k++) {
        smp[k] = P->smp[ord[k]];
        area_of[k] = P->area_of[ord[k]];
    }
    P->smp.swap(smp);
    P->area_of.swap(area_of);
    if (P->smp.size() > 65535) return fail("too many samplers in one launch set");
    if (P->smp.empty()) return 0;
    if (P->desc_cap < P->smp.size()) {
        if (P->desc_d) BSM_HIP(hipFree(P->desc_d));
        P->desc_d = nullptr;
        P->desc_cap = std::max<size_t>(P->smp.size(), 64);
        BSM_HIP(hipMalloc((void**)&P->desc_d, P->desc_cap * sizeof(SampleDesc)));
    }
    std::vector<SampleDesc> h(P->smp.size());
    for (size_t k = 0; k < h.size(); k++) {
        bsm_sampler* sp = P->smp[k];
        const uint32_t* da = dyn ? dyn + (size_t)P->area_of[k] * BSM_MAX_LAYERS : nullptr;
        if (make_desc(sp, da, sp->dev_slot, &h[k])) return 1;
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
    for (uint32_t k = 0; k < P->smp.size(); k++) {
        const uint32_t i = P->area_of[k];
        if (active && !active[i]) continue;
        bsm_sampler* sp = P->smp[k];
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) continue;
        if (sp->state != SMP_EMPTY) {
            if (!sp->backpressure) continue;   // frame skipped: the receiver keeps the older one
            sp->cv.wait(lk, [&] { return sp->state == SMP_EMPTY || sp->closed || !sp->backpressure; });
            if (sp->closed || sp->state != SMP_EMPTY) continue;
        }
        sp->state = SMP_FILLING;
        sp->frame_cycle = areas[i]->cycles_done;
        inc.push_back(k);
    }
    if (inc.empty()) return 0;
    auto reset = [&] {
        for (uint32_t k : inc) {
            bsm_sampler* sp = P->smp[k];
            std::lock_guard<std::mutex> g(sp->m);
            sp->state = SMP_EMPTY;
            sp->cv.notify_all();
        }
    };
    bsm_ctx* c = P->smp[inc[0]]->area->ctx;
    for (size_t g0 = 0; g0 < inc.size(); g0 += SAMPLE_MAX) {
        SampleArgs args{};
        args.n = (uint32_t)std::min<size_t>(SAMPLE_MAX, inc.size() - g0);
        uint64_t parts = 1;
        for (uint32_t j = 0; j < args.n; j++) {
            const uint32_t k = inc[g0 + j];
            const bsm_sampler* sp = P->smp[k];
            args.idx[j] = (uint16_t)k;
            const uint32_t par = sp->buf == BSM_AXN_STATES
                                     ? 0u : (uint32_t)(areas[P->area_of[k]]->L[sp->layer].cur & 1);
            args.par_bits[j >> 5] |= par << (j & 31u);
            parts = std::max<uint64_t>(parts, (sp->len + SAMPLE_SEG - 1) / SAMPLE_SEG);
        }
        k_sample<<<dim3(args.n, (uint32_t)parts), 256, 0, s>>>(P->desc_d, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            reset();
            return hip_fail(e, "k_sample");
        }
    }
    // one copy per run of adjacent slots (inc is in arena order)
    for (size_t r = 0; r < inc.size();) {
        const bsm_sampler* first = P->smp[inc[r]];
        size_t end = first->off + first->bytes, q = r + 1;
        while (q < inc.size()) {
            const bsm_sampler* nx = P->smp[inc[q]];
            if (nx->chunk != first->chunk || nx->off != end) break;
            end = nx->off + nx->bytes;
            q++;
        }
        hipError_t e = hipMemcpyAsync(first->host_slot, first->dev_slot, end - first->off,
                                      hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) {
            reset();
            return hip_fail(e, "hipMemcpyAsync (sampler frames)");
        }
        r = q;
    }
    std::shared_ptr<FrameEvent> fe;
    if (frame_event(c, s, &fe)) {
        reset();
        return 1;
    }
    for (uint32_t k : inc) {
        bsm_sampler* sp = P->smp[k];
        std::lock_guard<std::mutex> g(sp->m);
        sp->ev = fe;
        sp->state = SMP_PENDING;
        sp->cv.notify_all();
    }
    return 0;
}

void sample_plan_free(SamplePlan*& P) {
    if (!P) return;
    if (P->desc_d) (void)hipFree(P->desc_d);
    delete P;
    P = nullptr;
}

// Ends the sampler's frames: waits for a frame in flight, frees its slot.
static void sampler_close(bsm_sampler* sp) {
    std::shared_ptr<FrameEvent> ev;
    {
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) return;
        sp->closed = true;
        ev = std::move(sp->ev);
        sp->state = SMP_EMPTY;   // a pending frame ends unreceived
        sp->cv.notify_all();
        sp->cv.wait(lk, [&] { return !sp->reading; });   // the slot is freed below
    }
    if (ev) (void)hipEventSynchronize(ev->ev);
    bsm_area* a = sp->area;
    arena_free(a->ctx, sp->chunk, sp->off, sp->bytes);
    auto& v = a->samplers;
    v.erase(std::remove(v.begin(), v.end(), sp), v.end());
    a->sampler_gen = ++a->ctx->sampler_gen;
}

void samplers_detach(bsm_area* a) {
    while (!a->samplers.empty()) sampler_close(a->samplers.back());
    sample_plan_free(a->sample_plan);
}

void sample_arena_destroy(bsm_ctx* c) {
    if (!c->samples) return;
    for (auto& ch : c->samples->chunks) {
        (void)hipFree(ch.dev);
        (void)hipHostFree(ch.host);
    }
    for (hipEvent_t e : c->samples->ev_pool) (void)hipEventDestroy(e);
    delete c->samples;
    c->samples = nullptr;
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
    BSM_HIP(hipSetDevice(a->ctx->device));
    bsm_sampler* sp = new bsm_sampler();
    sp->area = a;
    sp->buf = buf;
    sp->layer = buf == BSM_AXN_STATES ? 0u : layer;
    sp->offset = offset;
    sp->len = len;
    sp->bytes = (len + SAMPLE_ALIGN - 1) & ~(uint64_t)(SAMPLE_ALIGN - 1);
    sp->backpressure = backpressure != 0;
    if (arena_alloc(a->ctx, sp->bytes, &sp->chunk, &sp->off, &sp->dev_slot, &sp->host_slot)) {
        delete sp;
        return 1;
    }
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
    std::shared_ptr<FrameEvent> ev;
    uint64_t fc;
    const uint8_t* host;
    {
        std::unique_lock<std::mutex> lk(sp->m);
        if (sp->closed) return 0;   // its slot is gone with the area
        if (sp->state != SMP_PENDING) {
            if (!wait_for_frame) return 0;
            sp->cv.wait(lk, [&] { return sp->state == SMP_PENDING || sp->closed; });
            if (sp->closed || sp->state != SMP_PENDING) return 0;
        }
        ev = sp->ev;
        fc = sp->frame_cycle;
        host = sp->host_slot;
        sp->reading = true;
    }
    hipError_t e = hipEventSynchronize(ev->ev);
    if (e == hipSuccess) std::memcpy(dst, host, sp->len);
    {
        std::lock_guard<std::mutex> g(sp->m);
        sp->reading = false;
        if (e == hipSuccess && !sp->closed) {
            sp->state = SMP_EMPTY;
            sp->ev.reset();
        }
        sp->cv.notify_all();
    }
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize (sampler frame)");
    if (cycle) *cycle = fc;
    *got = 1;
    return 0;
}

int bsm_sampler_recv_many(bsm_sampler* const* s, uint32_t n, int wait_for_frame,
                          void* const* dst, const uint64_t* dst_len, uint64_t* cycles, int* got) {
    if ((!s || !dst || !dst_len || !got) && n) return fail("null arg");
    for (uint32_t i = 0; i < n; i++)
        if (bsm_sampler_recv(s[i], wait_for_frame, dst[i], dst_len[i], cycles ? cycles + i : nullptr,
                             got + i))
            return 1;
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

int bsm_sampler_info(bsm_sampler* sp, uint64_t* len, int* backpressure, int* pending) {
    if (!sp) return fail("null sampler");
    std::lock_guard<std::mutex> g(sp->m);
    if (len) *len = sp->len;
    if (backpressure) *backpressure = sp->backpressure ? 1 : 0;
    if (pending) *pending = sp->state == SMP_PENDING ? 1 : 0;
    return 0;
}

int bsm_sampler_destroy(bsm_sampler* sp) {
    if (!sp) return 0;
    if (!sp->closed) {
        (void)hipSetDevice(sp->area->ctx->device);
        sampler_close(sp);
    }
    delete sp;
    return 0;
}

}  // extern "C"
