// bsm_filters.hip -- the sensory filter (retina) and the activity smoother.
//
// retina (cl/filters.cl:34-100, run by SensoryFilter, src/cortex/sensory_filter.rs:
// 24-174): a center-surround filter of the input frame into axon slices.  Cells
// alternate on/off-center (cell index bit 0) and small/large field (bits 0^1):
// radius 3 or 5; the surround is the mean over the whole hex field, off-grid
// positions counting as zeros; the output is (center - (mean + 48)) for on-center
// and (mean - (center + 48)) for off-center cells, floored at 0.  A workgroup
// filters a 16x16 tile of one slice from an LDS copy of the tile and its halo.
//
// smooth_activity (cl/control.cl:227-295, ActivitySmoother, src/cortex/smoother.rs):
// six groups of hex-field centers tile the layer, each group exactly once; each
// center finds the least and most active cell of its field (over all slices of
// the layer; ties go to the later cell in field order) and raises the least
// active cell's energy by one and lowers the most active one's.  The reference
// runs all centers in one launch, so where fields of different groups meet,
// updates race.  Here the groups apply in start order: k_smooth_select finds
// every center's pair, then k_smooth_apply walks each cell through the six
// groups in order (each group covers a cell once), which is the result of six
// race-free launches, one per group, in two launches.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "bsm_internal.h"

namespace bsm {

// Staging of bsm_area_intake_filtered_noalloc: a file-scope device array, so
// that filtering a host frame allocates nothing.  Each device holds its own copy
// for the life of the process.  Host code never names the array in a launch or
// copy: it uses the device address hipGetSymbolAddress returned for the device
// (scratch_ptr), taken once per device by filter_scratch_setup.
static __device__ __align__(256) uint8_t scratch[FILTER_SCRATCH_MAX];
static std::mutex scratch_mu;
static uint8_t* scratch_ptr[MAX_DEVICES] = {};

int filter_scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail("device index beyond the staging table");
    std::lock_guard<std::mutex> g(scratch_mu);
    if (scratch_ptr[device]) return 0;
    int cur = -1;
    BSM_HIP(hipGetDevice(&cur));
    if (cur != device) return fail("filter_scratch_setup: device is not current");
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail("hipGetSymbolAddress(scratch) returned null");
    scratch_ptr[device] = (uint8_t*)p;
    return 0;
}

constexpr int RETINA_THRESHOLD = 48;
constexpr int RETINA_SMALL_R = 3;
constexpr int RETINA_LARGE_R = 5;
constexpr int RT = 16;                        // retina tile side
constexpr int RH = RETINA_LARGE_R;            // halo
constexpr int RL = RT + 2 * RH;               // LDS tile side

// Blocks (u tile, v tile, slice); 256 threads, one output cell each.
__global__ __launch_bounds__(256) void k_retina(const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ out, uint32_t V,
                                                uint32_t U) {
    __shared__ uint8_t t[RL][RL];
    const uint32_t s = blockIdx.z;
    const int v0 = (int)(blockIdx.y * RT), u0 = (int)(blockIdx.x * RT);
    const uint8_t* src = in + (size_t)s * V * U;
    for (int i = threadIdx.x; i < RL * RL; i += blockDim.x) {
        const int v = v0 - RH + i / RL, u = u0 - RH + i % RL;
        t[i / RL][i % RL] = (v >= 0 && v < (int)V && u >= 0 && u < (int)U) ? src[v * U + u] : 0;
    }
    __syncthreads();
    const int tv = threadIdx.x / RT, tu = threadIdx.x % RT;
    const int v = v0 + tv, u = u0 + tu;
    if (v >= (int)V || u >= (int)U) return;
    const uint32_t cel = s * V * U + (uint32_t)v * U + (uint32_t)u;
    const uint32_t off_center = cel & 1u;
    const uint32_t large = ((cel & 3u) >> 1) ^ off_center;
    const int r = large ? RETINA_LARGE_R : RETINA_SMALL_R;
    int sum = 0;
    for (int dv = -r; dv <= r; dv++) {
        const int lo = max(-r, -dv - r), hi = min(r, -dv + r);
        const uint8_t* row = &t[RH + tv + dv][RH + tu];
        for (int du = lo; du <= hi; du++) sum += row[du];
    }
    const int avg = sum / (3 * r * (r + 1) + 1);
    const int center = t[RH + tv][RH + tu];
    int o;
    if (off_center) {
        const int cst = center + RETINA_THRESHOLD;
        o = avg > cst ? avg - cst : 0;
    } else {
        const int nat = avg + RETINA_THRESHOLD;
        o = center > nat ? center - nat : 0;
    }
    out[cel] = (uint8_t)o;
}

// One thread per center: the least and most active cell of its hex field.
__global__ __launch_bounds__(256) void k_smooth_select(const int32_t* __restrict__ centers,
                                                       uint32_t n, uint32_t V, uint32_t U,
                                                       uint32_t depth, int32_t r,
                                                       const uint8_t* __restrict__ act,
                                                       uint32_t* __restrict__ sel) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int cv = centers[2 * c], cu = centers[2 * c + 1];
    uint32_t least = 0, most = 0;
    uint32_t least_a = 255, most_a = 0;
    for (uint32_t s = 0; s < depth; s++) {
        for (int dv = -r; dv <= r; dv++) {
            const int v = cv + dv;
            if (v < 0 || v >= (int)V) continue;   // no cell of this row is on the grid
            const int lo = max(-r, -dv - r), hi = min(r, -dv + r);
            for (int du = lo; du <= hi; du++) {
                const int u = cu + du;
                if (u < 0 || u >= (int)U) continue;
                const uint32_t idx = s * V * U + (uint32_t)v * U + (uint32_t)u;
                const uint32_t x = act[idx];
                if (x <= least_a) { least = idx; least_a = x; }
                if (x >= most_a) { most = idx; most_a = x; }
            }
        }
    }
    sel[3 * c] = least;
    sel[3 * c + 1] = most;
    sel[3 * c + 2] = most_a;
}

// One thread per cell of the layer: the six groups' updates in group order.
__global__ __launch_bounds__(256) void k_smooth_apply(const uint32_t* __restrict__ cover,
                                                      const uint32_t* __restrict__ sel,
                                                      uint32_t VU, uint32_t cells,
                                                      uint8_t* __restrict__ energy) {
    const uint32_t cel = blockIdx.x * blockDim.x + threadIdx.x;
    if (cel >= cells) return;
    const uint32_t col = cel % VU;
    uint32_t e = energy[cel];
    const uint32_t e0 = e;
    for (uint32_t g = 0; g < SMOOTH_GROUPS; g++) {
        const uint32_t c = cover[g * VU + col];
        const uint32_t least = sel[3 * c], most = sel[3 * c + 1], most_a = sel[3 * c + 2];
        if (least == most) continue;
        if (least == cel) e += (e < 255u && most_a < 255u) ? 1u : 0u;
        if (most == cel) e -= (e > 0u && most_a > 0u) ? 1u : 0u;
    }
    if (e != e0) energy[cel] = (uint8_t)e;
}

// gen_grp_centers (src/cortex/smoother.rs:33-58) with HexGroupCenters::populate
// (src/cmn/mod.rs:892-921): per group, every point reachable from the group's
// start through the six surround steps inside [-r, dims + r); row-major order.
static void group_centers(int32_t r, int32_t V, int32_t U, std::vector<int32_t>& out,
                          uint32_t* counts) {
    const int32_t l = r + 1, s = r, ls = l + s, od = (r + 1) / 2;
    const int32_t starts[SMOOTH_GROUPS][2] = {{0, od}, {-od, od}, {-od, 0},
                                              {0, -od}, {od, -od}, {od, 0}};
    const int32_t steps[6][2] = {{-s, ls}, {l, s}, {ls, -l}, {s, -ls}, {-l, -s}, {-ls, l}};
    const int32_t W = V + 2 * r, H = U + 2 * r;   // bounded region, offset by r
    for (uint32_t g = 0; g < SMOOTH_GROUPS; g++) {
        std::vector<uint8_t> seen((size_t)W * H, 0);
        std::vector<std::pair<int32_t, int32_t>> q;
        auto add = [&](int32_t v, int32_t u) {
            if (v < -r || u < -r || v >= V + r || u >= U + r) return;
            uint8_t& m = seen[(size_t)(v + r) * H + (u + r)];
            if (!m) {
                m = 1;
                q.emplace_back(v, u);
            }
        };
        add(starts[g][0], starts[g][1]);
        for (size_t i = 0; i < q.size(); i++)
            for (auto& st : steps) add(q[i].first + st[0], q[i].second + st[1]);
        uint32_t n = 0;
        for (int32_t v = 0; v < W; v++)
            for (int32_t u = 0; u < H; u++)
                if (seen[(size_t)v * H + u]) {
                    out.push_back(v - r);
                    out.push_back(u - r);
                    n++;
                }
        counts[g] = n;
    }
}

void smoother_free(LayerRt& L) {
    (void)hipFree(L.smooth_centers);
    (void)hipFree(L.smooth_cover);
    (void)hipFree(L.smooth_sel);
    L.smooth_centers = nullptr;
    L.smooth_cover = nullptr;
    L.smooth_sel = nullptr;
    L.smooth_n = 0;
    L.smooth_r = 0;
}

int launch_smoother(bsm_area* a, uint32_t layer, hipStream_t s) {
    LayerRt& L = a->L[layer];
    if (!L.smooth_n) return 0;
    k_smooth_select<<<(L.smooth_n + 255) / 256, 256, 0, s>>>(
        L.smooth_centers, L.smooth_n, a->V, a->U, L.d.depth, L.smooth_r, L.cel_act, L.smooth_sel);
    BSM_HIP(hipGetLastError());
    k_smooth_apply<<<(L.cells + 255) / 256, 256, 0, s>>>(L.smooth_cover, L.smooth_sel,
                                                         a->V * a->U, L.cells, L.cel_energies);
    BSM_HIP(hipGetLastError());
    return 0;
}

int launch_area_stage(bsm_area* a, hipStream_t s, const Stage& st, uint32_t rnd_ofs,
                      const uint8_t* frame) {
    if (st.kind == 4) return launch_smoother(a, st.layer, s);
    if (frame && a->fuse_job >= 0)
        return launch_stage(s, st, a->tjobs_d, a->sjobs_d, a->pjobs_d, a->items_d,
                            (uint32_t)a->L[st.layer].cur, rnd_ofs, frame, a->axn + a->fuse_lo,
                            (uint32_t)a->fuse_job);
    return launch_stage(s, st, a->tjobs_d, a->sjobs_d, a->pjobs_d, a->items_d,
                        (uint32_t)a->L[st.layer].cur, rnd_ofs);
}

static int check_retina(bsm_area* a, uint32_t slc, uint32_t depth, uint32_t* V, uint32_t* U) {
    if (depth == 0 || slc >= a->n_slices || depth > a->n_slices - slc)
        return fail("filter slices out of the area's slice range");
    const SliceRt& s0 = a->slices_h[slc];
    for (uint32_t k = 1; k < depth; k++) {
        const SliceRt& sk = a->slices_h[slc + k];
        if (sk.v_size != s0.v_size || sk.u_size != s0.u_size)
            return fail("filtered slices must share their dimensions");
    }
    if (!s0.v_size || !s0.u_size) return fail("empty slice");
    *V = s0.v_size;
    *U = s0.u_size;
    return 0;
}

static int launch_retina(bsm_area* a, uint32_t slc, uint32_t depth, uint32_t V, uint32_t U,
                         const uint8_t* src) {
    dim3 grid((U + RT - 1) / RT, (V + RT - 1) / RT, depth);
    // slices [slc, slc + depth) share dims, so they lie back to back in the axon space
    k_retina<<<grid, 256, 0, a->stream>>>(src, a->axn + a->slices_h[slc].idz, V, U);
    BSM_HIP(hipGetLastError());
    return 0;
}

}  // namespace bsm

using namespace bsm;

extern "C" {

int bsm_smoother_centers(int32_t radius, uint32_t v_size, uint32_t u_size, int32_t* v,
                         int32_t* u, uint32_t* group_counts, uint32_t cap, uint32_t* n) {
    if (!n || !group_counts) return fail("null arg");
    if (radius < 1 || radius > 127) return fail("smoother radius must be in [1, 127]");
    if (!v_size || !u_size || v_size > 65535 || u_size > 65535) return fail("bad dimensions");
    std::vector<int32_t> c;
    group_centers(radius, (int32_t)v_size, (int32_t)u_size, c, group_counts);
    *n = (uint32_t)(c.size() / 2);
    if (*n > cap) return 0;   // caller asks again with room for *n
    for (uint32_t i = 0; i < *n; i++) {
        if (v) v[i] = c[2 * i];
        if (u) u[i] = c[2 * i + 1];
    }
    return 0;
}

int bsm_area_set_smoother(bsm_area* a, uint32_t layer, int32_t radius) {
    if (!a) return fail("null area");
    if (a->batch_stream) return fail("area belongs to a batch");
    if (layer >= a->n_layers) return fail("layer out of range");
    if (radius < 0 || radius > 127) return fail("smoother radius must be in [0, 127]");
    BSM_HIP(hipSetDevice(a->ctx->device));
    BSM_HIP(hipStreamSynchronize(a->stream));
    LayerRt& L = a->L[layer];
    smoother_free(L);
    a->plan_dirty = true;
    if (radius == 0) return 0;
    std::vector<int32_t> c;
    uint32_t counts[SMOOTH_GROUPS];
    group_centers(radius, (int32_t)a->V, (int32_t)a->U, c, counts);
    const uint32_t n = (uint32_t)(c.size() / 2), VU = a->V * a->U;
    // the center covering each column in each group: exactly one per group
    std::vector<uint32_t> cover((size_t)SMOOTH_GROUPS * VU, 0xFFFFFFFFu);
    uint32_t c0 = 0;
    for (uint32_t g = 0; g < SMOOTH_GROUPS; g++) {
        for (uint32_t k = c0; k < c0 + counts[g]; k++) {
            const int32_t cv = c[2 * k], cu = c[2 * k + 1];
            for (int32_t dv = -radius; dv <= radius; dv++) {
                const int32_t v = cv + dv;
                if (v < 0 || v >= (int32_t)a->V) continue;
                const int32_t lo = std::max(-radius, -dv - radius);
                const int32_t hi = std::min(radius, -dv + radius);
                for (int32_t du = lo; du <= hi; du++) {
                    const int32_t u = cu + du;
                    if (u < 0 || u >= (int32_t)a->U) continue;
                    uint32_t& x = cover[(size_t)g * VU + (uint32_t)v * a->U + (uint32_t)u];
                    if (x != 0xFFFFFFFFu) return fail("smoother fields of one group overlap");
                    x = k;
                }
            }
        }
        c0 += counts[g];
    }
    for (uint32_t x : cover)
        if (x == 0xFFFFFFFFu) return fail("smoother fields leave a cell uncovered");
    BSM_HIP(hipMalloc((void**)&L.smooth_centers, c.size() * sizeof(int32_t)));
    BSM_HIP(hipMalloc((void**)&L.smooth_cover, cover.size() * sizeof(uint32_t)));
    BSM_HIP(hipMalloc((void**)&L.smooth_sel, (size_t)3 * n * sizeof(uint32_t)));
    BSM_HIP(hipMemcpy(L.smooth_centers, c.data(), c.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    BSM_HIP(hipMemcpy(L.smooth_cover, cover.data(), cover.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    L.smooth_r = radius;
    L.smooth_n = n;
    return 0;
}

int bsm_k_smooth(bsm_area* a, uint32_t layer) {
    if (!a) return fail("null area");
    if (layer >= a->n_layers) return fail("layer out of range");
    if (!a->L[layer].smooth_n) return fail("layer has no smoother");
    BSM_HIP(hipSetDevice(a->ctx->device));
    return launch_smoother(a, layer, a->stream);
}

int bsm_k_retina(bsm_area* a, uint32_t slc, uint32_t depth, const uint8_t* dev_src) {
    if (!a || !dev_src) return fail("null arg");
    uint32_t V, U;
    if (check_retina(a, slc, depth, &V, &U)) return 1;
    BSM_HIP(hipSetDevice(a->ctx->device));
    return launch_retina(a, slc, depth, V, U, dev_src);
}

int bsm_area_intake_filtered_noalloc(bsm_area* a, int filter, uint32_t slc, uint32_t depth,
                                     const uint8_t* src) {
    if (!a || !src) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch");
    if (filter != BSM_FILTER_RETINA) return fail("unknown filter");
    uint32_t V, U;
    if (check_retina(a, slc, depth, &V, &U)) return 1;
    const size_t bytes = (size_t)depth * V * U;
    if (bytes > FILTER_SCRATCH_MAX)
        return fail("filtered frame of " + std::to_string(bytes) +
                    " bytes exceeds the static staging maximum of " +
                    std::to_string(FILTER_SCRATCH_MAX) + " bytes");
    const int dev = a->ctx->device;
    BSM_HIP(hipSetDevice(dev));
    uint8_t* stage = (dev >= 0 && dev < MAX_DEVICES) ? scratch_ptr[dev] : nullptr;
    if (!stage) return fail("static staging is not set up on this device");
    // work still queued on the area may read its axons; the staging copy is
    // rewritten whole, so nothing in it needs clearing
    BSM_HIP(hipStreamSynchronize(a->stream));
    BSM_HIP(hipMemcpy(stage, src, bytes, hipMemcpyHostToDevice));
    if (launch_retina(a, slc, depth, V, U, stage)) return 1;
    // the next call on this device may overwrite the staging copy
    BSM_HIP(hipStreamSynchronize(a->stream));
    return 0;
}

int bsm_area_intake_filtered(bsm_area* a, int filter, uint32_t slc, uint32_t depth,
                             const uint8_t* src, int src_is_device) {
    if (!a || !src) return fail("null arg");
    if (a->batch_stream) return fail("area belongs to a batch");
    if (filter != BSM_FILTER_RETINA) return fail("unknown filter");
    uint32_t V, U;
    if (check_retina(a, slc, depth, &V, &U)) return 1;
    BSM_HIP(hipSetDevice(a->ctx->device));
    const size_t bytes = (size_t)depth * V * U;
    if (!src_is_device) {
        // the previous filtered intake may still read the staging copy
        BSM_HIP(hipStreamSynchronize(a->stream));
        if (a->filt_cap < bytes) {
            (void)hipFree(a->filt_stage);
            a->filt_stage = nullptr;
            a->filt_cap = 0;
            BSM_HIP(hipMalloc((void**)&a->filt_stage, bytes));
            a->filt_cap = bytes;
        }
        BSM_HIP(hipMemcpy(a->filt_stage, src, bytes, hipMemcpyHostToDevice));
        src = a->filt_stage;
    }
    return launch_retina(a, slc, depth, V, U, src);
}

}  // extern "C"
