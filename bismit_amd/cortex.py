"""Host mirror of bismit's Cortex / CorticalArea / layer API over the C ABI.

Reference interface (src/cortex/cortex.rs:46-120, src/cortex/cortical_area.rs:376-1395):
  Cortex::cycle            -> Cortex.cycle(inputs)
  CorticalArea::cycle      -> CorticalArea.cycle(n, rnd)
  CorticalArea::sampler    -> CorticalArea.sampler(kind, buffer_kind, backpressure)
  CorticalArea::regrow     -> CorticalArea.regrow(seed, epoch, reach)
  *Test traits (cycle_solo, set_all_to_zero, read_idx_direct, write_to_axon, ...)
                           -> read/write/fill and the per-kernel k_* methods
Every call goes to libbismit_hip.so; nothing here computes the cycle.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .recipes import AreaScheme


class Context:
    """One HIP device (replaces the OpenCL context/ProQue, src/cortex/cortex.rs:56-61)."""

    def __init__(self, device=0):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.bsm_ctx_create(device, C.byref(self.h)))

    def info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        mem = C.c_size_t()
        L.check(self.lib.bsm_device_info(self.h, name, 256, C.byref(cu), C.byref(mem)))
        return {"name": name.value.decode(), "cus": cu.value, "mem_bytes": mem.value}

    def sync(self):
        L.check(self.lib.bsm_sync(self.h))

    def alloc(self, nbytes):
        p = C.c_void_p()
        L.check(self.lib.bsm_dev_alloc(self.h, nbytes, C.byref(p)))
        return p

    def free(self, p):
        L.check(self.lib.bsm_dev_free(self.h, p))

    def upload(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        L.check(self.lib.bsm_dev_upload(self.h, dptr, L.vptr(arr), arr.nbytes))

    def close(self):
        if self.h:
            self.lib.bsm_ctx_destroy(self.h)
            self.h = C.c_void_p()


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def footprint(scheme: AreaScheme):
    """Device bytes an area of this scheme holds (HBM admission)."""
    desc, keep = scheme.to_c()
    n = C.c_size_t()
    L.check(L.load().bsm_area_footprint(C.byref(desc), C.byref(n)))
    return n.value


class CorticalArea:
    """One cortical area resident on the device."""

    def __init__(self, scheme: AreaScheme, ctx: Context = None, rnd_seed=0, layout_seed=0,
                 reach=6, init_layout=True):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.scheme = scheme
        desc, self._keep = scheme.to_c()
        self.h = C.c_void_p()
        L.check(self.lib.bsm_area_create(self.ctx.h, C.byref(desc), C.c_uint64(rnd_seed),
                                         C.byref(self.h)))
        n = C.c_uint32()
        L.check(self.lib.bsm_area_rnd_slots(self.h, C.byref(n)))
        self.n_slots = n.value
        L.check(self.lib.bsm_area_axn_len(self.h, C.byref(n)))
        self.axn_len = n.value
        self.layer_names = [l.name for l in scheme.layers]
        if init_layout:
            self.init_layout(layout_seed, reach)
        for li, lay in enumerate(scheme.layers):
            if lay.smoother_radius:
                self.set_smoother(li, lay.smoother_radius)

    # ------------------------------------------------------------- layout
    def init_layout(self, seed=0, reach=6):
        """Initial synapse sources (replaces Synapses::grow(init=true))."""
        L.check(self.lib.bsm_area_init_layout(self.h, C.c_uint64(seed), reach))

    def layer_index(self, layer):
        return self.layer_names.index(layer) if isinstance(layer, str) else int(layer)

    # ------------------------------------------------------------- buffers
    def buf_len(self, buf, layer=0):
        n = C.c_size_t()
        L.check(self.lib.bsm_area_buf_len(self.h, L.BUF[buf], self.layer_index(layer), C.byref(n)))
        return n.value

    def read(self, buf, layer=0, offset=0, length=None):
        li = self.layer_index(layer)
        if length is None:
            length = self.buf_len(buf, li) - offset
        dt = np.int8 if buf in L.SIGNED_BUFS else np.uint8
        out = np.empty(length, dtype=dt)
        L.check(self.lib.bsm_area_read(self.h, L.BUF[buf], li, offset, length, L.vptr(out)))
        return out

    def write(self, buf, layer, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        if arr.dtype.itemsize != 1:
            raise TypeError("area buffers are 8-bit")
        L.check(self.lib.bsm_area_write(self.h, L.BUF[buf], self.layer_index(layer), offset,
                                        arr.size, L.vptr(arr)))

    def fill(self, buf, layer, val):
        L.check(self.lib.bsm_area_fill(self.h, L.BUF[buf], self.layer_index(layer), val & 0xFF))

    def slice_idz(self, slc):
        n = C.c_uint32()
        L.check(self.lib.bsm_area_slice_idz(self.h, slc, C.byref(n)))
        return n.value

    def axon_states(self):
        return self.read("AXN_STATES")

    # ------------------------------------------------------------- intake
    def intake(self, frame, offset=None):
        """Write a host frame into the input slices (AxonSpace::intake); through
        the scheme's sensory filter when it names one and no offset is given."""
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        if offset is None and self.scheme.input_filter:
            if frame.size != self.scheme.input_range()[1]:
                raise ValueError("a filtered frame covers the input slices exactly")
            self._filtered(frame.ctypes.data, 0)
            return
        if offset is None:
            offset = self.scheme.input_range()[0]
        L.check(self.lib.bsm_area_intake(self.h, offset, L.vptr(frame), frame.size, 0))

    def intake_device(self, dptr, length, offset=None):
        if offset is None and self.scheme.input_filter:
            if length != self.scheme.input_range()[1]:
                raise ValueError("a filtered frame covers the input slices exactly")
            self._filtered(dptr.value if isinstance(dptr, C.c_void_p) else dptr, 1)
            return
        if offset is None:
            offset = self.scheme.input_range()[0]
        L.check(self.lib.bsm_area_intake(self.h, offset, dptr, length, 1))

    def _filtered(self, src, on_device):
        if self.scheme.input_filter != "retina":
            raise ValueError(f"unknown sensory filter {self.scheme.input_filter!r}")
        s0 = min(self.scheme.input_slcs)
        L.check(self.lib.bsm_area_intake_filtered(self.h, L.FILTER_RETINA, s0,
                                                  len(self.scheme.input_slcs), src, on_device))

    # ------------------------------------------------------------- filter / smoother
    def set_smoother(self, layer, radius):
        """Attach an ActivitySmoother control layer (radius 0 removes it)."""
        L.check(self.lib.bsm_area_set_smoother(self.h, self.layer_index(layer), radius))

    def k_smooth(self, layer):
        L.check(self.lib.bsm_k_smooth(self.h, self.layer_index(layer)))

    def k_retina(self, slc, depth, dptr):
        L.check(self.lib.bsm_k_retina(self.h, slc, depth, dptr))

    # ------------------------------------------------------------- cycle
    def draw_rnd(self, n_cycles):
        out = np.empty(n_cycles * self.n_slots, dtype=np.int32)
        L.check(self.lib.bsm_area_draw_rnd(self.h, L.i32ptr(out), n_cycles))
        return out

    def cycle(self, n=1, rnd=None):
        if rnd is None:
            L.check(self.lib.bsm_area_cycle(self.h, None, n))
        else:
            rnd = np.ascontiguousarray(rnd, dtype=np.int32)
            if rnd.size != n * self.n_slots:
                raise ValueError("rnd must hold n * n_slots draws")
            L.check(self.lib.bsm_area_cycle(self.h, L.i32ptr(rnd), n))

    def run(self, dev_frames, frame_stride, frame_len, n, offset=None):
        if offset is None:
            offset = self.scheme.input_range()[0]
        L.check(self.lib.bsm_area_run(self.h, offset, dev_frames, frame_stride, frame_len, n))

    def time_cycles(self, dev_frames, frame_stride, frame_len, n, offset=None):
        if offset is None:
            offset = self.scheme.input_range()[0]
        ms = C.c_float()
        L.check(self.lib.bsm_area_time_cycles(self.h, offset, dev_frames, frame_stride, frame_len,
                                              n, C.byref(ms)))
        return ms.value

    def sync(self):
        L.check(self.lib.bsm_area_sync(self.h))

    def profile_stages(self, n_cycles):
        """[(kind, layer, total_ms)] of every fused stage over n_cycles cycles."""
        m = 64
        ms = (C.c_float * m)()
        kind = (C.c_int * m)()
        lay = (C.c_uint32 * m)()
        n = C.c_uint32()
        L.check(self.lib.bsm_area_profile_stages(self.h, n_cycles, m, ms, kind, lay, C.byref(n)))
        return [(kind[i], lay[i], ms[i]) for i in range(n.value)]

    def regrow(self, seed, epoch, reach=6):
        """Deterministic regrowth of every synapse with strength <= -25
        (CorticalArea::regrow, src/cortex/cortical_area.rs:792-806)."""
        L.check(self.lib.bsm_area_regrow(self.h, seed, epoch, reach))

    def stage_plan(self):
        """The fused launches of one cycle: kernel variant and tiling of each."""
        arr = (L.StageInfo * 64)()
        n = C.c_uint32()
        L.check(self.lib.bsm_area_stage_plan(self.h, 64, arr, C.byref(n)))
        return L.stage_dicts(arr, n.value)

    def set_tuning(self, tile_v=0, tile_u=0, halo=0, chunk=0, inhib_tile=0, overlap=0):
        """Zero keeps a field; inhib_tile=1 returns it to the planner's choice;
        overlap 1 runs independent pyramidal tufts on a second stream, 2 does not."""
        t = L.Tuning(tile_v, tile_u, halo, chunk, inhib_tile, overlap)
        L.check(self.lib.bsm_area_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = L.Tuning()
        L.check(self.lib.bsm_area_get_tuning(self.h, C.byref(t)))
        return {"tile_v": t.tile_v, "tile_u": t.tile_u, "halo": t.halo, "chunk": t.chunk,
                "inhib_tile": t.inhib_tile, "overlap": t.overlap}

    def sampler(self, kind, buffer_kind=None, backpressure=False):
        """Per-cycle readback of one buffer kind (CorticalArea::sampler,
        src/cortex/cortical_area.rs:1017-1284): returns a TractReceiver whose
        recv() yields the frame of a cycle.  kind is a sampler.SamplerKind."""
        from . import sampler as S
        return S.create(self, kind, buffer_kind or S.SamplerBufferKind.SINGLE, backpressure)

    # ------------------------------------------------------------- kernels
    def k_syn_flags(self, layer):
        L.check(self.lib.bsm_k_syn_flags(self.h, self.layer_index(layer)))

    def k_syn_cycle(self, layer, tuft):
        L.check(self.lib.bsm_k_syn_cycle(self.h, self.layer_index(layer), tuft))

    def k_den_cycle(self, layer, tuft, rnd):
        L.check(self.lib.bsm_k_den_cycle(self.h, self.layer_index(layer), tuft, rnd))

    def k_tft_cycle(self, layer, tuft):
        L.check(self.lib.bsm_k_tft_cycle(self.h, self.layer_index(layer), tuft))

    def k_ssc_cycle(self, layer):
        L.check(self.lib.bsm_k_ssc_cycle(self.h, self.layer_index(layer)))

    def k_inhib_simple(self, layer, rnd):
        L.check(self.lib.bsm_k_inhib_simple(self.h, self.layer_index(layer), rnd))

    def k_inhib_passthrough(self, layer, rnd):
        L.check(self.lib.bsm_k_inhib_passthrough(self.h, self.layer_index(layer), rnd))

    def k_ssc_mtp(self, layer, rnd):
        L.check(self.lib.bsm_k_ssc_mtp(self.h, self.layer_index(layer), rnd & 0xFFFFFFFF))

    def k_pyr_cycle(self, layer):
        L.check(self.lib.bsm_k_pyr_cycle(self.h, self.layer_index(layer)))

    def k_inhib_intra_column(self, layer):
        L.check(self.lib.bsm_k_inhib_intra_column(self.h, self.layer_index(layer)))

    def k_pyr_output(self, layer):
        L.check(self.lib.bsm_k_pyr_output(self.h, self.layer_index(layer)))

    def k_tft_dst_mtp(self, layer, tuft, rnd):
        L.check(self.lib.bsm_k_tft_dst_mtp(self.h, self.layer_index(layer), tuft, rnd))

    # ------------------------------------------------------------- teardown
    def close(self):
        if self.h:
            self.lib.bsm_area_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched form: many areas stepped by one persistent launch per stage.
    dynamic=True: each cycle steps only the areas of an active mask (serving)."""

    def __init__(self, areas, ctx: Context = None, dynamic=False):
        if any(a.scheme.input_filter for a in areas):
            raise ValueError("areas with a sensory filter take their input in the plain form")
        self.ctx = ctx or areas[0].ctx
        self.lib = self.ctx.lib
        self.areas = list(areas)
        self.dynamic = dynamic
        arr = (C.c_void_p * max(1, len(areas)))(*[a.h.value for a in areas])
        self.h = C.c_void_p()
        L.check(self.lib.bsm_batch_create_ex(self.ctx.h, arr, len(areas),
                                             L.BATCH_DYNAMIC if dynamic else 0, C.byref(self.h)))
        if dynamic and areas:
            n = len(areas)
            rng = [a.scheme.input_range() for a in areas]
            of = (C.c_uint32 * n)(*[r[0] for r in rng])
            ln = (C.c_uint32 * n)(*[r[1] for r in rng])
            L.check(self.lib.bsm_batch_set_input_ranges(self.h, of, ln))

    def add(self, areas):
        """Dynamic batches: the areas join (after the enqueued cycles); they take
        their frames into their schemes' input ranges."""
        areas = list(areas)
        if any(a.scheme.input_filter for a in areas):
            raise ValueError("areas with a sensory filter take their input in the plain form")
        n = len(areas)
        if not n:
            return
        arr = (C.c_void_p * n)(*[a.h.value for a in areas])
        rng = [a.scheme.input_range() for a in areas]
        of = (C.c_uint32 * n)(*[r[0] for r in rng])
        ln = (C.c_uint32 * n)(*[r[1] for r in rng])
        L.check(self.lib.bsm_batch_add_areas(self.h, arr, n, of, ln))
        self.areas.extend(areas)

    def remove(self, areas):
        """Dynamic batches: the areas leave (after the enqueued cycles) and
        continue in the plain form; the others keep their order."""
        areas = list(areas)
        if not areas:
            return
        arr = (C.c_void_p * len(areas))(*[a.h.value for a in areas])
        L.check(self.lib.bsm_batch_remove_areas(self.h, arr, len(areas)))
        gone = set(id(a) for a in areas)
        self.areas = [a for a in self.areas if id(a) not in gone]

    def cycle_masked(self, active, frames, frame_ofs, graph=True):
        """One shared cycle of the areas with active[i] != 0; frames is one packed
        uint8 array, area i's frame at frame_ofs[i] (-1: none; a uint64 offset
        with L.FRAME_SDR set names an SDR entry: uint32 count, then indices)."""
        active = np.ascontiguousarray(active, dtype=np.uint8)
        frames = np.ascontiguousarray(frames, dtype=np.uint8)
        fo = np.asarray(frame_ofs)
        fo = np.ascontiguousarray(fo if fo.dtype == np.uint64 else fo.astype(np.int64).view(np.uint64))
        L.check(self.lib.bsm_batch_cycle_masked(
            self.h, L.u8ptr(active), L.vptr(frames) if frames.size else None,
            fo.ctypes.data_as(C.POINTER(C.c_uint64)), frames.size, 1 if graph else 0))

    def query(self):
        d = C.c_int()
        L.check(self.lib.bsm_batch_query(self.h, C.byref(d)))
        return bool(d.value)

    def regrow(self, seeds, epochs, reach=6, which=None):
        """Deterministic regrowth of the batched areas (area i: seeds[i],
        epochs[i]; only where which[i] when which is given), the same draws as
        CorticalArea.regrow.  Waits for the enqueued cycles."""
        n = len(self.areas)
        sd = (C.c_uint64 * n)(*[int(x) for x in seeds])
        ep = (C.c_uint64 * n)(*[int(x) for x in epochs])
        wh = None
        if which is not None:
            w = np.ascontiguousarray(which, dtype=np.uint8)
            assert w.size == n
            wh = L.u8ptr(w)
        L.check(self.lib.bsm_batch_regrow(self.h, wh, sd, ep, reach))

    def set_inputs(self, dev_frames, strides, lens, offsets, n_frames):
        n = len(self.areas)
        fr = (C.c_void_p * n)(*[p.value if p is not None else None for p in dev_frames])
        st = (C.c_uint64 * n)(*strides)
        ln = (C.c_uint32 * n)(*lens)
        of = (C.c_uint32 * n)(*offsets)
        L.check(self.lib.bsm_batch_set_inputs(self.h, fr, st, ln, of, n_frames))

    def cycle(self, n=1, graph=True):
        L.check(self.lib.bsm_batch_cycle(self.h, n, 1 if graph else 0))

    def time_cycles(self, n, graph=True):
        ms = C.c_float()
        L.check(self.lib.bsm_batch_time_cycles(self.h, n, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def stage_plan(self):
        """The persistent launches of one batched cycle: variant and tiling of each."""
        arr = (L.StageInfo * 64)()
        n = C.c_uint32()
        L.check(self.lib.bsm_batch_stage_plan(self.h, 64, arr, C.byref(n)))
        return L.stage_dicts(arr, n.value)

    def profile_stages(self, n_cycles):
        """[(kind, total_ms)] of every persistent stage launch of a static batch
        over n_cycles ungraphed cycles (which advance the areas), launch order."""
        m = 64
        ms = (C.c_float * m)()
        kind = (C.c_int * m)()
        n = C.c_uint32()
        L.check(self.lib.bsm_batch_profile_stages(self.h, n_cycles, m, ms, kind, C.byref(n)))
        return [(kind[i], ms[i]) for i in range(n.value)]

    def set_tuning(self, tile_v=0, tile_u=0, halo=0, chunk=0, inhib_tile=0, overlap=0):
        """Zero keeps a field; inhib_tile=1 returns it to the planner's choice;
        overlap 1 runs independent pyramidal tufts on a second stream, 2 does not."""
        t = L.Tuning(tile_v, tile_u, halo, chunk, inhib_tile, overlap)
        L.check(self.lib.bsm_batch_set_tuning(self.h, C.byref(t)))

    def sync(self):
        L.check(self.lib.bsm_batch_sync(self.h))

    def close(self):
        if self.h:
            self.lib.bsm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cortex:
    """A set of areas cycled together (Cortex::cycle, src/cortex/cortex.rs:100-120):
    intake of each area's input frame, then the area's cycle."""

    def __init__(self, schemes, ctx: Context = None, seed=0):
        self.ctx = ctx or default_context()
        self.areas = {}
        for i, s in enumerate(schemes):
            self.areas[s.name] = CorticalArea(s, self.ctx, rnd_seed=seed + 1000 * i,
                                              layout_seed=seed + 1000 * i + 1)

    def cycle(self, inputs=None):
        inputs = inputs or {}
        for name, area in self.areas.items():
            if name in inputs:
                area.intake(inputs[name])
            area.cycle(1)

    def finish_queues(self):
        """Cortex::finish_queues: wait for every area's enqueued work."""
        for a in self.areas.values():
            a.sync()

    def cycle_device(self, dev_frames, lens, n_frames, n_cycles, offsets=None):
        """n_cycles host-driven cycles in C (bsm_cortex_cycle): each cycle, each
        area takes frame (c % n_frames) of its device frame ring and one cycle."""
        cycle_areas(list(self.areas.values()), dev_frames, lens, n_frames, n_cycles, offsets)

    def close(self):
        for a in self.areas.values():
            a.close()


def cycle_areas(areas, dev_frames, lens, n_frames, n_cycles, offsets=None):
    """bsm_cortex_cycle over CorticalArea objects (device frame rings, one per area)."""
    n = len(areas)
    if offsets is None:
        offsets = [a.scheme.input_range()[0] for a in areas]
    ah = (C.c_void_p * n)(*[a.h.value for a in areas])
    fr = (C.c_void_p * n)(*[p.value if p is not None else None for p in dev_frames])
    st = (C.c_uint64 * n)(*lens)
    ln = (C.c_uint32 * n)(*lens)
    of = (C.c_uint32 * n)(*offsets)
    L.check(areas[0].lib.bsm_cortex_cycle(ah, n, fr, st, ln, of, n_frames, n_cycles))
