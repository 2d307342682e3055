"""Samplers: per-cycle readback of an area's buffers to the host.

Mirrors the reference's sampler surface -- `SamplerKind`, `SamplerBufferKind`
and the `TractReceiver` a sampler request returns
(src/cortex/cortical_area.rs:51-90,1017-1284; src/subcortex/tract_channel.rs:
535-551) -- over the engine's bsm_sampler_* calls.  Every cycle the area steps
(plain or inside a batch) sends a frame when the previous one was received;
otherwise it waits for the receiver (backpressure on) or skips the frame
(backpressure off), so `recv` always returns the frame of the first cycle after
the last receive.  Frames are gathered on the device after the cycle and copied
to pinned host memory on the cycle's own stream.
"""
import ctypes as C
from dataclasses import dataclass
from enum import Enum
from typing import Optional

import numpy as np

from . import _lib as L


class SamplerBufferKind(Enum):
    """Frames a sampler holds: `SINGLE` (the reference's only implemented kind,
    src/cortex/cortical_area.rs:1038-1048), `DOUBLE` and `TRIPLE` (declared
    upstream, :52-58, unimplemented there): a cycle's frame is sent while fewer
    are held and received oldest first.  `NONE` and `MAP` are not implemented."""
    NONE = 0
    MAP = 1
    SINGLE = 2
    DOUBLE = 3
    TRIPLE = 4


# variant name -> engine buffer kind
_KINDS = {
    "Axons": "AXN_STATES",
    "SomaStates": "SOMA_STATES",
    "SomaEnergies": "SOMA_ENERGIES",
    "SomaActivities": "SOMA_ACTIVITIES",
    "SomaFlagSets": "SOMA_FLAG_SETS",
    "TuftStates": "TFT_STATES",
    "TuftBestDenIds": "TFT_BEST_DEN_IDS",
    "TuftBestDenStatesRaw": "TFT_BEST_DEN_STATES_RAW",
    "TuftBestDenStates": "TFT_BEST_DEN_STATES",
    "TuftPrevStates": "TFT_PREV_STATES",
    "TuftPrevBestDenIds": "TFT_PREV_BEST_DEN_IDS",
    "TuftPrevBestDenStatesRaw": "TFT_PREV_BEST_DEN_STATES_RAW",
    "TuftPrevBestDenStates": "TFT_PREV_BEST_DEN_STATES",
    "DenStates": "DEN_STATES",
    "DenStatesRaw": "DEN_STATES_RAW",
    "DenEnergies": "DEN_ENERGIES",
    "DenActivities": "DEN_ACTIVITIES",
    "DenThresholds": "DEN_THRESHOLDS",
    "SynStates": "SYN_STATES",
    "SynStrengths": "SYN_STRENGTHS",
    "SynSrcSlcIds": "SYN_SRC_SLC_IDS",
    "SynSrcColVOffs": "SYN_SRC_COL_V_OFFS",
    "SynSrcColUOffs": "SYN_SRC_COL_U_OFFS",
    "SynFlagSets": "SYN_FLAG_SETS",
}


@dataclass(frozen=True)
class SamplerKind:
    """One variant of the reference's `SamplerKind`: `name` is the variant,
    `layer` the layer (index or name; `Axons(None)` samples every slice)."""
    name: str
    layer: object = None

    @property
    def buf(self):
        return _KINDS[self.name]


def _variant(name):
    def make(layer=None):
        if layer is None and name != "Axons":
            raise ValueError(f"SamplerKind.{name} needs a layer")
        return SamplerKind(name, layer)
    make.__name__ = name
    return staticmethod(make)


for _n in _KINDS:
    setattr(SamplerKind, _n, _variant(_n))


@dataclass
class Frame:
    data: np.ndarray
    cycle: int          # the area's cycle count at this frame (1 = first cycle)


class TractReceiver:
    """Receiving end of one sampler (TractReceiver, src/subcortex/tract_channel.rs:
    535-551).  `recv` and the backpressure calls may come from any thread."""

    def __init__(self, area, handle, length, dtype, idx_range):
        self._area = area
        self.lib = area.lib
        self.h = handle
        self.len = length
        self.dtype = dtype
        self._range = idx_range

    def recv(self, wait_for_frame=True) -> Optional[Frame]:
        """The pending frame, or the next one when wait_for_frame (None if the
        sampler closes meanwhile); without wait_for_frame, None when nothing is
        pending."""
        out = np.empty(self.len, dtype=self.dtype)
        cyc = C.c_uint64()
        got = C.c_int()
        L.check(self.lib.bsm_sampler_recv(self.h, 1 if wait_for_frame else 0, L.vptr(out),
                                          self.len, C.byref(cyc), C.byref(got)))
        return Frame(out, cyc.value) if got.value else None

    def set_backpressure(self, bp) -> bool:
        prior = C.c_int()
        L.check(self.lib.bsm_sampler_set_backpressure(self.h, 1 if bp else 0, C.byref(prior)))
        return bool(prior.value)

    def backpressure_stale(self) -> bool:
        bp = C.c_int()
        L.check(self.lib.bsm_sampler_info(self.h, None, C.byref(bp), None))
        return bool(bp.value)

    def pending(self) -> bool:
        p = C.c_int()
        L.check(self.lib.bsm_sampler_info(self.h, None, None, C.byref(p)))
        return bool(p.value)

    def buffer_idx_range(self):
        return self._range

    def close(self):
        if self.h:
            self.lib.bsm_sampler_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def recv_many(rxs, wait_for_frame=True):
    """recv over several receivers in one engine call; a list of Frame or None."""
    n = len(rxs)
    if n == 0:
        return []
    outs = [np.empty(rx.len, dtype=rx.dtype) for rx in rxs]
    hs = (C.c_void_p * n)(*[rx.h.value for rx in rxs])
    ds = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    ls = (C.c_uint64 * n)(*[rx.len for rx in rxs])
    cyc = (C.c_uint64 * n)()
    got = (C.c_int * n)()
    L.check(rxs[0].lib.bsm_sampler_recv_many(hs, n, 1 if wait_for_frame else 0, ds, ls, cyc, got))
    return [Frame(outs[i], cyc[i]) if got[i] else None for i in range(n)]


class ReceiverSet:
    """Several receivers received together into buffers of their own, reused
    call after call: one engine call per recv, with the handle and destination
    arrays built once (a server receiving every resident session's output each
    cycle)."""

    def __init__(self, rxs, reuse=None):
        """reuse: an earlier set whose buffers the receivers it shares keep."""
        self.rxs = list(rxs)
        n = len(self.rxs)
        old = {id(rx): o for rx, o in zip(reuse.rxs, reuse.outs)} if reuse is not None else {}
        self.outs = [old.get(id(rx)) if id(rx) in old else np.empty(rx.len, dtype=rx.dtype)
                     for rx in self.rxs]
        self._hs = np.array([rx.h.value for rx in self.rxs], dtype=np.uint64)
        self._ds = np.array([o.ctypes.data for o in self.outs], dtype=np.uint64)
        self._ls = np.array([rx.len for rx in self.rxs], dtype=np.uint64)
        self.cycles = np.zeros(n, dtype=np.uint64)
        self.got = np.zeros(n, dtype=np.int32)

    def recv(self, wait_for_frame=False, idx=None):
        """Receives into self.outs; self.got[i] / self.cycles[i] say which got a
        frame and its cycle.  Without wait_for_frame, receivers with nothing
        pending or in flight get none; frames in flight are waited for.  idx
        (an integer array) limits the call to those receivers; the others keep
        their got / cycles entries."""
        if not self.rxs:
            return self.got
        lib = self.rxs[0].lib
        if idx is None:
            hs, ds, ls, cyc, got = self._hs, self._ds, self._ls, self.cycles, self.got
        else:
            idx = np.asarray(idx, dtype=np.int64)
            if idx.size == 0:
                return self.got
            hs, ds, ls = self._hs[idx], self._ds[idx], self._ls[idx]
            cyc = np.zeros(idx.size, dtype=np.uint64)
            got = np.zeros(idx.size, dtype=np.int32)
        L.check(lib.bsm_sampler_recv_many(
            C.cast(L.vptr(hs), C.POINTER(C.c_void_p)), len(hs),
            1 if wait_for_frame else 0, C.cast(L.vptr(ds), C.POINTER(C.c_void_p)),
            C.cast(L.vptr(ls), C.POINTER(C.c_uint64)),
            C.cast(L.vptr(cyc), C.POINTER(C.c_uint64)),
            C.cast(L.vptr(got), C.POINTER(C.c_int))))
        if idx is not None:
            self.cycles[idx] = cyc
            self.got[idx] = got
        return self.got


def create(area, kind: SamplerKind, buffer_kind=SamplerBufferKind.SINGLE, backpressure=False):
    """A sampler on `area` (CorticalArea::sampler, src/cortex/cortical_area.rs:1017-1284)."""
    held = {SamplerBufferKind.SINGLE: 1, SamplerBufferKind.DOUBLE: 2, SamplerBufferKind.TRIPLE: 3}
    if buffer_kind not in held:
        raise NotImplementedError(f"sampler buffer kind {buffer_kind} (SINGLE, DOUBLE or TRIPLE)")
    buf = kind.buf
    if buf == "AXN_STATES":
        if kind.layer is None:
            off, n = 0, area.axn_len
        else:
            off, n = area.scheme.layer_axn_range(area.layer_index(kind.layer))
        li = 0
    else:
        li = area.layer_index(kind.layer)
        off, n = 0, area.buf_len(buf, li)
    h = C.c_void_p()
    L.check(area.lib.bsm_sampler_create(area.h, L.BUF[buf], li, off, n, 1 if backpressure else 0,
                                        C.byref(h)))
    if held[buffer_kind] > 1:
        rc = area.lib.bsm_sampler_set_buffers(h, held[buffer_kind])
        if rc:
            area.lib.bsm_sampler_destroy(h)
            L.check(rc)
    dt = np.int8 if buf in L.SIGNED_BUFS else np.uint8
    return TractReceiver(area, h, n, dt, (off, off + n))
