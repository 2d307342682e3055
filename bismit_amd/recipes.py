"""Area schemes: the runtime form of bismit's layer-map builder DSL for the
cortical cycle (src/map/scheme/*.rs, src/map/slice_map.rs, src/cmn/slice_dims.rs).

An AreaScheme lists axon slices in slice-id order (the reference assigns slices
in layer-definition order, src/map/layer_map.rs) and cellular layers in the
area's cycle order (spatial layers first, src/cortex/cortical_area.rs:601-613).
It converts to the C descriptor of include/bismit_hip.h.
"""
import ctypes as C
import json
import math
from dataclasses import asdict, dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L

SLC_SCL_COEFF = 16


def calc_scale(src_dim, tar_dim):
    """Slice scale of a source dimension mapped onto a target area dimension
    (src/cmn/slice_dims.rs:200-223): 16 = 1:1."""
    src = src_dim * 1024
    tar = tar_dim * 1024
    if tar == 0:
        raise ValueError("Target area dimension cannot be zero.")
    incr = 1 if tar < SLC_SCL_COEFF else tar // SLC_SCL_COEFF
    s = src // incr
    if s == 0:
        raise ValueError("Source area dimension cannot be zero.")
    if s > 255:
        raise ValueError("Source area cannot have a dimension more than 16 times target area dimension.")
    return s


@dataclass
class SliceScheme:
    v_size: int
    u_size: int
    v_scale: int = 16
    u_scale: int = 16
    v_mid: int = 0
    u_mid: int = 0
    name: str = ""

    @staticmethod
    def spatial_input(src_v, src_u, area_v, area_u, name="aff_in"):
        """A spatial slice sourced from a (possibly differently sized) area."""
        return SliceScheme(src_v, src_u, calc_scale(src_v, area_v), calc_scale(src_u, area_u), 0, 0, name)

    @staticmethod
    def nonspatial(v_size, u_size, name="nsp_in"):
        """Horizontal (non-spatial) slice: scale 0, mid = size/2 (src/cmn/slice_dims.rs:56-93)."""
        return SliceScheme(v_size, u_size, 0, 0, v_size // 2, u_size // 2, name)


@dataclass
class TuftScheme:
    den_class: int
    den_kind: int
    dens_per_tft: int
    syns_per_den: int
    threshold: int = 0
    max_active_dens_l2: int = 0
    src_slcs: List[int] = field(default_factory=list)

    @property
    def syns_per_tft(self):
        return self.dens_per_tft * self.syns_per_den

    @staticmethod
    def basal_proximal(dens, spd, thresh, src, l2=0):
        return TuftScheme(L.DEN_BASAL, L.DEN_PROXIMAL, dens, spd, thresh, l2, list(src))

    @staticmethod
    def basal_distal(dens, spd, thresh, src, l2=0):
        return TuftScheme(L.DEN_BASAL, L.DEN_DISTAL, dens, spd, thresh, l2, list(src))

    @staticmethod
    def apical_distal(dens, spd, thresh, src, l2=0):
        return TuftScheme(L.DEN_APICAL, L.DEN_DISTAL, dens, spd, thresh, l2, list(src))


@dataclass
class LayerScheme:
    name: str
    kind: int
    depth: int
    base_slc: int
    tufts: List[TuftScheme]
    control: int = L.CTRL_NONE
    inhib_radius: int = 4
    smoother_radius: int = 0    # ActivitySmoother control layer (0: none)

    @property
    def spatial(self):
        return self.kind == L.LAYER_SSC


@dataclass
class AreaScheme:
    name: str
    v_size: int
    u_size: int
    slices: List[SliceScheme]
    layers: List[LayerScheme]
    settings: int = 0
    input_slcs: List[int] = field(default_factory=list)
    input_filter: Optional[str] = None   # "retina": intake through the sensory filter

    # ---------------------------------------------------------------- sizes
    def slice_idzs(self):
        out, tot = [], 0
        for s in self.slices:
            out.append(tot)
            tot += s.v_size * s.u_size
        return out, tot

    def axn_len(self):
        return self.slice_idzs()[1]

    def cells(self, li):
        return self.layers[li].depth * self.v_size * self.u_size

    def layer_synapses(self, li):
        return self.cells(li) * sum(t.syns_per_tft for t in self.layers[li].tufts)

    def n_synapses(self):
        return sum(self.layer_synapses(i) for i in range(len(self.layers)))

    def input_range(self):
        """(axon offset, length) covering the contiguous input slices."""
        if not self.input_slcs:
            return 0, 0
        idz, _ = self.slice_idzs()
        s0, s1 = min(self.input_slcs), max(self.input_slcs)
        end = idz[s1] + self.slices[s1].v_size * self.slices[s1].u_size
        return idz[s0], end - idz[s0]

    def layer_axn_range(self, li):
        idz, _ = self.slice_idzs()
        lay = self.layers[li]
        return idz[lay.base_slc], lay.depth * self.v_size * self.u_size

    def rnd_slots(self):
        return sum(2 * len(l.tufts) + 1 for l in self.layers)

    # ---------------------------------------------------------------- C form
    def to_c(self):
        """Return (AreaDesc, keepalive) for the C ABI."""
        slices = (L.SliceDesc * len(self.slices))()
        for i, s in enumerate(self.slices):
            slices[i] = L.SliceDesc(s.v_size, s.u_size, s.v_scale, s.u_scale, s.v_mid, s.u_mid)
        layers = (L.LayerDesc * max(1, len(self.layers)))()
        for i, lay in enumerate(self.layers):
            d = layers[i]
            d.kind = lay.kind
            d.spatial = 1 if lay.spatial else 0
            d.depth = lay.depth
            d.base_axn_slc = lay.base_slc
            d.n_tufts = len(lay.tufts)
            d.control = lay.control
            d.inhib_radius = lay.inhib_radius
            for t, ts in enumerate(lay.tufts):
                td = d.tufts[t]
                td.den_class = ts.den_class
                td.den_kind = ts.den_kind
                td.max_active_dens_l2 = ts.max_active_dens_l2
                td.n_src_slcs = len(ts.src_slcs)
                for k, s in enumerate(ts.src_slcs):
                    td.src_slc_ids[k] = s
                td.dens_per_tft = ts.dens_per_tft
                td.syns_per_den = ts.syns_per_den
                td.threshold = ts.threshold
        desc = L.AreaDesc(self.v_size, self.u_size, len(self.slices), len(self.layers),
                          self.settings, 0, C.cast(slices, C.POINTER(L.SliceDesc)),
                          C.cast(layers, C.POINTER(L.LayerDesc)))
        return desc, (slices, layers)

    def to_json(self):
        return json.dumps(asdict(self), sort_keys=True)


# Default thresholds of the build's recipes (not stated by the workloads; the
# reference test bed uses 400 for its SSC tuft and 1200 for its distal tuft,
# src/tests/testbed.rs:70-118, and 0 is the scheme default, src/cmn/mod.rs:228).
SSC_THRESH = 400
PYR_PRX_THRESH = 0
PYR_DST_THRESH = 1200


def cortical_area(name, side, input_side=None, ssc_depth=1, pyr_depth=2, ssc_dens=4, ssc_spd=32,
                  pyr_dens=4, pyr_spd=32, inhib_radius=4, settings=0, distal=True,
                  thresholds=(SSC_THRESH, PYR_PRX_THRESH, PYR_DST_THRESH), u_side=None):
    """The layer recipe of the workloads: one afferent input slice, a spiny
    stellate layer (1 proximal tuft, basket-cell inhibition) and a pyramidal
    layer (basal proximal tuft from the SSC axons + basal distal tuft from its
    own axons, intra-column inhibition)."""
    v = side
    u = u_side or side
    iv = input_side or side
    slices = [SliceScheme.spatial_input(iv, iv if u_side is None else u, v, u, "aff_in")]
    ssc_base = 1
    for k in range(ssc_depth):
        slices.append(SliceScheme(v, u, name=f"iv_{k}"))
    pyr_base = ssc_base + ssc_depth
    for k in range(pyr_depth):
        slices.append(SliceScheme(v, u, name=f"iii_{k}"))
    ssc_src = [0]
    pyr_prx_src = list(range(ssc_base, ssc_base + ssc_depth))
    pyr_dst_src = list(range(pyr_base, pyr_base + pyr_depth))
    t_ssc, t_prx, t_dst = thresholds
    layers = [
        LayerScheme("iv", L.LAYER_SSC, ssc_depth, ssc_base,
                    [TuftScheme.basal_proximal(ssc_dens, ssc_spd, t_ssc, ssc_src)],
                    L.CTRL_INHIB_SIMPLE, inhib_radius),
    ]
    ptufts = [TuftScheme.basal_proximal(pyr_dens, pyr_spd, t_prx, pyr_prx_src)]
    if distal:
        ptufts.append(TuftScheme.basal_distal(pyr_dens, pyr_spd, t_dst, pyr_dst_src))
    layers.append(LayerScheme("iii", L.LAYER_PYR, pyr_depth, pyr_base, ptufts, L.CTRL_INTRA_COLUMN))
    return AreaScheme(name, v, u, slices, layers, settings, input_slcs=[0])


def config1_area():
    """64x64 columns, 24x24 glyph input, SSC 1x4x32, pyramidal 2 slices x 2 tufts x 4x32."""
    return cortical_area("c1", 64, input_side=24)


def config2_area():
    """256x256 columns, 256x256 SDR input, SSC 1x4x32, pyramidal 2 slices x 2 tufts x 8x64."""
    return cortical_area("c2", 256, pyr_dens=8, pyr_spd=64)


def config3_areas(n=64, seed=0, sides=(32, 192)):
    """64 uneven areas: sides log-uniform in [32, 192] rounded to multiples of 8,
    1-3 slices per layer, 16-64 synapses per dendrite (multiples of 4).  `sides`
    narrows the side range (the same recipe at oracle-checkable sizes)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(8 * round(math.exp(rng.uniform(math.log(sides[0]), math.log(sides[1]))) / 8))
        side = max(8, side)
        sd = int(rng.integers(1, 4))
        pd = int(rng.integers(1, 4))
        s_spd = int(4 * rng.integers(4, 17))
        p_spd = int(4 * rng.integers(4, 17))
        out.append(cortical_area(f"c3_{i}", side, ssc_depth=sd, pyr_depth=pd, ssc_spd=s_spd,
                                 pyr_spd=p_spd))
    return out


def small_area(side=16, input_side=None, **kw):
    return cortical_area("small", side, input_side=input_side, **kw)
