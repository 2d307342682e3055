"""Hierarchical sequence workload (Config 4): three 128x128 areas in a chain
with feed-forward and feedback tracts, driven by a Zipf/Markov symbol stream.

Area layout (slice order): ff_in (feed-forward input), fb_in (feedback input,
absent on the top area), iv (spiny stellate), iii (pyramidal).  The SSC tuft
sources ff_in; the pyramidal cells carry a basal proximal tuft (from iv), a
basal distal tuft (their own axons: temporal context) and, below the top, an
apical distal tuft from fb_in.  Tracts copy each area's first pyramidal slice
to the next area's ff_in and the previous area's fb_in, one cycle later, as the
reference's thalamus does (src/subcortex/thalamus/thalamus.rs:71-285).
"""
import numpy as np

from . import _lib as L
from . import recipes
from .cortex import Batch, CorticalArea
from .inputs import SymbolStream
from .recipes import AreaScheme, LayerScheme, SliceScheme, TuftScheme


def hier_area(name, side, ff_side, has_fb, pyr_depth=2, dens=4, spd=32,
              thresholds=(400, 0, 1200, 800)):
    slices = [SliceScheme.spatial_input(ff_side, ff_side, side, side, "ff_in")]
    if has_fb:
        slices.append(SliceScheme(side, side, name="fb_in"))
    iv = len(slices)
    slices.append(SliceScheme(side, side, name="iv"))
    iii = len(slices)
    for k in range(pyr_depth):
        slices.append(SliceScheme(side, side, name=f"iii_{k}"))
    t_ssc, t_prx, t_dst, t_apc = thresholds
    ptufts = [TuftScheme.basal_proximal(dens, spd, t_prx, [iv]),
              TuftScheme.basal_distal(dens, spd, t_dst, list(range(iii, iii + pyr_depth)))]
    if has_fb:
        ptufts.append(TuftScheme.apical_distal(dens, spd, t_apc, [1]))
    layers = [LayerScheme("iv", L.LAYER_SSC, 1, iv,
                          [TuftScheme.basal_proximal(dens, spd, t_ssc, [0])], L.CTRL_INHIB_SIMPLE, 4),
              LayerScheme("iii", L.LAYER_PYR, pyr_depth, iii, ptufts, L.CTRL_INTRA_COLUMN)]
    return AreaScheme(name, side, side, slices, layers, 0, input_slcs=[0])


def config4(side=128, sym_side=32):
    """Schemes A (bottom), B, C (top) and the tracts between them as
    (src, src_slice, dst, dst_slice) tuples."""
    a = hier_area("A", side, sym_side, True)
    b = hier_area("B", side, side, True)
    c = hier_area("C", side, side, False)
    schemes = [a, b, c]

    def out_slc(s):
        return s.layers[1].base_slc

    tracts = [(0, out_slc(a), 1, 0),   # A -> B feed-forward
              (1, out_slc(b), 2, 0),   # B -> C feed-forward
              (2, out_slc(c), 1, 1),   # C -> B feedback
              (1, out_slc(b), 0, 1)]   # B -> A feedback
    return schemes, tracts


def tract_ranges(schemes, tracts):
    out = []
    for s_i, s_slc, d_i, d_slc in tracts:
        sidz, _ = schemes[s_i].slice_idzs()
        didz, _ = schemes[d_i].slice_idzs()
        ss = schemes[s_i].slices[s_slc]
        ds = schemes[d_i].slices[d_slc]
        n = ss.v_size * ss.u_size
        assert n == ds.v_size * ds.u_size
        out.append((s_i, sidz[s_slc], d_i, didz[d_slc], n))
    return out


class Hierarchy:
    """Three areas with tracts, run in the plain form or the batched form."""

    def __init__(self, ctx, seed=0, batched=False, side=128):
        self.schemes, tracts = config4(side)
        self.ranges = tract_ranges(self.schemes, tracts)
        self.areas = [CorticalArea(s, ctx, rnd_seed=seed + 11 * i, layout_seed=seed + 13 * i)
                      for i, s in enumerate(self.schemes)]
        self.ctx = ctx
        self.batch = None
        self.symbols = SymbolStream(seed=seed)
        if batched:
            self.batch = Batch(self.areas, ctx)
            arr = (L.Tract * len(self.ranges))(*[L.Tract(*r) for r in self.ranges])
            L.check(ctx.lib.bsm_batch_set_tracts(self.batch.h, arr, len(self.ranges)))

    def frames(self, n):
        return self.symbols.frames(n)

    def run(self, frames, graph=True):
        """Cycle once per frame (frame i feeds area A's ff_in)."""
        n = len(frames)
        if self.batch is not None:
            dev = self.ctx.alloc(frames.nbytes)
            self.ctx.upload(dev, frames)
            k = frames.shape[1]
            ofs = self.schemes[0].input_range()[0]
            self.batch.set_inputs([dev, None, None], [k, 0, 0], [k, 0, 0], [ofs, 0, 0], n)
            self.batch.cycle(n, graph=graph)
            self.batch.sync()
            self.ctx.free(dev)
            return
        for c in range(n):
            # tracts read the previous cycle's outputs before any area cycles
            for s_i, s_ofs, d_i, d_ofs, ln in self.ranges:
                L.check(self.ctx.lib.bsm_area_copy_axons(self.areas[d_i].h, d_ofs,
                                                         self.areas[s_i].h, s_ofs, ln))
            self.areas[0].intake(frames[c])
            for a in self.areas:
                a.cycle(1)

    def close(self):
        if self.batch is not None:
            self.batch.close()
        for a in self.areas:
            a.close()
