"""Synthetic sensory input streams of the workloads (the reference's encoders and
input generators, src/encode/, src/subcortex/input_generator.rs, are replaced by
these seeded generators).  Every frame is a u8 array of 0/255 axon states.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L

U32 = float(1 << 32)


class GlyphStream:
    """24x24 binary bitmaps, each pixel on with probability p, drawn from the
    engine's xorshift128 stream; a new glyph every `hold` cycles."""

    def __init__(self, side=24, p=0.15, hold=10, seed=0):
        self.side, self.p, self.hold, self.seed = side, p, hold, seed

    def frames(self, n_cycles):
        n_glyphs = (n_cycles + self.hold - 1) // self.hold
        draws = L.xorshift128(self.seed, n_glyphs * self.side * self.side)
        on = draws.astype(np.float64) < self.p * U32
        glyphs = np.where(on, 255, 0).astype(np.uint8).reshape(n_glyphs, self.side * self.side)
        return np.repeat(glyphs, self.hold, axis=0)[:n_cycles]


class SdrStream:
    """Sparse distributed frames over n_axons: exactly round(density*n) active
    (255) axons, with `overlap` of the previous frame's active set carried into
    the next frame (a uniform subset) and the rest drawn fresh from outside it.
    The stream runs in the engine library (bsm_sdr_*, xorshift128 seeded by
    splitmix64(seed)), so that the serving loop fills every stepped session's
    frame with one call (next_packed)."""

    def __init__(self, n_axons, density=0.02, overlap=0.5, seed=0):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.bsm_sdr_create(int(n_axons), float(density), float(overlap),
                                        C.c_uint64(int(seed) & (2**64 - 1)), C.byref(self.h)))
        n, k = C.c_uint32(), C.c_uint32()
        L.check(self.lib.bsm_sdr_info(self.h, C.byref(n), C.byref(k)))
        self.n, self.k = n.value, k.value

    def next_frame(self):
        out = np.empty(self.n, dtype=np.uint8)
        L.check(self.lib.bsm_sdr_next(self.h, L.vptr(out)))
        return out

    def frames(self, n_cycles):
        out = np.empty((n_cycles, self.n), dtype=np.uint8)
        if n_cycles:
            L.check(self.lib.bsm_sdr_frames(self.h, n_cycles, L.vptr(out)))
        return out

    def next_indices(self):
        """The next frame as its active axon indices (the frame next_frame
        would return, as 255 at these indices)."""
        out = np.empty(self.k, dtype=np.uint32)
        L.check(self.lib.bsm_sdr_next_indices(self.h, L.vptr(out)))
        return out

    def close(self):
        if self.h:
            self.lib.bsm_sdr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def next_packed(streams, buf, offsets, sdr=False):
    """The next frame of each stream, stream i's written at buf[offsets[i]:]:
    dense n bytes, or with sdr=True an SDR entry of bsm_batch_cycle_masked
    (uint32 k, then the k active indices; offsets 4-byte aligned)."""
    if not streams:
        return
    hs = np.array([s.h.value for s in streams], dtype=np.uint64)
    ofs = np.ascontiguousarray(offsets, dtype=np.uint64)
    assert buf.dtype == np.uint8 and buf.flags.c_contiguous
    size = (lambda s: 4 * (s.k + 1)) if sdr else (lambda s: s.n)
    assert all(int(o) + size(s) <= buf.size for o, s in zip(ofs, streams))
    fn = streams[0].lib.bsm_sdr_next_packed_sdr if sdr else streams[0].lib.bsm_sdr_next_packed
    L.check(fn(L.vptr(hs), len(streams), L.vptr(buf), L.vptr(ofs)))


class SymbolStream:
    """Zipf(s) marginal over n_symbols with first-order Markov transitions
    (each symbol has `succ` successors from a seeded sparse transition matrix);
    each symbol is a fixed random SDR of `bits` active inputs over n_inputs."""

    def __init__(self, n_symbols=96, zipf_s=1.1, succ=8, bits=40, n_inputs=1024, seed=0):
        rng = np.random.default_rng(seed)
        self.n_symbols, self.n_inputs = n_symbols, n_inputs
        ranks = np.arange(1, n_symbols + 1, dtype=np.float64)
        zipf = ranks ** -zipf_s
        zipf /= zipf.sum()
        self.zipf = zipf
        self.succ = np.empty((n_symbols, succ), dtype=np.int64)
        self.prob = np.empty((n_symbols, succ), dtype=np.float64)
        for s in range(n_symbols):
            nxt = rng.choice(n_symbols, succ, replace=False, p=zipf)
            w = zipf[nxt] * rng.uniform(0.5, 1.5, succ)
            self.succ[s] = nxt
            self.prob[s] = w / w.sum()
        self.codes = np.stack([rng.choice(n_inputs, bits, replace=False) for _ in range(n_symbols)])
        self.rng = rng
        self.sym = int(rng.choice(n_symbols, p=zipf))

    def symbols(self, n):
        out = np.empty(n, dtype=np.int64)
        for i in range(n):
            out[i] = self.sym
            j = self.rng.choice(self.succ.shape[1], p=self.prob[self.sym])
            self.sym = int(self.succ[self.sym, j])
        return out

    def frames(self, n_cycles):
        syms = self.symbols(n_cycles)
        out = np.zeros((n_cycles, self.n_inputs), dtype=np.uint8)
        for i, s in enumerate(syms):
            out[i, self.codes[s]] = 255
        return out


def uniform_glyph_check(frames):
    """Fraction of on pixels (helper for tests)."""
    return float((frames != 0).mean()) if frames.size else math.nan
