"""Synthetic sensory input streams of the workloads (the reference's encoders and
input generators, src/encode/, src/subcortex/input_generator.rs, are replaced by
these seeded generators).  Every frame is a u8 array of 0/255 axon states.
"""
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
    the next frame and the rest drawn fresh."""

    def __init__(self, n_axons, density=0.02, overlap=0.5, seed=0):
        self.n = n_axons
        self.k = max(1, int(round(density * n_axons)))
        self.keep = int(round(overlap * self.k))
        if self.n - self.k < self.k - self.keep:
            raise ValueError("density too high for the requested overlap")
        self.rng = np.random.default_rng(seed)
        self.active = self.rng.choice(self.n, self.k, replace=False)
        self.member = np.zeros(self.n, dtype=bool)   # membership of self.active
        self.member[self.active] = True

    def next_active(self):
        """The current active set; the next one keeps `keep` of it (a uniform
        random subset) and draws the rest uniformly from outside it."""
        kept = self.active[self.rng.permutation(self.k)[:self.keep]]
        need = self.k - self.keep
        fresh = np.empty(0, dtype=np.int64)
        taken = self.member.copy()
        # rejection sampling against the previous set and the draws so far
        while fresh.size < need:
            cand = self.rng.integers(0, self.n, 2 * (need - fresh.size) + 8)
            cand = cand[~taken[cand]]
            _, first = np.unique(cand, return_index=True)
            cand = cand[np.sort(first)][:need - fresh.size]
            taken[cand] = True
            fresh = np.concatenate([fresh, cand])
        cur = self.active
        self.member[cur] = False
        self.active = np.concatenate([kept, fresh])
        self.member[self.active] = True
        return cur

    def next_frame(self):
        out = np.zeros(self.n, dtype=np.uint8)
        out[self.next_active()] = 255
        return out

    def frames(self, n_cycles):
        out = np.zeros((n_cycles, self.n), dtype=np.uint8)
        for c in range(n_cycles):
            out[c, self.next_active()] = 255
        return out


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
