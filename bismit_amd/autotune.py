"""Project autotuner of the fused kernels' launch geometry.

Picks columns per workgroup tile (tile_v x tile_u), the synapse chunk width a
lane streams (32/16/8 synapses of each metadata stream) and the LDS halo staged
around each tile, by timing a few cycles of each candidate on scratch copies of
the areas (the real areas' state is never advanced).  Every candidate computes
identical results (tests/test_gpu_parity.py::test_cycle_tunings).
"""
import itertools

from .cortex import Batch, CorticalArea

TILES = [(8, 8), (8, 16), (16, 8), (16, 16), (4, 16), (8, 32)]
CHUNKS = [32, 16, 8]
HALOS = [6, 8]


def candidates(tiles=TILES, chunks=CHUNKS, halos=HALOS):
    for (tv, tu), ch, h in itertools.product(tiles, chunks, halos):
        yield dict(tile_v=tv, tile_u=tu, chunk=ch, halo=h)


def tune_area(scheme, ctx, cycles=20, cands=None, seed=0):
    """Best tuning for the plain form of one area; returns (best, table)."""
    a = CorticalArea(scheme, ctx, rnd_seed=seed, layout_seed=seed)
    table = []
    try:
        a.cycle(3)
        for t in (cands or list(candidates())):
            a.set_tuning(**t)
            a.cycle(2)
            a.sync()
            prof = a.profile_stages(cycles)
            ms = sum(p[2] for p in prof) / cycles
            table.append((ms, t))
    finally:
        a.close()
    table.sort(key=lambda x: x[0])
    return table[0][1], table


def tune_batch(schemes, ctx, cycles=10, cands=None, seed=0, graph=True):
    """Best tuning for a batch of areas of these schemes; returns (best, table)."""
    areas = [CorticalArea(s, ctx, rnd_seed=seed + i, layout_seed=seed + i)
             for i, s in enumerate(schemes)]
    b = Batch(areas, ctx)
    table = []
    try:
        b.cycle(2, graph=graph)
        for t in (cands or list(candidates())):
            b.set_tuning(**t)
            b.cycle(2, graph=graph)
            ms = b.time_cycles(cycles, graph=graph) / cycles
            table.append((ms, t))
    finally:
        b.close()
        for a in areas:
            a.close()
    table.sort(key=lambda x: x[0])
    return table[0][1], table
