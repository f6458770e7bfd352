"""Autotuner of the batched plan: picks the uv tile edge, the records per tile work item and the
BDA blocks per gridding work item by timing whole major cycles (grid + degrid, all facets) of
the chunk at hand, then rebuilds the plan with the fastest options.

The trade-offs it settles are workload-dependent: a larger tile amortises the halo (the tile
image is (T + 2 h)^2 for T^2 owned cells) but holds more LDS per workgroup and leaves fewer
tiles to spread over the device; a larger item amortises the tile zero / flush but balances
uneven bins worse (visibility counts per facet and w-plane vary by 50x under baseline-dependent
averaging); more blocks per gridding item cut work-queue traffic but lengthen the tail.
"""
import itertools

import numpy as np
import torch

DEFAULT_SPACE = {
    "tile": (16, 32, 64),
    "chunk": (1024, 2048, 4096),
    "blocks_per_item": (64, 128, 256),
}


def candidates(space=None, halo=0):
    """Every combination of the search space; tiles no larger than the stamp halo are skipped
    (a tile must be wider than the halo so that a grid cell lies in at most two tiles per axis)."""
    space = dict(DEFAULT_SPACE if space is None else space)
    keys = sorted(space)
    out = []
    for vals in itertools.product(*(space[k] for k in keys)):
        c = dict(zip(keys, vals))
        if c.get("tile", 32) <= halo:
            continue
        out.append(c)
    return out


def time_cycle(fs, reps=2, warmup=1, stream=None):
    """Mean milliseconds of one batched cycle (zero grids, grid, degrid) over `reps` runs, on the
    current plan of `fs`; the visibilities are restored afterwards."""
    s = torch.cuda.current_stream() if stream is None else stream
    vis0 = fs.t["vis"].clone()
    for _ in range(warmup):
        fs.zero_grids()
        fs.grid_batched(stream=s)
        fs.degrid_batched(stream=s)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fs.zero_grids()
        fs.grid_batched(stream=s)
        fs.degrid_batched(stream=s)
    e1.record(s)
    e1.synchronize()
    fs.t["vis"].copy_(vis0)
    return e0.elapsed_time(e1) / reps


def autotune(fs, space=None, reps=2, precision="mixed", stream=None, log=None):
    """Time every candidate plan of `fs` (a DeviceFacetSet) and keep the fastest.  Returns
    (best options, [(options, ms per cycle), ...]).  Plans are built one at a time, so the
    device memory peak is one plan."""
    f0 = fs.P.facets[0]
    overs = int(f0.winfos[3])
    smax = max(int(round(np.sqrt(np.asarray(p).size))) // overs for p in f0.wplanes)
    halo = (smax - 1) // 2
    results = []
    best, best_ms = None, float("inf")
    for c in candidates(space, halo):
        fs.plan_batched(precision=precision, stream=stream, **c)
        ms = time_cycle(fs, reps=reps, stream=stream)
        results.append((c, ms))
        if log:
            log("autotune %s: %.2f ms" % (c, ms))
        if ms < best_ms:
            best, best_ms = c, ms
    fs.plan_batched(precision=precision, stream=stream, **best)
    return best, results
