"""Throughput of one grid + degrid major cycle over all facets (vis/s), cfg2 (81-facet MeerKAT-like).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain] [--no-cpu]

A step = grid the chunk's visibilities onto all 81 facet grids, then degrid all 81 model grids
and subtract them from the visibilities (vis -= sum over facets).  Inputs are resident in HBM
before timing starts.  value = Nvis / t_step with Nvis = rows x channels (4-correlation
visibilities presented, flagged ones included).

Roofline: HBM-bound scatter/gather.  Algorithmic bytes per visibility-facet pair (one 4-corr
visibility gridded to and degridded from one facet): grid 32 (vis) + 4 (flag) + 4 (weight)
+ 24/C (uvw); degrid 32 + 32 (vis read + write) + 24/C; plus per facet per cycle the grid
(zero + write), the model grid read, the BDA maps and the CF tables.  achieved = those bytes
of the dominant kernel / its mean launch time (HIP events on the launch stream).
The CPU baseline is the reference gridder (oracle/_ref, built from the reference sources) on
a stride sample of the same BDA blocks and a few facets, scaled to the full cycle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(P, nf, kind):
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    gbytes = P.nband * P.npol * P.npix * P.npix * 8
    bda = 4 * (P.bda_grid.size if kind == "grid" else P.bda_degrid.size)
    cfb = 0
    for f in P.facets[:nf]:
        cfb += sum(p.size * 8 for p in f.wplanes) + sum(p.size * 8 for p in f.wplanes_conj)
    cfb //= max(1, len(P.facets[:nf]))
    if kind == "grid":
        per_vis = 40 + 24.0 / C
        per_facet = 2 * gbytes + bda + cfb  # zero + accumulate
    else:
        per_vis = 64 + 24.0 / C
        per_facet = gbytes + bda + cfb
    return nf * (nvis * per_vis + per_facet)


def cpu_baseline(P, target_s=12.0):
    """Reference gridder on a stride sample of the BDA blocks, `nfs` facets, one core."""
    from ddfacet_amd import bda as bdamod
    from ddfacet_amd.problems import Problem
    from oracle import refmod
    if not refmod.available():
        return None
    mod = refmod.module()
    vis_full = P.vis
    import torch
    nb = bdamod.nblocks_of(P.bda_grid)
    starts = bdamod.block_starts(P.bda_grid)
    sizes = P.bda_grid[2:2 + nb]

    def sample(stride):
        sel = np.arange(0, nb, stride)
        rows = np.unique(np.concatenate([P.bda_grid[starts[b] + 2:starts[b] + sizes[b]] for b in sel]))
        remap = -np.ones(P.uvw.shape[0], np.int64)
        remap[rows] = np.arange(rows.size)
        recs, szs = [], []
        for b in sel:
            r = P.bda_grid[starts[b]:starts[b] + sizes[b]].astype(np.int64).copy()
            r[2:] = remap[r[2:]]
            recs.append(r)
            szs.append(sizes[b])
        m = np.concatenate([np.array([sel.size, 0], np.int64), np.array(szs, np.int64)] + recs).astype(np.int32)
        nvis_s = int(sum((r[1] - r[0]) * (r.size - 2) for r in recs))
        if isinstance(vis_full, torch.Tensor):
            vs = vis_full[torch.as_tensor(rows, device=vis_full.device)].cpu().numpy()
        else:
            vs = vis_full[rows]
        j = None
        if P.jones is not None:
            j = dict(P.jones)
            j["kms_time"] = P.jones["kms_time"][rows]
        S = Problem(uvw=np.ascontiguousarray(P.uvw[rows]), vis=np.ascontiguousarray(vs),
                    flags=np.ascontiguousarray(P.flags[rows]), weights=np.ascontiguousarray(P.weights[rows]),
                    freqs=P.freqs, a0=np.int32(P.a0[rows]), a1=np.int32(P.a1[rows]), chan_map=P.chan_map,
                    chan_map_degrid=P.chan_map_degrid, bda_grid=m, bda_degrid=m, facets=P.facets,
                    npix=P.npix, nband=P.nband, npol=P.npol, incr=P.incr, corr_codes=P.corr_codes,
                    stokes_codes=P.stokes_codes, jones=j)
        return S, nvis_s

    def run(S, facets):
        t0 = time.perf_counter()
        for fi in facets:
            g = S.new_grid()
            s = np.zeros((S.nband, S.npol))
            mod.pyGridderWPol32(*S.grid_args(fi, g, s))
            v = S.vis.copy()
            mod.pyDeGridderWPol(*S.degrid_args(fi, g, v))
        return time.perf_counter() - t0

    S, nv = sample(400)
    nf = len(P.facets)
    facets = [0, nf // 2, nf - 1]
    t = run(S, facets)
    scale = max(1, int(target_s / max(t, 1e-3)))
    stride = max(1, 400 // scale)
    S, nv = sample(stride)
    t = run(S, facets)
    # vis-facet pairs per second on one core, expressed as full-cycle visibilities per second
    vf_per_s = nv * len(facets) / t
    return dict(value=vf_per_s / nf, unit="vis/s", cores=1, kind="reference",
                sample="reference _pyGridderSmearPols3x (oracle/_ref, built from DDFacet/Gridder sources), "
                       "grid+degrid of every %d-th BDA block (%d vis) for facets %s, 1 core, %.1f s; "
                       "value = vis-facet pairs/s / %d facets" % (stride, nv, facets, t, nf))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="plain", choices=["plain", "batched"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--facets", type=int, default=0, help="use only the first N facets (debug)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a device copy of 1 GiB (reads 1 GiB, writes 1 GiB), to scale the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        x = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3}))
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (see DESIGN.md for how "
              "the path would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from ddfacet_amd import configs
    from ddfacet_amd.facets import DeviceFacetSet
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.time()
    P = configs.cfg2(device="cuda", log=log)
    if a.facets:
        P.facets = P.facets[:a.facets]
    fs = DeviceFacetSet(P, device=dev)
    nf = fs.nf
    gen = torch.Generator(device=dev).manual_seed(7)
    fs.models.copy_(torch.randn(fs.models.shape, generator=gen, device=dev, dtype=torch.complex64) * 1e-3)
    vis0 = fs.t["vis"].clone()
    log("setup %.1fs" % (time.time() - t0))

    stream = torch.cuda.current_stream()
    ev = {"grid": [], "degrid": []}

    def step(record):
        fs.zero_grids()
        if record:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e2 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        if a.mode == "plain":
            fs.grid_plain(stream=stream)
        else:
            fs.grid_batched(stream=stream)
        if record:
            e1.record(stream)
        if a.mode == "plain":
            fs.degrid_plain(stream=stream)
        else:
            fs.degrid_batched(stream=stream)
        if record:
            e2.record(stream)
            ev["grid"].append((e0, e1))
            ev["degrid"].append((e1, e2))

    for _ in range(a.warmup):
        step(False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        step(True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ms = (t2 - t1) * 1e3 / a.steps
    fs.t["vis"].copy_(vis0)
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    value = nvis / (ms * 1e-3)
    tg = np.mean([x.elapsed_time(y) for x, y in ev["grid"]])
    td = np.mean([x.elapsed_time(y) for x, y in ev["degrid"]])
    launches = nf if a.mode == "plain" else 1
    per = {"grid": tg / launches, "degrid": td / launches}
    dom = "grid" if tg >= td else "degrid"
    nfl = 1 if a.mode == "plain" else nf
    ach = algorithmic_bytes(P, nfl, dom) / (per[dom] * 1e-3)
    log("grid %.2f ms, degrid %.2f ms per cycle (%s)" % (tg, td, a.mode))
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(P)
        except Exception as e:  # the baseline must not hide the GPU result
            log("cpu baseline failed: %r" % (e,))
    out = {
        "metric": "vis/sec grid+degrid, 81-facet MeerKAT-like",
        "value": value, "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": {"workload": "cfg2: 81 facets (9x9, 455->735 px at 1.5\"), 64 ant MeerKAT-like, "
                               "600x8s, 64 ch 0.9-1.67 GHz, 4 corr, per-vis DD Jones, default BDA (0.02)",
                   "rows": R, "channels": C, "nvis": nvis, "facets": nf,
                   "bda_blocks": fs.nblocks_grid, "mode": a.mode, "grid_ms": tg, "degrid_ms": td},
        "roofline": {"bound": "hbm", "kernel": ("grid_plain" if dom == "grid" else "degrid_plain")
                     if a.mode == "plain" else dom + "_batched",
                     "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": None,
                     "launch_ms": per[dom], "algorithmic_bytes_per_launch": algorithmic_bytes(P, nfl, dom)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
