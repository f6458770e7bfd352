// Batched all-facet major cycle for gfx950: every facet of a visibility chunk in a few launches.
//
// Plan (once per chunk; depends on uvw, flags, freqs, maps and facet geometry only):
//   place   per (BDA block, facet) placement record: band, grid cell, CF plane and oversampling
//           offset.  Lanes are facets; each lane forms the block means in the reference's
//           sequential order (gridder.h:211-289, degridder.h:141-167), so the integer maps equal
//           the reference's.
//   bins    counting sort of the valid records by (facet, band, uv tile of T x T cells), split
//           into work items of at most `chunk` records.
//   vlist   the (block, row, channel) of every visibility of the degrid map, in map order.
// Cycle (per major cycle, all facets):
//   grid A  persistent, work queue over (facet group, block range); lanes are facets, each lane
//           walks its block's visibilities once (read once for all facets of the group) and forms
//           the phased, Jones-corrected block sum in FP64 -> Stokes record value V[b][f].
//           sumwt / SumJones / SumJonesChan go through per-workgroup LDS accumulators.
//   grid B  persistent over bins: the S x S stamps of a bin accumulate in an LDS tile (T + 2 h)^2,
//           flushed once into the facet grid with global float atomics (non-zero cells only).
//   degrid A persistent over bins: the model tile is staged in LDS; one wavefront gathers each
//           record's stamp (FP32 products summed with a DPP wave reduction, in FP64 as
//           degridder.h:202-212 in the FP64 precision, in FP32 otherwise) -> M[b][f].
//   degrid B lanes are visibilities; each lane sums the model of every facet (Stokes -> corr,
//           Jones, phasor, decorrelation) in registers and writes its visibility once.
// Work queues are device counters reset by a small kernel (no memset nodes, capture-safe).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/ddfgpu.h"
#include "ddf_device.h"
#include "ddf_internal.h"

using namespace ddf;

namespace {

// packed placement record: locx 16 | locy 16 | plane 8 | io 8 | jo 8 | band 7 | valid 1
__device__ __forceinline__ uint64_t pack_place(int locx, int locy, int plane, int io, int jo, int band) {
  return uint64_t(uint32_t(locx) & 0xFFFF) | (uint64_t(uint32_t(locy) & 0xFFFF) << 16) |
         (uint64_t(plane & 0xFF) << 32) | (uint64_t(io & 0xFF) << 40) | (uint64_t(jo & 0xFF) << 48) |
         (uint64_t(band & 0x7F) << 56) | (uint64_t(1) << 63);
}
struct Unpacked {
  int locx, locy, plane, io, jo, band;
};
__device__ __forceinline__ Unpacked unpack_place(uint64_t r) {
  return Unpacked{int(r & 0xFFFF), int((r >> 16) & 0xFFFF), int((r >> 32) & 0xFF), int((r >> 40) & 0xFF),
                  int((r >> 48) & 0xFF), int((r >> 56) & 0x7F)};
}
__device__ __forceinline__ bool place_valid(uint64_t r) { return (r >> 63) != 0; }

// Placement with the integer results kept (iw -> plane index in the put order of the lists).
__device__ inline bool place_full(const BatchParams& p, double um, double vm, double wm, double recip_wvl,
                                  bool degrid, int* out_locx, int* out_locy, int* out_plane, int* out_io,
                                  int* out_jo) {
#pragma clang fp contract(off)
  double fiw = (double(p.nw) - 1) * fabs(wm) * (p.ref_wave * recip_wvl) / p.wmax;
  if (!(fiw >= 0.0 && fiw < 2147483647.0)) return false;
  int iw = int(rint(fiw));
  if (iw >= p.nw) return false;
  // put(): W > 0 -> Wplanes[iw] (plane iw), else WplanesConj[iw] (plane nw + iw);
  // get() passes the lists swapped, so W > 0 -> WplanesConj[iw].
  bool first = wm > 0;
  int plane = (first != degrid) ? iw : p.nw + iw;
  int S = p.support[iw];
  int O = p.overs;
  int sup = (S - 1) / 2;
  double posx = double(p.nx) * p.incr0 * um * recip_wvl + double(p.nx / 2);
  double posy = double(p.nx) * p.incr1 * vm * recip_wvl + double(p.ny / 2);
  if (!(fabs(posx) < 1e9 && fabs(posy) < 1e9)) return false;
  int locx = int(rint(posx)), locy = int(rint(posy));
  if (locx - sup < 0 || locx + sup >= p.nx || locy - sup < 0 || locy + sup >= p.ny) return false;
  int offx = int(rint((double(locx) - posx) * double(O))) + (S * O - 1) / 2;
  int offy = int(rint((double(locy) - posy) * double(O))) + (S * O - 1) / 2;
  *out_io = clampi(offy - sup * O, 0, O - 1);
  *out_jo = clampi(offx - sup * O, 0, O - 1);
  *out_locx = locx;
  *out_locy = locy;
  *out_plane = plane;
  return true;
}

__device__ __forceinline__ int plane_support(const BatchParams& p, int plane) {
  return p.support[plane < p.nw ? plane : plane - p.nw];
}

// ---- plan: placements -------------------------------------------------------------------------
// grid: one thread per (block, facet); threads of a block's facets are adjacent, so the block's
// visibilities are read by neighbouring lanes (broadcast loads).
template <bool DEGRID>
__global__ __launch_bounds__(256) void plan_place(BatchParams p, const VisP v, uint64_t* place) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t total = v.nblocks * p.nf;
  for (int64_t t = tid; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = t / p.nf;
    const int f = int(t - b * p.nf);
    const FacetDev fd = gload(p.facets + f);
    uint64_t rec = 0;
    BlockRec br = read_block(v, b);
    const int nch = br.ch1 - br.ch0;
    bool skip = (br.nr <= 0 || nch <= 0);
    if (!DEGRID && v.sparse && !v.sparse[b]) skip = true;
    if (!skip) {
      const int band = v.chan_map[br.ch0];
      double su = 0, sv = 0, sw = 0, fm = 0;
      int cnt = 0;
      if (!DEGRID) {
        for (int r = 0; r < br.nr; ++r) {
          const int64_t row = br.rows[r];
          if (row < 0 || row >= v.nrow) continue;
          const double* uvw = v.uvw + 3 * row;
        const DDF_GLOBAL double* guvw = gptr(v.uvw) + 3 * row;
          for (int c = br.ch0; c < br.ch1; ++c) {
            if (gptr(v.flags)[(size_t(row) * v.nchan + c) * v.ncorr]) continue;
            su += shifted(guvw[0], guvw[2], fd.cu);
            sv += shifted(guvw[1], guvw[2], fd.cv);
            sw += guvw[2];
            fm += gptr(v.freqs)[c];
            ++cnt;
          }
        }
        if (cnt > 0) {
          su /= cnt; sv /= cnt; sw /= cnt; fm /= cnt;
        }
      } else {
        for (int c = br.ch0; c < br.ch1; ++c) fm += gptr(v.freqs)[c];
        fm /= double(nch);
        for (int r = 0; r < br.nr; ++r) {
          const int64_t row = br.rows[r];
          if (row < 0 || row >= v.nrow) continue;
          const double* uvw = v.uvw + 3 * row;
        const DDF_GLOBAL double* guvw = gptr(v.uvw) + 3 * row;
          su += shifted(guvw[0], guvw[2], fd.cu);
          sv += shifted(guvw[1], guvw[2], fd.cv);
          sw += guvw[2];
          ++cnt;
        }
        if (cnt > 0) {
          su /= cnt; sv /= cnt; sw /= cnt;
        }
      }
      int lx, ly, pl, io, jo;
      if (cnt > 0 && band >= 0 && band < p.nband && band < 128 &&
          place_full(p, su, sv, sw, fm / kC, DEGRID, &lx, &ly, &pl, &io, &jo))
        rec = pack_place(lx, ly, pl, io, jo, band);
    }
    place[t] = rec;
  }
}

// Pair codes of a map: one byte per (block, facet), 0x80 | band for a placed record, 0 for a
// dropped one; the cycle passes that need no more of a placement (grid pass A, the adjoint of
// degrid pass B) read these instead of the 8-byte records.
__global__ __launch_bounds__(256) void plan_pair_codes(const uint64_t* place, int64_t n, uint8_t* code) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t rec = place[t];
    code[t] = place_valid(rec) ? uint8_t(0x80u | uint32_t(unpack_place(rec).band)) : uint8_t(0);
  }
}

// ---- plan: binning (counting sort) --------------------------------------------------------------
// Bin = (facet, band, tile y, tile x, CF plane): the records of one tile are contiguous and, inside
// the tile, grouped by CF plane so consecutive stamps reuse the same plane (cache locality).
__device__ __forceinline__ int64_t bin_of(const BatchParams& p, int64_t t, uint64_t rec) {
  const int f = int(t % p.nf);
  Unpacked u = unpack_place(rec);
  const int tx = u.locx / p.tile, ty = u.locy / p.tile;
  return (((int64_t(f) * p.nband + u.band) * p.ntiles_y + ty) * p.ntiles_x + tx) * (2 * p.nw) + u.plane;
}

// Binning with one atomic per run of equal bins: a wavefront takes 64 consecutive blocks of one
// facet (lane = block), whose placements mostly share a bin (the blocks of a baseline are
// consecutive and move slowly across the uv tiles), and the lanes of one run of equal bins share one
// counter update (plan_count) or one slot reservation (plan_fill).  The waves of one block range
// read the same lines of the block-major placement array.
struct BinRun {
  int64_t bin;          // this lane's bin (-1: no valid record)
  int64_t t;            // this lane's record
  bool ok, head;        // valid record; first lane of its run
  uint64_t valid, heads;  // masks of valid lanes and run heads
};
__device__ __forceinline__ BinRun bin_run(const BatchParams& p, int64_t nblocks, const uint64_t* place, int64_t w,
                                          int lane) {
  BinRun r;
  const int f = int(w % p.nf);
  const int64_t b = (w / p.nf) * 64 + lane;
  r.t = b * p.nf + f;
  r.ok = false;
  r.bin = -1;
  if (b < nblocks) {
    const uint64_t rec = place[r.t];
    if (place_valid(rec)) {
      r.ok = true;
      r.bin = bin_of(p, r.t, rec);
    }
  }
  const int lo = __shfl_up(int(uint32_t(r.bin)), 1, 64), hi = __shfl_up(int(uint32_t(uint64_t(r.bin) >> 32)), 1, 64);
  const int64_t pbin = int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
  // a valid lane after an invalid one (bin -1) always opens a run
  r.head = r.ok && (lane == 0 || pbin != r.bin);
  r.valid = __ballot(r.ok);
  r.heads = __ballot(r.head);
  return r;
}

__global__ __launch_bounds__(256) void plan_count(BatchParams p, int64_t nblocks, const uint64_t* place,
                                                  uint32_t* counts) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((nblocks + 63) / 64) * p.nf;
  const int64_t wstride = int64_t(gridDim.x) * (blockDim.x >> 6);
  for (int64_t w = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nwaves; w += wstride) {
    const BinRun r = bin_run(p, nblocks, place, w, lane);
    if (r.head) {
      // the run: this lane up to the next head (lanes past a run's end are invalid)
      const uint64_t above = r.heads & ~((2ull << lane) - 1);
      const uint64_t upto = above ? (above & (~above + 1)) - 1 : ~0ull;
      const uint64_t mine = upto & ~((1ull << lane) - 1);
      atomicAdd(&counts[r.bin], uint32_t(__popcll(r.valid & mine)));
    }
  }
}

__global__ __launch_bounds__(256) void plan_fill(BatchParams p, int64_t nblocks, const uint64_t* place,
                                                 const int64_t* bin_start, uint32_t* fill, uint32_t* recs) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((nblocks + 63) / 64) * p.nf;
  const int64_t wstride = int64_t(gridDim.x) * (blockDim.x >> 6);
  for (int64_t w = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nwaves; w += wstride) {
    const BinRun r = bin_run(p, nblocks, place, w, lane);
    uint32_t base = 0;
    if (r.head) {
      const uint64_t above = r.heads & ~((2ull << lane) - 1);
      const uint64_t upto = above ? (above & (~above + 1)) - 1 : ~0ull;
      const uint64_t mine = upto & ~((1ull << lane) - 1);
      base = atomicAdd(&fill[r.bin], uint32_t(__popcll(r.valid & mine)));
    }
    // each valid lane: its run's head h (the highest head at or below it) and its rank in the run
    const uint64_t below = r.heads & ((2ull << lane) - 1);
    const int h = below ? 63 - __clzll(below) : 0;
    const uint32_t hbase = uint32_t(__shfl(int(base), h, 64));
    if (r.ok) {
      const uint64_t between = ((1ull << lane) - 1) & ~((1ull << h) - 1);
      const uint32_t slot = hbase + uint32_t(__popcll(r.valid & between));
      recs[bin_start[r.bin] + slot] = uint32_t(r.t);
    }
  }
}

// Record counts of every (tile, plane group): the sum of the group's plane bins (groups of
// consecutive planes, the bounds g * nsub / ngrp), one thread per (tile, group).
__global__ __launch_bounds__(256) void plan_group_counts(const uint32_t* counts, int64_t ntile, int nsub, int ngrp,
                                                         uint32_t* gcount) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < ntile * ngrp; k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = k / ngrp;
    const int g = int(k - i * ngrp);
    uint32_t c = 0;
    for (int q = g * nsub / ngrp; q < (g + 1) * nsub / ngrp; ++q) c += counts[i * nsub + q];
    gcount[k] = c;
  }
}

// First record of every bin: the tile's first record plus the counts of its earlier plane bins, one
// thread per tile.
__global__ __launch_bounds__(256) void plan_bin_starts(const uint32_t* counts, const int64_t* tile_start, int64_t ntile,
                                                       int nsub, int64_t* bin_start) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < ntile; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t s = tile_start[i];
    for (int q = 0; q < nsub; ++q) {
      bin_start[i * nsub + q] = s;
      s += counts[i * nsub + q];
    }
  }
}

// Order of the records inside a work item, independent of the atomics of plan_fill.  Degridding
// (STAMP false): by CF plane, then oversampled sub-kernel, then record, so that the lanes of a
// wavefront (consecutive records) read the same sub-kernels (one cache line for a cell instead of
// one per lane); key = plane << 43 | sub << 32 | record.  Gridding (STAMP true): by support, then
// stamp position in the tile image, then record, so that the stamps of one position come in runs
// that grid pass B sums in registers; key = S << 46 | base << 32 | record (base taken to 14 bits:
// tile images that fit LDS have fewer cells, and grid pass B finds runs from the full stamp
// record, so the key only decides the order).
template <bool STAMP>
__global__ __launch_bounds__(256) void plan_sort_keys(BatchParams p, const uint64_t* place, const uint32_t* recs,
                                                      int64_t n, uint64_t* keys) {
  const int W = p.tile + 2 * p.halo;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t t = recs[k];
    const Unpacked u = unpack_place(place[t]);
    if (STAMP) {
      const int S = plane_support(p, u.plane);
      const int sup = (S - 1) / 2;
      const int x0 = (u.locx / p.tile) * p.tile - p.halo, y0 = (u.locy / p.tile) * p.tile - p.halo;
      const uint64_t base = uint64_t((u.locy - sup - y0) * W + (u.locx - sup - x0)) & 0x3FFF;
      keys[k] = (uint64_t(S & 0x1F) << 46) | (base << 32) | uint64_t(t);
    } else {
      const uint64_t sub = uint64_t(u.io * p.overs + u.jo);
      keys[k] = (uint64_t(u.plane) << 43) | (sub << 32) | uint64_t(t);
    }
  }
}

__global__ __launch_bounds__(256) void plan_item_bounds(const BinItem* items, uint32_t n, uint32_t* begin, uint32_t* end) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    begin[i] = items[i].begin;
    end[i] = items[i].end;
  }
}

__global__ __launch_bounds__(256) void plan_keys_to_recs(const uint64_t* keys, int64_t n, uint32_t* recs) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
    recs[k] = uint32_t(keys[k]);
}

// Stamp records of the binned records (record k of the bin list): see Stamp.
__global__ __launch_bounds__(256) void plan_stamp_records(BatchParams p, const uint64_t* place, const uint32_t* recs,
                                                          int64_t n, uint64_t* bstamp) {
  const int W = p.tile + 2 * p.halo;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t t = recs[k];
    const Unpacked u = unpack_place(place[t]);
    const int f = int(t % uint32_t(p.nf));
    const int S = plane_support(p, u.plane);
    const int sup = (S - 1) / 2;
    const int x0 = (u.locx / p.tile) * p.tile - p.halo, y0 = (u.locy / p.tile) * p.tile - p.halo;
    const int base = (u.locy - sup - y0) * W + (u.locx - sup - x0);
    const int64_t cfoff = gptr(gload(p.facets + f).cf_offset)[u.plane] + int64_t(u.io * p.overs + u.jo) * S * S;
    const uint32_t sub = uint32_t(u.io * p.overs + u.jo);
    bstamp[k] = uint64_t(uint32_t(base) & 0xFFFF) | (uint64_t(S & 0x1F) << 16) | (uint64_t(sub & 0x7FF) << 21) |
                (uint64_t(uint32_t(cfoff)) << 32);
  }
}

// CF copy for degrid pass A: each sub-kernel's S x S cells with its rows padded to an even pitch
// Sp = S + (S & 1) (the pad cell is zero), plane pl at twice its put-order offset (a padded plane,
// O2 S Sp cells, fits below twice the next plane's offset), so that every row starts on a 16-byte
// boundary and a lane reads two cells per load.
__global__ __launch_bounds__(256) void plan_cf_pad(const float2* in, float2* out, const int64_t* off,
                                                   const int32_t* support, int nw, int overs) {
  const int pl = blockIdx.y;
  const int S = support[pl < nw ? pl : pl - nw];
  const int Sp = S + (S & 1);
  const int S2 = S * S, SSp = S * Sp, O2 = overs * overs;
  const int64_t o = off[pl];
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < int64_t(O2) * SSp;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int k = int(i / SSp);
    const int rem = int(i - int64_t(k) * SSp);
    const int r = rem / Sp, c = rem - r * Sp;
    out[2 * o + i] = c < S ? in[o + int64_t(k) * S2 + r * S + c] : make_float2(0.f, 0.f);
  }
}

// Prepared Jones table of the batched plan: entry (t, dir, ant, chan) of a raw table
// [t][dir][ant][chan][4] through GiveJones (interpolation mode) and, for the DD table, the
// amplitude / phase part of NormJones, in FP64 at [t][ant][chan][dir][4] (facet-fastest: the
// lanes of a wavefront are facets).  ScaleJones depends on the visibility and stays in the passes.
__global__ __launch_bounds__(256) void plan_prepare_jones(const float2* src, double2* dst, float2* dstf, JonesP jp,
                                                          int dd) {
  const int* dims = dd ? jp.kdims : jp.bdims;
  const int nt = dims[0], nd = dims[1], na = dims[2], nch = dims[3];
  JonesP q = jp;
  q.scale = 0;
  const int64_t total = int64_t(nt) * nd * na * nch;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i;
    const int ch = int(r % nch); r /= nch;
    const int a = int(r % na); r /= na;
    const int d = int(r % nd); r /= nd;
    const int t = int(r);
    M2 J = give_jones(src, dims, t, a, d, ch, jp.interp, jp.coefs, jp.ncoefs);
    if (dd) norm_jones(q, J, nullptr, 0.);
    const int64_t o = ((int64_t(t) * na + a) * nch + ch) * nd + d;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dst[o * 4 + k] = make_double2(J.a[k].re, J.a[k].im);
      dstf[o * 4 + k] = make_float2(float(J.a[k].re), float(J.a[k].im));
    }
  }
}

// (block, row, channel) of every visibility of the degrid map, block-major: entry = row | chan
// << 32 in one uint64 plus the block in a second array.
__global__ __launch_bounds__(256) void plan_vlist(const VisP v, const int64_t* vis_start, uint64_t* vrc,
                                                  uint32_t* vblk) {
  for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < v.nblocks;
       b += int64_t(gridDim.x) * blockDim.x) {
    BlockRec br = read_block(v, b);
    const int nch = br.ch1 - br.ch0;
    int64_t k = vis_start[b];
    if (br.nr <= 0 || nch <= 0) continue;
    for (int r = 0; r < br.nr; ++r) {
      const int64_t row = br.rows[r];
      for (int c = br.ch0; c < br.ch1; ++c, ++k) {
        vrc[k] = uint64_t(uint32_t(row)) | (uint64_t(uint32_t(c)) << 32);
        vblk[k] = uint32_t(b);
      }
    }
  }
}

__global__ void block_vis_counts(const VisP v, int32_t* cnt) {
  for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < v.nblocks;
       b += int64_t(gridDim.x) * blockDim.x) {
    BlockRec br = read_block(v, b);
    const int nch = br.ch1 - br.ch0;
    cnt[b] = (br.nr > 0 && nch > 0) ? br.nr * nch : 0;
  }
}

// Per-block descriptor (VisP::bdesc) for the per-block passes: record offset, row count, channel
// range and, when the block's rows are consecutive rows of the chunk (a baseline's time block in
// a baseline-time sorted chunk), the first row, so that a pass addresses the visibilities without
// reading the map's row list.
__global__ void plan_block_desc(const VisP v, uint4* desc) {
  for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < v.nblocks;
       b += int64_t(gridDim.x) * blockDim.x) {
    const BlockRec br = read_block(v, b);
    const int64_t s = br.rows - gptr(v.bda) - 2;
    bool consec = br.nr > 0;
    const int64_t r0 = br.nr > 0 ? int64_t(br.rows[0]) : 0;
    for (int r = 0; r < br.nr && consec; ++r) {
      const int64_t row = br.rows[r];
      consec = row == r0 + r && row >= 0 && row < v.nrow;
    }
    desc[b] = make_uint4(uint32_t(s), uint32_t(br.nr) | (consec ? 0x80000000u : 0u),
                         uint32_t(br.ch0) | (uint32_t(br.ch1) << 16), uint32_t(consec ? r0 : 0));
  }
}

__global__ void reset_counter(uint32_t* c, int n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// Next work item of a tile pass for a workgroup, taken by one thread whose state (x, left) starts
// at (blockIdx.x % n, n) for n item ranges.  Workgroups are dispatched to the 8 XCDs round robin,
// so with 8 ranges the workgroups of one XCD take the items of their own range x (one counter per
// range) in order; a workgroup whose range is used up moves on to the next ranges.  Returns
// 0xFFFFFFFF once it has found every range used up (at most n failed atomics per workgroup).
__device__ __forceinline__ uint32_t next_tile_item(uint32_t* queue, const uint32_t* range, int n, int& x,
                                                   int& left) {
  while (left > 0) {
    const uint32_t it = range[x] + atomicAdd(queue + x, 1u);
    if (it < range[x + 1]) return it;
    x = x + 1 == n ? 0 : x + 1;
    --left;
  }
  return 0xFFFFFFFFu;
}

// a += (er + i ei) w: four FMAs in FP32; in FP64 the product then the sum, as the reference forms it
__device__ __forceinline__ void cmac_run(cx<float>& a, float er, float ei, const cx<float>& w) {
  a.re = fmaf(er, w.re, a.re);
  a.re = fmaf(-ei, w.im, a.re);
  a.im = fmaf(er, w.im, a.im);
  a.im = fmaf(ei, w.re, a.im);
}
__device__ __forceinline__ void cmac_run(cx<double>& a, float er, float ei, const cx<double>& w) {
  a = a + cx<double>{double(er), double(ei)} * w;
}

// ---- cycle: grid pass A --------------------------------------------------------------------------
// One workgroup handles one facet group; lane = slot * G + facet-in-group.
// JMODE: 0 no Jones, 1 per-visibility Jones (applied once per run of one Jones key), 2 per-block
// Jones, 3 per-visibility Jones with ScaleJones (applied per visibility).
// Order in which grid pass A walks the visibilities of a block: the block's channel range is cut
// where the DD or beam Jones channel changes (at most kWalkSeg segments, else one segment) and the
// segments are walked one after the other, row-major within a segment.  The Jones key then stays
// constant over the rows of a segment that share a Jones time (a row-major walk of a block across a
// Jones-channel edge alternates between keys at every row), and the channel phasor still steps
// along the channels of a segment's row.  Visibility g of the walk lies in the segment s with
// nr (cs[s] - ch0) <= g < nr (cs[s + 1] - ch0).
constexpr int kWalkSeg = 4;
struct WalkOrder {
  int nseg, nr, ch0;
  int cs[kWalkSeg + 1];  // segment start channels; cs[nseg] .. cs[kWalkSeg] = ch1
  __device__ __forceinline__ void at(int g, int& rr, int& cc) const {
    int c0 = cs[0], w = cs[1] - cs[0];
#pragma unroll
    for (int k = 1; k < kWalkSeg; ++k)
      if (k < nseg && g >= nr * (cs[k] - ch0)) {
        c0 = cs[k];
        w = cs[k + 1] - cs[k];
      }
    const int loc = g - nr * (c0 - ch0);
    rr = loc / w;
    cc = c0 + (loc - rr * w);
  }
};
__device__ __forceinline__ WalkOrder walk_order(const JonesP& jp, const BlockRec& br, bool by_key) {
  WalkOrder w;
  w.nr = br.nr;
  w.ch0 = br.ch0;
  w.nseg = 1;
#pragma unroll
  for (int k = 0; k <= kWalkSeg; ++k) w.cs[k] = br.ch1;
  w.cs[0] = br.ch0;
  if (!by_key) return w;
  for (int c = br.ch0 + 1; c < br.ch1; ++c) {
    const bool cut = (jp.kms && gptr(jp.kchan)[c] != gptr(jp.kchan)[c - 1]) ||
                     (jp.beam && gptr(jp.bchan)[c] != gptr(jp.bchan)[c - 1]);
    if (!cut) continue;
    if (w.nseg == kWalkSeg) {
      // more edges than segments: the block is walked row-major
      w.nseg = 1;
#pragma unroll
      for (int k = 1; k <= kWalkSeg; ++k) w.cs[k] = br.ch1;
      break;
    }
#pragma unroll
    for (int k = 1; k < kWalkSeg; ++k)
      if (k == w.nseg) w.cs[k] = c;
    ++w.nseg;
  }
  return w;
}

template <int READ, int JMODE, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void grid_pass_a(BatchParams p, const VisP v, const uint8_t* pcode, float2* V,
                                                   uint32_t* queue) {
  // LDS: per-facet sums of every facet [nf][nband*npol + 2*nband], flushed once at the end; the
  // per-channel Jones sums of the current facet group [G][2][nchan] (when they fit), flushed when
  // the workgroup moves to another group (work items are group-major, so a workgroup changes
  // group at most a few times); and one staging buffer per slot.
  //
  // Slots are aligned to wavefronts (64 / G slots of G lanes per wavefront): the lanes of a slot
  // are the facets of the group for one block; they stage the block's visibilities (values,
  // weight, flag) in LDS, G at a time, one per lane, and then every lane walks the staged
  // visibilities in the order of WalkOrder (the reference order, row-major, unless a Jones-channel
  // edge cuts the block).  A wavefront's LDS operations complete in order, so the staging needs no
  // workgroup barrier.
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const int G = p.group;
  const int spw = 64 / G;
  const int nbp = p.nband * p.npol;
  const int per_f = nbp + 2 * p.nband;
  double* s_acc = lds;
  double* s_chan = lds + p.nf * per_f;
  // one [2][nchan] row per facet of the group, padded to an odd number of doubles so that the
  // lanes (facets) of a wavefront fall on different LDS banks
  const int cs = 2 * v.nchan + 1;
  const int nchan_acc = p.chan_lds ? G * cs : 0;
  // staged entry (floats): weight, flag word, row, channel, READ complex values, the Jones key
  // (JMODE 1: DD time, DD chan, beam time, beam chan, antenna 0, antenna 1) and uvw (3 doubles);
  // (weight, flag word, row, channel) first, so that the walk reads an entry with a few 16-byte LDS
  // reads.  Flag word: bit 0 the visibility is skipped (flagged, or a row outside the data), bit 1
  // (JMODE 1) its Jones key differs from that of the visibility before it in the walk or that one
  // is skipped, i.e. it opens a run of one key.
  constexpr int kW = 0, kF = 1, kRow = 2, kCh = 3, kV = 4;
  constexpr int kKey = grid_a_key_offset(READ), kU = kKey + 6, kStride = grid_a_stride(READ);
  // mixed precision (except with ScaleJones, which needs the uvw in metres): the staged uvw is
  // scaled by kPhaseScale nu / 2 pi, so that the phase of a (visibility, facet) pair in
  // revolutions is uvw' . (l0, m0, n0), three FMAs with no per-row state
  constexpr bool kScaled = PREC == 0 && JMODE != 3;
  // 16-byte aligned staging buffers
  float* s_stage = reinterpret_cast<float*>(lds + ((p.nf * per_f + nchan_acc + 1) & ~1));
  for (int i = threadIdx.x; i < p.nf * per_f + nchan_acc; i += blockDim.x) lds[i] = 0.;
  __syncthreads();
  const int wl = threadIdx.x & 63;
  const int sw = wl / G;
  const int slot = (threadIdx.x >> 6) * spw + sw;
  const int slots = (blockDim.x >> 6) * spw;
  const int lf = wl - sw * G;
  const bool active = sw < spw;
  const int K = G;  // visibilities staged per slot and round
  float* stg = s_stage + size_t(active ? slot : 0) * K * kStride;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  __shared__ uint32_t s_item;
  __shared__ uint32_t s_next;
  int cur_group = -1;

  auto flush_chan = [&](int group) {
    if (!p.chan_lds || group < 0) return;
    for (int i = threadIdx.x; i < nchan_acc; i += blockDim.x) {
      const double x = s_chan[i];
      if (x == 0.) continue;
      const int fg = i / cs;
      const int k = i - fg * cs;
      const int f = group * G + fg;
      if (f < p.nf && k < 2 * v.nchan) atomic_add_g(gptr_w(gload(p.facets + f).sum_jones_chan) + k, x);
      s_chan[i] = 0.;
    }
  };

  while (true) {
    if (threadIdx.x == 0) {
      s_item = atomicAdd(queue, 1u);
      s_next = 0;
    }
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= p.nitems_a) break;
    const int group = int(int64_t(item) / p.nchunks);
    const int64_t chunk = int64_t(item) - int64_t(group) * p.nchunks;
    if (group != cur_group) {
      flush_chan(cur_group);
      __syncthreads();
      cur_group = group;
    }
    const int f = group * G + lf;
    const int fc = f < p.nf ? f : p.nf - 1;
    const int64_t b0 = chunk * p.blocks_per_item;
    const int64_t b1 = min(b0 + p.blocks_per_item, v.nblocks);
    // the slots take the item's blocks one at a time from an LDS counter (blocks differ in size)
    auto next_block = [&]() -> int64_t {
      uint32_t kb_ = 0;
      if (lf == 0) kb_ = atomicAdd(&s_next, 1u);
      return b0 + int64_t(__shfl(int(kb_), sw * G, 64));
    };
    for (int64_t b = active ? next_block() : b1; active && b < b1; b = next_block()) {
      // every lane of the slot stages; lanes whose record is dropped (or past the last facet)
      // only stage
      const uint32_t code = pcode[b * p.nf + fc];
      const bool doit = f < p.nf && (code & 0x80u) != 0;
      const FacetDev fd = gload(p.facets + fc);
      const BlockDescRec bdr = read_block_desc(v, b);
      const BlockRec& br = bdr.br;
      const int64_t mid = br.nr > 0 ? bdr.row(br.nr / 2) : -1;
      const bool mid_ok = mid >= 0 && mid < v.nrow;
      JonesP jp = p.jp;
      jp.kdir = fd.kms_dir;
      jp.bdir = fd.beam_dir;
      const WalkOrder wo = walk_order(jp, br, JMODE == 1);
      DecorrP dp = p.dp;
      dp.l0 = fd.dec_l0;
      dp.m0 = fd.dec_m0;
      const double dec = (doit && mid_ok) ? decorr_factor(dp, v.uvw, p.tscale0, mid) : 1.;
      double swt = 0, ssw = 0, sj = 0;
      // block sums: with per-visibility Jones (JMODE 1) a sum over the block's runs (about one per
      // block), in the arithmetic type of the run sums; otherwise over visibilities, in FP64
      using AT = typename std::conditional<JMODE == 1 && PREC != 1, float, double>::type;
      cx<AT> acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      const int band = int(code & 0x7Fu);
      double BBblk = 0;
      JonesPair jblk;
      if (JMODE == 2 && doit) {
        // block Jones at the block means (gridder.h:291-301), means in the reference order
        double su = 0, sv = 0, sw2 = 0;
        int cnt = 0;
        for (int r = 0; r < br.nr; ++r) {
          const int64_t row = bdr.row(r);
          if (row < 0 || row >= v.nrow) continue;
          const DDF_GLOBAL double* q = gptr(v.uvw) + 3 * row;
          for (int c = br.ch0; c < br.ch1; ++c) {
            if (gptr(v.flags)[(size_t(row) * v.nchan + c) * v.ncorr]) continue;
            su += shifted(q[0], q[2], fd.cu);
            sv += shifted(q[1], q[2], fd.cv);
            sw2 += q[2];
            ++cnt;
          }
        }
        double uvwm[3] = {su / cnt, sv / cnt, sw2 / cnt};
        jblk = jones_at<true>(jp, mid_ok ? mid : 0, (br.ch0 + br.ch1) / 2, uvwm, p.wl_mean);
        BBblk = jblk.BB;
      }
      double* sch = s_chan + lf * cs;
      double* sch1 = sch + v.nchan;  // second half: the weight sums
      using RT = typename std::conditional<PREC == 1, double, float>::type;
      // Per-visibility Jones without ScaleJones (JMODE 1) depends only on the key (baseline,
      // Jones time, Jones channel), so J0^H (sum_v meas_v w_v) J1 over a run of one key equals
      // the sum of the per-visibility products (gridder.h:255-260): the run sum is accumulated
      // in the precision's arithmetic type and the Jones pair applied once per run.
      cx<RT> racc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      double rfw = 0;  // sum of the run's weights
      double sjr = 0;  // JMODE 1: sum of BB w over the block's runs
      const double dec2 = dec * dec, ndec2 = double(npg) * dec2;
      M2T<RT> cJ0, cJ1;
      double cBB = 0;
      bool any = false;
      auto flush = [&]() {
        M2T<RT> m;
        if (p.dopsf) {
          m = mul(mul(herm(cJ0), mul(cJ0, herm(cJ1))), cJ1);
#pragma unroll
          for (int q = 0; q < 4; ++q) m.a[q] = m.a[q] * RT(dec * rfw);
        } else {
          m = mul(mul(herm(cJ0), M2T<RT>{{racc[0], racc[1], racc[2], racc[3]}}), cJ1);
        }
        if (READ == 4) {
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = acc[q] + cvt<AT>(m.a[q]);
        } else if (READ == 2) {
          acc[0] = acc[0] + cvt<AT>(m.a[0]);
          acc[3] = acc[3] + cvt<AT>(m.a[3]);
        } else {
          acc[0] = acc[0] + cvt<AT>(m.a[0]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) racc[q] = cx<RT>{0, 0};
        // weight sums of the run: sum w, and sum BB w (dec^2 applied at the block end)
        swt += rfw;
        sjr += cBB * rfw;
        rfw = 0;
      };
      JonesCache<true> jc;
      jc.reset();
      // row state of the walk
      int64_t row = -1;
      double uvwr[3] = {0, 0, 0};
      // phasors: with the scaled uvw through phasor_rev, otherwise ChanPhasor along each row
      ChanPhasor<PREC> ph;
      const int nch = br.ch1 - br.ch0;
      const int nvis = br.nr * nch;
      // JMODE 1 run starts: the Jones key and validity (a row of the data, not flagged) of this
      // lane's staged entry, and of the slot's last entry of the round before
      int mk[6] = {0, 0, 0, 0, 0, 0}, lk[6] = {0, 0, 0, 0, 0, 0};
      int mval = 0, lval = 0;
      for (int i0 = 0; i0 < nvis; i0 += K) {
        const int n = min(K, nvis - i0);
        // stage visibilities i0 .. i0 + n - 1 of the walk (values, weight, flag word, row, channel,
        // uvw and the Jones key), so that the walk below reads LDS only
        auto key_of = [&](int64_t r, int c, int* k) {
          k[0] = jp.kms ? gptr(jp.ktime)[r] : -1;
          k[1] = jp.kms ? gptr(jp.kchan)[c] : -1;
          k[2] = jp.beam ? gptr(jp.btime)[r] : -1;
          k[3] = jp.beam ? gptr(jp.bchan)[c] : -1;
          k[4] = gptr(jp.a0)[r];
          k[5] = gptr(jp.a1)[r];
        };
        float* me = nullptr;  // this lane's staged entry (K = G: at most one per lane and round)
        int mfl = 0;
        mval = 0;
        for (int i = lf; i < n; i += G) {
          const int g = i0 + i;
          int rr, cc;
          wo.at(g, rr, cc);
          const int64_t srow = bdr.row(rr);
          float* e = stg + i * kStride;
          if (srow < 0 || srow >= v.nrow) {
            e[kF] = __int_as_float(1);
            if (kScaled) {
              // walked without a skip test: a zero entry on a valid channel
              e[kW] = 0.f;
#pragma unroll
              for (int k = 0; k < 2 * READ; ++k) e[kV + k] = 0.f;
              e[kCh] = __int_as_float(cc);
              double* eu = reinterpret_cast<double*>(e + kU);
              eu[0] = eu[1] = eu[2] = 0.;
            }
            continue;
          }
          const size_t vc = size_t(srow) * v.nchan + cc;
          const DDF_GLOBAL float2* vp = gptr(v.vis) + vc * v.ncorr;
          const DDF_GLOBAL double* q = gptr(v.uvw) + 3 * srow;
          const float w = gptr(v.weights)[vc];
          // the mixed precision stages the weighted values (w V, rounded once to FP32), so that
          // the walk forms phasor x (w V) per facet without its own weight product
#pragma unroll
          for (int k = 0; k < READ; ++k) {
            e[kV + 2 * k] = kScaled ? w * vp[k].x : vp[k].x;
            e[kV + 2 * k + 1] = kScaled ? w * vp[k].y : vp[k].y;
          }
          e[kW] = w;
          int fl = gptr(v.flags)[vc * v.ncorr] ? 1 : 0;
          e[kRow] = __int_as_float(int(srow));
          e[kCh] = __int_as_float(cc);
          double* eu = reinterpret_cast<double*>(e + kU);
          if (kScaled) {
            const double sc = gptr(v.freqs)[cc] * (kPhaseScale * kInvTwoPi);
            eu[0] = q[0] * sc;
            eu[1] = q[1] * sc;
            eu[2] = q[2] * sc;
          } else {
            eu[0] = q[0];
            eu[1] = q[1];
            eu[2] = q[2];
          }
          if (JMODE == 1) {
            int* ek = reinterpret_cast<int*>(e + kKey);
            key_of(srow, cc, mk);
#pragma unroll
            for (int j = 0; j < 6; ++j) ek[j] = mk[j];
            mval = (fl & 1) ? 0 : 1;
            me = e;
            mfl = fl;
          }
          e[kF] = __int_as_float(fl);
          if (kScaled && (fl & 1)) {
            // a flagged visibility is walked as a zero entry (no skip test, no run of its own)
            e[kW] = 0.f;
#pragma unroll
            for (int k = 0; k < 2 * READ; ++k) e[kV + k] = 0.f;
          }
        }
        if (JMODE == 1) {
          // An entry opens a run of one key unless its predecessor in the walk (the slot's previous
          // lane, or for the first lane the last entry of the round before, kept by the slot's last
          // lane) is a valid visibility of the same key; with the scaled uvw a flagged entry opens
          // none.  The slot's lanes are all active here.
          const int src = lf > 0 ? wl - 1 : sw * G + G - 1;
          int pk[6];
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            const int a = __shfl(mk[j], src, 64), b = __shfl(lk[j], src, 64);
            pk[j] = lf > 0 ? a : b;
          }
          const int pa = __shfl(mval, src, 64), pb = __shfl(lval, src, 64);
          const int pval = lf > 0 ? pa : pb;
          if (me) {
            bool opens = true;
            if (i0 + lf > 0 && pval) {
              opens = false;
#pragma unroll
              for (int j = 0; j < 6; ++j) opens = opens || pk[j] != mk[j];
            }
            if (opens && !(kScaled && (mfl & 1))) me[kF] = __int_as_float(mfl | 2);
          }
#pragma unroll
          for (int j = 0; j < 6; ++j) lk[j] = mk[j];
          lval = mval;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (doit) {
          for (int i = 0; i < n; ++i) {
            const float* e = stg + i * kStride;
            // the whole entry in one batch of LDS reads (one wait)
            const float4 hd = *reinterpret_cast<const float4*>(e);
            float ev[8];
            if (READ == 4) {
              const float4 a = *reinterpret_cast<const float4*>(e + kV);
              const float4 b4 = *reinterpret_cast<const float4*>(e + kV + 4);
              ev[0] = a.x; ev[1] = a.y; ev[2] = a.z; ev[3] = a.w;
              ev[4] = b4.x; ev[5] = b4.y; ev[6] = b4.z; ev[7] = b4.w;
            } else if (READ == 2) {
              const float4 a = *reinterpret_cast<const float4*>(e + kV);
              ev[0] = a.x; ev[1] = a.y; ev[2] = a.z; ev[3] = a.w;
            } else {
              const float2 a = *reinterpret_cast<const float2*>(e + kV);
              ev[0] = a.x; ev[1] = a.y;
            }
            double su[3] = {0, 0, 0};
            if (kScaled) {
              const double* eu = reinterpret_cast<const double*>(e + kU);
              su[0] = eu[0];
              su[1] = eu[1];
              su[2] = eu[2];
            }
            const int fl = __float_as_int(hd.y);
            // with the scaled uvw, skipped visibilities are staged as zero entries and walked like
            // the others (their weight, values and run bit are zero)
            if (!kScaled && (fl & 1)) continue;
            const int cv_ = __float_as_int(hd.w);
            if (!kScaled) {
              const int erow = __float_as_int(hd.z);
              if (erow != row) {
                row = erow;
                const double* eu = reinterpret_cast<const double*>(e + kU);
                uvwr[0] = eu[0];
                uvwr[1] = eu[1];
                uvwr[2] = eu[2];
                ph.row(kPhaseScale * (uvwr[0] * fd.l0 + uvwr[1] * fd.m0 + uvwr[2] * fd.n0));
              }
            }
            const double fw = double(hd.x);
            const double fwd = fw * dec2;
            if (JMODE == 1) {
              if (fl & 2) {
                // a new run: the previous run's sum goes through its Jones pair first
                const int4 ka = *reinterpret_cast<const int4*>(e + kKey);
                const int2 kb = *reinterpret_cast<const int2*>(e + kKey + 4);
                if (any) flush();
                const JonesPair jj = jones_keyed(jp, kb.x, kb.y, ka.x, ka.y, ka.z, ka.w);
                cJ0 = cvt<RT>(jj.J0);
                cJ1 = cvt<RT>(jj.J1);
                cBB = jj.BB;
                any = true;
              }
              rfw += fw;
              if (!p.dopsf) {
                cx<RT> wt;
                if (kScaled) {
                  // the staged values carry the weight
                  const float2 pz = phasor_rev(su[0] * fd.l0 + su[1] * fd.m0 + su[2] * fd.n0);
                  wt = cx<RT>{RT(pz.x), RT(pz.y)};
                } else {
                  wt = ph.template weighted<RT>(v, cv_, hd.x);
                }
                if (READ == 4) {
#pragma unroll
                  for (int q = 0; q < 4; ++q) cmac_run(racc[q], ev[2 * q], ev[2 * q + 1], wt);
                } else if (READ == 2) {
                  cmac_run(racc[0], ev[0], ev[1], wt);
                  cmac_run(racc[3], ev[2], ev[3], wt);
                } else {
                  const cx<RT> x = cx<RT>{RT(ev[0]), RT(ev[1])} * wt;
                  racc[0] = racc[0] + x;
                  racc[3] = racc[3] + x;
                }
              }
              const double x1 = ndec2 * fw, x0 = cBB * x1;
              if (p.chan_lds) {
                atomicAdd(&sch[cv_], x0);
                atomicAdd(&sch1[cv_], x1);
              } else {
                atomic_add_g(gptr_w(fd.sum_jones_chan) + cv_, x0);
                atomic_add_g(gptr_w(fd.sum_jones_chan) + (v.nchan + cv_), x1);
              }
            } else {
              cd meas[4];
              if (p.dopsf) {
                meas[0] = cd{dec, 0}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = cd{dec, 0};
              } else if (READ == 4) {
#pragma unroll
                for (int q = 0; q < 4; ++q) meas[q] = cd{ev[2 * q], ev[2 * q + 1]};
              } else if (READ == 2) {
                meas[0] = cd{ev[0], ev[1]}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = cd{ev[2], ev[3]};
              } else {
                meas[0] = cd{ev[0], ev[1]}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = meas[0];
              }
              cd wt = cd{fw, 0};
              if (!p.dopsf) {
                if (kScaled) {
                  // the staged values carry the weight
                  const float2 pz = phasor_rev(su[0] * fd.l0 + su[1] * fd.m0 + su[2] * fd.n0);
                  wt = cd{double(pz.x), double(pz.y)};
                } else {
                  wt = ph.at(v, cv_) * fw;
                }
              }
              if (JMODE == 3) {
                const JonesPair& jj = jc.get(jp, row, cv_, uvwr, p.wl_mean);
                M2 m;
                if (p.dopsf) {
                  m = mul(mul(herm(jj.J0), mul(jj.J0, herm(jj.J1))), jj.J1);
#pragma unroll
                  for (int q = 0; q < 4; ++q) m.a[q] = m.a[q] * dec;
                } else {
                  m = mul(mul(herm(jj.J0), M2{{meas[0], meas[1], meas[2], meas[3]}}), jj.J1);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) meas[q] = m.a[q];
                sj += jj.BB * fwd;
                const double x = double(npg) * (jj.BB * fwd);
                if (p.chan_lds) atomicAdd(&sch[cv_], x);
                else atomic_add_g(gptr_w(fd.sum_jones_chan) + cv_, x);
              }
              if (JMODE == 2) {
                const double x = double(npg) * (fwd * BBblk);
                if (p.chan_lds) atomicAdd(&sch[cv_], x);
                else atomic_add_g(gptr_w(fd.sum_jones_chan) + cv_, x);
              }
              if (JMODE != 0) {
                const double x = double(npg) * fwd;
                if (p.chan_lds) atomicAdd(&sch1[cv_], x);
                else atomic_add_g(gptr_w(fd.sum_jones_chan) + (v.nchan + cv_), x);
              }
              if (READ == 4) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = acc[q] + cvt<AT>(meas[q] * wt);
              } else if (READ == 2) {
                acc[0] = acc[0] + cvt<AT>(meas[0] * wt);
                acc[3] = acc[3] + cvt<AT>(meas[3] * wt);
              } else {
                acc[0] = acc[0] + cvt<AT>(meas[0] * wt);
              }
            }
            if (JMODE != 1) {
              swt += fw;
              ssw += fwd;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (doit) {
        if (JMODE == 1) {
          if (any) flush();
          ssw = dec2 * swt;
          sj = dec2 * sjr;
        }
        if (JMODE == 2) {
          M2 m{{cvt<double>(acc[0]), cvt<double>(acc[1]), cvt<double>(acc[2]), cvt<double>(acc[3])}};
          if (p.dopsf) m = mul(mul(jblk.J0, m), herm(jblk.J1));
          m = mul(mul(herm(jblk.J0), m), jblk.J1);
  #pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = cvt<AT>(m.a[q]);
          sj = ssw * BBblk;
        }
        // facet-major [f][block]: the tile pass reads one facet's records in block order
        float2* out = V + (size_t(f) * v.nblocks + b) * p.nstokes;
        for (int k = 0; k < p.nstokes; ++k) {
          cd s = cd{0, 0};
  #pragma unroll
          for (int q = 0; q < 4; ++q) s = s + cd{p.sm[k][q][0], p.sm[k][q][1]} * cvt<double>(acc[q]);
          out[k] = make_float2(float(s.re), float(s.im));
        }
        double* a = s_acc + f * per_f;
        for (int ip = 0; ip < npg; ++ip) {
          atomicAdd(&a[band * p.npol + ip], swt);
          if (JMODE != 0) {
            atomicAdd(&a[nbp + band], sj);
            atomicAdd(&a[nbp + p.nband + band], ssw);
          }
        }
      }
    }
    __syncthreads();
  }
  flush_chan(cur_group);
  __syncthreads();
  for (int i = threadIdx.x; i < p.nf * per_f; i += blockDim.x) {
    const double x = s_acc[i];
    if (x == 0.) continue;
    const int f = i / per_f;
    const int k = i - f * per_f;
    if (k < nbp) atomic_add_g(gptr_w(gload(p.facets + f).sumwt) + k, x);
    else atomic_add_g(gptr_w(gload(p.facets + f).sum_jones) + (k - nbp), x);
  }
}

// ---- wave-level helpers for the tile passes ---------------------------------------------------------
// Values of lane j (j uniform) in scalar registers.
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int j) { return uint32_t(__builtin_amdgcn_readlane(int(x), j)); }
__device__ __forceinline__ uint64_t lane_u64(uint64_t x, int j) {
  return uint64_t(lane_u32(uint32_t(x), j)) | (uint64_t(lane_u32(uint32_t(x >> 32), j)) << 32);
}
__device__ __forceinline__ float lane_f32(float x, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j)); }

// Stamp record of a binned (block, facet) record (plan_stamp_records): base | S << 16 | sub << 21 |
// cfoff << 32, with base the tile offset of the stamp's first cell (row (locy - sup) - y0, column
// (locx - sup) - x0 of the (T + 2 h)^2 tile image), sub the oversampled sub-kernel index
// io * OverS + jo (used with the padded CF copy of degrid pass A) and cfoff the element offset of the record's
// oversampled S x S sub-kernel in its facet's CF buffer.  Cell c of the stamp (row c / S, column
// c % S) sits at tile offset base + (c / S) W + c % S and takes CF value cf[cfoff + c].
__device__ __forceinline__ int stamp_support(uint64_t srec) { return int((srec >> 16) & 0x1F); }
__device__ __forceinline__ int stamp_base(uint64_t srec) { return int(srec & 0xFFFF); }

// Per-lane cell layout of a stamp of support S for the first 256 cells (cell c = 64 r + lane):
// off[r] its tile offset from the stamp's first cell (kOffOut past the stamp).  It depends on S
// only, and the records of a work item are sorted by support, so a pass recomputes it only when S
// changes.  (c + 0.5) / S is at least 0.5 / S from an integer, far above the error of
// (c + 0.5) * rcp(S) in FP32 for c < 1024, S < 32 (below 1e-3), so the row is exact.
constexpr uint32_t kOffOut = 1u << 30;
template <int G>
struct StampLayout {
  int S;
  uint32_t off[G];
};
template <int G>
__device__ __forceinline__ void stamp_layout(int S, int lane, int W, StampLayout<G>& L) {
  L.S = S;
  const int S2 = S * S;
  const float invS = __builtin_amdgcn_rcpf(float(S));
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const int c = 64 * r + lane;
    const int sy = int((float(c) + 0.5f) * invS);
    L.off[r] = c < S2 ? uint32_t(sy * W + (c - sy * S)) : kOffOut;
  }
}

// x += (vr + i vi) (wx + i wy), in the tile's precision: FP32 FMAs, or in FP64 the exact products
// of the FP32 operands summed in FP64.
__device__ __forceinline__ void cmac(float2& x, float vr, float vi, float wx, float wy) {
  x.x = fmaf(vr, wx, x.x);
  x.x = fmaf(-vi, wy, x.x);
  x.y = fmaf(vr, wy, x.y);
  x.y = fmaf(vi, wx, x.y);
}
__device__ __forceinline__ void cmac(double2& x, float vr, float vi, float wx, float wy) {
  x.x = fma(double(vr), double(wx), x.x);
  x.x = fma(-double(vi), double(wy), x.x);
  x.y = fma(double(vr), double(wy), x.y);
  x.y = fma(double(vi), double(wx), x.y);
}

// ---- cycle: grid pass B (LDS-tiled scatter) ------------------------------------------------------
// One single-wavefront workgroup per bin item, with a private LDS tile image (T + 2 h)^2 per
// polarisation, (re, im) interleaved.  The records of an item are sorted by (support, stamp
// position, record), so records with the same stamp come in runs (~30 on the 81-facet workload).
// The wavefront loads 64 records (stamp record, value) at a time, then takes them one after the
// other with the record in scalar registers; lanes are stamp cells, G groups of 64 cells per
// record, with G the largest ceil(S^2 / 64) of the 64 records.  A record's stamp (value times its
// CF sub-kernel) is added to the run's sums in registers; when the stamp position changes, the
// sums go into the LDS tile with one read-modify-write per cell.  Every record issues exactly G CF
// loads, so the number of loads in flight is known at compile time and the wait before a record's
// adds covers its own loads only: the CF values of the next D - 1 records (7, 3 or 2 for G = 1, 2,
// 3-4) are in flight while the current one is added.  Addresses: the record's sub-kernel and tile position are scalars, the
// lane's cell offsets come from the support's StampLayout (kept across records of one support).
// The read-modify-writes need no atomics: the cells of one stamp are distinct lanes' and a
// wavefront's LDS operations complete in issue order.  The tile then goes to its slot (or, without
// slots, into the facet grid with global float atomics on the non-zero cells).
template <bool CONJ, bool ONE, bool BIG, int G, bool FULL, typename T2>
__device__ __forceinline__ void grid_b_batch(T2* tile, int W, int WW, int npg, int lane, const DDF_GLOBAL char* cf,
                                             uint32_t my_o8, uint32_t my_lim8, uint32_t my_key,
                                             const float2 (&my_v)[4], int cnt) {
  constexpr int NP = ONE ? 1 : 4;
  // records in flight: D - 1 ahead of the one being added (D divides 64)
  constexpr int D = G == 1 ? 8 : 4;
  StampLayout<G> Lrun;
  Lrun.S = -1;
  uint32_t c8[G];  // byte offset of the lane's cell in a sub-kernel, per group
#pragma unroll
  for (int r = 0; r < G; ++r) c8[r] = uint32_t(64 * r + lane) * uint32_t(sizeof(float2));
  uint32_t oq[D];
  float2 vq[D][G];
  // the open run: its stamp (support << 16 | base, -1 for none) and its register sums
  int run = -1;
  T2 acc[G][NP];
  // Record k's CF values: cell c of its sub-kernel (byte offset o8 in the facet's CF buffer), a
  // lane past the stamp reading the last cell (byte offset lim8) again.  Records past a partial
  // batch load the last one again (their values are not used).
  auto load = [&](int k, uint32_t& o, float2 (&v)[G]) {
    const int kk = FULL ? k : min(k, cnt - 1);
    o = lane_u32(my_o8, kk);
    const uint32_t lim8 = lane_u32(my_lim8, kk);
#pragma unroll
    for (int r = 0; r < G; ++r) v[r] = ld2((const DDF_GLOBAL float2*)(cf + (min(c8[r], lim8) + o)));
  };
  // the open run's sums into the tile: one read-modify-write per cell
  auto flush = [&]() {
    if (run < 0) return;
    const int base = run & 0xFFFF;
    if (ONE) {
      // a lane past the stamp adds to its own spare cell after the tile (WW + lane, never read
      // back) instead of being masked off; the G reads are issued before the adds
      uint32_t a[G];
      T2 x[G];
#pragma unroll
      for (int r = 0; r < G; ++r) {
        a[r] = min(uint32_t(base) + Lrun.off[r], uint32_t(WW + lane));
        x[r] = tile[a[r]];
      }
#pragma unroll
      for (int r = 0; r < G; ++r) {
        x[r].x += acc[r][0].x;
        x[r].y += acc[r][0].y;
        tile[a[r]] = x[r];
      }
    } else {
      const int S2 = Lrun.S * Lrun.S;
      T2* tb = tile + base;
#pragma unroll
      for (int r = 0; r < G; ++r) {
        if (64 * r >= S2) break;
        if (Lrun.off[r] != kOffOut) {
          T2* t = tb + Lrun.off[r];
#pragma unroll
          for (int ip = 0; ip < NP; ++ip) {
            if (ip < npg) {
              T2 x = t[ip * WW];
              x.x += acc[r][ip].x;
              x.y += acc[r][ip].y;
              t[ip * WW] = x;
            }
          }
        }
      }
    }
  };
  auto add = [&](int j, uint32_t o8, const float2 (&v)[G]) {
    if (!FULL && j >= cnt) return;
    float vr[NP], vi[NP];
#pragma unroll
    for (int ip = 0; ip < NP; ++ip) {
      if (ip < npg) {
        vr[ip] = lane_f32(my_v[ip].x, j);
        vi[ip] = lane_f32(my_v[ip].y, j);
      }
    }
    const int key = int(lane_u32(my_key, j));
    const int S = key >> 16, base = key & 0xFFFF;
    if (key != run) {
      flush();
      run = key;
      if (S != Lrun.S) stamp_layout<G>(S, lane, W, Lrun);
#pragma unroll
      for (int r = 0; r < G; ++r)
#pragma unroll
        for (int ip = 0; ip < NP; ++ip) acc[r][ip] = T2{0, 0};
    }
    // this record's stamp into the run's sums (lanes past the stamp add values that their spare
    // or masked cells discard)
#pragma unroll
    for (int r = 0; r < G; ++r) {
      const float wx = v[r].x, wy = CONJ ? -v[r].y : v[r].y;
#pragma unroll
      for (int ip = 0; ip < NP; ++ip)
        if (ip < npg) cmac(acc[r][ip], vr[ip], vi[ip], wx, wy);
    }
    // stamps wider than 16 x 16: the cells past the first 256 straight into the tile, without
    // prefetch (distinct cells from the run's sums, which cover the first 256)
    const int S2 = S * S;
    if (BIG && S2 > 256) {
      T2* tb = tile + base;
      const DDF_GLOBAL float2* cfp = (const DDF_GLOBAL float2*)(cf + o8);
      const float invS = __builtin_amdgcn_rcpf(float(S));
      for (int c = 256 + lane; c < S2; c += 64) {
        const int sy = int((float(c) + 0.5f) * invS);
        const float2 w = ld2(cfp + c);
        const float wx = w.x, wy = CONJ ? -w.y : w.y;
        T2* t = tb + sy * W + (c - sy * S);
#pragma unroll
        for (int ip = 0; ip < NP; ++ip) {
          if (ip < npg) {
            T2 x = t[ip * WW];
            cmac(x, vr[ip], vi[ip], wx, wy);
            t[ip * WW] = x;
          }
        }
      }
    }
  };
  // D records in flight, in statically rotating buffers, so no register copy has to wait for a
  // load.  A full batch (64 records) runs without index checks: the last D records are peeled off
  // and issue only the one load still inside the batch.
#pragma unroll
  for (int s = 0; s < D - 1; ++s) load(s, oq[s], vq[s]);
  if (FULL) {
#pragma unroll 1
    for (int j = 0; j < 64 - D; j += D) {
#pragma unroll
      for (int s = 0; s < D; ++s) {
        load(j + s + D - 1, oq[(s + D - 1) % D], vq[(s + D - 1) % D]);
        add(j + s, oq[s], vq[s]);
      }
    }
    load(63, oq[D - 1], vq[D - 1]);
#pragma unroll
    for (int s = 0; s < D; ++s) add(64 - D + s, oq[s], vq[s]);
  } else {
    for (int j = 0; j < cnt; j += D) {
#pragma unroll
      for (int s = 0; s < D; ++s) {
        load(j + s + D - 1, oq[(s + D - 1) % D], vq[(s + D - 1) % D]);
        add(j + s, oq[s], vq[s]);
      }
    }
  }
  flush();
}

// ONE: a single polarisation (npg == 1), with the per-polarisation loops resolved at compile time.
// BIG: stamps may have more than 256 cells (halo > 7).
// WIDE: the tile sums in FP64 (the adjoint of degrid pass A in the FP64 precision, whose FP32 tile
// sums would round at ~sqrt(records per cell) ulps of the result); tiles are flushed with atomics
// then.
template <bool CONJ, bool ONE, bool BIG, bool WIDE>
__global__ __launch_bounds__(64) void grid_pass_b(BatchParams p, const uint64_t* bstamp, const float2* V,
                                                  const BinItem* items, const uint32_t* recs, uint32_t* queue,
                                                  float2* slots) {
  using T2 = typename std::conditional<WIDE, double2, float2>::type;
  // 16-byte aligned: the read-modify-writes must not straddle their own size
  extern __shared__ __attribute__((aligned(16))) float lds_f[];
  T2* tile = reinterpret_cast<T2*>(lds_f);  // [npg][W * W], then 64 spare cells when npg == 1
  __shared__ uint32_t s_item;
  const int W = p.tile + 2 * p.halo;
  const int WW = W * W;
  const int npg = ONE ? 1 : (p.nstokes < p.npol ? p.nstokes : p.npol);
  const int lane = threadIdx.x;
  int qx = int(blockIdx.x % uint32_t(p.nranges)), qleft = p.nranges;
  while (true) {
    if (lane == 0) s_item = next_tile_item(queue, p.item_range, p.nranges, qx, qleft);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= p.nitems_b) break;
    const BinItem bi = items[it];
    const FacetDev fd = gload(p.facets + bi.facet);
    // the CF base in scalar registers (the loads address it with a 32-bit vector offset)
    const uint64_t cfa = reinterpret_cast<uint64_t>(fd.cf);
    const DDF_GLOBAL char* cf = (const DDF_GLOBAL char*)((uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(cfa >> 32)))
                                                          << 32) |
                                                         uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(cfa))));
    for (int i = lane; i < npg * WW; i += 64) tile[i] = T2{0, 0};
    const int x0 = bi.tx * p.tile - p.halo, y0 = bi.ty * p.tile - p.halo;
    for (uint32_t k0 = bi.begin; k0 < bi.end; k0 += 64u) {
      const uint32_t k = k0 + lane;
      const bool ok = k < bi.end;
      const uint64_t my_rec = ok ? bstamp[k] : 0;
      // record t = block * nf + facet; its value at [facet][block]
      const uint32_t my_t = ok ? recs[k] : 0;
      const size_t my_v0 = (size_t(bi.facet) * p.nblk + my_t / uint32_t(p.nf)) * p.nstokes;
      float2 my_v[4];
#pragma unroll
      for (int ip = 0; ip < 4; ++ip) my_v[ip] = (ok && ip < npg) ? V[my_v0 + ip] : make_float2(0.f, 0.f);
      const int cnt = int(min(64u, bi.end - k0));
      // each lane decodes its record once: the byte offsets of its sub-kernel and of the
      // sub-kernel's last cell, and its run key (support << 16 | base)
      const int myS = stamp_support(my_rec);
      const uint32_t my_o8 = uint32_t(my_rec >> 32) << 3;
      const uint32_t my_lim8 = ok ? uint32_t(myS * myS - 1) << 3 : 0u;
      const uint32_t my_key = uint32_t(my_rec & 0x1FFFFF);
      // G = the largest ceil(S^2 / 64) of the batch, capped at 4
      const int myG = (myS * myS + 63) >> 6;
      const int G = __ballot(myG > 3) ? 4 : __ballot(myG > 2) ? 3 : __ballot(myG > 1) ? 2 : 1;
      // (the full-batch form for one polarisation only: it doubles the code of a batch)
#define DDF_GRID_B_BATCH(GG)                                                                                   \
  (ONE && cnt == 64 ? grid_b_batch<CONJ, ONE, BIG, GG, true>(tile, W, WW, npg, lane, cf, my_o8, my_lim8, my_key, my_v, cnt) \
             : grid_b_batch<CONJ, ONE, BIG, GG, false>(tile, W, WW, npg, lane, cf, my_o8, my_lim8, my_key, my_v, cnt))
      switch (G) {
        case 1: DDF_GRID_B_BATCH(1); break;
        case 2: DDF_GRID_B_BATCH(2); break;
        case 3: DDF_GRID_B_BATCH(3); break;
        default: DDF_GRID_B_BATCH(4); break;
      }
#undef DDF_GRID_B_BATCH
    }
    if (!WIDE && slots) {
      // the whole tile into this item's slot with plain stores (grid_slot_reduce adds the slots)
      float4* dst = reinterpret_cast<float4*>(slots + size_t(bi.slot) * npg * WW);
      const float4* src = reinterpret_cast<const float4*>(tile);
      const int n4 = (npg * WW) / 2;
      for (int i = lane; i < n4; i += 64) dst[i] = src[i];
      if ((npg * WW) & 1 && lane == 0) {
        const T2 x = tile[npg * WW - 1];
        slots[size_t(bi.slot) * npg * WW + npg * WW - 1] = make_float2(float(x.x), float(x.y));
      }
    } else {
      DDF_GLOBAL float* grid = gptr_w(reinterpret_cast<float*>(fd.grid));
      for (int i = lane; i < npg * WW; i += 64) {
        const int ip = i / WW;
        const int rem = i - ip * WW;
        const T2 x = tile[i];
        if (x.x == 0 && x.y == 0) continue;
        const int gy = y0 + rem / W, gx = x0 + rem % W;
        if (gx < 0 || gy < 0 || gx >= p.nx || gy >= p.ny) continue;
        const size_t gi = ((size_t(bi.band) * p.npol + ip) * p.ny + gy) * p.nx + gx;
        atomic_add_g(grid + 2 * gi, float(x.x));
        atomic_add_g(grid + 2 * gi + 1, float(x.y));
      }
    }
    __syncthreads();
  }
}

// grid[f][band][pol][y][x] += the tile slots covering (y, x), one thread per grid cell: the
// tiles ty with ty * T - h <= y < ty * T + T + h (at most two per axis when h < T), each with
// its consecutive work items, in item order.  No atomics; the same sum order on every run.
__global__ __launch_bounds__(256) void grid_slot_reduce(BatchParams p, const float2* slots, const uint32_t* tile_first,
                                                        const uint32_t* tile_count) {
  const int T = p.tile, h = p.halo;
  const int W = T + 2 * h;
  const int WW = W * W;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  const int64_t total = int64_t(p.nf) * p.nband * npg * p.ny * p.nx;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i;
    const int x = int(r % p.nx); r /= p.nx;
    const int y = int(r % p.ny); r /= p.ny;
    const int ip = int(r % npg); r /= npg;
    const int band = int(r % p.nband); r /= p.nband;
    const int f = int(r);
    float re = 0.f, im = 0.f;
    const int ty0 = max(0, (y - h) / T), ty1 = min(p.ntiles_y - 1, (y + h) / T);
    const int tx0 = max(0, (x - h) / T), tx1 = min(p.ntiles_x - 1, (x + h) / T);
    for (int ty = ty0; ty <= ty1; ++ty) {
      const int ly = y - (ty * T - h);
      for (int tx = tx0; tx <= tx1; ++tx) {
        const int lx = x - (tx * T - h);
        const int64_t tid = ((int64_t(f) * p.nband + band) * p.ntiles_y + ty) * p.ntiles_x + tx;
        const uint32_t first = tile_first[tid], n = tile_count[tid];
        const float2* t = slots + size_t(first) * npg * WW + size_t(ip) * WW + size_t(ly) * W + lx;
        for (uint32_t k = 0; k < n; ++k, t += size_t(npg) * WW) {
          const float2 v = *t;
          re += v.x;
          im += v.y;
        }
      }
    }
    if (re != 0.f || im != 0.f) {
      DDF_GLOBAL float2* g = gptr_w(reinterpret_cast<float2*>(gload(p.facets + f).grid)) +
                             ((size_t(band) * p.npol + ip) * p.ny + y) * p.nx + x;
      const float2 o = ld2(g);
      st2(g, make_float2(o.x + re, o.y + im));
    }
  }
}

// ---- cycle: degrid pass A (LDS-staged gather) --------------------------------------------------
// The model tile of a bin item is staged in LDS (complex64, 16-byte aligned); lanes are records:
// each lane gathers its own record's S x S stamp, row by row, from the padded CF copy (two cells
// per 16-byte load: the pass is bound by the texture-address path of its gathers, so the number of
// load instructions sets its time), with FP32 products (grid x CF)
// summed in ACC (degridder.h:202-212 sums the FP32 products in double; ACC = double in the FP64
// and mixed precisions, float in the FP32 precision).  No cross-lane work: the records of a bin
// share their CF plane, so the lanes of a wavefront mostly share S and run in step.
// ONE: a single Stokes plane (nstokes == 1), resolved at compile time.
// PM: the precision (0 mixed: FP32 products summed in FP64; 1 FP64: FP64 products and sums, as
// the reference degridder; 2 FP32 throughout).  ONE: a single Stokes plane.
template <int PM, bool ONE>
__global__ __launch_bounds__(256) void degrid_pass_a(BatchParams p, const uint64_t* bstamp, float2* M,
                                                     const BinItem* items, const uint32_t* recs, uint32_t* queue) {
  using ACC = typename std::conditional<PM == 2, float, double>::type;
  extern __shared__ __attribute__((aligned(16))) float tile_f[];
  float2* tile = reinterpret_cast<float2*>(tile_f);  // [nstokes][W * W]
  __shared__ uint32_t s_item;
  const int W = p.tile + 2 * p.halo;
  const int WW = W * W;
  const int ns = ONE ? 1 : p.nstokes;
  int qx = int(blockIdx.x % uint32_t(p.nranges)), qleft = p.nranges;
  while (true) {
    if (threadIdx.x == 0) s_item = next_tile_item(queue, p.item_range, p.nranges, qx, qleft);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= p.nitems_b) break;
    const BinItem bi = items[it];
    const FacetDev fd = gload(p.facets + bi.facet);
    const DDF_GLOBAL float2* cfp = gptr(fd.cft);
    const int x0 = bi.tx * p.tile - p.halo, y0 = bi.ty * p.tile - p.halo;
    const DDF_GLOBAL float2* model = gptr(reinterpret_cast<const float2*>(fd.model));
    for (int i = threadIdx.x; i < ns * WW; i += blockDim.x) {
      const int ip = i / WW;
      const int rem = i - ip * WW;
      const int gy = y0 + rem / W, gx = x0 + rem % W;
      float2 m = make_float2(0.f, 0.f);
      if (gx >= 0 && gy >= 0 && gx < p.nx && gy < p.ny)
        m = ld2(model + ((size_t(bi.band) * p.npol + ip) * p.ny + gy) * p.nx + gx);
      tile[i] = m;
    }
    __syncthreads();
    const int O2 = p.overs * p.overs;
    for (uint32_t k = bi.begin + threadIdx.x; k < bi.end; k += blockDim.x) {
      const uint64_t rec = bstamp[k];
      const uint32_t t = recs[k];
      const int S = int((rec >> 16) & 0x1F);
      const int Sp = S + (S & 1);
      // row r of sub-kernel `sub` at 2 x (plane offset) + sub S Sp + r Sp of the padded copy; the
      // record holds the put-order offset plane offset + sub S^2
      const int sub = int((rec >> 21) & 0x7FF);
      const int64_t poff = int64_t(uint32_t(rec >> 32)) - int64_t(sub) * S * S;
      const DDF_GLOBAL v4f* cp = reinterpret_cast<const DDF_GLOBAL v4f*>(cfp + (2 * poff + int64_t(sub) * S * Sp));
      const int toff0 = int(rec & 0xFFFF);  // tile offset of the stamp's first cell
      ACC sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
      for (int r = 0; r < S; ++r) {
        const DDF_GLOBAL v4f* crow = cp + (r * Sp >> 1);
        const float2* trow = tile + toff0 + r * W;
        // four cells at a time (two 16-byte loads), their CF loads issued together
        for (int c = 0; c < S; c += 4) {
          float2 w[4];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            v4f x = {0.f, 0.f, 0.f, 0.f};
            if (c + 2 * h < S) x = crow[(c >> 1) + h];
            w[2 * h] = make_float2(x.x, x.y);
            w[2 * h + 1] = make_float2(x.z, x.w);
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if (c + j >= S) break;
#pragma unroll
            for (int ip = 0; ip < 4; ++ip) {
              if (ip < ns) {
                const float2 m = trow[ip * WW + c + j];
                if (PM == 0) {
                  sr[ip] += ACC(m.x * w[j].x - m.y * w[j].y);
                  si[ip] += ACC(m.x * w[j].y + m.y * w[j].x);
                } else if (PM == 1) {
                  sr[ip] += double(m.x) * double(w[j].x) - double(m.y) * double(w[j].y);
                  si[ip] += double(m.x) * double(w[j].y) + double(m.y) * double(w[j].x);
                } else {
                  sr[ip] = fmaf(m.x, w[j].x, sr[ip]);
                  sr[ip] = fmaf(-m.y, w[j].y, sr[ip]);
                  si[ip] = fmaf(m.x, w[j].y, si[ip]);
                  si[ip] = fmaf(m.y, w[j].x, si[ip]);
                }
              }
            }
          }
        }
      }
      // value of record t = block * nf + facet, stored at [facet][block]
      float2* out = M + (size_t(bi.facet) * p.nblk + t / uint32_t(p.nf)) * ns;
#pragma unroll
      for (int ip = 0; ip < 4; ++ip)
        if (ip < ns) out[ip] = make_float2(float(sr[ip]), float(si[ip]));
    }
    __syncthreads();
  }
}

// ---- cycle: degrid pass B (lanes over visibilities) ----------------------------------------------
// A visibility's prediction is sum_f corr_f(vis) Y_f with Y_f = dec_f J0_f S(M_f) J1_f^H, where
// S(M_f) is the Stokes-to-correlation map of the facet's pass-A value for the block.  Y_f depends
// on the visibility only through its Jones key, so it is shared by a run of consecutive
// visibilities of the list with one block and one key (plan_run_ids).  A workgroup takes 256
// consecutive visibilities; facets go in batches of FB: the workgroup first forms Y_f of every
// (run of the tile, facet of the batch) into LDS, one pair per lane at a time, and then every
// lane adds corr_f Y_f of its visibility for the batch's facets.  The sum over facets is FP64; Y_f
// and the phasor product are in the precision's arithmetic type.
struct RunInfo {
  int row, chan, mid;  // representative row and channel; middle row of the block (-1: none)
  uint32_t blk;
  double tscale;       // decorr_tscale of the mean frequency of the block's channels
  JonesKey key;        // the Jones key of the run (complex64 Jones tables only)
};

template <int NCORR, int JMODE, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void degrid_pass_b(BatchParams p, const VisP v, const float2* M,
                                                     const uint64_t* vrc, const uint32_t* vblk,
                                                     const uint32_t* vrun, const uint32_t* run_first, int64_t nvl) {
  using RT = typename std::conditional<PREC == 1, double, float>::type;
  // Y of (run, facet) for a batch of facets, [facet][run][correlation]: as many facets per batch
  // as fit for the tile's run count (a tile of 256 visibilities holds ~20 runs with BDA, so 2
  // batches cover 81 facets; 4 facets when every visibility is a run of its own)
  constexpr int kYCap = 4 * 256 * 4;
  constexpr int kFMax = 64;
  __shared__ cx<RT> sY[kYCap];
  __shared__ RunInfo sRun[256];
  // l0, m0, n0 of the batch's facets: the visibility loop reads them from LDS, not from the facet
  // table in global memory (whose loads the compiler cannot hoist or move to the scalar unit); the
  // facets' DD Jones and beam directions, so that forming Y reads no facet record
  __shared__ double sGeo[kFMax][3];
  __shared__ int2 sDir[kFMax];
  // mixed precision: each lane's pre-scaled uvw, read back after every batch's barrier so that it
  // is not held in registers (or spilled) across the forming of Y
  __shared__ double sU[3][256];
  const int t = threadIdx.x;
  // with the complex64 Jones tables the run's Jones key is read with the run table, once per run,
  // so that forming Y of a (run, facet) waits on the model value and the Jones entries only
  const bool keyed = JMODE != 0 && jones_f32_path<RT>(p.jp);
  for (int64_t k0 = int64_t(blockIdx.x) * 256; k0 < nvl; k0 += int64_t(gridDim.x) * 256) {
    const int64_t k = k0 + t;
    const int64_t klast = min(k0 + 255, nvl - 1);
    const uint32_t r0 = vrun[k0];
    const int nr = int(vrun[klast] - r0) + 1;  // runs of the tile (at most 256)
    // run table of the tile
    if (t < nr) {
      RunInfo ri;
      ri.row = -1;
      ri.chan = 0;
      ri.mid = -1;
      ri.blk = 0;
      ri.tscale = 0;
      ri.key = JonesKey{0, 0, 0, 0, 0, 0};
      const uint32_t kr = run_first[r0 + t];
      if (int64_t(kr) < nvl) {
        const uint64_t e = vrc[kr];
        const int64_t rrow = int64_t(uint32_t(e));
        const int rc_ = int(e >> 32);
        const uint32_t rb = vblk[kr];
        if (rrow >= 0 && rrow < v.nrow && rc_ >= 0 && rc_ < v.nchan && int64_t(rb) < v.nblocks) {
          ri.row = int(rrow);
          ri.chan = rc_;
          ri.blk = rb;
          if (JMODE == 2 || p.dp.enabled) {
            const BlockRec br = read_block(v, rb);
            double fm = 0;
            for (int cc = br.ch0; cc < br.ch1; ++cc) fm += gptr(v.freqs)[cc];
            if (br.ch1 > br.ch0) fm /= double(br.ch1 - br.ch0);
            ri.tscale = decorr_tscale(fm);
            if (br.nr > 0) {
              const int64_t mid = br.rows[br.nr / 2];
              if (mid >= 0 && mid < v.nrow) ri.mid = int(mid);
            }
            if (JMODE == 2) ri.chan = (br.ch0 + br.ch1) / 2;
          }
          const int jrow = JMODE == 2 ? ri.mid : ri.row;
          if (keyed && jrow >= 0) ri.key = jones_key_f(p.jp, jrow, ri.chan);
        }
      }
      sRun[t] = ri;
    }
    // this lane's visibility
    bool live = k < nvl;
    int64_t row = 0;
    int c = 0, myrun = 0;
    if (live) {
      const uint64_t e = vrc[k];
      row = int64_t(uint32_t(e));
      c = int(e >> 32);
      myrun = int(vrun[k] - r0);
      live = row >= 0 && row < v.nrow && c >= 0 && c < v.nchan && myrun >= 0 && myrun < nr;
    }
    double u0 = 0, u1 = 0, u2 = 0, fr = 0;
    if (live) {
      const DDF_GLOBAL double* q = gptr(v.uvw) + 3 * row;
      u0 = q[0];
      u1 = q[1];
      u2 = q[2];
      fr = gptr(v.freqs)[c];
    }
    if (PREC == 0) {
      // mixed precision: uvw pre-scaled by -kPhaseScale freq / 2 pi, so that the phase of facet f
      // in revolutions is u0 l0 + u1 m0 + u2 n0 (three FP64 operations per facet; the rounding
      // differs from forming the angle first by a few ulp of the phase), the phasor phasor_rev of
      // it, as in the gridding pass A (the adjoint)
      const double sc = -kPhaseScale * kInvTwoPi * fr;
      sU[0][t] = u0 * sc;
      sU[1][t] = u1 * sc;
      sU[2][t] = u2 * sc;
    }
    // the sum over facets, in FP64
    cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    __syncthreads();
    const int FBd = min(min(p.nf, kFMax), max(1, kYCap / (4 * nr)));
    for (int f0 = 0; f0 < p.nf; f0 += FBd) {
      const int nfb = min(FBd, p.nf - f0);
      if (t < nfb) {
        const FacetDev fd = gload(p.facets + f0 + t);
        sGeo[t][0] = fd.l0;
        sGeo[t][1] = fd.m0;
        sGeo[t][2] = fd.n0;
        sDir[t] = make_int2(fd.kms_dir, fd.beam_dir);
      }
      __syncthreads();
      // Y of every (run, facet of the batch); facets vary fastest over the threads, so the Jones
      // loads of one run's facets (directions) are contiguous in the prepared tables
      for (int j = t; j < nr * nfb; j += 256) {
        const int ir = j / nfb;
        const int fi = j - ir * nfb;
        const int f = f0 + fi;
        const RunInfo ri = sRun[ir];
        M2T<RT> cv;
#pragma unroll
        for (int q = 0; q < 4; ++q) cv.a[q] = cx<RT>{0, 0};
        if (ri.row >= 0) {
          const DDF_GLOBAL float2* m = gptr(M) + (size_t(f) * v.nblocks + ri.blk) * p.nstokes;
          cd st[4];
          int nz = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (q < p.nstokes) {
              const float2 x = ld2(m + q);
              st[q] = cd{x.x, x.y};
              nz |= (x.x != 0.f || x.y != 0.f);
            } else {
              st[q] = cd{0, 0};
            }
          }
          if (nz) {
            // Stokes -> correlations and the Jones products in the precision's arithmetic type
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
              for (int s = 0; s < 4; ++s)
                if (s < p.nstokes) cv.a[q] = cv.a[q] + cx<RT>{RT(p.sm[q][s][0]), RT(p.sm[q][s][1])} * cvt<RT>(st[s]);
            }
            JonesP jp = p.jp;
            const int2 dir = sDir[fi];
            jp.kdir = dir.x;
            jp.bdir = dir.y;
            const int jrow = JMODE == 2 ? ri.mid : ri.row;
            if (JMODE != 0 && jrow >= 0) {
              M2T<RT> J0, J1;
              if (keyed) {
                M2T<float> F0, F1;
                jones_pair_f(jp, ri.key, F0, F1);
                J0 = cvt<RT>(F0);
                J1 = cvt<RT>(F1);
              } else {
                jones_pair_t<RT>(jp, jrow, ri.chan, v.uvw + 3 * int64_t(jrow), p.wl_mean, J0, J1);
              }
              cv = mul(mul(J0, cv), herm(J1));
            }
            if (p.dp.enabled && ri.mid >= 0) {
              const FacetDev fd = gload(p.facets + f);
              DecorrP dp = p.dp;
              dp.l0 = fd.dec_l0;
              dp.m0 = fd.dec_m0;
              const RT dec = RT(decorr_factor(dp, v.uvw, ri.tscale, ri.mid));
#pragma unroll
              for (int q = 0; q < 4; ++q) cv.a[q] = cv.a[q] * dec;
            }
          }
        }
        cx<RT>* yo = sY + (size_t(fi) * nr + ir) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) yo[q] = cv.a[q];
      }
      __syncthreads();
      if (live) {
        if (PREC == 0) {
          u0 = sU[0][t];
          u1 = sU[1][t];
          u2 = sU[2][t];
        }
        // facets whose Y is zero for this run (an empty model there) are skipped; in the mixed
        // precision groups of 8 facets are summed by FP32 fused multiply-adds first, then added
        // to the FP64 sum
        for (int g0 = 0; g0 < nfb; g0 += 8) {
          const int g1 = min(g0 + 8, nfb);
          cf bacc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
          for (int fi = g0; fi < g1; ++fi) {
            const cx<RT>* yp = sY + (size_t(fi) * nr + myrun) * 4;
            cx<RT> y[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) y[q] = yp[q];
            bool nz = false;
#pragma unroll
            for (int q = 0; q < 4; ++q) nz = nz || y[q].re != RT(0) || y[q].im != RT(0);
            if (!nz) continue;
            const double dot = u0 * sGeo[fi][0] + u1 * sGeo[fi][1] + u2 * sGeo[fi][2];
            cx<RT> corr;
            if (PREC == 0) {
              const float2 pz = phasor_rev(dot);
              corr = cx<RT>{RT(pz.x), RT(pz.y)};
            } else {
              corr = cvt<RT>(phasor_prec<PREC>(fr, -kPhaseScale * dot));
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (PREC == 0) {
                bacc[q].re = fmaf(float(y[q].re), float(corr.re), bacc[q].re);
                bacc[q].re = fmaf(-float(y[q].im), float(corr.im), bacc[q].re);
                bacc[q].im = fmaf(float(y[q].re), float(corr.im), bacc[q].im);
                bacc[q].im = fmaf(float(y[q].im), float(corr.re), bacc[q].im);
              } else {
                acc[q] = acc[q] + cvt<double>(y[q] * corr);
              }
            }
          }
          if (PREC == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = acc[q] + cvt<double>(bacc[q]);
          }
        }
      }
      __syncthreads();
    }
    if (live) {
      DDF_GLOBAL float2* vp = gptr_w(v.vis) + (size_t(row) * v.nchan + c) * NCORR;
      if (NCORR == 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float2 x = ld2(vp + q);
          st2(vp + q, make_float2(float(double(x.x) - acc[q].re), float(double(x.y) - acc[q].im)));
        }
      } else if (NCORR == 2) {
        const float2 x0 = ld2(vp), x1 = ld2(vp + 1);
        st2(vp, make_float2(float(double(x0.x) - acc[0].re), float(double(x0.y) - acc[0].im)));
        st2(vp + 1, make_float2(float(double(x1.x) - acc[3].re), float(double(x1.y) - acc[3].im)));
      } else {
        const float2 x0 = ld2(vp);
        st2(vp, make_float2(float(double(x0.x) - acc[0].re), float(double(x0.y) - acc[0].im)));
      }
    }
  }
}

// Run flags of the degrid visibility list: 1 where visibility k starts a new run (another block,
// or for per-visibility Jones another Jones key: antennas, DD / beam time and channel, and the
// row itself when ScaleJones makes the Jones depend on uvw).
__global__ __launch_bounds__(256) void plan_run_flags(const BatchParams p, const VisP v, const uint64_t* vrc,
                                                      const uint32_t* vblk, int64_t nvl, int32_t* flag) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nvl; k += int64_t(gridDim.x) * blockDim.x) {
    int fl = 1;
    if (k > 0 && vblk[k] == vblk[k - 1]) {
      fl = 0;
      if (p.jmode == 1) {
        const uint64_t e1 = vrc[k], e0 = vrc[k - 1];
        const int64_t r1 = int64_t(uint32_t(e1)), r0 = int64_t(uint32_t(e0));
        const int c1 = int(e1 >> 32), c0 = int(e0 >> 32);
        const bool ok = r1 >= 0 && r1 < v.nrow && r0 >= 0 && r0 < v.nrow && c1 >= 0 && c1 < v.nchan && c0 >= 0 &&
                        c0 < v.nchan;
        if (!ok) {
          fl = 1;
        } else {
          const JonesP& jp = p.jp;
          bool diff = jp.a0[r1] != jp.a0[r0] || jp.a1[r1] != jp.a1[r0] || (jp.scale && r1 != r0);
          if (jp.kms) diff = diff || jp.ktime[r1] != jp.ktime[r0] || jp.kchan[c1] != jp.kchan[c0];
          if (jp.beam) diff = diff || jp.btime[r1] != jp.btime[r0] || jp.bchan[c1] != jp.bchan[c0];
          fl = diff ? 1 : 0;
        }
      }
    }
    flag[k] = fl;
  }
}

// vrun[k] = run of visibility k (inclusive count of run starts - 1); run_first[run] = its first k.
__global__ __launch_bounds__(256) void plan_run_ids(const int32_t* flag, const int64_t* excl, int64_t nvl,
                                                    uint32_t* vrun, uint32_t* run_first) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nvl; k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = excl[k] + flag[k] - 1;
    vrun[k] = uint32_t(r);
    if (flag[k]) run_first[r] = uint32_t(k);
  }
}


// ---- adjoint of degrid pass B ------------------------------------------------------------------------
// Forward (degrid_pass_b): vis_k = mask(corr_k * J0 (D M) J1^H) summed over facets.  Adjoint for
// one (block, facet): Mbar = D^H (sum_k conj(corr_k) J0^H vbar_k J1), mode-2 Jones applied to the sum.
// Lanes are facets; each lane walks its block's visibilities (all of them: degridding ignores flags).
template <int NCORR, int JMODE, int PREC>
__global__ __launch_bounds__(256) void degrid_adj_b(BatchParams p, const VisP v, const uint8_t* pcode,
                                                    const float2* vbar, float2* Mbar, uint32_t* queue) {
  const int G = p.group;
  const int slots = p.slots;
  const int slot = threadIdx.x / G;
  const int lf = threadIdx.x - slot * G;
  const bool active = slot < slots;
  __shared__ uint32_t s_item;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= p.nitems_a) break;
    const int group = int(int64_t(item) / p.nchunks);
    const int64_t chunk = int64_t(item) - int64_t(group) * p.nchunks;
    const int f = group * G + lf;
    const int64_t b0 = chunk * p.blocks_per_item;
    const int64_t b1 = min(b0 + p.blocks_per_item, v.nblocks);
    for (int64_t b = b0 + slot; active && f < p.nf && b < b1; b += slots) {
      if (!(pcode[b * p.nf + f] & 0x80u)) continue;
      const FacetDev fd = gload(p.facets + f);
      BlockRec br = read_block(v, b);
      double fm = 0;
      for (int c = br.ch0; c < br.ch1; ++c) fm += gptr(v.freqs)[c];
      fm /= double(br.ch1 - br.ch0);
      const int64_t mid = br.rows[br.nr / 2];
      const bool mid_ok = mid >= 0 && mid < v.nrow;
      DecorrP dp = p.dp;
      dp.l0 = fd.dec_l0;
      dp.m0 = fd.dec_m0;
      const double dec = mid_ok ? decorr_factor(dp, v.uvw, decorr_tscale(fm), mid) : 1.;
      JonesP jp = p.jp;
      jp.kdir = fd.kms_dir;
      jp.bdir = fd.beam_dir;
      cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      for (int r = 0; r < br.nr; ++r) {
        const int64_t row = br.rows[r];
        if (row < 0 || row >= v.nrow) continue;
        const double* uvw = v.uvw + 3 * row;
        const DDF_GLOBAL double* guvw = gptr(v.uvw) + 3 * row;
        const double angle = -kPhaseScale * (guvw[0] * fd.l0 + guvw[1] * fd.m0 + guvw[2] * fd.n0);
        for (int c = br.ch0; c < br.ch1; ++c) {
          const float2* vp = vbar + (size_t(row) * v.nchan + c) * NCORR;
          M2 m;
          if (NCORR == 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) m.a[q] = cd{vp[q].x, vp[q].y};
          } else if (NCORR == 2) {
            m.a[0] = cd{vp[0].x, vp[0].y}; m.a[1] = cd{0, 0}; m.a[2] = cd{0, 0}; m.a[3] = cd{vp[1].x, vp[1].y};
          } else {
            m.a[0] = cd{vp[0].x, vp[0].y}; m.a[1] = cd{0, 0}; m.a[2] = cd{0, 0}; m.a[3] = cd{0, 0};
          }
          const cd cc = conj(phasor_prec<PREC>(gptr(v.freqs)[c], angle)) * dec;
          if (JMODE == 1) {
            JonesPair jj = jones_at<true>(jp, row, c, uvw, p.wl_mean);
            m = mul(mul(herm(jj.J0), m), jj.J1);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = acc[q] + m.a[q] * cc;
        }
      }
      if (JMODE == 2 && mid_ok) {
        JonesPair jj = jones_at<true>(jp, mid, (br.ch0 + br.ch1) / 2, v.uvw + 3 * mid, p.wl_mean);
        M2 m{{acc[0], acc[1], acc[2], acc[3]}};
        m = mul(mul(herm(jj.J0), m), jj.J1);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = m.a[q];
      }
      float2* out = Mbar + (size_t(f) * v.nblocks + b) * p.nstokes;
      for (int s = 0; s < p.nstokes; ++s) {
        cd t = cd{0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) t = t + conj(cd{p.sm[q][s][0], p.sm[q][s][1]}) * acc[q];
        out[s] = make_float2(float(t.re), float(t.im));
      }
    }
    __syncthreads();
  }
}

}  // namespace

// ---- host side -----------------------------------------------------------------------------------
namespace ddf {

int batched_block_desc(const VisP& v, uint4* desc, hipStream_t st) {
  int64_t nb = (v.nblocks + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_block_desc, dim3(unsigned(nb)), dim3(256), 0, st, v, desc);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_place(const BatchParams& p, const VisP& v, bool degrid, uint64_t* place, hipStream_t st) {
  int64_t total = v.nblocks * p.nf;
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  if (degrid) hipLaunchKernelGGL(plan_place<true>, dim3(unsigned(nb)), dim3(256), 0, st, p, v, place);
  else hipLaunchKernelGGL(plan_place<false>, dim3(unsigned(nb)), dim3(256), 0, st, p, v, place);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_group_counts(const uint32_t* counts, int64_t ntile, int nsub, int ngrp, uint32_t* gcount, hipStream_t st) {
  int64_t nb = (ntile * ngrp + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_group_counts, dim3(unsigned(nb)), dim3(256), 0, st, counts, ntile, nsub, ngrp, gcount);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_bin_starts(const uint32_t* counts, const int64_t* tile_start, int64_t ntile, int nsub, int64_t* bin_start,
                       hipStream_t st) {
  int64_t nb = (ntile + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_bin_starts, dim3(unsigned(nb)), dim3(256), 0, st, counts, tile_start, ntile, nsub, bin_start);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_pair_codes(const uint64_t* place, int64_t n, uint8_t* code, hipStream_t st) {
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_pair_codes, dim3(unsigned(nb)), dim3(256), 0, st, place, n, code);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_count(const BatchParams& p, int64_t nblocks, const uint64_t* place, uint32_t* counts, hipStream_t st) {
  // four wavefronts of 64 blocks of one facet per workgroup
  const int64_t nwaves = ((nblocks + 63) / 64) * p.nf;
  int64_t nb = (nwaves + 3) / 4;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_count, dim3(unsigned(nb)), dim3(256), 0, st, p, nblocks, place, counts);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_fill(const BatchParams& p, int64_t nblocks, const uint64_t* place, const int64_t* bin_start,
                 uint32_t* fill, uint32_t* recs, hipStream_t st) {
  const int64_t nwaves = ((nblocks + 63) / 64) * p.nf;
  int64_t nb = (nwaves + 3) / 4;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_fill, dim3(unsigned(nb)), dim3(256), 0, st, p, nblocks, place, bin_start, fill, recs);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

// Records of every work item sorted by (CF plane, sub-kernel, record): one segmented radix sort
// of 51-bit keys over the items' ranges (a plane needs 8 bits, a sub-kernel 11, a record 32).
int batched_sort_items(const BatchParams& p, const uint64_t* place, uint32_t* recs, int64_t n, const BinItem* items,
                       uint32_t nitems, bool stamp_order, hipStream_t st) {
  if (n <= 0 || nitems == 0) return DDF_OK;
  if (n >= (int64_t(1) << 31)) return DDF_EINVAL;
  uint64_t *k0 = nullptr, *k1 = nullptr;
  uint32_t *b0 = nullptr, *b1 = nullptr;
  void* tmp = nullptr;
  size_t tb = 0;
  int rc = DDF_OK;
  auto grid = [](int64_t m) { int64_t g = (m + 255) / 256; return unsigned(g > 65536 ? 65536 : (g < 1 ? 1 : g)); };
  if (hipMallocAsync(reinterpret_cast<void**>(&k0), size_t(n) * 8, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&k1), size_t(n) * 8, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&b0), size_t(nitems) * 4, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&b1), size_t(nitems) * 4, st) != hipSuccess) {
    rc = DDF_ENOMEM;
  } else {
    if (stamp_order) hipLaunchKernelGGL(plan_sort_keys<true>, dim3(grid(n)), dim3(256), 0, st, p, place, recs, n, k0);
    else hipLaunchKernelGGL(plan_sort_keys<false>, dim3(grid(n)), dim3(256), 0, st, p, place, recs, n, k0);
    hipLaunchKernelGGL(plan_item_bounds, dim3(grid(nitems)), dim3(256), 0, st, items, nitems, b0, b1);
    if (hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, k0, k1, int(n), int(nitems), b0, b1, 0, 51, st) !=
            hipSuccess ||
        hipMallocAsync(&tmp, tb ? tb : 1, st) != hipSuccess ||
        hipcub::DeviceSegmentedRadixSort::SortKeys(tmp, tb, k0, k1, int(n), int(nitems), b0, b1, 0, 51, st) !=
            hipSuccess) {
      rc = DDF_EHIP;
    } else {
      hipLaunchKernelGGL(plan_keys_to_recs, dim3(grid(n)), dim3(256), 0, st, k1, n, recs);
      if (hipGetLastError() != hipSuccess) rc = DDF_EHIP;
    }
  }
  if (tmp) hipFreeAsync(tmp, st);
  if (k0) hipFreeAsync(k0, st);
  if (k1) hipFreeAsync(k1, st);
  if (b0) hipFreeAsync(b0, st);
  if (b1) hipFreeAsync(b1, st);
  return rc;
}

int batched_vis_counts(const VisP& v, int32_t* cnt, hipStream_t st) {
  int64_t nb = (v.nblocks + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(block_vis_counts, dim3(unsigned(nb)), dim3(256), 0, st, v, cnt);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_vlist(const VisP& v, const int64_t* vis_start, uint64_t* vrc, uint32_t* vblk, hipStream_t st) {
  int64_t nb = (v.nblocks + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_vlist, dim3(unsigned(nb)), dim3(256), 0, st, v, vis_start, vrc, vblk);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_reset(uint32_t* counters, int n, hipStream_t st) {
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, st, counters, n);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

template <int READ, int JMODE, int PREC>
static void launch_a(const BatchParams& p, const VisP& v, const uint8_t* pcode, float2* V, uint32_t* q,
                     size_t lds, unsigned nwg, hipStream_t st) {
  hipLaunchKernelGGL((grid_pass_a<READ, JMODE, PREC>), dim3(nwg), dim3(256), lds, st, p, v, pcode, V, q);
}

template <int READ, int JMODE>
static void launch_a_p(const BatchParams& p, const VisP& v, const uint8_t* pcode, float2* V, uint32_t* q,
                       size_t lds, unsigned nwg, hipStream_t st) {
  if (p.prec == 1) launch_a<READ, JMODE, 1>(p, v, pcode, V, q, lds, nwg, st);
  else if (p.prec == 2) launch_a<READ, JMODE, 2>(p, v, pcode, V, q, lds, nwg, st);
  else launch_a<READ, JMODE, 0>(p, v, pcode, V, q, lds, nwg, st);
}

template <int READ>
static void launch_a_j(const BatchParams& p, const VisP& v, const uint8_t* pcode, float2* V, uint32_t* q,
                       size_t lds, unsigned nwg, hipStream_t st) {
  // per-visibility Jones: run-accumulated (1), or per visibility with ScaleJones (3), whose Jones
  // also depends on the uvw of the visibility that refreshed it
  if (p.jmode == 1 && p.jp.scale) launch_a_p<READ, 3>(p, v, pcode, V, q, lds, nwg, st);
  else if (p.jmode == 1) launch_a_p<READ, 1>(p, v, pcode, V, q, lds, nwg, st);
  else if (p.jmode == 2) launch_a_p<READ, 2>(p, v, pcode, V, q, lds, nwg, st);
  else launch_a_p<READ, 0>(p, v, pcode, V, q, lds, nwg, st);
}

size_t batched_grid_a_stage_bytes(int group, int read) {
  // (256 / 64) wavefronts x (64 / group) slots x group entries: at most one entry per lane
  // plus 16 bytes for the alignment of the staging area
  return size_t(256 / 64) * size_t(64 / group) * size_t(group) * size_t(grid_a_stride(read)) * sizeof(float) + 16;
}

int batched_grid_a(const BatchParams& p, const VisP& v, const uint8_t* pcode, float2* V, uint32_t* queue,
                   unsigned nwg, hipStream_t st) {
  size_t lds = (size_t(p.nf) * (p.nband * p.npol + 2 * p.nband) + (p.chan_lds ? size_t(p.group) * (2 * v.nchan + 1) : 0)) * 8 +
               batched_grid_a_stage_bytes(p.group, p.read);
  if (p.read == 4) launch_a_j<4>(p, v, pcode, V, queue, lds, nwg, st);
  else if (p.read == 2) launch_a_j<2>(p, v, pcode, V, queue, lds, nwg, st);
  else launch_a_j<1>(p, v, pcode, V, queue, lds, nwg, st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

// Persistent grid size: `nwg` when given, else as many workgroups as fit on the device at once.
template <typename K>
static unsigned persistent_wgs(K kernel, size_t lds, unsigned nwg, int block = 256) {
  if (nwg) return nwg;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, lds) != hipSuccess || per < 1) per = 1;
  return unsigned(ncu * per);
}

// grid_pass_b instance for a tile pass: CONJ, one polarisation or several, stamps beyond 256 cells.
typedef void (*GridPassB)(BatchParams, const uint64_t*, const float2*, const BinItem*, const uint32_t*, uint32_t*,
                          float2*);
// LDS of a grid_pass_b workgroup: the tile image, plus 64 spare cells with one polarisation.
static size_t grid_pass_b_lds(int npg, int W, bool wide) {
  return (size_t(npg) * W * W + (npg == 1 ? 64 : 0)) * (wide ? sizeof(double2) : sizeof(float2));
}
// The adjoint of degrid pass A sums in FP64 in the FP64 precision, wherever those tiles fit 64 KB
// of LDS.
static bool adjoint_wide(int npg, int W, int prec) { return prec == 1 && grid_pass_b_lds(npg, W, true) <= 64 * 1024; }
template <bool CONJ, bool WIDE>
static GridPassB grid_pass_b_for(int npg, int halo) {
  const bool big = halo > 7;
  if (npg == 1) return big ? grid_pass_b<CONJ, true, true, WIDE> : grid_pass_b<CONJ, true, false, WIDE>;
  return big ? grid_pass_b<CONJ, false, true, WIDE> : grid_pass_b<CONJ, false, false, WIDE>;
}
static GridPassB grid_pass_b_adjoint_for(int npg, int halo, int W, int prec) {
  return adjoint_wide(npg, W, prec) ? grid_pass_b_for<true, true>(npg, halo) : grid_pass_b_for<true, false>(npg, halo);
}

// degrid_pass_a instance for a precision (0 mixed, 1 FP64, 2 FP32) and one or several Stokes planes.
typedef void (*DegridPassA)(BatchParams, const uint64_t*, float2*, const BinItem*, const uint32_t*, uint32_t*);
static DegridPassA degrid_pass_a_for(int prec, bool one) {
  if (prec == 1) return one ? degrid_pass_a<1, true> : degrid_pass_a<1, false>;
  if (prec == 2) return one ? degrid_pass_a<2, true> : degrid_pass_a<2, false>;
  return one ? degrid_pass_a<0, true> : degrid_pass_a<0, false>;
}

// Persistent grid sizes of the tile passes (grid B, degrid A, adjoint of degrid A), computed once
// per plan so that the launches make no device queries (they may run under stream capture).
int batched_tile_wgs(const BatchParams& pg, const BatchParams& pd, unsigned out[3]) {
  const int Wg = pg.tile + 2 * pg.halo, Wd = pd.tile + 2 * pd.halo;
  const int npg = pg.nstokes < pg.npol ? pg.nstokes : pg.npol;
  const int npd = pd.nstokes < pd.npol ? pd.nstokes : pd.npol;
  out[0] = persistent_wgs(grid_pass_b_for<false, false>(npg, pg.halo), grid_pass_b_lds(npg, Wg, false), 0, 64);
  const size_t lds_d = size_t(pd.nstokes) * Wd * Wd * 2 * sizeof(float);
  out[1] = persistent_wgs(degrid_pass_a_for(pd.prec, pd.nstokes == 1), lds_d, 0);
  out[2] = persistent_wgs(grid_pass_b_adjoint_for(npd, pd.halo, Wd, pd.prec),
                          grid_pass_b_lds(npd, Wd, adjoint_wide(npd, Wd, pd.prec)), 0, 64);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_grid_b(const BatchParams& p, const uint64_t* place, const float2* V, const BinItem* items,
                   const uint32_t* recs, uint32_t* queue, unsigned nwg, float* slots, const uint32_t* tile_first,
                   const uint32_t* tile_count, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  const size_t lds = grid_pass_b_lds(npg, W, false);
  hipLaunchKernelGGL((grid_pass_b_for<false, false>(npg, p.halo)), dim3(nwg), dim3(64), lds, st, p, place, V, items, recs,
                     queue, reinterpret_cast<float2*>(slots));
  if (hipGetLastError() != hipSuccess) return DDF_EHIP;
  if (slots) {
    const int64_t total = int64_t(p.nf) * p.nband * npg * p.ny * p.nx;
    int64_t nb = (total + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(grid_slot_reduce, dim3(unsigned(nb)), dim3(256), 0, st, p, reinterpret_cast<const float2*>(slots),
                       tile_first, tile_count);
  }
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

// Adjoint of degrid pass A: conj(CF) stamps of M-bar, scattered through the degrid bins.
int batched_degrid_a_adjoint(const BatchParams& p, const uint64_t* place, const float2* Mbar, const BinItem* items,
                             const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  BatchParams q = p;
  q.npol = p.npol;
  const int npg = q.nstokes < q.npol ? q.nstokes : q.npol;
  const size_t lds = grid_pass_b_lds(npg, W, adjoint_wide(npg, W, q.prec));
  hipLaunchKernelGGL(grid_pass_b_adjoint_for(npg, q.halo, W, q.prec), dim3(nwg), dim3(64), lds, st, q, place, Mbar, items, recs,
                     queue, static_cast<float2*>(nullptr));
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_degrid_a(const BatchParams& p, const uint64_t* place, float2* M, const BinItem* items,
                     const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  const size_t lds = size_t(p.nstokes) * W * W * 2 * sizeof(float);
  const DegridPassA k = degrid_pass_a_for(p.prec, p.nstokes == 1);
  nwg = persistent_wgs(k, lds, nwg);
  hipLaunchKernelGGL(k, dim3(nwg), dim3(256), lds, st, p, place, M, items, recs, queue);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

template <int NCORR, int JMODE>
static void launch_db_p(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc,
                        const uint32_t* vblk, const uint32_t* vrun, const uint32_t* run_first, int64_t nvl, unsigned nb,
                        hipStream_t st) {
  if (p.prec == 1)
    hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 1>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, vrun, run_first, nvl);
  else if (p.prec == 2)
    hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 2>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, vrun, run_first, nvl);
  else
    hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 0>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, vrun, run_first, nvl);
}

template <int NCORR>
static void launch_db_j(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc,
                        const uint32_t* vblk, const uint32_t* vrun, const uint32_t* run_first, int64_t nvl, unsigned nb,
                        hipStream_t st) {
  if (p.jmode == 1) launch_db_p<NCORR, 1>(p, v, M, vrc, vblk, vrun, run_first, nvl, nb, st);
  else if (p.jmode == 2) launch_db_p<NCORR, 2>(p, v, M, vrc, vblk, vrun, run_first, nvl, nb, st);
  else launch_db_p<NCORR, 0>(p, v, M, vrc, vblk, vrun, run_first, nvl, nb, st);
}

int batched_degrid_b(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc, const uint32_t* vblk,
                     const uint32_t* vrun, const uint32_t* run_first, int64_t nvl, hipStream_t st) {
  int64_t nb = (nvl + 255) / 256;
  if (nb > 65536 * 4) nb = 65536 * 4;
  if (nb < 1) nb = 1;
  if (v.ncorr == 4) launch_db_j<4>(p, v, M, vrc, vblk, vrun, run_first, nvl, unsigned(nb), st);
  else if (v.ncorr == 2) launch_db_j<2>(p, v, M, vrc, vblk, vrun, run_first, nvl, unsigned(nb), st);
  else launch_db_j<1>(p, v, M, vrc, vblk, vrun, run_first, nvl, unsigned(nb), st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_run_ids(const BatchParams& p, const VisP& v, const uint64_t* vrc, const uint32_t* vblk, int64_t nvl,
                    int32_t* flag, const int64_t* excl, uint32_t* vrun, uint32_t* run_first, int phase, hipStream_t st) {
  int64_t nb = (nvl + 255) / 256;
  if (nb > 65536 * 4) nb = 65536 * 4;
  if (nb < 1) nb = 1;
  if (phase == 0) hipLaunchKernelGGL(plan_run_flags, dim3(unsigned(nb)), dim3(256), 0, st, p, v, vrc, vblk, nvl, flag);
  else hipLaunchKernelGGL(plan_run_ids, dim3(unsigned(nb)), dim3(256), 0, st, flag, excl, nvl, vrun, run_first);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf

namespace ddf {

template <int NCORR, int JMODE>
static void launch_adj_b_p(const BatchParams& p, const VisP& v, const uint8_t* pcode, const float2* vbar,
                           float2* Mbar, uint32_t* q, unsigned nwg, hipStream_t st) {
  if (p.prec == 1)
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 1>), dim3(nwg), dim3(256), 0, st, p, v, pcode, vbar, Mbar, q);
  else if (p.prec == 2)
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 2>), dim3(nwg), dim3(256), 0, st, p, v, pcode, vbar, Mbar, q);
  else
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 0>), dim3(nwg), dim3(256), 0, st, p, v, pcode, vbar, Mbar, q);
}

template <int NCORR>
static void launch_adj_b_j(const BatchParams& p, const VisP& v, const uint8_t* pcode, const float2* vbar,
                           float2* Mbar, uint32_t* q, unsigned nwg, hipStream_t st) {
  if (p.jmode == 1) launch_adj_b_p<NCORR, 1>(p, v, pcode, vbar, Mbar, q, nwg, st);
  else if (p.jmode == 2) launch_adj_b_p<NCORR, 2>(p, v, pcode, vbar, Mbar, q, nwg, st);
  else launch_adj_b_p<NCORR, 0>(p, v, pcode, vbar, Mbar, q, nwg, st);
}

int batched_degrid_b_adjoint(const BatchParams& p, const VisP& v, const uint8_t* pcode, const float2* vbar,
                             float2* Mbar, uint32_t* queue, unsigned nwg, hipStream_t st) {
  if (v.ncorr == 4) launch_adj_b_j<4>(p, v, pcode, vbar, Mbar, queue, nwg, st);
  else if (v.ncorr == 2) launch_adj_b_j<2>(p, v, pcode, vbar, Mbar, queue, nwg, st);
  else launch_adj_b_j<1>(p, v, pcode, vbar, Mbar, queue, nwg, st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf

namespace ddf {

int batched_cf_pad(const float2* in, float2* out, const int64_t* off, const int32_t* support, int nw, int overs,
                   int max_support, hipStream_t st) {
  const int64_t n = int64_t(overs) * overs * max_support * (max_support + 1);
  int64_t nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_cf_pad, dim3(unsigned(nb), unsigned(2 * nw)), dim3(256), 0, st, in, out, off, support, nw,
                     overs);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_stamp_records(const BatchParams& p, const uint64_t* place, const uint32_t* recs, int64_t n,
                          uint64_t* bstamp, hipStream_t st) {
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_stamp_records, dim3(unsigned(nb)), dim3(256), 0, st, p, place, recs, n, bstamp);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_prepare_jones(const BatchParams& p, const float2* src, double2* dst, float2* dstf, int dd, hipStream_t st) {
  const int* dims = dd ? p.jp.kdims : p.jp.bdims;
  int64_t total = int64_t(dims[0]) * dims[1] * dims[2] * dims[3];
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_prepare_jones, dim3(unsigned(nb)), dim3(256), 0, st, src, dst, dstf, p.jp, dd);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf
