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
//           record's stamp (FP32 products summed in FP64, degridder.h:202-212) -> M[b][f].
//   degrid B lanes are visibilities; each lane sums the model of every facet (Stokes -> corr,
//           Jones, phasor, decorrelation) in registers and writes its visibility once.
// Work queues are device counters reset by a small kernel (no memset nodes, capture-safe).
#include <hip/hip_runtime.h>

#include <cstring>
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
    const FacetDev& fd = p.facets[f];
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
          for (int c = br.ch0; c < br.ch1; ++c) {
            if (v.flags[(size_t(row) * v.nchan + c) * v.ncorr]) continue;
            su += shifted(uvw[0], uvw[2], fd.cu);
            sv += shifted(uvw[1], uvw[2], fd.cv);
            sw += uvw[2];
            fm += v.freqs[c];
            ++cnt;
          }
        }
        if (cnt > 0) {
          su /= cnt; sv /= cnt; sw /= cnt; fm /= cnt;
        }
      } else {
        for (int c = br.ch0; c < br.ch1; ++c) fm += v.freqs[c];
        fm /= double(nch);
        for (int r = 0; r < br.nr; ++r) {
          const int64_t row = br.rows[r];
          if (row < 0 || row >= v.nrow) continue;
          const double* uvw = v.uvw + 3 * row;
          su += shifted(uvw[0], uvw[2], fd.cu);
          sv += shifted(uvw[1], uvw[2], fd.cv);
          sw += uvw[2];
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

// ---- plan: binning (counting sort) --------------------------------------------------------------
// Bin = (facet, band, tile y, tile x, CF plane): the records of one tile are contiguous and, inside
// the tile, grouped by CF plane so consecutive stamps reuse the same plane (cache locality).
__device__ __forceinline__ int64_t bin_of(const BatchParams& p, int64_t t, uint64_t rec) {
  const int f = int(t % p.nf);
  Unpacked u = unpack_place(rec);
  const int tx = u.locx / p.tile, ty = u.locy / p.tile;
  return (((int64_t(f) * p.nband + u.band) * p.ntiles_y + ty) * p.ntiles_x + tx) * (2 * p.nw) + u.plane;
}

__global__ __launch_bounds__(256) void plan_count(BatchParams p, int64_t nblocks, const uint64_t* place,
                                                  uint32_t* counts) {
  const int64_t total = nblocks * p.nf;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    uint64_t rec = place[t];
    if (place_valid(rec)) atomicAdd(&counts[bin_of(p, t, rec)], 1u);
  }
}

__global__ __launch_bounds__(256) void plan_fill(BatchParams p, int64_t nblocks, const uint64_t* place,
                                                 const int64_t* bin_start, uint32_t* fill, uint32_t* recs) {
  const int64_t total = nblocks * p.nf;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    uint64_t rec = place[t];
    if (!place_valid(rec)) continue;
    int64_t bin = bin_of(p, t, rec);
    uint32_t slot = atomicAdd(&fill[bin], 1u);
    recs[bin_start[bin] + slot] = uint32_t(t);
  }
}

__global__ __launch_bounds__(256) void plan_gather_place(const uint64_t* place, const uint32_t* recs, int64_t n,
                                                         uint64_t* bplace) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
    bplace[k] = place[recs[k]];
}

// Jones table [t][dir][ant][chan][4] -> [t][ant][chan][dir][4]
__global__ __launch_bounds__(256) void plan_transpose_jones(const float2* src, float2* dst, int nt, int nd, int na,
                                                            int nch) {
  const int64_t total = int64_t(nt) * nd * na * nch;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i;
    const int ch = int(r % nch); r /= nch;
    const int a = int(r % na); r /= na;
    const int d = int(r % nd); r /= nd;
    const int t = int(r);
    const int64_t o = ((int64_t(t) * na + a) * nch + ch) * nd + d;
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[o * 4 + q] = src[i * 4 + q];
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

__global__ void reset_counter(uint32_t* c, int n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// ---- cycle: grid pass A --------------------------------------------------------------------------
// One workgroup handles one facet group; lane = slot * G + facet-in-group.
template <int READ, int JMODE, int PREC>
__global__ __launch_bounds__(256) void grid_pass_a(BatchParams p, const VisP v, const uint64_t* place, float2* V,
                                                   uint32_t* queue) {
  // LDS: per-facet sums of every facet [nf][nband*npol + 2*nband], flushed once at the end, and
  // the per-channel Jones sums of the current facet group [G][2][nchan] (when they fit),
  // flushed when the workgroup moves to another group.  Work items are group-major, so a
  // workgroup changes group at most a few times.
  extern __shared__ double lds[];
  const int G = p.group;
  const int slots = p.slots;
  const int nbp = p.nband * p.npol;
  const int per_f = nbp + 2 * p.nband;
  double* s_acc = lds;
  double* s_chan = lds + p.nf * per_f;
  const int nchan_acc = p.chan_lds ? G * 2 * v.nchan : 0;
  for (int i = threadIdx.x; i < p.nf * per_f + nchan_acc; i += blockDim.x) lds[i] = 0.;
  __syncthreads();
  const int slot = threadIdx.x / G;
  const int lf = threadIdx.x - slot * G;
  const bool active = slot < slots;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  __shared__ uint32_t s_item;
  int cur_group = -1;

  auto flush_chan = [&](int group) {
    if (!p.chan_lds || group < 0) return;
    for (int i = threadIdx.x; i < nchan_acc; i += blockDim.x) {
      const double x = s_chan[i];
      if (x == 0.) continue;
      const int fg = i / (2 * v.nchan);
      const int f = group * G + fg;
      if (f < p.nf) atomicAdd(&p.facets[f].sum_jones_chan[i - fg * 2 * v.nchan], x);
      s_chan[i] = 0.;
    }
  };

  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
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
    const bool fok = active && f < p.nf;
    const int64_t b0 = chunk * p.blocks_per_item;
    const int64_t b1 = min(b0 + p.blocks_per_item, v.nblocks);
    for (int64_t b = b0 + slot; fok && b < b1; b += slots) {
      const uint64_t rec = place[b * p.nf + f];
      if (!place_valid(rec)) continue;
      const FacetDev& fd = p.facets[f];
      BlockRec br = read_block(v, b);
      const int64_t mid = br.rows[br.nr / 2];
      const bool mid_ok = mid >= 0 && mid < v.nrow;
      DecorrP dp = p.dp;
      dp.l0 = fd.dec_l0;
      dp.m0 = fd.dec_m0;
      const double dec = mid_ok ? decorr_factor(dp, v.uvw, p.fmean0, mid) : 1.;
      JonesP jp = p.jp;
      jp.kdir = fd.kms_dir;
      jp.bdir = fd.beam_dir;
      double swt = 0, ssw = 0, sj = 0;
      cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      const Unpacked u = unpack_place(rec);
      double BBblk = 0;
      JonesPair jblk;
      if (JMODE == 2) {
        // block Jones at the block means (gridder.h:291-301), means in the reference order
        double su = 0, sv = 0, sw = 0;
        int cnt = 0;
        for (int r = 0; r < br.nr; ++r) {
          const int64_t row = br.rows[r];
          if (row < 0 || row >= v.nrow) continue;
          const double* q = v.uvw + 3 * row;
          for (int c = br.ch0; c < br.ch1; ++c) {
            if (v.flags[(size_t(row) * v.nchan + c) * v.ncorr]) continue;
            su += shifted(q[0], q[2], fd.cu);
            sv += shifted(q[1], q[2], fd.cv);
            sw += q[2];
            ++cnt;
          }
        }
        double uvwm[3] = {su / cnt, sv / cnt, sw / cnt};
        jblk = jones_at(jp, mid_ok ? mid : 0, (br.ch0 + br.ch1) / 2, uvwm, p.wl_mean);
        BBblk = jblk.BB;
      }
      double* sch = s_chan + lf * 2 * v.nchan;
      JonesCache jc;
      jc.reset();
      for (int r = 0; r < br.nr; ++r) {
        const int64_t row = br.rows[r];
        if (row < 0 || row >= v.nrow) continue;
        const double* uvw = v.uvw + 3 * row;
        const double angle = -2. * kPi * (uvw[0] * fd.l0 + uvw[1] * fd.m0 + uvw[2] * fd.n0) / kC;
        for (int c = br.ch0; c < br.ch1; ++c) {
          const size_t vc = size_t(row) * v.nchan + c;
          if (v.flags[vc * v.ncorr]) continue;
          const double fw = double(v.weights[vc]);
          const double fwd = fw * dec * dec;
          cd meas[4];
          if (p.dopsf) {
            meas[0] = cd{dec, 0}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = cd{dec, 0};
          } else {
            const float2* vp = v.vis + vc * v.ncorr;
            if (READ == 4) {
#pragma unroll
              for (int q = 0; q < 4; ++q) meas[q] = cd{vp[q].x, vp[q].y};
            } else if (READ == 2) {
              meas[0] = cd{vp[0].x, vp[0].y}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = cd{vp[1].x, vp[1].y};
            } else {
              meas[0] = cd{vp[0].x, vp[0].y}; meas[1] = cd{0, 0}; meas[2] = cd{0, 0}; meas[3] = meas[0];
            }
          }
          cd wt = cd{fw, 0};
          if (!p.dopsf) wt = phasor_prec<PREC>(v.freqs[c], angle) * fw;
          if (JMODE == 1) {
            const JonesPair& jj = jc.get(jp, row, c, uvw, p.wl_mean);
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
            if (p.chan_lds) atomicAdd(&sch[c], x);
            else atomicAdd(&fd.sum_jones_chan[c], x);
          }
          if (JMODE == 2) {
            const double x = double(npg) * (fwd * BBblk);
            if (p.chan_lds) atomicAdd(&sch[c], x);
            else atomicAdd(&fd.sum_jones_chan[c], x);
          }
          if (JMODE != 0) {
            const double x = double(npg) * fwd;
            if (p.chan_lds) atomicAdd(&sch[v.nchan + c], x);
            else atomicAdd(&fd.sum_jones_chan[v.nchan + c], x);
          }
          if (READ == 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = acc[q] + meas[q] * wt;
          } else if (READ == 2) {
            acc[0] = acc[0] + meas[0] * wt;
            acc[3] = acc[3] + meas[3] * wt;
          } else {
            acc[0] = acc[0] + meas[0] * wt;
          }
          swt += fw;
          ssw += fwd;
        }
      }
      if (JMODE == 2) {
        M2 m{{acc[0], acc[1], acc[2], acc[3]}};
        if (p.dopsf) m = mul(mul(jblk.J0, m), herm(jblk.J1));
        m = mul(mul(herm(jblk.J0), m), jblk.J1);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = m.a[q];
        sj = ssw * BBblk;
      }
      float2* out = V + (size_t(b) * p.nf + f) * p.nstokes;
      for (int k = 0; k < p.nstokes; ++k) {
        cd s = cd{0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) s = s + cd{p.sm[k][q][0], p.sm[k][q][1]} * acc[q];
        out[k] = make_float2(float(s.re), float(s.im));
      }
      double* a = s_acc + f * per_f;
      for (int ip = 0; ip < npg; ++ip) {
        atomicAdd(&a[u.band * p.npol + ip], swt);
        if (JMODE != 0) {
          atomicAdd(&a[nbp + u.band], sj);
          atomicAdd(&a[nbp + p.nband + u.band], ssw);
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
    if (k < nbp) atomicAdd(&p.facets[f].sumwt[k], x);
    else atomicAdd(&p.facets[f].sum_jones[k - nbp], x);
  }
}

// ---- cycle: grid pass B (LDS tiles) ---------------------------------------------------------------
template <bool CONJ>
__global__ __launch_bounds__(256) void grid_pass_b(BatchParams p, const uint64_t* bplace, const float2* V,
                                                   const BinItem* items, const uint32_t* recs, uint32_t* queue) {
  extern __shared__ float tile[];  // [npg][W][W][2]
  const int W = p.tile + 2 * p.halo;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  __shared__ uint32_t s_item;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    if (it >= p.nitems_b) break;
    const BinItem bi = items[it];
    for (int i = threadIdx.x; i < npg * W * W * 2; i += blockDim.x) tile[i] = 0.f;
    __syncthreads();
    const int f = bi.facet;
    const FacetDev& fd = p.facets[f];
    const int x0 = bi.tx * p.tile - p.halo, y0 = bi.ty * p.tile - p.halo;
    // each wavefront takes 64 records at a time: lane j fetches record j's placement and values,
    // then the wavefront stamps them one after the other (broadcast from lane j)
    for (uint32_t k0 = bi.begin + 64u * wave; k0 < bi.end; k0 += 64u * nwave) {
      const uint32_t k = k0 + lane;
      const bool ok = k < bi.end;
      const uint64_t my_rec = ok ? bplace[k] : 0;
      const uint32_t my_t = ok ? recs[k] : 0;
      float2 my_v[4];
#pragma unroll
      for (int ip = 0; ip < 4; ++ip) my_v[ip] = (ok && ip < npg) ? V[size_t(my_t) * p.nstokes + ip] : make_float2(0.f, 0.f);
      const int cnt = int(min(64u, bi.end - k0));
      for (int j = 0; j < cnt; ++j) {
        const uint64_t rec = __shfl(my_rec, j, 64);
        const Unpacked u = unpack_place(rec);
        const int S = plane_support(p, u.plane);
        const int S2 = S * S, sup = (S - 1) / 2;
        const float2* cfp = fd.cf + fd.cf_offset[u.plane] + int64_t(u.io * p.overs + u.jo) * S2;
        float2 vv[4];
#pragma unroll
        for (int ip = 0; ip < 4; ++ip) vv[ip] = make_float2(__shfl(my_v[ip].x, j, 64), __shfl(my_v[ip].y, j, 64));
        for (int c = lane; c < S2 * npg; c += 64) {
          const int ip = c / S2;
          const int kk = c - ip * S2;
          const int sy = kk / S, sx = kk - sy * S;
          float2 w = cfp[kk];
          if (CONJ) w.y = -w.y;
          const float2 val = ip == 0 ? vv[0] : (ip == 1 ? vv[1] : (ip == 2 ? vv[2] : vv[3]));
          const float re = val.x * w.x - val.y * w.y;
          const float im = val.x * w.y + val.y * w.x;
          const int ty = u.locy - sup + sy - y0, tx = u.locx - sup + sx - x0;
          float* cell = tile + ((size_t(ip) * W + ty) * W + tx) * 2;
          atomicAdd(cell, re);
          atomicAdd(cell + 1, im);
        }
      }
    }
    __syncthreads();
    float* grid = reinterpret_cast<float*>(fd.grid);
    for (int i = threadIdx.x; i < npg * W * W; i += blockDim.x) {
      const float re = tile[2 * i], im = tile[2 * i + 1];
      if (re == 0.f && im == 0.f) continue;
      const int ip = i / (W * W);
      const int rem = i - ip * W * W;
      const int gy = y0 + rem / W, gx = x0 + rem % W;
      if (gx < 0 || gy < 0 || gx >= p.nx || gy >= p.ny) continue;
      const size_t gi = ((size_t(bi.band) * p.npol + ip) * p.ny + gy) * p.nx + gx;
      atomicAdd(&grid[2 * gi], re);
      atomicAdd(&grid[2 * gi + 1], im);
    }
    __syncthreads();
  }
}

// ---- cycle: degrid pass A (LDS-staged gather) --------------------------------------------------
__global__ __launch_bounds__(256) void degrid_pass_a(BatchParams p, const uint64_t* bplace, float2* M,
                                                     const BinItem* items, const uint32_t* recs, uint32_t* queue) {
  extern __shared__ float tile[];  // [nstokes][W][W][2]
  const int W = p.tile + 2 * p.halo;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  __shared__ uint32_t s_item;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    if (it >= p.nitems_b) break;
    const BinItem bi = items[it];
    const int f = bi.facet;
    const FacetDev& fd = p.facets[f];
    const int x0 = bi.tx * p.tile - p.halo, y0 = bi.ty * p.tile - p.halo;
    const float* model = reinterpret_cast<const float*>(fd.model);
    for (int i = threadIdx.x; i < p.nstokes * W * W; i += blockDim.x) {
      const int ip = i / (W * W);
      const int rem = i - ip * W * W;
      const int gy = y0 + rem / W, gx = x0 + rem % W;
      float re = 0.f, im = 0.f;
      if (gx >= 0 && gy >= 0 && gx < p.nx && gy < p.ny) {
        const size_t gi = ((size_t(bi.band) * p.npol + ip) * p.ny + gy) * p.nx + gx;
        re = model[2 * gi];
        im = model[2 * gi + 1];
      }
      tile[2 * i] = re;
      tile[2 * i + 1] = im;
    }
    __syncthreads();
    for (uint32_t k0 = bi.begin + 64u * wave; k0 < bi.end; k0 += 64u * nwave) {
      const uint32_t k = k0 + lane;
      const bool ok = k < bi.end;
      const uint64_t my_rec = ok ? bplace[k] : 0;
      const uint32_t my_t = ok ? recs[k] : 0;
      float2 res[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      const int cnt = int(min(64u, bi.end - k0));
      for (int j = 0; j < cnt; ++j) {
        const uint64_t rec = __shfl(my_rec, j, 64);
        const Unpacked u = unpack_place(rec);
        const int S = plane_support(p, u.plane);
        const int S2 = S * S, sup = (S - 1) / 2;
        const float2* cfp = fd.cf + fd.cf_offset[u.plane] + int64_t(u.io * p.overs + u.jo) * S2;
        double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
        for (int c = lane; c < S2 * p.nstokes; c += 64) {
          const int ip = c / S2;
          const int kk = c - ip * S2;
          const int sy = kk / S, sx = kk - sy * S;
          const int ty = u.locy - sup + sy - y0, tx = u.locx - sup + sx - x0;
          const float* cell = tile + ((size_t(ip) * W + ty) * W + tx) * 2;
          const float2 w = cfp[kk];
          const float pr = cell[0] * w.x - cell[1] * w.y;
          const float pi = cell[0] * w.y + cell[1] * w.x;
          if (ip == 0) { sr[0] += pr; si[0] += pi; }
          else if (ip == 1) { sr[1] += pr; si[1] += pi; }
          else if (ip == 2) { sr[2] += pr; si[2] += pi; }
          else { sr[3] += pr; si[3] += pi; }
        }
#pragma unroll
        for (int ip = 0; ip < 4; ++ip) {
          if (ip < p.nstokes) {
            const double a = wave_sum(sr[ip]), b = wave_sum(si[ip]);
            if (lane == j) res[ip] = make_float2(float(a), float(b));
          }
        }
      }
      if (ok) {
#pragma unroll
        for (int ip = 0; ip < 4; ++ip)
          if (ip < p.nstokes) M[size_t(my_t) * p.nstokes + ip] = res[ip];
      }
    }
    __syncthreads();
  }
}

// ---- cycle: degrid pass B (lanes over visibilities) ----------------------------------------------
template <int NCORR, int JMODE, int PREC>
__global__ __launch_bounds__(256) void degrid_pass_b(BatchParams p, const VisP v, const float2* M,
                                                     const uint64_t* vrc, const uint32_t* vblk, int64_t nvl) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nvl; k += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t e = vrc[k];
    const int64_t row = int64_t(uint32_t(e));
    const int c = int(e >> 32);
    const int64_t b = vblk[k];
    if (row < 0 || row >= v.nrow || c < 0 || c >= v.nchan) continue;
    const double* uvw = v.uvw + 3 * row;
    const double fr = v.freqs[c];
    cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    BlockRec br;
    bool have_br = false;
    double fm = 0;
    for (int f = 0; f < p.nf; ++f) {
      const float2* m = M + (size_t(b) * p.nf + f) * p.nstokes;
      cd st[4];
      bool any = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < p.nstokes) {
          const float2 x = m[q];
          st[q] = cd{x.x, x.y};
          any |= (x.x != 0.f || x.y != 0.f);
        } else {
          st[q] = cd{0, 0};
        }
      }
      if (!any) continue;  // dropped record, or a zero model: contributes exactly 0
      const FacetDev& fd = p.facets[f];
      M2 cv;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cv.a[q] = cd{0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) cv.a[q] = cv.a[q] + cd{p.sm[q][s][0], p.sm[q][s][1]} * st[s];
      }
      JonesP jp = p.jp;
      jp.kdir = fd.kms_dir;
      jp.bdir = fd.beam_dir;
      if ((JMODE == 2 || p.dp.enabled) && !have_br) {
        br = read_block(v, b);
        for (int cc = br.ch0; cc < br.ch1; ++cc) fm += v.freqs[cc];
        fm /= double(br.ch1 - br.ch0);
        have_br = true;
      }
      if (JMODE == 2) {
        const int64_t mid = br.rows[br.nr / 2];
        if (mid >= 0 && mid < v.nrow) {
          JonesPair jj = jones_at(jp, mid, (br.ch0 + br.ch1) / 2, v.uvw + 3 * mid, p.wl_mean);
          cv = mul(mul(jj.J0, cv), herm(jj.J1));
        }
      }
      double dec = 1.;
      if (p.dp.enabled) {
        const int64_t mid = br.rows[br.nr / 2];
        DecorrP dp = p.dp;
        dp.l0 = fd.dec_l0;
        dp.m0 = fd.dec_m0;
        if (mid >= 0 && mid < v.nrow) dec = decorr_factor(dp, v.uvw, fm, mid);
      }
      const double angle = 2. * kPi * (uvw[0] * fd.l0 + uvw[1] * fd.m0 + uvw[2] * fd.n0) / kC;
      const cd corr = phasor_prec<PREC>(fr, angle) * dec;
      if (JMODE == 1) {
        JonesPair jj = jones_at(jp, row, c, uvw, p.wl_mean);
        cv = mul(mul(jj.J0, cv), herm(jj.J1));
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = acc[q] + cv.a[q] * corr;
    }
    float2* vp = v.vis + (size_t(row) * v.nchan + c) * NCORR;
    if (NCORR == 4) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 x = vp[q];
        vp[q] = make_float2(float(double(x.x) - acc[q].re), float(double(x.y) - acc[q].im));
      }
    } else if (NCORR == 2) {
      const float2 x0 = vp[0], x1 = vp[1];
      vp[0] = make_float2(float(double(x0.x) - acc[0].re), float(double(x0.y) - acc[0].im));
      vp[1] = make_float2(float(double(x1.x) - acc[3].re), float(double(x1.y) - acc[3].im));
    } else {
      const float2 x0 = vp[0];
      vp[0] = make_float2(float(double(x0.x) - acc[0].re), float(double(x0.y) - acc[0].im));
    }
  }
}


// ---- adjoint of degrid pass B ------------------------------------------------------------------------
// Forward (degrid_pass_b): vis_k = mask(corr_k * J0 (D M) J1^H) summed over facets.  Adjoint for
// one (block, facet): Mbar = D^H (sum_k conj(corr_k) J0^H vbar_k J1), mode-2 Jones applied to the sum.
// Lanes are facets; each lane walks its block's visibilities (all of them: degridding ignores flags).
template <int NCORR, int JMODE, int PREC>
__global__ __launch_bounds__(256) void degrid_adj_b(BatchParams p, const VisP v, const uint64_t* place,
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
      if (!place_valid(place[b * p.nf + f])) continue;
      const FacetDev& fd = p.facets[f];
      BlockRec br = read_block(v, b);
      double fm = 0;
      for (int c = br.ch0; c < br.ch1; ++c) fm += v.freqs[c];
      fm /= double(br.ch1 - br.ch0);
      const int64_t mid = br.rows[br.nr / 2];
      const bool mid_ok = mid >= 0 && mid < v.nrow;
      DecorrP dp = p.dp;
      dp.l0 = fd.dec_l0;
      dp.m0 = fd.dec_m0;
      const double dec = mid_ok ? decorr_factor(dp, v.uvw, fm, mid) : 1.;
      JonesP jp = p.jp;
      jp.kdir = fd.kms_dir;
      jp.bdir = fd.beam_dir;
      cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      for (int r = 0; r < br.nr; ++r) {
        const int64_t row = br.rows[r];
        if (row < 0 || row >= v.nrow) continue;
        const double* uvw = v.uvw + 3 * row;
        const double angle = 2. * kPi * (uvw[0] * fd.l0 + uvw[1] * fd.m0 + uvw[2] * fd.n0) / kC;
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
          const cd cc = conj(phasor_prec<PREC>(v.freqs[c], angle)) * dec;
          if (JMODE == 1) {
            JonesPair jj = jones_at(jp, row, c, uvw, p.wl_mean);
            m = mul(mul(herm(jj.J0), m), jj.J1);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = acc[q] + m.a[q] * cc;
        }
      }
      if (JMODE == 2 && mid_ok) {
        JonesPair jj = jones_at(jp, mid, (br.ch0 + br.ch1) / 2, v.uvw + 3 * mid, p.wl_mean);
        M2 m{{acc[0], acc[1], acc[2], acc[3]}};
        m = mul(mul(herm(jj.J0), m), jj.J1);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = m.a[q];
      }
      float2* out = Mbar + (size_t(b) * p.nf + f) * p.nstokes;
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

int batched_place(const BatchParams& p, const VisP& v, bool degrid, uint64_t* place, hipStream_t st) {
  int64_t total = v.nblocks * p.nf;
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  if (degrid) hipLaunchKernelGGL(plan_place<true>, dim3(unsigned(nb)), dim3(256), 0, st, p, v, place);
  else hipLaunchKernelGGL(plan_place<false>, dim3(unsigned(nb)), dim3(256), 0, st, p, v, place);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_count(const BatchParams& p, int64_t nblocks, const uint64_t* place, uint32_t* counts, hipStream_t st) {
  int64_t total = nblocks * p.nf;
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_count, dim3(unsigned(nb)), dim3(256), 0, st, p, nblocks, place, counts);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_fill(const BatchParams& p, int64_t nblocks, const uint64_t* place, const int64_t* bin_start,
                 uint32_t* fill, uint32_t* recs, hipStream_t st) {
  int64_t total = nblocks * p.nf;
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_fill, dim3(unsigned(nb)), dim3(256), 0, st, p, nblocks, place, bin_start, fill, recs);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
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
static void launch_a(const BatchParams& p, const VisP& v, const uint64_t* place, float2* V, uint32_t* q,
                     size_t lds, unsigned nwg, hipStream_t st) {
  hipLaunchKernelGGL((grid_pass_a<READ, JMODE, PREC>), dim3(nwg), dim3(256), lds, st, p, v, place, V, q);
}

template <int READ, int JMODE>
static void launch_a_p(const BatchParams& p, const VisP& v, const uint64_t* place, float2* V, uint32_t* q,
                       size_t lds, unsigned nwg, hipStream_t st) {
  if (p.prec == 1) launch_a<READ, JMODE, 1>(p, v, place, V, q, lds, nwg, st);
  else if (p.prec == 2) launch_a<READ, JMODE, 2>(p, v, place, V, q, lds, nwg, st);
  else launch_a<READ, JMODE, 0>(p, v, place, V, q, lds, nwg, st);
}

template <int READ>
static void launch_a_j(const BatchParams& p, const VisP& v, const uint64_t* place, float2* V, uint32_t* q,
                       size_t lds, unsigned nwg, hipStream_t st) {
  if (p.jmode == 1) launch_a_p<READ, 1>(p, v, place, V, q, lds, nwg, st);
  else if (p.jmode == 2) launch_a_p<READ, 2>(p, v, place, V, q, lds, nwg, st);
  else launch_a_p<READ, 0>(p, v, place, V, q, lds, nwg, st);
}

int batched_grid_a(const BatchParams& p, const VisP& v, const uint64_t* place, float2* V, uint32_t* queue,
                   unsigned nwg, hipStream_t st) {
  size_t lds = (size_t(p.nf) * (p.nband * p.npol + 2 * p.nband) + (p.chan_lds ? size_t(p.group) * 2 * v.nchan : 0)) * 8;
  if (p.read == 4) launch_a_j<4>(p, v, place, V, queue, lds, nwg, st);
  else if (p.read == 2) launch_a_j<2>(p, v, place, V, queue, lds, nwg, st);
  else launch_a_j<1>(p, v, place, V, queue, lds, nwg, st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_grid_b(const BatchParams& p, const uint64_t* place, const float2* V, const BinItem* items,
                   const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  size_t lds = size_t(npg) * W * W * 2 * sizeof(float);
  hipLaunchKernelGGL(grid_pass_b<false>, dim3(nwg), dim3(256), lds, st, p, place, V, items, recs, queue);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

// Adjoint of degrid pass A: conj(CF) stamps of M-bar, scattered through the degrid bins.
int batched_degrid_a_adjoint(const BatchParams& p, const uint64_t* place, const float2* Mbar, const BinItem* items,
                             const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  size_t lds = size_t(p.nstokes) * W * W * 2 * sizeof(float);
  BatchParams q = p;
  q.npol = p.npol;
  hipLaunchKernelGGL(grid_pass_b<true>, dim3(nwg), dim3(256), lds, st, q, place, Mbar, items, recs, queue);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_degrid_a(const BatchParams& p, const uint64_t* place, float2* M, const BinItem* items,
                     const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st) {
  const int W = p.tile + 2 * p.halo;
  size_t lds = size_t(p.nstokes) * W * W * 2 * sizeof(float);
  hipLaunchKernelGGL(degrid_pass_a, dim3(nwg), dim3(256), lds, st, p, place, M, items, recs, queue);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

template <int NCORR, int JMODE>
static void launch_db_p(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc,
                        const uint32_t* vblk, int64_t nvl, unsigned nb, hipStream_t st) {
  if (p.prec == 1) hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 1>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, nvl);
  else if (p.prec == 2) hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 2>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, nvl);
  else hipLaunchKernelGGL((degrid_pass_b<NCORR, JMODE, 0>), dim3(nb), dim3(256), 0, st, p, v, M, vrc, vblk, nvl);
}

template <int NCORR>
static void launch_db_j(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc,
                        const uint32_t* vblk, int64_t nvl, unsigned nb, hipStream_t st) {
  if (p.jmode == 1) launch_db_p<NCORR, 1>(p, v, M, vrc, vblk, nvl, nb, st);
  else if (p.jmode == 2) launch_db_p<NCORR, 2>(p, v, M, vrc, vblk, nvl, nb, st);
  else launch_db_p<NCORR, 0>(p, v, M, vrc, vblk, nvl, nb, st);
}

int batched_degrid_b(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc, const uint32_t* vblk,
                     int64_t nvl, hipStream_t st) {
  int64_t nb = (nvl + 255) / 256;
  if (nb > 65536 * 4) nb = 65536 * 4;
  if (nb < 1) nb = 1;
  if (v.ncorr == 4) launch_db_j<4>(p, v, M, vrc, vblk, nvl, unsigned(nb), st);
  else if (v.ncorr == 2) launch_db_j<2>(p, v, M, vrc, vblk, nvl, unsigned(nb), st);
  else launch_db_j<1>(p, v, M, vrc, vblk, nvl, unsigned(nb), st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf

namespace ddf {

template <int NCORR, int JMODE>
static void launch_adj_b_p(const BatchParams& p, const VisP& v, const uint64_t* place, const float2* vbar,
                           float2* Mbar, uint32_t* q, unsigned nwg, hipStream_t st) {
  if (p.prec == 1)
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 1>), dim3(nwg), dim3(256), 0, st, p, v, place, vbar, Mbar, q);
  else if (p.prec == 2)
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 2>), dim3(nwg), dim3(256), 0, st, p, v, place, vbar, Mbar, q);
  else
    hipLaunchKernelGGL((degrid_adj_b<NCORR, JMODE, 0>), dim3(nwg), dim3(256), 0, st, p, v, place, vbar, Mbar, q);
}

template <int NCORR>
static void launch_adj_b_j(const BatchParams& p, const VisP& v, const uint64_t* place, const float2* vbar,
                           float2* Mbar, uint32_t* q, unsigned nwg, hipStream_t st) {
  if (p.jmode == 1) launch_adj_b_p<NCORR, 1>(p, v, place, vbar, Mbar, q, nwg, st);
  else if (p.jmode == 2) launch_adj_b_p<NCORR, 2>(p, v, place, vbar, Mbar, q, nwg, st);
  else launch_adj_b_p<NCORR, 0>(p, v, place, vbar, Mbar, q, nwg, st);
}

int batched_degrid_b_adjoint(const BatchParams& p, const VisP& v, const uint64_t* place, const float2* vbar,
                             float2* Mbar, uint32_t* queue, unsigned nwg, hipStream_t st) {
  if (v.ncorr == 4) launch_adj_b_j<4>(p, v, place, vbar, Mbar, queue, nwg, st);
  else if (v.ncorr == 2) launch_adj_b_j<2>(p, v, place, vbar, Mbar, queue, nwg, st);
  else launch_adj_b_j<1>(p, v, place, vbar, Mbar, queue, nwg, st);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf

namespace ddf {

int batched_gather_place(const uint64_t* place, const uint32_t* recs, int64_t n, uint64_t* bplace, hipStream_t st) {
  if (n <= 0) return DDF_OK;
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(plan_gather_place, dim3(unsigned(nb)), dim3(256), 0, st, place, recs, n, bplace);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

int batched_transpose_jones(const float2* src, float2* dst, const int* dims, hipStream_t st) {
  int64_t total = int64_t(dims[0]) * dims[1] * dims[2] * dims[3];
  int64_t nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(plan_transpose_jones, dim3(unsigned(nb)), dim3(256), 0, st, src, dst, dims[0], dims[1], dims[2],
                     dims[3]);
  return hipGetLastError() == hipSuccess ? DDF_OK : DDF_EHIP;
}

}  // namespace ddf
