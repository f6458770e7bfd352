// Definitions shared by the plain (ddfgpu.hip) and batched (ddf_batched.hip) translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ddf_device.h"

namespace ddf {

// Visibility-side kernel arguments (device pointers).
struct VisP {
  float2* vis;
  const double* uvw;
  const uint8_t* flags;
  const float* weights;
  const double* freqs;
  const int32_t* chan_map;
  int64_t nrow;
  int nchan, ncorr;
  const int32_t* bda;
  const int64_t* bda_start;
  int64_t nblocks, bda_len;
  const uint8_t* sparse;
  int equi;  // channels equidistant (LOptimisation[1])
  double f0, df;  // equidistant channels: first frequency and step (set by the batched plan)
  // per-block descriptors of the batched plan (null elsewhere): x record offset in bda, y row
  // count | 1 << 31 when the rows are row0, row0 + 1, ..., z ch0 | ch1 << 16, w row0
  const uint4* bdesc;
};

struct BlockRec {
  int ch0, ch1, nr;
  const DDF_GLOBAL int32_t* rows;
};

// Block b of the BDA map, with its channel range clamped to [0, nchan] and its row count to the
// map length, so a malformed map cannot send a kernel outside its buffers.
__device__ __forceinline__ BlockRec read_block(const VisP& v, int64_t b) {
  int64_t s = gptr(v.bda_start)[b];
  BlockRec r;
  int nr = gptr(v.bda)[2 + b] - 2;
  if (s < 0 || s + 2 > v.bda_len) {
    r.ch0 = r.ch1 = 0;
    r.nr = 0;
    r.rows = gptr(v.bda);
    return r;
  }
  if (nr < 0) nr = 0;
  if (s + 2 + nr > v.bda_len) nr = int(v.bda_len - s - 2);
  r.ch0 = clampi(gptr(v.bda)[s], 0, v.nchan);
  r.ch1 = clampi(gptr(v.bda)[s + 1], r.ch0, v.nchan);
  r.nr = nr;
  r.rows = gptr(v.bda) + s + 2;
  return r;
}

// Block b from its plan descriptor (one 16-byte load, no dependent loads): the BlockRec of
// read_block, plus whether its rows are consecutive (row r of the block is then row0 + r).
struct BlockDescRec {
  BlockRec br;
  bool consec;
  int64_t row0;
  __device__ __forceinline__ int64_t row(int r) const { return consec ? row0 + r : int64_t(br.rows[r]); }
};
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ BlockDescRec read_block_desc(const VisP& v, int64_t b) {
  const u32x4 d = reinterpret_cast<const DDF_GLOBAL u32x4*>(gptr(reinterpret_cast<const uint32_t*>(v.bdesc)))[b];
  BlockDescRec o;
  o.br.rows = gptr(v.bda) + int64_t(d.x) + 2;
  o.br.nr = int(d.y & 0x7FFFFFFFu);
  o.br.ch0 = int(d.z & 0xFFFFu);
  o.br.ch1 = int(d.z >> 16);
  o.consec = (d.y >> 31) != 0;
  o.row0 = int64_t(d.w);
  return o;
}

// Phasor exp(i * freq * angle) at three precisions:
//   0 mixed: FP64 phase and range reduction, FP32 sin/cos (phasor: ~1e-7 for any phase);
//   1 FP64 throughout;
//   2 FP32 throughout (phase formed and reduced in FP32: the precision-study variant).
template <int PREC>
__device__ __forceinline__ cd phasor_prec(double freq, double angle) {
  if (PREC == 1) {
    double s, c;
    sincos(freq * angle, &s, &c);
    return cd{c, s};
  } else if (PREC == 2) {
    float s, c;
    sincosf(float(freq) * float(angle), &s, &c);
    return cd{double(c), double(s)};
  } else {
    return phasor(freq * angle);
  }
}

// Phasors exp(i freq_c angle) along the channels of one row, for a lane that walks them in
// increasing order.  PREC 0 with equidistant channels: FP64 phase base (freq_c = f0 + c df, from
// VisP, so no frequency is loaded), FP32 sin/cos at an anchor (the first channel walked, a channel
// after a gap, every multiple of 16), and in between the phasor advanced in FP32 by the per-row
// step exp(i df angle), formed at the first step of the row (rows of a single channel never
// form it).  Other precisions, or non-equidistant channels: every phasor formed directly by
// phasor_prec.
template <int PREC>
struct ChanPhasor {
  double angle;
  int last;
  bool have_step;
  float cr, ci, sr, si;
  __device__ __forceinline__ void row(double a) {
    angle = a;
    last = -2;
    have_step = false;
  }
  __device__ __forceinline__ cd at(const VisP& v, int c) {
    if (PREC != 0 || !v.equi) return phasor_prec<PREC>(gptr(v.freqs)[c], angle);
    if (c == last + 1 && (c & 15) != 0) {
      if (!have_step) {
        const cd st = phasor(v.df * angle);
        sr = float(st.re);
        si = float(st.im);
        have_step = true;
      }
      const float r = cr * sr - ci * si, i = cr * si + ci * sr;
      cr = r;
      ci = i;
    } else {
      const cd a = phasor(fma(double(c), v.df, v.f0) * angle);
      cr = float(a.re);
      ci = float(a.im);
    }
    last = c;
    return cd{cr, ci};
  }
  // phasor x weight in the arithmetic type T (float: the FP32 phasor times the FP32 weight, the
  // same correctly rounded product as forming it in FP64 and rounding)
  template <typename T>
  __device__ __forceinline__ cx<T> weighted(const VisP& v, int c, float w);
};
template <int PREC>
template <typename T>
__device__ __forceinline__ cx<T> ChanPhasor<PREC>::weighted(const VisP& v, int c, float w) {
  if (std::is_same<T, double>::value || PREC == 1 || (PREC == 0 && !v.equi)) return cvt<T>(at(v, c) * double(w));
  if (PREC == 2) {
    float s, co;
    sincosf(float(gptr(v.freqs)[c]) * float(angle), &s, &co);
    return cx<T>{T(co * w), T(s * w)};
  }
  (void)at(v, c);
  return cx<T>{T(cr * w), T(ci * w)};
}

// One facet as the batched kernels see it.
struct FacetDev {
  void* grid;              // complex64 [nband][npol][ny][nx] (gridding output)
  const void* model;       // complex64 model grid (degridding input)
  const float2* cf;        // CF planes, put order (Wplanes then WplanesConj)
  const float2* cft;       // the same planes with rows padded to an even pitch, at twice the offsets (degrid pass A)
  const int64_t* cf_offset;// [2*nw]
  double cu, cv, l0, m0, n0;
  double dec_l0, dec_m0;   // LSmearing[5:7]
  int kms_dir, beam_dir;
  double* sumwt;           // [nband][npol]
  double* sum_jones;       // [2][nband]
  double* sum_jones_chan;  // [2][nchan]
};

// Work-item ranges of a tile pass, one per XCD (workgroup b is dispatched to XCD b % 8), for plans
// of at least kRangeMinItems work items.
constexpr int kItemRanges = 8;
constexpr size_t kRangeMinItems = 65536;

struct BinItem {
  uint32_t begin, end;     // record range in the binned record list
  int16_t facet, band, tx, ty;
  uint32_t slot;           // tile slot (the slots of one tile are consecutive)
};

struct BatchParams {
  const FacetDev* facets;
  int nf;
  int nband, npol, ny, nx;
  double incr0, incr1;
  int dopsf;
  // CF geometry shared by every facet
  const int32_t* support;  // [nw]
  int nw, overs;
  double ref_wave, wmax;
  // Stokes map (grid: stokes <- corr; degrid: corr <- stokes)
  int nstokes, read;
  double sm[4][4][2];
  // Jones tables (direction taken per facet) and decorrelation (l0, m0 per facet)
  JonesP jp;
  int jmode;
  DecorrP dp;
  double fmean0, wl_mean;
  double tscale0;  // decorr_tscale(fmean0)
  int prec;
  // pass A layout
  int group, slots, ngroups;
  int64_t blocks_per_item, nchunks;
  uint32_t nitems_a;
  int chan_lds;
  // tiles
  int tile, halo, ntiles_x, ntiles_y;
  uint32_t nitems_b;
  // the work items split into nranges (1 or kItemRanges) contiguous ranges of about equal record
  // counts: range x is items [item_range[x], item_range[x + 1])
  const uint32_t* item_range;
  int nranges;
  // blocks of the pass's BDA map: the per-(block, facet) values between the passes are stored
  // facet-major, [facet][block][nstokes]
  int64_t nblk;
};

// Staged visibility entry of gridding pass A (floats): weight, flag word, row, channel; READ complex
// values from offset 4; the Jones key (6 ints) from grid_a_key_offset; uvw (3 doubles) after it
// (in the mixed precision scaled to revolutions per unit direction cosine); the stride a multiple
// of 4 floats so that every entry is 16-byte aligned.
constexpr int grid_a_key_offset(int read) { return (4 + 2 * read + 3) & ~3; }
constexpr int grid_a_stride(int read) { return (grid_a_key_offset(read) + 12 + 3) & ~3; }

// host launchers (ddf_batched.hip)
int batched_place(const BatchParams& p, const VisP& v, bool degrid, uint64_t* place, hipStream_t st);
int batched_pair_codes(const uint64_t* place, int64_t n, uint8_t* code, hipStream_t st);
int batched_group_counts(const uint32_t* counts, int64_t ntile, int nsub, int ngrp, uint32_t* gcount, hipStream_t st);
int batched_bin_starts(const uint32_t* counts, const int64_t* tile_start, int64_t ntile, int nsub, int64_t* bin_start,
                       hipStream_t st);
int batched_block_desc(const VisP& v, uint4* desc, hipStream_t st);
int batched_count(const BatchParams& p, int64_t nblocks, const uint64_t* place, uint32_t* counts, hipStream_t st);
int batched_fill(const BatchParams& p, int64_t nblocks, const uint64_t* place, const int64_t* bin_start,
                 uint32_t* fill, uint32_t* recs, hipStream_t st);
int batched_sort_items(const BatchParams& p, const uint64_t* place, uint32_t* recs, int64_t n, const BinItem* items,
                       uint32_t nitems, bool stamp_order, hipStream_t st);
int batched_vis_counts(const VisP& v, int32_t* cnt, hipStream_t st);
int batched_vlist(const VisP& v, const int64_t* vis_start, uint64_t* vrc, uint32_t* vblk, hipStream_t st);
int batched_reset(uint32_t* counters, int n, hipStream_t st);
// LDS bytes of the visibility staging buffers of gridding pass A (facet group size, corrs read)
size_t batched_grid_a_stage_bytes(int group, int read);
int batched_grid_a(const BatchParams& p, const VisP& v, const uint8_t* pcode, float2* V, uint32_t* queue,
                   unsigned nwg, hipStream_t st);
// slots: null -> tiles added into the grids with global atomics; else one tile slot per work item
// ([nitems][npg][2][W*W]) and grid_slot_reduce (tile_first / tile_count: the items of each tile)
int batched_grid_b(const BatchParams& p, const uint64_t* place, const float2* V, const BinItem* items,
                   const uint32_t* recs, uint32_t* queue, unsigned nwg, float* slots, const uint32_t* tile_first,
                   const uint32_t* tile_count, hipStream_t st);
int batched_degrid_a(const BatchParams& p, const uint64_t* place, float2* M, const BinItem* items,
                     const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st);
// vrun / run_first: the runs of the degrid visibility list (batched_run_ids)
int batched_degrid_b(const BatchParams& p, const VisP& v, const float2* M, const uint64_t* vrc, const uint32_t* vblk,
                     const uint32_t* vrun, const uint32_t* run_first, int64_t nvl, hipStream_t st);
// phase 0: run-start flags of the degrid visibility list; phase 1 (after an exclusive scan of the
// flags into excl): the run of every visibility and the first visibility of every run
int batched_run_ids(const BatchParams& p, const VisP& v, const uint64_t* vrc, const uint32_t* vblk, int64_t nvl,
                    int32_t* flag, const int64_t* excl, uint32_t* vrun, uint32_t* run_first, int phase, hipStream_t st);

int batched_degrid_b_adjoint(const BatchParams& p, const VisP& v, const uint8_t* pcode, const float2* vbar,
                             float2* Mbar, uint32_t* queue, unsigned nwg, hipStream_t st);
int batched_degrid_a_adjoint(const BatchParams& p, const uint64_t* place, const float2* Mbar, const BinItem* items,
                             const uint32_t* recs, uint32_t* queue, unsigned nwg, hipStream_t st);

int batched_tile_wgs(const BatchParams& pg, const BatchParams& pd, unsigned out[3]);
// stamp records (tile offset | support << 16 | CF offset << 32) of the binned records
int batched_stamp_records(const BatchParams& p, const uint64_t* place, const uint32_t* recs, int64_t n,
                          uint64_t* bstamp, hipStream_t st);
// CF planes of one facet copied for degrid pass A: plane p (offset off[p], support S =
// support[p mod nw]) at offset 2 off[p], each sub-kernel's S rows padded to an even pitch with zeros
int batched_cf_pad(const float2* in, float2* out, const int64_t* off, const int32_t* support, int nw, int overs,
                   int max_support, hipStream_t st);
int batched_prepare_jones(const BatchParams& p, const float2* src, double2* dst, float2* dstf, int dd, hipStream_t st);

}  // namespace ddf
