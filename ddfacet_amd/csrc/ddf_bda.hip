// Baseline-dependent-averaging (BDA) block maps built on the device, in the reference int32 word
// format (DDFacet/Data/ClassSmearMapping.py:83-135):
//   [NTot_lo, NTot_hi, size_0 .. size_{NTot-1}, rec_0, rec_1, ...],  rec_b = [chStart, chEnd, rows...]
// with the per-baseline block lists concatenated in lexicographic order of the string key
// "a0:a1" (ClassSmearMapping.py:113) and each baseline's rows in increasing row order
// (np.where((A0 == a0) & (A1 == a1)), :498).
//
// Steps (the reference runs one process-pool job per baseline, ClassSmearMapping.py:64-79):
//   keys     per row the rank of its baseline's string key (a host table), autocorrelations last
//   sort     stable radix sort of the rows by that rank: each baseline's rows become contiguous, in
//            row order
//   segments first / last sorted position of every baseline
//   count    one thread per baseline walks its rows (the builder's sequential scan: time blocks
//            cut on the accumulated |duvw| or the phase, channel blocks from the block's smallest
//            decorrelation span) and counts blocks and words
//   scan     exclusive sums over the baselines in key order -> block and word offsets
//   walk     the same walk again, writing one descriptor per block
//   emit     one thread per block writes its size word and record
// The builder's FP64 arithmetic is evaluated in the reference's operation order with
// contraction off, so every cut and every channel boundary equals the host builder's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ddfgpu.h"

namespace ddf {
void set_last_error(const std::string& msg);  // ddfgpu.hip: the message ddf_last_error() returns
}

namespace {

int bfail(int code, const std::string& m) {
  ddf::set_last_error(m);
  return code;
}

struct BlockDesc {
  int64_t word;    // offset of the record inside the records region
  int64_t row0;    // first sorted-row position of the block
  int32_t nrows;
  int32_t ch0, ch1;
  int32_t pad;
};

struct BdaParams {
  int32_t mode;
  int32_t nchan;
  int32_t nbl;
  // mode 1 (GiveBlocksRowsListBL_old, ClassSmearMapping.py:496-625)
  double duv_cut;  // C dPhi / (pi l nu0)
  double k_dnu;    // (C / pi) dPhi
  double l;
  double dfreq;
  // mode 2 (GiveBlocksRowsListBL as restated in ddfacet_amd/bda.py)
  double k_phase;  // 2 pi nu0 / C
  double lmax, nmax;
  double k_span;   // (C / (2 pi)) dPhi
  double dphi;
  int32_t per_row;  // C dPhi / (2 pi nu0) == 0: every row its own time block
};

__global__ void bda_keys_k(const int32_t* a0, const int32_t* a1, int64_t nrow, const int32_t* rank, int na, int nbl,
                           uint32_t* keys, uint32_t* rows) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nrow; i += int64_t(gridDim.x) * blockDim.x) {
    const int p = a0[i], q = a1[i];
    uint32_t k = uint32_t(nbl);
    if (p != q && p >= 0 && q >= 0 && p < na && q < na) k = uint32_t(rank[p * na + q]);
    keys[i] = k;
    rows[i] = uint32_t(i);
  }
}

__global__ void bda_fill64_k(int64_t* x, int64_t n, int64_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) x[i] = v;
}

// seg[2 k] = first sorted position of baseline k, seg[2 k + 1] = one past its last (0, 0: absent)
__global__ void bda_segments_k(const uint32_t* keys, int64_t nrow, int nbl, int64_t* seg) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nrow; i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t k = keys[i];
    if (k >= uint32_t(nbl)) continue;
    if (i == 0 || keys[i - 1] != k) seg[2 * k] = i;
    if (i == nrow - 1 || keys[i + 1] != k) seg[2 * k + 1] = i + 1;
  }
}

// Channel cuts of one time block: blocks [cut_j, cut_{j+1}) of the (block number, band) changes.
// MODE 1: the block number of channel c is that of the np.int32(np.linspace(0, NChan, NChBlocks))
// boundaries with NChBlocks the length of np.arange(0, NChan, nchb) plus the trailing NChan.
// MODE 2: block number c // size with size = ceil(NChan / ceil(NChan / m)).
struct ChanCuts {
  int32_t nblk;     // MODE 1: number of linspace boundaries; MODE 2: channel block size
  double lin_step;  // MODE 1: NChan / (nblk - 1)
};

__device__ inline ChanCuts chan_cuts_mode1(double nchb, int nchan) {
#pragma clang fp contract(off)
  ChanCuts cc;
  const double step = nchb < 1.0 ? 1.0 : nchb;  // np.max([NChanBlockMax, 1])
  // len(np.arange(0, NChan, step)) = ceil(NChan / step); element i = i * step
  const double n = ceil(double(nchan) / step);
  int64_t len = int64_t(n);
  if (len < 0) len = 0;
  bool has_nchan = false;
  if (len > 0) has_nchan = (double(len - 1) * step == double(nchan));
  cc.nblk = int32_t(len + (has_nchan ? 0 : 1));
  cc.lin_step = cc.nblk > 1 ? double(nchan) / double(cc.nblk - 1) : 0.0;
  return cc;
}

// np.int32(np.linspace(0, nchan, nblk))[k]
__device__ inline int32_t lin_bound(const ChanCuts& cc, int k, int nchan) {
#pragma clang fp contract(off)
  if (k >= cc.nblk - 1) return nchan;
  return int32_t(double(k) * cc.lin_step);
}

// Block number of channel c (the rampe of ClassSmearMapping.py:582-587): the largest k with
// bound_k <= c < bound_{k+1}; channels past the last bound keep 0.
__device__ inline int32_t rampe_of(const ChanCuts& cc, int c, int nchan, int32_t* kcur) {
  // bounds are nondecreasing; advance the cursor
  int32_t k = *kcur;
  while (k + 1 < cc.nblk - 1 && lin_bound(cc, k + 1, nchan) <= c) ++k;
  *kcur = k;
  const int32_t b0 = lin_bound(cc, k, nchan), b1 = lin_bound(cc, k + 1, nchan);
  return (c >= b0 && c < b1) ? k : 0;
}

template <int MODE>
__device__ inline int32_t chan_block_of(const ChanCuts& cc, int c, int nchan, int32_t* kcur) {
  if (MODE == 1) return rampe_of(cc, c, nchan, kcur);
  return c / cc.nblk;
}

// Walks one time block's channel cuts; emit(ch0, ch1) per channel block.
template <int MODE, typename F>
__device__ inline int32_t for_channel_blocks(const ChanCuts& cc, int nchan, const int32_t* chan_map, F emit) {
  int32_t kcur = 0;
  int32_t prev_blk = -1, prev_band = -1;
  int32_t start = 0, n = 0;
  bool have = false;
  for (int c = 0; c < nchan; ++c) {
    const int32_t blk = chan_block_of<MODE>(cc, c, nchan, &kcur);
    const int32_t band = chan_map[c];
    if (!have || blk != prev_blk || band != prev_band) {
      if (have) {
        emit(start, c);
        ++n;
      }
      start = c;
      have = true;
    }
    prev_blk = blk;
    prev_band = band;
  }
  if (have) {
    // MODE 1 appends ChBlock[-1] = NChan as the last cut; MODE 2 appends nchan
    emit(start, nchan);
    ++n;
  }
  return n;
}

// One baseline's walk (rows r[0..n) in row order).  WRITE: descriptors at desc + blk0, record
// words counted from word0; otherwise only the totals.
template <int MODE, bool WRITE>
__device__ void walk_baseline(const BdaParams& P, const double* __restrict__ uvw, const uint32_t* __restrict__ r,
                              int64_t row_base, int64_t n, const int32_t* __restrict__ chan_map, int64_t* nblk_out,
                              int64_t* nwords_out, BlockDesc* desc, int64_t blk0, int64_t word0) {
#pragma clang fp contract(off)
  int64_t nblk = 0, nwords = 0;
  auto close_block = [&](int64_t t0, int64_t t1, const ChanCuts& cc) {
    const int32_t nr = int32_t(t1 - t0);
    for_channel_blocks<MODE>(cc, P.nchan, chan_map, [&](int32_t ch0, int32_t ch1) {
      if (WRITE) {
        BlockDesc d;
        d.word = word0 + nwords;
        d.row0 = row_base + t0;
        d.nrows = nr;
        d.ch0 = ch0;
        d.ch1 = ch1;
        d.pad = 0;
        desc[blk0 + nblk] = d;
      }
      ++nblk;
      nwords += 2 + nr;
    });
  };
  if (MODE == 1) {
    // du[i] = u[i+1] - u[i], the last step repeated; a single row gets the 1e-30 step
    double duvtot = 0.0;
    double nchb_max = 1e3;
    int64_t t0 = 0;
    for (int64_t i = 0; i < n; ++i) {
      const double* x = uvw + 3 * size_t(r[i]);
      double du, dv, dw;
      if (n == 1) {
        du = dv = dw = 1e-30;
      } else {
        const int64_t j = (i + 1 < n) ? i : n - 2;
        const double* a = uvw + 3 * size_t(r[j]);
        const double* b = uvw + 3 * size_t(r[j + 1]);
        du = b[0] - a[0];
        dv = b[1] - a[1];
        dw = b[2] - a[2];
      }
      double uv = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      uv = uv > 1e-10 ? uv : 1e-10;
      const double dnu = P.k_dnu / (uv * P.l);
      const double nchb = dnu / P.dfreq;
      if (nchb < nchb_max) nchb_max = nchb;
      duvtot += sqrt(du * du + dv * dv + dw * dw);
      if (duvtot > P.duv_cut || i == n - 1) {
        close_block(t0, i + 1, chan_cuts_mode1(nchb_max, P.nchan));
        nchb_max = 1e3;
        duvtot = 0.0;
        t0 = i + 1;
      }
    }
  } else {
    // phase-based blocks: blk_i = int32(cumsum(dph)_i / dPhi), cut where it changes
    double csum = 0.0;
    int64_t t0 = 0;
    int32_t prev = 0;
    double span_min = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      const double* x = uvw + 3 * size_t(r[i]);
      double d0 = 0.0, d1 = 0.0, d2 = 0.0;
      if (n > 1) {
        const int64_t j = (i + 1 < n) ? i : n - 2;
        const double* a = uvw + 3 * size_t(r[j]);
        const double* b = uvw + 3 * size_t(r[j + 1]);
        d0 = a[0] - b[0];
        d1 = a[1] - b[1];
        d2 = a[2] - b[2];
      }
      const double dph = P.k_phase * (sqrt(d0 * d0 + d1 * d1) * P.lmax + fabs(d2) * P.nmax);
      const double ext = sqrt(x[0] * x[0] + x[1] * x[1]) * P.lmax + fabs(x[2]) * P.nmax;
      const double span = (P.k_span / ext) / P.dfreq;
      csum += dph;
      const int32_t blk = int32_t(csum / P.dphi);
      const bool cut = P.per_row ? (i > 0) : (i > 0 && blk != prev);
      if (cut) {
        ChanCuts cc;
        double m = span_min < 1.0 ? 1.0 : span_min;
        double nb = isfinite(m) ? ceil(double(P.nchan) / m) : 1.0;
        cc.nblk = int32_t(ceil(double(P.nchan) / (nb < 1.0 ? 1.0 : nb)));
        cc.lin_step = 0.0;
        close_block(t0, i, cc);
        t0 = i;
        span_min = INFINITY;
      }
      prev = blk;
      if (span < span_min || isnan(span_min)) span_min = span;
    }
    if (n > 0) {
      ChanCuts cc;
      double m = span_min < 1.0 ? 1.0 : span_min;
      double nb = isfinite(m) ? ceil(double(P.nchan) / m) : 1.0;
      cc.nblk = int32_t(ceil(double(P.nchan) / (nb < 1.0 ? 1.0 : nb)));
      cc.lin_step = 0.0;
      close_block(t0, n, cc);
    }
  }
  *nblk_out = nblk;
  *nwords_out = nwords;
}

template <int MODE>
__global__ void bda_count_k(BdaParams P, const double* uvw, const uint32_t* rows, const int64_t* seg,
                            const int32_t* chan_map, int64_t* nblk, int64_t* nwords) {
  const int k = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= P.nbl) return;
  const int64_t s0 = seg[2 * k], s1 = seg[2 * k + 1];
  int64_t b = 0, w = 0;
  if (s1 > s0) walk_baseline<MODE, false>(P, uvw, rows + s0, s0, s1 - s0, chan_map, &b, &w, nullptr, 0, 0);
  nblk[k] = b;
  nwords[k] = w;
}

template <int MODE>
__global__ void bda_walk_k(BdaParams P, const double* uvw, const uint32_t* rows, const int64_t* seg,
                           const int32_t* chan_map, const int64_t* blk_off, const int64_t* word_off, BlockDesc* desc) {
  const int k = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= P.nbl) return;
  const int64_t s0 = seg[2 * k], s1 = seg[2 * k + 1];
  int64_t b = 0, w = 0;
  if (s1 > s0)
    walk_baseline<MODE, true>(P, uvw, rows + s0, s0, s1 - s0, chan_map, &b, &w, desc, blk_off[k], word_off[k]);
}

__global__ void bda_emit_k(const BlockDesc* desc, int64_t nblk, const uint32_t* rows, int32_t* out) {
  int32_t* sizes = out + 2;
  int32_t* recs = out + 2 + nblk;
  for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < nblk; b += int64_t(gridDim.x) * blockDim.x) {
    const BlockDesc d = desc[b];
    sizes[b] = 2 + d.nrows;
    int32_t* rec = recs + d.word;
    rec[0] = d.ch0;
    rec[1] = d.ch1;
    for (int32_t j = 0; j < d.nrows; ++j) rec[2 + j] = int32_t(rows[d.row0 + j]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[0] = int32_t(uint32_t(uint64_t(nblk) & 0xFFFFFFFFu));
    out[1] = int32_t(nblk >> 32);
  }
}

unsigned grid_for(int64_t n, int per = 256) {
  int64_t b = (n + per - 1) / per;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return unsigned(b);
}

}  // namespace

struct ddf_bda_builder {
  BdaParams P{};
  hipStream_t st = nullptr;
  int64_t nrow = 0, nblk = 0, nwords = 0;
  const double* uvw = nullptr;
  const int32_t* chan_map = nullptr;
  std::vector<void*> allocs;
  uint32_t* rows = nullptr;   // sorted rows
  int64_t* seg = nullptr;
  int64_t* blk_off = nullptr;
  int64_t* word_off = nullptr;
  ~ddf_bda_builder() {
    for (void* p : allocs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
};

#define BDA_TRY(x)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) {                                                                     \
      delete B;                                                                                 \
      return bfail(DDF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));               \
    }                                                                                           \
  } while (0)

extern "C" int ddf_bda_build_create(const ddf_bda_spec* spec, const double* uvw, const int32_t* a0, const int32_t* a1,
                                    int64_t nrow, const int32_t* chan_map, const int32_t* rank, ddf_bda_builder** out,
                                    int64_t* nwords, void* stream) {
  if (!spec || !out || !nwords || (nrow > 0 && (!uvw || !a0 || !a1)) || !chan_map || !rank)
    return bfail(DDF_EINVAL, "null argument");
  if (spec->mode != 1 && spec->mode != 2) return bfail(DDF_EINVAL, "BDA mode must be 1 or 2");
  if (spec->nchan <= 0 || spec->na <= 0 || nrow < 0 || nrow >= (int64_t(1) << 32))
    return bfail(DDF_EINVAL, "bad sizes");
  ddf_bda_builder* B = new ddf_bda_builder();
  hipStream_t st = static_cast<hipStream_t>(stream);
  B->st = st;
  B->nrow = nrow;
  B->uvw = uvw;
  B->chan_map = chan_map;
  const int na = spec->na;
  const int nbl = na * (na - 1);
  BdaParams& P = B->P;
  P.mode = spec->mode;
  P.nchan = spec->nchan;
  P.nbl = nbl;
  P.duv_cut = spec->duv_cut;
  P.k_dnu = spec->k_dnu;
  P.l = spec->l;
  P.dfreq = spec->dfreq;
  P.k_phase = spec->k_phase;
  P.lmax = spec->lmax;
  P.nmax = spec->nmax;
  P.k_span = spec->k_span;
  P.dphi = spec->dphi;
  P.per_row = spec->per_row;
  uint32_t* keys = B->alloc<uint32_t>(size_t(nrow));
  uint32_t* keys2 = B->alloc<uint32_t>(size_t(nrow));
  uint32_t* rows = B->alloc<uint32_t>(size_t(nrow));
  uint32_t* rows2 = B->alloc<uint32_t>(size_t(nrow));
  int32_t* d_rank = B->alloc<int32_t>(size_t(na) * na);
  B->seg = B->alloc<int64_t>(2 * size_t(nbl));
  int64_t* cnt = B->alloc<int64_t>(2 * size_t(nbl) + 2);
  B->blk_off = B->alloc<int64_t>(size_t(nbl) + 1);
  B->word_off = B->alloc<int64_t>(size_t(nbl) + 1);
  if (!keys || !keys2 || !rows || !rows2 || !d_rank || !B->seg || !cnt || !B->blk_off || !B->word_off) {
    delete B;
    return bfail(DDF_ENOMEM, "BDA builder allocation failed");
  }
  BDA_TRY(hipMemcpyAsync(d_rank, rank, size_t(na) * na * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (nrow > 0) {
    hipLaunchKernelGGL(bda_keys_k, dim3(grid_for(nrow)), dim3(256), 0, st, a0, a1, nrow, d_rank, na, nbl, keys, rows);
    BDA_TRY(hipGetLastError());
    int end_bit = 1;
    while ((int64_t(1) << end_bit) <= nbl) ++end_bit;
    hipcub::DoubleBuffer<uint32_t> dk(keys, keys2), dv(rows, rows2);
    size_t tmp_bytes = 0;
    BDA_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, dk, dv, int(nrow), 0, end_bit, st));
    void* tmp = B->alloc<char>(tmp_bytes);
    if (!tmp) {
      delete B;
      return bfail(DDF_ENOMEM, "BDA sort workspace allocation failed");
    }
    BDA_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, dk, dv, int(nrow), 0, end_bit, st));
    keys = dk.Current();
    B->rows = dv.Current();
  } else {
    B->rows = rows;
  }
  hipLaunchKernelGGL(bda_fill64_k, dim3(grid_for(2 * int64_t(nbl))), dim3(256), 0, st, B->seg, 2 * int64_t(nbl),
                     int64_t(0));
  if (nrow > 0)
    hipLaunchKernelGGL(bda_segments_k, dim3(grid_for(nrow)), dim3(256), 0, st, keys, nrow, nbl, B->seg);
  int64_t* nb = cnt;
  int64_t* nw = cnt + nbl + 1;
  if (P.mode == 1)
    hipLaunchKernelGGL(bda_count_k<1>, dim3(grid_for(nbl, 64)), dim3(64), 0, st, P, uvw, B->rows, B->seg, chan_map, nb, nw);
  else
    hipLaunchKernelGGL(bda_count_k<2>, dim3(grid_for(nbl, 64)), dim3(64), 0, st, P, uvw, B->rows, B->seg, chan_map, nb, nw);
  BDA_TRY(hipGetLastError());
  hipLaunchKernelGGL(bda_fill64_k, dim3(1), dim3(64), 0, st, nb + nbl, int64_t(1), int64_t(0));
  hipLaunchKernelGGL(bda_fill64_k, dim3(1), dim3(64), 0, st, nw + nbl, int64_t(1), int64_t(0));
  size_t scan_bytes = 0;
  BDA_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, nb, B->blk_off, nbl + 1, st));
  void* stmp = B->alloc<char>(scan_bytes);
  if (!stmp) {
    delete B;
    return bfail(DDF_ENOMEM, "BDA scan workspace allocation failed");
  }
  BDA_TRY(hipcub::DeviceScan::ExclusiveSum(stmp, scan_bytes, nb, B->blk_off, nbl + 1, st));
  BDA_TRY(hipcub::DeviceScan::ExclusiveSum(stmp, scan_bytes, nw, B->word_off, nbl + 1, st));
  int64_t tot[2] = {0, 0};
  BDA_TRY(hipMemcpyAsync(&tot[0], B->blk_off + nbl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  BDA_TRY(hipMemcpyAsync(&tot[1], B->word_off + nbl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  BDA_TRY(hipStreamSynchronize(st));
  B->nblk = tot[0];
  B->nwords = 2 + tot[0] + tot[1];
  *nwords = B->nwords;
  *out = B;
  return DDF_OK;
}

extern "C" int ddf_bda_build_write(ddf_bda_builder* B, int32_t* out, int64_t n, void* stream) {
  if (!B || !out) return bfail(DDF_EINVAL, "null argument");
  if (n < B->nwords) return bfail(DDF_EINVAL, "output shorter than the map");
  hipStream_t st = static_cast<hipStream_t>(stream);
  BlockDesc* desc = nullptr;
  hipError_t ea = hipMallocAsync(reinterpret_cast<void**>(&desc), size_t(B->nblk ? B->nblk : 1) * sizeof(BlockDesc), st);
  if (ea != hipSuccess) return bfail(DDF_EHIP, std::string("BDA descriptors: ") + hipGetErrorString(ea));
  const BdaParams& P = B->P;
  if (P.mode == 1)
    hipLaunchKernelGGL(bda_walk_k<1>, dim3(grid_for(P.nbl, 64)), dim3(64), 0, st, P, B->uvw, B->rows, B->seg,
                       B->chan_map, B->blk_off, B->word_off, desc);
  else
    hipLaunchKernelGGL(bda_walk_k<2>, dim3(grid_for(P.nbl, 64)), dim3(64), 0, st, P, B->uvw, B->rows, B->seg,
                       B->chan_map, B->blk_off, B->word_off, desc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(bda_emit_k, dim3(grid_for(B->nblk)), dim3(256), 0, st, desc, B->nblk, B->rows, out);
    e = hipGetLastError();
  }
  (void)hipFreeAsync(desc, st);
  if (e != hipSuccess) return bfail(DDF_EHIP, std::string("BDA map kernels: ") + hipGetErrorString(e));
  return DDF_OK;
}

extern "C" int ddf_bda_build_destroy(ddf_bda_builder* B) {
  delete B;
  return DDF_OK;
}
