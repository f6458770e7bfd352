// gfx950 facet gridder / degridder: plain single-facet kernels and the C ABI of include/ddfgpu.h.
//
// Work decomposition (one facet per launch, the "plain" path):
//   * one 64-lane wavefront per BDA block (grid-stride over blocks);
//   * the block's (row, channel) visibilities are spread over the lanes, channel fastest, so a
//     wavefront reads a row's consecutive channels as one coalesced segment;
//   * block sums (visibility, uvw means, weights) are reduced across the wavefront by butterfly
//     shuffles; the w-plane / cell / oversampling decision is then lane-uniform;
//   * gridding scatters the S x S x npol stamp with one lane per cell through global float
//     atomics; degridding gathers the same stamp, reduces it, and writes every (row, channel)
//     of the block back once (blocks of a degrid map are disjoint).
//   * sumwt / SumJones / SumJonesChan go through per-workgroup LDS accumulators flushed once per
//     workgroup with global FP64 atomics.
// Follows gridder::gridder (DDFacet/Gridder/gridder.h:78-377) and degridder::degridder
// (degridder.h:55-266); argument dispatch follows GridderSmearPols.cc:76-348.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ddfgpu.h"
#include "ddf_device.h"
#include "ddf_internal.h"

using namespace ddf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(DDF_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- correlation <-> Stokes maps (Stokes.h), as 4x4 complex matrices -----------------------
struct StokesMap {
  int read;        // 4, 2, 1: Read_4Corr / Read_2Corr_Pad / Read_1Corr_Pad
  int nstokes;
  double m[4][4][2];
};

bool same(const uint16_t* c, int n, std::initializer_list<int> w) {
  if (n != int(w.size())) return false;
  int i = 0;
  for (int x : w)
    if (c[i++] != x) return false;
  return true;
}

std::string stokes_string(const ddf_vis* v) {
  std::string s;
  for (int i = 0; i < v->nstokes; ++i) s += (v->stokes_codes[i] >= 1 && v->stokes_codes[i] <= 4) ? "?IQUV"[v->stokes_codes[i]] : '!';
  return s;
}

void set(double (&m)[4][4][2], int r, int c, double re, double im) {
  m[r][c][0] = re;
  m[r][c][1] = im;
}

// Grid side, GridderSmearPols.cc:128-206 with the Stokes.h:18-34 formulas.
int grid_map(const ddf_vis* v, StokesMap* sm) {
  const uint16_t* c = v->corr_codes;
  int nc = v->ncorr_codes;
  for (int i = 0; i < nc; ++i)
    if (!((c[i] >= 5 && c[i] <= 12) || c[i] == 1))
      return fail(DDF_EINVAL, "Only accepts I,RR,RL,LR,LL,XX,XY,YX,YY as correlation input type");
  for (int i = 0; i < v->nstokes; ++i)
    if (v->stokes_codes[i] < 1 || v->stokes_codes[i] > 4)
      return fail(DDF_EINVAL, "Only accepts I,Q,U,V as polarization output type");
  std::memset(sm, 0, sizeof(*sm));
  std::string want = stokes_string(v), eff = want;
  int kind;  // 0 circular 4-corr, 1 linear 4-corr, 2 linear 2-corr, 3 circular 2-corr, 4 Stokes I
  std::vector<std::string> ok;
  if (same(c, nc, {9, 10, 11, 12})) { kind = 1; sm->read = 4; ok = {"I", "Q", "U", "V", "IQ", "IU", "IV", "QU", "IQU", "IQUV"}; }
  else if (same(c, nc, {5, 6, 7, 8})) { kind = 0; sm->read = 4; ok = {"I", "IQ", "IU", "IV", "QU", "IQU", "IQUV"}; }
  else if (same(c, nc, {9, 12})) { kind = 2; sm->read = 2; ok = {"I", "IQ"}; }
  else if (same(c, nc, {5, 8})) { kind = 3; sm->read = 2; ok = {"I", "IV"}; }
  else if (same(c, nc, {1})) { kind = 4; sm->read = 1; ok = {"I"}; }
  else return fail(DDF_EINVAL, "Cannot convert input correlations to desired output Stokes parameters.");
  bool found = false;
  for (auto& o : ok) found |= (o == want);
  if (!found) return fail(DDF_EINVAL, "Cannot convert input correlations to desired output Stokes parameters.");
  if (kind == 1 && want == "IU") eff = "IQ";  // the {I,U} request runs the IQ policy (GridderSmearPols.cc:142-143)
  sm->nstokes = v->nstokes;
  for (int k = 0; k < sm->nstokes; ++k) {
    char s = eff[k];
    auto& m = sm->m;
    if (kind == 4) { set(m, k, 0, 1, 0); continue; }
    if (s == 'I') { set(m, k, 0, .5, 0); set(m, k, 3, .5, 0); continue; }
    bool lin = (kind == 1 || kind == 2);
    if ((lin && s == 'Q') || (!lin && s == 'V')) { set(m, k, 0, .5, 0); set(m, k, 3, -.5, 0); }
    else if ((lin && s == 'U') || (!lin && s == 'Q')) { set(m, k, 1, .5, 0); set(m, k, 2, .5, 0); }
    else { set(m, k, 1, 0, -.5); set(m, k, 2, 0, .5); }  // V (linear) / U (circular): -(c1-c2)*0.5i
  }
  return DDF_OK;
}

// Degrid side, GridderSmearPols.cc:261-342 with the Stokes.h:90-192 formulas.
int degrid_map(const ddf_vis* v, StokesMap* sm) {
  const uint16_t* c = v->corr_codes;
  int nc = v->ncorr_codes;
  for (int i = 0; i < nc; ++i)
    if (!((c[i] >= 5 && c[i] <= 12) || c[i] == 1))
      return fail(DDF_EINVAL, "Only accepts I,RR,RL,LR,LL,XX,XY,YX,YY as correlation output types");
  std::memset(sm, 0, sizeof(*sm));
  std::string w = stokes_string(v);
  bool X4 = same(c, nc, {9, 10, 11, 12}), R4 = same(c, nc, {5, 6, 7, 8});
  bool X2 = same(c, nc, {9, 12}), R2 = same(c, nc, {5, 8}), S1 = same(c, nc, {1});
  auto& m = sm->m;
  // rows: correlation slot (2x2 flattened), columns: Stokes index
  if (w == "I" && (X4 || X2 || R4 || R2 || S1)) { set(m, 0, 0, 1, 0); set(m, 3, 0, 1, 0); }
  else if (w == "Q" && X4) { set(m, 0, 0, 1, 0); set(m, 3, 0, -1, 0); }
  else if (w == "U" && X4) { set(m, 1, 0, 1, 0); set(m, 2, 0, 1, 0); }
  else if (w == "V" && X4) { set(m, 1, 0, 0, 1); set(m, 2, 0, 0, -1); }
  else if (w == "IQUV" && X4) { set(m, 0, 0, 1, 0); set(m, 0, 1, 1, 0); set(m, 1, 2, 1, 0); set(m, 1, 3, 0, 1);
                                set(m, 2, 2, 1, 0); set(m, 2, 3, 0, -1); set(m, 3, 0, 1, 0); set(m, 3, 1, -1, 0); }
  else if (w == "IQUV" && R4) { set(m, 0, 0, 1, 0); set(m, 0, 3, 1, 0); set(m, 1, 1, 1, 0); set(m, 1, 2, 0, 1);
                                set(m, 2, 1, 1, 0); set(m, 2, 2, 0, -1); set(m, 3, 0, 1, 0); set(m, 3, 3, -1, 0); }
  else if (w == "IQ" && (X4 || X2)) { set(m, 0, 0, 1, 0); set(m, 0, 1, 1, 0); set(m, 3, 0, 1, 0); set(m, 3, 1, -1, 0); }
  else if (w == "IQ" && R4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 1, 0); set(m, 2, 1, 1, 0); set(m, 3, 0, 1, 0); }
  else if (w == "IU" && X4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 1, 0); set(m, 2, 1, 1, 0); set(m, 3, 0, 1, 0); }
  else if (w == "IU" && R4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 0, 1); set(m, 2, 1, 0, -1); set(m, 3, 0, 1, 0); }
  else if (w == "IV" && (R4 || R2)) { set(m, 0, 0, 1, 0); set(m, 0, 1, 1, 0); set(m, 3, 0, 1, 0); set(m, 3, 1, -1, 0); }
  else if (w == "IV" && X4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 0, 1); set(m, 2, 1, 0, -1); set(m, 3, 0, 1, 0); }
  else if (w == "QU" && X4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 1, 0); set(m, 2, 1, 1, 0); set(m, 3, 0, -1, 0); }
  else if (w == "IQU" && X4) { set(m, 0, 0, 1, 0); set(m, 0, 1, 1, 0); set(m, 1, 2, 1, 0); set(m, 2, 2, 1, 0);
                               set(m, 3, 0, 1, 0); set(m, 3, 1, -1, 0); }
  else if (w == "IQU" && R4) { set(m, 0, 0, 1, 0); set(m, 1, 1, 1, 0); set(m, 1, 2, 0, 1); set(m, 2, 1, 1, 0);
                               set(m, 2, 2, 0, -1); set(m, 3, 0, 1, 0); }
  else return fail(DDF_EINVAL, "Cannot convert input Stokes parameter to desired output correlations.");
  sm->read = (X4 || R4) ? 4 : ((X2 || R2) ? 2 : 1);
  sm->nstokes = v->nstokes;
  return DDF_OK;
}

// ---- kernel parameters -----------------------------------------------------------------------
struct KParams {
  VisP v;
  void* grid;
  int nband, npol, ny, nx;
  double incr0, incr1;
  int dopsf;
  double cu, cv, l0, m0, n0;
  double fmean0, wl_mean;
  double tscale0;  // decorr_tscale(fmean0)
  CFP cf;
  JonesP jp;
  DecorrP dp;
  double* sumwt;
  int nstokes;
  double sm[4][4][2];
  int chan_lds;   // per-channel Jones sums accumulated in LDS (else straight to global)
};

__device__ __forceinline__ cd smc(const KParams& p, int r, int c) { return cd{p.sm[r][c][0], p.sm[r][c][1]}; }

// ---- gridding kernel -----------------------------------------------------------------------
// Block means over the unflagged (row, channel) visibilities (gridder.h:272-289): lanes take the
// visibilities channel-fastest and the partial sums are reduced across the wavefront.
struct Means {
  double su, sv, sw, sf;
  int cnt;
};

__device__ inline Means block_means(const KParams& p, const BlockRec& br, int lane) {
  const VisP& v = p.v;
  const int nch = br.ch1 - br.ch0;
  const int64_t nvis = int64_t(br.nr) * nch;
  Means m{0, 0, 0, 0, 0};
  for (int64_t k = lane; k < nvis; k += kWave) {
    const int r = int(k / nch);
    const int c = br.ch0 + int(k - int64_t(r) * nch);
    const int64_t row = br.rows[r];
    if (row < 0 || row >= v.nrow) continue;
    if (v.flags[(size_t(row) * v.nchan + c) * v.ncorr]) continue;
    const double* uvw = v.uvw + 3 * row;
    m.su += shifted(uvw[0], uvw[2], p.cu);
    m.sv += shifted(uvw[1], uvw[2], p.cv);
    m.sw += uvw[2];
    m.sf += v.freqs[c];
    ++m.cnt;
  }
  m.cnt = wave_sum(m.cnt);
  m.su = wave_sum(m.su);
  m.sv = wave_sum(m.sv);
  m.sw = wave_sum(m.sw);
  m.sf = wave_sum(m.sf);
  return m;
}

template <int READ, int JMODE, typename GT>
__global__ __launch_bounds__(256) void grid_plain(KParams p) {
  extern __shared__ double lds[];
  const int nbp = p.nband * p.npol;
  double* s_sumwt = lds;                       // [nband*npol]
  double* s_sj = lds + nbp;                    // [2*nband]
  double* s_sjc = lds + nbp + 2 * p.nband;     // [2*nchan] when chan_lds
  const int nacc = nbp + 2 * p.nband + (p.chan_lds ? 2 * p.v.nchan : 0);
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) lds[i] = 0.;
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const VisP& v = p.v;
  GT* grid = reinterpret_cast<GT*>(p.grid);
  const int npg = p.nstokes < p.npol ? p.nstokes : p.npol;
  double* sjc = p.chan_lds ? s_sjc : p.jp.sum_jones_chan;

  for (int64_t b = wave; b < v.nblocks; b += nwaves) {
    if (v.sparse && !v.sparse[b]) continue;
    BlockRec br = read_block(v, b);
    const int nch = br.ch1 - br.ch0;
    const int64_t nvis = int64_t(br.nr) * nch;
    if (nvis <= 0) continue;
    const int64_t mid = br.rows[br.nr / 2];
    const bool mid_ok = mid >= 0 && mid < v.nrow;
    const double dec = mid_ok ? decorr_factor(p.dp, v.uvw, p.tscale0, mid) : 1.;
    const int gchan = v.chan_map[br.ch0];

    // With Jones the per-channel sums are only added for gridded blocks, so the placement is
    // decided first from a means-only pass (the same partial sums the main pass forms).
    Place pl;
    bool decided = false;
    double BBblk = 0;
    JonesPair jblk;
    if (JMODE != 0) {
      Means m = block_means(p, br, lane);
      if (m.cnt == 0) continue;
      const double um = m.su / m.cnt, vm = m.sv / m.cnt, wm = m.sw / m.cnt, fm = m.sf / m.cnt;
      if (!(gchan >= 0 && gchan < p.nband &&
            place_block(p.cf, p.nx, p.ny, p.incr0, p.incr1, um, vm, wm, fm / kC, &pl)))
        continue;
      decided = true;
      if (JMODE == 2) {
        double uvwm[3] = {um, vm, wm};
        jblk = jones_at(p.jp, mid_ok ? mid : 0, (br.ch0 + br.ch1) / 2, uvwm, p.wl_mean);
        BBblk = jblk.BB;
      }
    }

    double su = 0, sv = 0, sw = 0, sf = 0, swt = 0, ssw = 0, sj = 0;
    int cnt = 0;
    cd acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int64_t k = lane; k < nvis; k += kWave) {
      const int r = int(k / nch);
      const int c = br.ch0 + int(k - int64_t(r) * nch);
      const int64_t row = br.rows[r];
      if (row < 0 || row >= v.nrow) continue;
      const size_t vc = size_t(row) * v.nchan + c;
      if (v.flags[vc * v.ncorr]) continue;
      const double* uvw = v.uvw + 3 * row;
      const double U = uvw[0], V = uvw[1], W = uvw[2];
      const double f = v.freqs[c];
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
      if (!p.dopsf) {
        const double angle = -2. * kPi * (U * p.l0 + V * p.m0 + W * p.n0) / kC;
        wt = phasor(f * angle) * fw;
      }
      if (JMODE == 1) {
        JonesPair jj = jones_at(p.jp, row, c, uvw, p.wl_mean);
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
        atomicAdd(&sjc[c], double(npg) * (jj.BB * fwd));
      }
      if (JMODE == 2) atomicAdd(&sjc[c], double(npg) * (fwd * BBblk));
      if (JMODE != 0) atomicAdd(&sjc[v.nchan + c], double(npg) * fwd);
      if (READ == 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = acc[q] + meas[q] * wt;
      } else if (READ == 2) {
        acc[0] = acc[0] + meas[0] * wt;
        acc[3] = acc[3] + meas[3] * wt;
      } else {
        acc[0] = acc[0] + meas[0] * wt;
      }
      if (JMODE == 0) {
        su += shifted(U, W, p.cu);
        sv += shifted(V, W, p.cv);
        sw += W;
        sf += f;
        ++cnt;
      }
      swt += fw;
      ssw += fwd;
    }
    if (JMODE == 0) {
      cnt = wave_sum(cnt);
      if (cnt == 0) continue;
      su = wave_sum(su); sv = wave_sum(sv); sw = wave_sum(sw); sf = wave_sum(sf);
      decided = gchan >= 0 && gchan < p.nband &&
                place_block(p.cf, p.nx, p.ny, p.incr0, p.incr1, su / cnt, sv / cnt, sw / cnt, (sf / cnt) / kC, &pl);
      if (!decided) continue;
    }
    swt = wave_sum(swt);
    ssw = wave_sum(ssw);
    if (JMODE == 1) sj = wave_sum(sj);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (READ == 4 || q == 0 || (READ == 2 && q == 3)) {
        acc[q].re = wave_sum(acc[q].re);
        acc[q].im = wave_sum(acc[q].im);
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
    // Stokes conversion
    cd st[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      st[k] = cd{0, 0};
      if (k < p.nstokes)
#pragma unroll
        for (int q = 0; q < 4; ++q) st[k] = st[k] + smc(p, k, q) * acc[q];
    }
    const int S = pl.S, S2 = S * S;
    const float2* cfp = p.cf.data + pl.cfoff;
    for (int t = lane; t < S2 * npg; t += kWave) {
      const int ip = t / S2;
      const int k = t - ip * S2;
      const int sy = k / S, sx = k - sy * S;
      const float2 w = cfp[k];
      cd val = ip == 0 ? st[0] : (ip == 1 ? st[1] : (ip == 2 ? st[2] : st[3]));
      cd prod = val * cd{double(w.x), double(w.y)};
      size_t gi = ((size_t(gchan) * p.npol + ip) * p.ny + size_t(pl.locy - pl.sup + sy)) * p.nx +
                  size_t(pl.locx - pl.sup + sx);
      atomicAdd(&grid[2 * gi], GT(prod.re));
      atomicAdd(&grid[2 * gi + 1], GT(prod.im));
    }
    if (lane == 0) {
      for (int ip = 0; ip < npg; ++ip) {
        atomicAdd(&s_sumwt[gchan * p.npol + ip], swt);
        if (JMODE != 0) {
          atomicAdd(&s_sj[gchan], sj);
          atomicAdd(&s_sj[gchan + p.nband], ssw);
        }
      }
    }
    (void)decided;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbp; i += blockDim.x)
    if (s_sumwt[i] != 0.) atomicAdd(&p.sumwt[i], s_sumwt[i]);
  if (JMODE != 0) {
    for (int i = threadIdx.x; i < 2 * p.nband; i += blockDim.x)
      if (s_sj[i] != 0.) atomicAdd(&p.jp.sum_jones[i], s_sj[i]);
    if (p.chan_lds)
      for (int i = threadIdx.x; i < 2 * p.v.nchan; i += blockDim.x)
        if (s_sjc[i] != 0.) atomicAdd(&p.jp.sum_jones_chan[i], s_sjc[i]);
  }
}

// ---- degridding kernel ---------------------------------------------------------------------
template <int NCORR, int JMODE>
__global__ __launch_bounds__(256) void degrid_plain(KParams p) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  const VisP& v = p.v;
  const float2* grid = reinterpret_cast<const float2*>(p.grid);

  for (int64_t b = wave; b < v.nblocks; b += nwaves) {
    BlockRec br = read_block(v, b);
    const int nch = br.ch1 - br.ch0;
    if (br.nr <= 0 || nch <= 0) continue;
    const int gchan = v.chan_map[br.ch0];
    if (gchan < 0 || gchan >= p.nband) continue;
    double fm = 0;
    for (int c = br.ch0; c < br.ch1; ++c) fm += v.freqs[c];
    fm /= double(nch);
    double su = 0, sv = 0, sw = 0;
    int cnt = 0;
    for (int r = lane; r < br.nr; r += kWave) {
      const int64_t row = br.rows[r];
      if (row < 0 || row >= v.nrow) continue;
      const double* uvw = v.uvw + 3 * row;
      su += shifted(uvw[0], uvw[2], p.cu);
      sv += shifted(uvw[1], uvw[2], p.cv);
      sw += uvw[2];
      ++cnt;
    }
    cnt = wave_sum(cnt);
    if (cnt == 0) continue;
    su = wave_sum(su); sv = wave_sum(sv); sw = wave_sum(sw);
    Place pl;
    if (!place_block(p.cf, p.nx, p.ny, p.incr0, p.incr1, su / cnt, sv / cnt, sw / cnt, fm / kC, &pl)) continue;
    // gather the S x S stamp of every Stokes plane: FP32 products summed in FP64 (degridder.h:202-212)
    const int S = pl.S, S2 = S * S;
    const float2* cfp = p.cf.data + pl.cfoff;
    cd st[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int t = lane; t < S2 * p.nstokes; t += kWave) {
      const int ip = t / S2;
      const int k = t - ip * S2;
      const int sy = k / S, sx = k - sy * S;
      const size_t gi = ((size_t(gchan) * p.npol + ip) * p.ny + size_t(pl.locy - pl.sup + sy)) * p.nx +
                        size_t(pl.locx - pl.sup + sx);
      const float2 g = grid[gi];
      const float2 w = cfp[k];
      const float pr = g.x * w.x - g.y * w.y;
      const float pi = g.x * w.y + g.y * w.x;
      if (ip == 0) st[0] = st[0] + cd{pr, pi};
      else if (ip == 1) st[1] = st[1] + cd{pr, pi};
      else if (ip == 2) st[2] = st[2] + cd{pr, pi};
      else st[3] = st[3] + cd{pr, pi};
    }
    for (int k = 0; k < p.nstokes; ++k) {
      st[k].re = wave_sum(st[k].re);
      st[k].im = wave_sum(st[k].im);
    }
    M2 cv;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      cv.a[k] = cd{0, 0};
#pragma unroll
      for (int q = 0; q < 4; ++q) cv.a[k] = cv.a[k] + smc(p, k, q) * st[q];
    }
    const int64_t mid = br.rows[br.nr / 2];
    const bool mid_ok = mid >= 0 && mid < v.nrow;
    if (JMODE == 2 && mid_ok) {
      JonesPair jj = jones_at(p.jp, mid, (br.ch0 + br.ch1) / 2, v.uvw + 3 * mid, p.wl_mean);
      cv = mul(mul(jj.J0, cv), herm(jj.J1));
    }
    const double dec = mid_ok ? decorr_factor(p.dp, v.uvw, decorr_tscale(fm), mid) : 1.;
    const int64_t nvis = int64_t(br.nr) * nch;
    for (int64_t k = lane; k < nvis; k += kWave) {
      const int r = int(k / nch);
      const int c = br.ch0 + int(k - int64_t(r) * nch);
      const int64_t row = br.rows[r];
      if (row < 0 || row >= v.nrow) continue;
      const double* uvw = v.uvw + 3 * row;
      const double angle = 2. * kPi * (uvw[0] * p.l0 + uvw[1] * p.m0 + uvw[2] * p.n0) / kC;
      const cd corr = phasor(v.freqs[c] * angle) * dec;
      M2 vb;
      if (JMODE == 1) {
        JonesPair jj = jones_at(p.jp, row, c, uvw, p.wl_mean);
        vb = mul(mul(jj.J0, cv), herm(jj.J1));
#pragma unroll
        for (int q = 0; q < 4; ++q) vb.a[q] = vb.a[q] * corr;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) vb.a[q] = cv.a[q] * corr;
      }
      float2* vp = v.vis + (size_t(row) * v.nchan + c) * NCORR;
      if (NCORR == 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float2 x = vp[q];
          vp[q] = make_float2(float(double(x.x) - vb.a[q].re), float(double(x.y) - vb.a[q].im));
        }
      } else if (NCORR == 2) {
        float2 x0 = vp[0], x1 = vp[1];
        vp[0] = make_float2(float(double(x0.x) - vb.a[0].re), float(double(x0.y) - vb.a[0].im));
        vp[1] = make_float2(float(double(x1.x) - vb.a[3].re), float(double(x1.y) - vb.a[3].im));
      } else {
        float2 x0 = vp[0];
        vp[0] = make_float2(float(double(x0.x) - vb.a[0].re), float(double(x0.y) - vb.a[0].im));
      }
    }
  }
}

// ---- BDA block index ---------------------------------------------------------------------------
// Word offset of every block record: start[b] = 2 + nblocks + sum_{b' < b} size[b'].
// Three phases: per-chunk sums, an exclusive scan of the chunk sums (one workgroup), and a
// per-chunk scan plus chunk offset.
constexpr int kScanChunk = 2048;

__global__ __launch_bounds__(256) void scan_chunk_sums(const int32_t* sizes, int64_t n, int64_t* chunk_sum) {
  __shared__ int64_t red[256];
  int64_t base = int64_t(blockIdx.x) * kScanChunk;
  int64_t s = 0;
  for (int i = threadIdx.x; i < kScanChunk; i += 256)
    if (base + i < n) s += sizes[base + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) chunk_sum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void scan_chunk_offsets(int64_t* chunk_sum, int64_t nchunk, int64_t base) {
  // exclusive scan of chunk_sum in place, one workgroup, sequential over tiles of 256
  __shared__ int64_t tile[256];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = base;
  __syncthreads();
  for (int64_t t0 = 0; t0 < nchunk; t0 += 256) {
    int64_t i = t0 + threadIdx.x;
    int64_t x = i < nchunk ? chunk_sum[i] : 0;
    tile[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      int64_t y = threadIdx.x >= o ? tile[threadIdx.x - o] : 0;
      __syncthreads();
      tile[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < nchunk) chunk_sum[i] = carry + tile[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 255) carry += tile[255];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void scan_chunk_final(const int32_t* sizes, int64_t n, const int64_t* chunk_off,
                                                        int64_t* out) {
  __shared__ int64_t tile[256];
  __shared__ int64_t carry;
  int64_t base = int64_t(blockIdx.x) * kScanChunk;
  if (threadIdx.x == 0) carry = chunk_off[blockIdx.x];
  __syncthreads();
  for (int t0 = 0; t0 < kScanChunk; t0 += 256) {
    int64_t i = base + t0 + threadIdx.x;
    int64_t x = i < n ? sizes[i] : 0;
    tile[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      int64_t y = threadIdx.x >= o ? tile[threadIdx.x - o] : 0;
      __syncthreads();
      tile[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < n) out[i] = carry + tile[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 255) carry += tile[255];
    __syncthreads();
  }
}

// out[i] = base + sum_{j < i} x[j] (int32 in, int64 out); tmp holds ceil(n / kScanChunk) words.
int launch_scan(const int32_t* x, int64_t n, int64_t base, int64_t* out, int64_t* chunk_tmp, hipStream_t st) {
  if (n <= 0) return DDF_OK;
  int64_t nchunk = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(scan_chunk_sums, dim3(unsigned(nchunk)), dim3(256), 0, st, x, n, chunk_tmp);
  hipLaunchKernelGGL(scan_chunk_offsets, dim3(1), dim3(256), 0, st, chunk_tmp, nchunk, base);
  hipLaunchKernelGGL(scan_chunk_final, dim3(unsigned(nchunk)), dim3(256), 0, st, x, n, chunk_tmp, out);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

int launch_bda_index(const int32_t* bda, int64_t nblocks, int64_t* out, int64_t* chunk_tmp, hipStream_t st) {
  return launch_scan(bda + 2, nblocks, 2 + nblocks, out, chunk_tmp, st);
}

// ---- host-side assembly ---------------------------------------------------------------------
int check_common(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f) {
  if (!g || !v || !f) return fail(DDF_EINVAL, "null argument struct");
  if (g->nband <= 0 || g->npol <= 0 || g->nx <= 0 || g->ny <= 0) return fail(DDF_EINVAL, "bad grid shape");
  if (!f->grid) return fail(DDF_EINVAL, "grid pointer is null");
  if (v->nrow < 0 || v->nchan <= 0 || v->ncorr <= 0 || v->ncorr > 4) return fail(DDF_EINVAL, "bad visibility shape");
  if (v->ncorr != v->ncorr_codes) return fail(DDF_EINVAL, "DataCorrFormat must list one code per correlation");
  if (v->nstokes <= 0 || v->nstokes > 4) return fail(DDF_EINVAL, "bad number of Stokes parameters");
  if (!v->uvw || !v->vis || !v->flags || !v->freqs || !v->chan_map || !v->bda)
    return fail(DDF_EINVAL, "null visibility-side pointer");
  if (v->bda_len < 2 || v->nblocks < 0 || 2 + v->nblocks > v->bda_len) return fail(DDF_EINVAL, "inconsistent BDA map");
  const ddf_cf& cf = f->cf;
  if (!cf.data || !cf.offset || !cf.support || cf.nw <= 0 || cf.overs <= 0 || cf.max_support <= 0)
    return fail(DDF_EINVAL, "bad CF table");
  if (f->jones.enabled) {
    const ddf_jones& j = f->jones;
    if (!j.kms && !j.beam)
      return fail(DDF_ERUNTIME, "Jones matricies specified but neither E or DD Jones are applied. This is a bug!");
    if (j.mode != 1 && j.mode != 2) return fail(DDF_EINVAL, "Jones mode must be 1 or 2");
    if (!j.a0 || !j.a1) return fail(DDF_EINVAL, "Jones antenna maps are null");
    if (j.kms && (!j.kms_time || !j.kms_chan)) return fail(DDF_EINVAL, "DD Jones maps are null");
    if (j.beam && (!j.beam_time || !j.beam_chan)) return fail(DDF_EINVAL, "beam Jones maps are null");
    for (int i = 0; i < 4; ++i)
      if ((j.kms && j.kms_dims[i] <= 0) || (j.beam && j.beam_dims[i] <= 0)) return fail(DDF_EINVAL, "empty Jones table");
    if (j.interp_mode == 1 && j.ncoefs > 0 && !j.coefs) return fail(DDF_EINVAL, "interpolation coefficients are null");
  }
  if (f->decorr.enabled && f->decorr.tsmear && !f->decorr.uvw_dt) return fail(DDF_EINVAL, "uvw_dt is null");
  return DDF_OK;
}

void fill_params(KParams& p, const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, const StokesMap& sm,
                 const int64_t* bda_start, double fmean0, double wl_mean) {
  std::memset(&p, 0, sizeof(p));
  p.v.vis = reinterpret_cast<float2*>(v->vis);
  p.v.uvw = v->uvw;
  p.v.flags = v->flags;
  p.v.weights = v->weights;
  p.v.freqs = v->freqs;
  p.v.chan_map = v->chan_map;
  p.v.nrow = v->nrow;
  p.v.nchan = v->nchan;
  p.v.ncorr = v->ncorr;
  p.v.bda = v->bda;
  p.v.bda_start = bda_start;
  p.v.nblocks = v->nblocks;
  p.v.bda_len = v->bda_len;
  p.v.sparse = v->sparsification;
  p.grid = f->grid;
  p.nband = g->nband;
  p.npol = g->npol;
  p.ny = g->ny;
  p.nx = g->nx;
  p.incr0 = g->incr[0];
  p.incr1 = g->incr[1];
  p.dopsf = g->dopsf;
  p.cu = f->cu;
  p.cv = f->cv;
  p.l0 = f->l0;
  p.m0 = f->m0;
  p.n0 = std::sqrt(1 - f->l0 * f->l0 - f->m0 * f->m0) - 1;
  p.wl_mean = wl_mean;
  p.fmean0 = fmean0;
  p.tscale0 = decorr_tscale(fmean0);
  p.cf.data = reinterpret_cast<const float2*>(f->cf.data);
  p.cf.offset = f->cf.offset;
  p.cf.support = f->cf.support;
  p.cf.nw = f->cf.nw;
  p.cf.overs = f->cf.overs;
  p.cf.ref_wave = f->cf.ref_wave;
  p.cf.wmax = f->cf.wmax;
  const ddf_jones& j = f->jones;
  p.jp.enabled = j.enabled;
  p.jp.mode = j.mode;
  p.jp.interp = j.interp_mode;
  p.jp.kms = reinterpret_cast<const float2*>(j.kms);
  p.jp.beam = reinterpret_cast<const float2*>(j.beam);
  for (int i = 0; i < 4; ++i) {
    p.jp.kdims[i] = j.kms_dims[i];
    p.jp.bdims[i] = j.beam_dims[i];
  }
  p.jp.ktime = j.kms_time;
  p.jp.kchan = j.kms_chan;
  p.jp.kdir = j.kms_dir;
  p.jp.btime = j.beam_time;
  p.jp.bchan = j.beam_chan;
  p.jp.bdir = j.beam_dir;
  p.jp.a0 = j.a0;
  p.jp.a1 = j.a1;
  p.jp.coefs = j.coefs;
  p.jp.ncoefs = j.coefs ? j.ncoefs : 0;
  p.jp.apply_amp = j.apply_amp;
  p.jp.apply_phase = j.apply_phase;
  p.jp.scale = j.scale_jones;
  p.jp.calib_error = j.calib_error;
  p.jp.sum_jones = j.sum_jones;
  p.jp.sum_jones_chan = j.sum_jones_chan;
  p.dp.enabled = f->decorr.enabled;
  p.dp.tsmear = f->decorr.tsmear;
  p.dp.fsmear = f->decorr.fsmear;
  p.dp.uvw_dt = f->decorr.uvw_dt;
  p.dp.dt = f->decorr.dt;
  p.dp.dnu = f->decorr.dnu;
  p.dp.fscale = decorr_fscale(f->decorr.dnu);
  p.dp.l0 = f->decorr.l0;
  p.dp.m0 = f->decorr.m0;
  p.sumwt = f->sumwt;
  p.nstokes = sm.nstokes;
  std::memcpy(p.sm, sm.m, sizeof(p.sm));
}

// Per-call temporary device memory (stream ordered).
struct Temp {
  hipStream_t st;
  std::vector<void*> ptrs;
  explicit Temp(hipStream_t s) : st(s) {}
  void* get(size_t bytes) {
    void* p = nullptr;
    if (hipMallocAsync(&p, bytes, st) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return p;
  }
  ~Temp() {
    for (void* p : ptrs) (void)hipFreeAsync(p, st);
  }
};

int resolve_block_index(const ddf_vis* v, Temp& tmp, const int64_t** out) {
  if (v->bda_start) {
    *out = v->bda_start;
    return DDF_OK;
  }
  if (v->nblocks == 0) {
    *out = nullptr;
    return DDF_OK;
  }
  int64_t nchunk = (v->nblocks + kScanChunk - 1) / kScanChunk;
  int64_t* starts = static_cast<int64_t*>(tmp.get(size_t(v->nblocks) * 8));
  int64_t* chunks = static_cast<int64_t*>(tmp.get(size_t(nchunk) * 8));
  if (!starts || !chunks) return fail(DDF_ENOMEM, "device allocation of the BDA block index failed");
  int rc = launch_bda_index(v->bda, v->nblocks, starts, chunks, tmp.st);
  *out = starts;
  return rc;
}

// Mean frequency and wavelength of the channels (gridder.h:164-171): taken from the host fields
// when set, else read back from the device array (a synchronising copy).
int freq_means(const ddf_vis* v, double* fmean, double* wlmean, hipStream_t st) {
  if (v->freq_mean > 0 && v->wavelength_mean > 0) {
    *fmean = v->freq_mean;
    *wlmean = v->wavelength_mean;
    return DDF_OK;
  }
  std::vector<double> f(size_t(v->nchan));
  HIP_TRY(hipMemcpyAsync(f.data(), v->freqs, f.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  double wl = 0, fm = 0;
  for (double x : f) {
    wl += kC / x;
    fm += x;
  }
  *fmean = fm / double(f.size());
  *wlmean = wl / double(f.size());
  return DDF_OK;
}

unsigned grid_blocks(int64_t nblocks) {
  int64_t need = (nblocks + 3) / 4;
  int64_t cap = 256 * 16;
  if (need < 1) need = 1;
  return unsigned(need < cap ? need : cap);
}

template <int READ, int JMODE>
int launch_grid_t(const KParams& p, int grid_f64, size_t lds, unsigned nb, hipStream_t st) {
  if (grid_f64)
    hipLaunchKernelGGL((grid_plain<READ, JMODE, double>), dim3(nb), dim3(256), lds, st, p);
  else
    hipLaunchKernelGGL((grid_plain<READ, JMODE, float>), dim3(nb), dim3(256), lds, st, p);
  return DDF_OK;
}

template <int READ>
int launch_grid_r(const KParams& p, int jmode, int grid_f64, size_t lds, unsigned nb, hipStream_t st) {
  if (jmode == 1) return launch_grid_t<READ, 1>(p, grid_f64, lds, nb, st);
  if (jmode == 2) return launch_grid_t<READ, 2>(p, grid_f64, lds, nb, st);
  return launch_grid_t<READ, 0>(p, grid_f64, lds, nb, st);
}

template <int NCORR>
int launch_degrid_r(const KParams& p, int jmode, unsigned nb, hipStream_t st) {
  if (jmode == 1) hipLaunchKernelGGL((degrid_plain<NCORR, 1>), dim3(nb), dim3(256), 0, st, p);
  else if (jmode == 2) hipLaunchKernelGGL((degrid_plain<NCORR, 2>), dim3(nb), dim3(256), 0, st, p);
  else hipLaunchKernelGGL((degrid_plain<NCORR, 0>), dim3(nb), dim3(256), 0, st, p);
  return DDF_OK;
}

}  // namespace

namespace ddf {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace ddf

// ---- Briggs transform without allocation ------------------------------------------------------
// The per-workgroup partial sums of the robust transform live in a file-scope device array, one
// copy per GPU, instead of a caller workspace: for callers that cannot allocate (graph capture,
// serving loops).  Its address on each device comes from hipGetSymbolAddress, once per device
// (ddf_init, or the first call on a device); host code uses that address, never the symbol.
constexpr int kBriggsBlocks = 256;  // workgroups per band of the robust transform
constexpr int kBriggsMaxBands = 64;
static __device__ __align__(256) double scratch[kBriggsMaxBands * kBriggsBlocks * 2];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the current device (resolved once per device).
int scratch_address(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DDF_EINVAL, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(DDF_EHIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return DDF_OK;
}
}  // namespace

extern "C" {

int ddf_version(void) { return DDF_ABI_VERSION; }

const char* ddf_last_error(void) { return g_err.c_str(); }

int ddf_init(int device) {
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipFree(nullptr));
  double* ptr = nullptr;
  return scratch_address(&ptr);
}

int ddf_bda_index(const int32_t* bda, int64_t bda_len, int64_t nblocks, int64_t* out, void* stream) {
  if (!bda || !out || nblocks < 0 || 2 + nblocks > bda_len) return fail(DDF_EINVAL, "inconsistent BDA map");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Temp tmp(st);
  int64_t nchunk = (nblocks + kScanChunk - 1) / kScanChunk;
  int64_t* chunks = nchunk ? static_cast<int64_t*>(tmp.get(size_t(nchunk) * 8)) : nullptr;
  if (nchunk && !chunks) return fail(DDF_ENOMEM, "device allocation failed");
  return launch_bda_index(bda, nblocks, out, chunks, st);
}

int ddf_grid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, void* stream) {
  int rc = check_common(g, v, f);
  if (rc) return rc;
  if (!v->weights || !f->sumwt) return fail(DDF_EINVAL, "weights / sumwt pointer is null");
  if (f->jones.enabled && (!f->jones.sum_jones || !f->jones.sum_jones_chan))
    return fail(DDF_EINVAL, "SumJones / SumJonesChan pointer is null");
  StokesMap sm;
  rc = grid_map(v, &sm);
  if (rc) return rc;
  if (v->nblocks == 0 || v->nrow == 0) return DDF_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double fmean0 = 0, wl_mean = 0;
  rc = freq_means(v, &fmean0, &wl_mean, st);
  if (rc) return rc;
  Temp tmp(st);
  const int64_t* starts = nullptr;
  rc = resolve_block_index(v, tmp, &starts);
  if (rc) return rc;
  KParams p;
  fill_params(p, g, v, f, sm, starts, fmean0, wl_mean);
  const int jmode = f->jones.enabled ? f->jones.mode : 0;
  size_t lds_base = size_t(g->nband) * g->npol + 2 * size_t(g->nband);
  p.chan_lds = (jmode != 0) && (lds_base + 2 * size_t(v->nchan)) * 8 <= 48 * 1024;
  size_t lds = (lds_base + (p.chan_lds ? 2 * size_t(v->nchan) : 0)) * 8;
  if (lds > 60 * 1024) return fail(DDF_EINVAL, "nband * npol too large for the gridding accumulators");
  unsigned nb = grid_blocks(v->nblocks);
  if (sm.read == 4) rc = launch_grid_r<4>(p, jmode, g->grid_f64, lds, nb, st);
  else if (sm.read == 2) rc = launch_grid_r<2>(p, jmode, g->grid_f64, lds, nb, st);
  else rc = launch_grid_r<1>(p, jmode, g->grid_f64, lds, nb, st);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

int ddf_degrid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, void* stream) {
  int rc = check_common(g, v, f);
  if (rc) return rc;
  if (g->grid_f64) return fail(DDF_EINVAL, "the degridder reads complex64 grids");
  StokesMap sm;
  rc = degrid_map(v, &sm);
  if (rc) return rc;
  if (sm.nstokes > g->npol) return fail(DDF_EINVAL, "grid has fewer polarisations than the requested Stokes parameters");
  if (v->nblocks == 0 || v->nrow == 0) return DDF_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double fmean0 = 0, wl_mean = 0;
  rc = freq_means(v, &fmean0, &wl_mean, st);
  if (rc) return rc;
  Temp tmp(st);
  const int64_t* starts = nullptr;
  rc = resolve_block_index(v, tmp, &starts);
  if (rc) return rc;
  KParams p;
  fill_params(p, g, v, f, sm, starts, fmean0, wl_mean);
  const int jmode = f->jones.enabled ? f->jones.mode : 0;
  unsigned nb = grid_blocks(v->nblocks);
  if (v->ncorr == 4) rc = launch_degrid_r<4>(p, jmode, nb, st);
  else if (v->ncorr == 2) rc = launch_degrid_r<2>(p, jmode, nb, st);
  else if (v->ncorr == 1) rc = launch_degrid_r<1>(p, jmode, nb, st);
  else return fail(DDF_EINVAL, "ncorr must be 1, 2 or 4");
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

}  // extern "C"

// ---- weights scatter-add ---------------------------------------------------------------------
namespace {
__global__ __launch_bounds__(256) void accumulate_weights(double* grid, int64_t glen, const float* w, const int64_t* idx,
                                                          int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    float x = w[i];
    if (x != 0) {
      int64_t k = idx[i];
      if (k >= 0 && k < glen) atomicAdd(&grid[k], double(x));
    }
  }
}
}  // namespace

extern "C" int ddf_accumulate_weights(double* grid, int64_t grid_len, const float* weights, const int64_t* index,
                                      int64_t n, void* stream) {
  if (n < 0 || grid_len < 0) return fail(DDF_EINVAL, "negative size");
  if (n == 0) return DDF_OK;
  if (!grid || !weights || !index) return fail(DDF_EINVAL, "null pointer");
  int64_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(accumulate_weights, dim3(unsigned(nb)), dim3(256), 0, static_cast<hipStream_t>(stream), grid,
                     grid_len, weights, index, n);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

// ---- imaging weights: uv-grid index, Briggs transform, renormalisation ---------------------------
namespace {
constexpr double kLightSpeedWeights = 299792458.0;  // ClassVisServer._cc

// One thread per (row, channel): the uv-grid cell of ClassVisServer._uv_to_index
// (ClassVisServer.py:983-1004), in its operation order and without contraction: (u freq) / c, then
// / cell, floored; the int64 cell number is narrowed to int32 as numpy's assignment into an int32
// array does (two's complement), and the band offset added and narrowed again.
__global__ __launch_bounds__(256) void weights_index_k(const double* uvw, const float* w, int64_t nrow, int nchan,
                                                       const double* freqs, const int32_t* chan_band, double cell,
                                                       int64_t npix, int64_t npixx, int nbands, int64_t xymax,
                                                       int32_t* index) {
#pragma clang fp contract(off)
  const int64_t n = nrow * int64_t(nchan);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i / nchan;
    const int c = int(i - r * nchan);
    double u = uvw[3 * r], v = uvw[3 * r + 1];
    if (v < 0) {
      u = -u;
      v = -v;
    }
    const double f = freqs[c];
    const double ul = (u * f) / kLightSpeedWeights, vl = (v * f) / kLightSpeedWeights;
    const int64_t x = int64_t(floor(ul / cell)) + xymax;
    const int64_t y = int64_t(floor(vl / cell));
    int32_t k = int32_t(uint32_t(uint64_t(y * npixx + x)));
    if (nbands > 1) k = int32_t(uint32_t(uint64_t(int64_t(k) + int64_t(chan_band[c]) * npix)));
    if (w[i] == 0.f) k = 0;
    index[i] = k;
  }
}

// Scatter-add of a [nrow][nchan] weight array through its int32 cell index (the sum of
// pyAccumulateWeightsOntoGrid, GridderSmearPols.cc:32-74, zero weights skipped).  A wavefront takes
// one channel of 64 consecutive rows: rows of one baseline at successive times mostly fall in one
// cell, so the lanes first sum each run of equal cells (segmented scan over the wavefront) and only
// a run's last lane issues the FP64 atomic.
__global__ __launch_bounds__(256) void weights_accumulate_k(double* grid, int64_t glen, const float* w,
                                                            const int32_t* idx, int64_t nrow, int nchan) {
  const int lane = threadIdx.x & 63;
  const int64_t ntile = (nrow + 63) / 64;
  const int64_t ntask = ntile * nchan;
  const int64_t wave0 = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwave = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t task = wave0; task < ntask; task += nwave) {
    const int c = int(task % nchan);
    const int64_t r = (task / nchan) * 64 + lane;
    int k = -1;
    double x = 0;
    if (r < nrow) {
      const int64_t i = r * nchan + c;
      const float wi = w[i];
      const int ki = idx[i];
      if (wi != 0.f && ki >= 0 && int64_t(ki) < glen) {
        k = ki;
        x = double(wi);
      }
    }
    // head of this lane's run of equal cells
    const int kp = __shfl_up(k, 1, 64);
    int start = (lane == 0 || kp != k) ? lane : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(start, d, 64);
      if (lane >= d && o > start) start = o;
    }
    // sum from the run's head to this lane
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double y = __shfl_up(x, d, 64);
      if (lane - d >= start) x += y;
    }
    const int kn = __shfl_down(k, 1, 64);
    const bool tail = lane == 63 || kn != k;
    if (tail && k >= 0) atomicAdd(&grid[k], x);
  }
}

// Robust transform of the weight grid (ClassVisServer.py:1188-1197) in two launches over
// kBriggsBlocks workgroups per band: the first writes each workgroup's partial sums of g^2 and g
// over a strided share of the band's cells; the second has every workgroup reduce the band's
// partials in the same order (so each obtains the same avgW = sum g^2 / sum g), then apply
// g <- 1 + g sSq, sSq = num2 / avgW, to its share, the product rounded before the sum.

__global__ __launch_bounds__(256) void weights_briggs_sums_k(const double* grid, int64_t npix, double* part) {
  __shared__ double s2[256], s1[256];
  const int band = blockIdx.y;
  const double* g = grid + int64_t(band) * npix;
  double a2 = 0, a1 = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += int64_t(gridDim.x) * blockDim.x) {
    const double x = g[i];
    a2 = fma(x, x, a2);
    a1 += x;
  }
  s2[threadIdx.x] = a2;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      s2[threadIdx.x] += s2[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[(int64_t(band) * gridDim.x + blockIdx.x) * 2] = s2[0];
    part[(int64_t(band) * gridDim.x + blockIdx.x) * 2 + 1] = s1[0];
  }
}

__global__ __launch_bounds__(256) void weights_briggs_apply_k(double* grid, int64_t npix, const double* part,
                                                              double num2) {
#pragma clang fp contract(off)
  __shared__ double s2[256], s1[256];
  const int band = blockIdx.y;
  const double* pb = part + int64_t(band) * gridDim.x * 2;
  double a2 = 0, a1 = 0;
  for (int k = threadIdx.x; k < int(gridDim.x); k += blockDim.x) {
    a2 += pb[2 * k];
    a1 += pb[2 * k + 1];
  }
  s2[threadIdx.x] = a2;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      s2[threadIdx.x] += s2[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  const double avg = s2[0] / s1[0];
  const double ssq = num2 / avg;
  double* g = grid + int64_t(band) * npix;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += int64_t(gridDim.x) * blockDim.x) {
    const double t = g[i] * ssq;
    g[i] = 1. + t;
  }
}

// weights[i] /= grid[index[i]] (ClassVisServer.py:1035-1040): indices >= grid_len read cell 0 as
// the reference sets them; negative ones count from the end as numpy indexing does (below
// -grid_len, where numpy raises, cell 0).  The quotient is formed in FP64 and rounded to FP32, as
// numpy's in-place division of a float32 array by float64 values is.
__global__ __launch_bounds__(256) void weights_finalize_k(float* w, const int32_t* idx, int64_t n, const double* grid,
                                                          int64_t glen) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t k = idx[i];
    if (k >= glen) k = 0;
    if (k < 0) k += glen;
    if (k < 0) k = 0;
    w[i] = float(double(w[i]) / grid[k]);
  }
}

unsigned weight_blocks(int64_t n) {
  int64_t nb = (n + 255) / 256;
  return unsigned(nb > 65536 ? 65536 : (nb < 1 ? 1 : nb));
}
}  // namespace

extern "C" int ddf_weights_index(const double* uvw, const float* weights, int64_t nrow, int nchan, const double* freqs,
                                 const int32_t* chan_band, double cell, int64_t npix, int64_t npixx, int nbands,
                                 int64_t xymax, int32_t* index, void* stream) {
  if (nrow < 0 || nchan < 0 || nbands < 1) return fail(DDF_EINVAL, "negative size or nbands < 1");
  if (nrow == 0 || nchan == 0) return DDF_OK;
  if (!uvw || !weights || !freqs || !index) return fail(DDF_EINVAL, "null pointer");
  if (nbands > 1 && !chan_band) return fail(DDF_EINVAL, "chan_band is required when nbands > 1");
  if (!(cell > 0)) return fail(DDF_EINVAL, "cell must be positive");
  hipLaunchKernelGGL(weights_index_k, dim3(weight_blocks(nrow * int64_t(nchan))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), uvw, weights, nrow, nchan, freqs, chan_band, cell, npix, npixx,
                     nbands, xymax, index);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_weights_accumulate(double* grid, int64_t grid_len, const float* weights, const int32_t* index,
                                      int64_t nrow, int nchan, void* stream) {
  if (nrow < 0 || nchan < 0 || grid_len < 0) return fail(DDF_EINVAL, "negative size");
  if (nrow == 0 || nchan == 0) return DDF_OK;
  if (!grid || !weights || !index) return fail(DDF_EINVAL, "null pointer");
  const int64_t waves = ((nrow + 63) / 64) * nchan;
  int64_t nb = (waves + 3) / 4;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(weights_accumulate_k, dim3(unsigned(nb)), dim3(256), 0, static_cast<hipStream_t>(stream), grid,
                     grid_len, weights, index, nrow, nchan);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

extern "C" int64_t ddf_weights_briggs_work_size(int nbands) {
  return nbands > 0 ? int64_t(nbands) * kBriggsBlocks * 2 : 0;
}

extern "C" int ddf_weights_briggs(double* grid, int nbands, int64_t npix, double robust, double* work,
                                  int64_t work_len, void* stream) {
  if (nbands < 0 || npix < 0) return fail(DDF_EINVAL, "negative size");
  if (nbands == 0 || npix == 0) return DDF_OK;
  if (nbands > 65535) return fail(DDF_EINVAL, "more than 65535 bands");
  if (!grid || !work) return fail(DDF_EINVAL, "null pointer");
  if (work_len < ddf_weights_briggs_work_size(nbands)) return fail(DDF_EINVAL, "workspace too small");
  // numeratorSqrt = 5 * 10 ** (-Robust); its square as Python's float ** 2 (an exact square, rounded)
  const double num = 5.0 * std::pow(10.0, -robust);
  const double num2 = num * num;
  const dim3 grid2(kBriggsBlocks, unsigned(nbands));
  hipLaunchKernelGGL(weights_briggs_sums_k, grid2, dim3(256), 0, static_cast<hipStream_t>(stream), grid, npix, work);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(weights_briggs_apply_k, grid2, dim3(256), 0, static_cast<hipStream_t>(stream), grid, npix, work,
                     num2);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_weights_briggs_noalloc(double* grid, int nbands, int64_t npix, double robust, void* stream) {
  if (nbands < 0 || npix < 0) return fail(DDF_EINVAL, "negative size");
  if (nbands > kBriggsMaxBands)
    return fail(DDF_EINVAL, "ddf_weights_briggs_noalloc handles at most 64 bands (use ddf_weights_briggs)");
  if (nbands == 0 || npix == 0) return DDF_OK;
  if (!grid) return fail(DDF_EINVAL, "null pointer");
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  // every partial the second launch reads is written by the first in this call: nothing to clear
  const double num = 5.0 * std::pow(10.0, -robust);
  const double num2 = num * num;
  const dim3 grid2(kBriggsBlocks, unsigned(nbands));
  hipLaunchKernelGGL(weights_briggs_sums_k, grid2, dim3(256), 0, static_cast<hipStream_t>(stream), grid, npix, part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(weights_briggs_apply_k, grid2, dim3(256), 0, static_cast<hipStream_t>(stream), grid, npix, part,
                     num2);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_weights_finalize(float* weights, const int32_t* index, int64_t n, const double* grid,
                                    int64_t grid_len, void* stream) {
  if (n < 0 || grid_len < 0) return fail(DDF_EINVAL, "negative size");
  if (n == 0) return DDF_OK;
  if (!weights || !index || !grid || grid_len == 0) return fail(DDF_EINVAL, "null pointer or empty grid");
  hipLaunchKernelGGL(weights_finalize_k, dim3(weight_blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), weights,
                     index, n, grid, grid_len);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

// ---- chunk flag preparation (ClassMS.py:1155-1167) -------------------------------------------------
namespace {
// One thread per (row, channel): autocorrelation rows flagged, NaN visibilities (either part)
// flagged, the values flagged so far overwritten with 1e9 + 0i, then every correlation flagged when
// any is, in the reference's order.
__global__ __launch_bounds__(256) void prepare_flags_k(uint8_t* flags, float2* data, const int32_t* a0,
                                                       const int32_t* a1, int64_t nrow, int nchan, int ncorr,
                                                       int vec4) {
  const int64_t n = nrow * int64_t(nchan);
  if (vec4) {
    // four correlations with 4-byte aligned flags and 16-byte aligned values: the flags of a
    // (row, channel) as one 32-bit word, the values as two 16-byte loads
    uint32_t* fw = reinterpret_cast<uint32_t*>(flags);
    float4* d4 = reinterpret_cast<float4*>(data);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
      const int64_t r = i / nchan;
      const bool autoc = a0[r] == a1[r];
      const uint32_t f0 = fw[i];
      bool f[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) f[k] = ((f0 >> (8 * k)) & 0xFFu) != 0 || autoc;
      if (data) {
        float4 x0 = d4[2 * i], x1 = d4[2 * i + 1];
        f[0] = f[0] || isnan(x0.x) || isnan(x0.y);
        f[1] = f[1] || isnan(x0.z) || isnan(x0.w);
        f[2] = f[2] || isnan(x1.x) || isnan(x1.y);
        f[3] = f[3] || isnan(x1.z) || isnan(x1.w);
        if (f[0] || f[1]) {
          if (f[0]) { x0.x = 1e9f; x0.y = 0.f; }
          if (f[1]) { x0.z = 1e9f; x0.w = 0.f; }
          d4[2 * i] = x0;
        }
        if (f[2] || f[3]) {
          if (f[2]) { x1.x = 1e9f; x1.y = 0.f; }
          if (f[3]) { x1.z = 1e9f; x1.w = 0.f; }
          d4[2 * i + 1] = x1;
        }
      }
      const bool any = f[0] || f[1] || f[2] || f[3];
      fw[i] = any ? 0x01010101u : 0u;
    }
    return;
  }
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i / nchan;
    const bool autoc = a0[r] == a1[r];
    bool any = false;
    for (int k = 0; k < ncorr; ++k) {
      const int64_t j = i * ncorr + k;
      bool f = flags[j] != 0 || autoc;
      if (data) {
        const float2 x = data[j];
        f = f || isnan(x.x) || isnan(x.y);
        if (f) data[j] = make_float2(1e9f, 0.f);
      }
      any = any || f;
      flags[j] = f ? 1 : 0;
    }
    if (any)
      for (int k = 0; k < ncorr; ++k) flags[i * ncorr + k] = 1;
  }
}
}  // namespace

extern "C" int ddf_prepare_flags(uint8_t* flags, float* data, const int32_t* a0, const int32_t* a1, int64_t nrow,
                                 int nchan, int ncorr, void* stream) {
  if (nrow < 0 || nchan < 0 || ncorr < 1 || ncorr > 4) return fail(DDF_EINVAL, "bad sizes (ncorr must be 1..4)");
  if (nrow == 0 || nchan == 0) return DDF_OK;
  if (!flags || !a0 || !a1) return fail(DDF_EINVAL, "null pointer");
  const int vec4 = ncorr == 4 && reinterpret_cast<uintptr_t>(flags) % 4 == 0 &&
                   reinterpret_cast<uintptr_t>(data) % 16 == 0;
  hipLaunchKernelGGL(prepare_flags_k, dim3(weight_blocks(nrow * int64_t(nchan))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), flags, reinterpret_cast<float2*>(data), a0, a1, nrow, nchan,
                     ncorr, vec4);
  HIP_TRY(hipGetLastError());
  return DDF_OK;
}

// ======================================================================================================
// Batched plan (host side).
// ======================================================================================================
struct ddf_plan {
  BatchParams pg, pd;
  VisP vg, vd;
  FacetDev* d_facets = nullptr;
  int32_t* d_support = nullptr;
  // placement records of both maps while the plan is built (ddf_plan_placements recomputes them)
  uint64_t* place_g = nullptr;
  uint64_t* place_d = nullptr;
  // pair codes (0x80 | band for a placed (block, facet) record) read by the cycle passes
  uint8_t* pcode_g = nullptr;
  uint8_t* pcode_d = nullptr;
  uint32_t* recs_g = nullptr;
  uint32_t* recs_d = nullptr;
  uint64_t* bstamp_g = nullptr;  // stamp records in bin order (tile offset, support, CF offset)
  float* slots_g = nullptr;      // one LDS tile image per grid work item (null: atomic flush)
  uint32_t* tile_first_g = nullptr;
  uint32_t* tile_count_g = nullptr;
  uint64_t* bstamp_d = nullptr;
  BinItem* items_g = nullptr;
  BinItem* items_d = nullptr;
  float2* V = nullptr;
  float2* M = nullptr;
  uint64_t* vrc = nullptr;
  uint32_t* vblk = nullptr;
  int64_t nvl = 0;
  uint32_t* vrun = nullptr;       // run of each degrid list entry (block and Jones key shared)
  uint32_t* run_first = nullptr;  // first list entry of each run
  int64_t nruns = 0;
  // work-queue counters: [0] grid pass A, [1, 9) grid pass B (one per item range), [9, 17) degrid
  // pass A, [17] the adjoint of degrid pass B, [18, 26) the adjoint of degrid pass A
  uint32_t* queues = nullptr;
  const float2* jsrc[2] = {nullptr, nullptr};  // caller's DD / beam Jones tables
  double2* jprep[2] = {nullptr, nullptr};      // prepared copies (JonesP::pk / pb)
  float2* jprepf[2] = {nullptr, nullptr};      // complex64 prepared copies (JonesP::pkf / pbf)
  unsigned nwg = 0;
  unsigned nwg_tiles[3] = {0, 0, 0};  // tile passes (grid B, degrid A, adjoint): as many as fit
  int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<void*> allocs;
  // per-pass timing (ddf_plan_timing): events around each pass on the launch stream
  int timing = 0;
  hipEvent_t ev[DDF_PLAN_NPASS][2] = {};
  int ev_set[DDF_PLAN_NPASS] = {};
  ~ddf_plan() {
    for (void* p : allocs) (void)hipFree(p);
    for (auto& e : ev)
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
  }
  void mark(int pass, int end, hipStream_t st) {
    if (!timing) return;
    (void)hipEventRecord(ev[pass][end], st);
    if (end) ev_set[pass] = 1;
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    size_t bytes = n * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    allocs.push_back(p);
    stats[5] += int64_t(bytes);
    return static_cast<T*>(p);
  }
  // frees one allocation of this plan (bytes as allocated)
  template <typename T>
  void release(T*& p, size_t n) {
    for (size_t k = 0; k < allocs.size(); ++k) {
      if (allocs[k] != static_cast<void*>(p)) continue;
      (void)hipFree(allocs[k]);
      allocs[k] = allocs.back();
      allocs.pop_back();
      stats[5] -= int64_t(n ? n * sizeof(T) : sizeof(T));
      break;
    }
    p = nullptr;
  }
};

namespace {

int prepare_jones(ddf_plan* P, hipStream_t st) {
  for (int k = 0; k < 2; ++k)
    if (P->jprep[k] && batched_prepare_jones(P->pg, P->jsrc[k], P->jprep[k], P->jprepf[k], k == 0, st)) return DDF_EHIP;
  return DDF_OK;
}

int fill_vis(VisP& out, const ddf_vis* v, const int64_t* starts) {
  std::memset(&out, 0, sizeof(out));
  out.vis = reinterpret_cast<float2*>(v->vis);
  out.uvw = v->uvw;
  out.flags = v->flags;
  out.weights = v->weights;
  out.freqs = v->freqs;
  out.chan_map = v->chan_map;
  out.nrow = v->nrow;
  out.nchan = v->nchan;
  out.ncorr = v->ncorr;
  out.bda = v->bda;
  out.bda_start = starts;
  out.nblocks = v->nblocks;
  out.bda_len = v->bda_len;
  out.sparse = v->sparsification;
  out.equi = v->chan_equidistant && v->nchan > 1;
  return DDF_OK;
}

void fill_batch(BatchParams& p, const ddf_grid_desc* g, const ddf_facet* f0, const StokesMap& sm,
                double fmean0, double wl_mean) {
  std::memset(&p, 0, sizeof(p));
  p.nband = g->nband;
  p.npol = g->npol;
  p.ny = g->ny;
  p.nx = g->nx;
  p.incr0 = g->incr[0];
  p.incr1 = g->incr[1];
  p.nw = f0->cf.nw;
  p.overs = f0->cf.overs;
  p.ref_wave = f0->cf.ref_wave;
  p.wmax = f0->cf.wmax;
  p.nstokes = sm.nstokes;
  p.read = sm.read;
  std::memcpy(p.sm, sm.m, sizeof(p.sm));
  const ddf_jones& j = f0->jones;
  p.jmode = j.enabled ? j.mode : 0;
  p.jp.enabled = j.enabled;
  p.jp.mode = j.mode;
  p.jp.interp = j.interp_mode;
  p.jp.kms = reinterpret_cast<const float2*>(j.kms);
  p.jp.beam = reinterpret_cast<const float2*>(j.beam);
  for (int i = 0; i < 4; ++i) {
    p.jp.kdims[i] = j.kms_dims[i];
    p.jp.bdims[i] = j.beam_dims[i];
  }
  p.jp.ktime = j.kms_time;
  p.jp.kchan = j.kms_chan;
  p.jp.btime = j.beam_time;
  p.jp.bchan = j.beam_chan;
  p.jp.a0 = j.a0;
  p.jp.a1 = j.a1;
  p.jp.coefs = j.coefs;
  p.jp.ncoefs = j.coefs ? j.ncoefs : 0;
  p.jp.apply_amp = j.apply_amp;
  p.jp.apply_phase = j.apply_phase;
  p.jp.scale = j.scale_jones;
  p.jp.calib_error = j.calib_error;
  p.dp.enabled = f0->decorr.enabled;
  p.dp.tsmear = f0->decorr.tsmear;
  p.dp.fsmear = f0->decorr.fsmear;
  p.dp.uvw_dt = f0->decorr.uvw_dt;
  p.dp.dt = f0->decorr.dt;
  p.dp.dnu = f0->decorr.dnu;
  p.dp.fscale = decorr_fscale(f0->decorr.dnu);
  p.fmean0 = fmean0;
  p.tscale0 = decorr_tscale(fmean0);
  p.wl_mean = wl_mean;
}

// Count -> scan -> fill of the valid records, then the work items (largest first).
int build_bins(ddf_plan* P, BatchParams& p, const VisP& v, const uint64_t* place, int chunk, uint32_t** recs_out,
               BinItem** items_out, uint32_t* nitems, int64_t* nrec, const uint32_t** ranges_out, int* nranges_out,
               bool stamp_order,
               hipStream_t st, uint32_t** tile_first_out = nullptr, uint32_t** tile_count_out = nullptr) {
  const int64_t nsub = 2 * int64_t(p.nw);  // CF-plane sub-bins of each tile
  const int64_t ntile_all = int64_t(p.nf) * p.nband * p.ntiles_y * p.ntiles_x;
  const int64_t nbins = ntile_all * nsub;
  uint32_t* counts = P->alloc<uint32_t>(size_t(nbins));
  uint32_t* fill = P->alloc<uint32_t>(size_t(nbins));
  int64_t* bstart = P->alloc<int64_t>(size_t(nbins));
  // A tile's sub-bins (CF planes) are split into ngrp groups of consecutive planes, each chunked
  // into items.  Slots are numbered tile by tile (a tile's slots consecutive, for
  // grid_slot_reduce); the launch order is (facet, band, plane group, tile), so that the
  // workgroups running at one time read the CF planes of one group and those stay in L2.
  const int64_t ngrp = nsub < 4 ? nsub : 4;
  uint32_t* gcount = P->alloc<uint32_t>(size_t(ntile_all * ngrp));
  int64_t* tstart = P->alloc<int64_t>(size_t(ntile_all));
  if (!counts || !fill || !bstart || !gcount || !tstart) return fail(DDF_ENOMEM, "plan allocation failed");
  HIP_TRY(hipMemsetAsync(counts, 0, size_t(nbins) * 4, st));
  HIP_TRY(hipMemsetAsync(fill, 0, size_t(nbins) * 4, st));
  if (batched_count(p, v.nblocks, place, counts, st)) return fail(DDF_EHIP, "plan_count launch failed");
  // only the (tile, group) counts come to the host; the bin starts are formed on the device
  if (batched_group_counts(counts, ntile_all, int(nsub), int(ngrp), gcount, st))
    return fail(DDF_EHIP, "plan_group_counts launch failed");
  std::vector<uint32_t> hg(static_cast<size_t>(ntile_all * ngrp));
  HIP_TRY(hipMemcpyAsync(hg.data(), gcount, hg.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // tile-major pass: each tile's first record and first slot, each (tile, group)'s item count
  std::vector<int64_t> ts(static_cast<size_t>(ntile_all));
  std::vector<uint32_t> tfirst(static_cast<size_t>(ntile_all)), tcount(static_cast<size_t>(ntile_all));
  int64_t tot = 0, nitem = 0;
  for (int64_t i = 0; i < ntile_all; ++i) {
    ts[size_t(i)] = tot;
    tfirst[size_t(i)] = uint32_t(nitem);
    for (int64_t g = 0; g < ngrp; ++g) {
      const int64_t cg = hg[size_t(i * ngrp + g)];
      tot += cg;
      nitem += (cg + chunk - 1) / chunk;
    }
    tcount[size_t(i)] = uint32_t(nitem - tfirst[size_t(i)]);
  }
  if (tot >= (int64_t(1) << 32)) return fail(DDF_EINVAL, "too many (block, facet) records for one plan");
  // items in launch order
  std::vector<BinItem> items;
  items.reserve(size_t(nitem));
  const int64_t ntile_fb = int64_t(p.ntiles_x) * p.ntiles_y;
  for (int64_t fb = 0; fb < int64_t(p.nf) * p.nband; ++fb) {
    for (int64_t g = 0; g < ngrp; ++g) {
      for (int64_t tin = 0; tin < ntile_fb; ++tin) {
        const int64_t i = fb * ntile_fb + tin;
        int64_t g0 = 0, s0 = 0;  // records and items of the tile's earlier groups
        for (int64_t h = 0; h < g; ++h) {
          const int64_t ch = hg[size_t(i * ngrp + h)];
          g0 += ch;
          s0 += (ch + chunk - 1) / chunk;
        }
        const int64_t cg = hg[size_t(i * ngrp + g)];
        for (int64_t b = 0, c = 0; b < cg; b += chunk, ++c) {
          BinItem it;
          it.begin = uint32_t(ts[size_t(i)] + g0 + b);
          it.end = uint32_t(ts[size_t(i)] + g0 + (b + chunk < cg ? b + chunk : cg));
          it.facet = int16_t(fb / p.nband);
          it.band = int16_t(fb % p.nband);
          it.tx = int16_t(tin % p.ntiles_x);
          it.ty = int16_t(tin / p.ntiles_x);
          it.slot = uint32_t(tfirst[size_t(i)] + s0 + c);
          items.push_back(it);
        }
      }
    }
  }
  HIP_TRY(hipMemcpyAsync(tstart, ts.data(), ts.size() * 8, hipMemcpyHostToDevice, st));
  if (batched_bin_starts(counts, tstart, ntile_all, int(nsub), bstart, st))
    return fail(DDF_EHIP, "plan_bin_starts launch failed");
  // kItemRanges contiguous ranges of the launch order with about equal record counts: the
  // workgroups of one XCD take the items of their own range first, so the items in flight on an
  // XCD come from few (facet, plane group)s and their CF planes stay in that XCD's L2.  Only with
  // many items: with few, the heaviest items (neighbouring uv tiles) would share one XCD's CUs, so
  // all items then form one range.
  const int nranges = items.size() >= kRangeMinItems ? kItemRanges : 1;
  std::vector<uint32_t> range(size_t(nranges) + 1, uint32_t(items.size()));
  range[0] = 0;
  if (nranges > 1) {
    int64_t acc = 0;
    int x = 1;
    for (size_t k = 0; k < items.size() && x < nranges; ++k) {
      acc += int64_t(items[k].end) - items[k].begin;
      while (x < nranges && acc * nranges >= tot * x) range[size_t(x++)] = uint32_t(k + 1);
    }
  }
  uint32_t* recs = P->alloc<uint32_t>(size_t(tot));
  BinItem* d_items = P->alloc<BinItem>(items.size());
  uint32_t* d_range = P->alloc<uint32_t>(range.size());
  if (!recs || !d_items || !d_range) return fail(DDF_ENOMEM, "plan allocation failed");
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(BinItem), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_range, range.data(), range.size() * 4, hipMemcpyHostToDevice, st));
  *ranges_out = d_range;
  *nranges_out = nranges;
  if (tile_first_out) {
    uint32_t* tf = P->alloc<uint32_t>(size_t(ntile_all));
    uint32_t* tc = P->alloc<uint32_t>(size_t(ntile_all));
    if (!tf || !tc) return fail(DDF_ENOMEM, "plan allocation failed");
    HIP_TRY(hipMemcpyAsync(tf, tfirst.data(), size_t(ntile_all) * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(tc, tcount.data(), size_t(ntile_all) * 4, hipMemcpyHostToDevice, st));
    *tile_first_out = tf;
    *tile_count_out = tc;
  }
  if (batched_fill(p, v.nblocks, place, bstart, fill, recs, st)) return fail(DDF_EHIP, "plan_fill launch failed");
  if (batched_sort_items(p, place, recs, tot, d_items, uint32_t(items.size()), stamp_order, st))
    return fail(DDF_EHIP, "sorting the records of the work items failed");
  HIP_TRY(hipStreamSynchronize(st));
  // the binning buffers are not needed by the cycle
  P->release(counts, size_t(nbins));
  P->release(fill, size_t(nbins));
  P->release(bstart, size_t(nbins));
  P->release(gcount, size_t(ntile_all * ngrp));
  P->release(tstart, size_t(ntile_all));
  *recs_out = recs;
  *items_out = d_items;
  *nitems = uint32_t(items.size());
  *nrec = tot;
  return DDF_OK;
}

}  // namespace

extern "C" int ddf_plan_create(const ddf_grid_desc* g, const ddf_vis* vgrid, const ddf_vis* vdegrid,
                               const ddf_facet* facets, void* const* models, int32_t nfacets,
                               const ddf_plan_opts* opts, ddf_plan** out, void* stream) {
  if (!out) return fail(DDF_EINVAL, "null plan output");
  *out = nullptr;
  if (nfacets <= 0 || !facets || !models) return fail(DDF_EINVAL, "no facets");
  int rc = check_common(g, vgrid, &facets[0]);
  if (rc) return rc;
  rc = check_common(g, vdegrid, &facets[0]);
  if (rc) return rc;
  if (g->grid_f64) return fail(DDF_EINVAL, "the batched path accumulates complex64 grids");
  if (g->nx >= 65536 || g->ny >= 65536 || g->nband >= 128) return fail(DDF_EINVAL, "grid too large for the plan records");
  if (!vgrid->weights) return fail(DDF_EINVAL, "weights pointer is null");
  const ddf_cf& c0 = facets[0].cf;
  if (2 * c0.nw > 255 || c0.overs > 255) return fail(DDF_EINVAL, "too many w-planes / oversampling for the plan records");
  StokesMap smg, smd;
  rc = grid_map(vgrid, &smg);
  if (rc) return rc;
  rc = degrid_map(vdegrid, &smd);
  if (rc) return rc;
  if (smd.nstokes > g->npol) return fail(DDF_EINVAL, "grid has fewer polarisations than the requested Stokes parameters");
  if ((vgrid->nblocks + 1) * int64_t(nfacets) >= (int64_t(1) << 32) ||
      (vdegrid->nblocks + 1) * int64_t(nfacets) >= (int64_t(1) << 32))
    return fail(DDF_EINVAL, "blocks x facets exceeds the 32-bit record index");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // shared CF geometry
  std::vector<int32_t> sup0(static_cast<size_t>(c0.nw)), supi(static_cast<size_t>(c0.nw));
  HIP_TRY(hipMemcpyAsync(sup0.data(), c0.support, sup0.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int maxsup = 0;
  for (int32_t x : sup0) {
    if (x <= 0 || x % 2 == 0) return fail(DDF_EINVAL, "CF supports must be odd and positive");
    maxsup = x > maxsup ? x : maxsup;
  }
  for (int i = 1; i < nfacets; ++i) {
    const ddf_cf& ci = facets[i].cf;
    if (ci.nw != c0.nw || ci.overs != c0.overs || ci.ref_wave != c0.ref_wave || ci.wmax != c0.wmax)
      return fail(DDF_EINVAL, "all facets of a plan must share nw, OverS, RefWave and wmax");
    HIP_TRY(hipMemcpyAsync(supi.data(), ci.support, supi.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (supi != sup0) return fail(DDF_EINVAL, "all facets of a plan must share the CF supports");
    if (!facets[i].grid || !models[i] || !facets[i].sumwt) return fail(DDF_EINVAL, "facet buffer is null");
    if (facets[0].jones.enabled && (!facets[i].jones.sum_jones || !facets[i].jones.sum_jones_chan))
      return fail(DDF_EINVAL, "SumJones buffers are null");
  }
  if (!models[0] || !facets[0].sumwt) return fail(DDF_EINVAL, "facet buffer is null");
  if (facets[0].jones.enabled && (!facets[0].jones.sum_jones || !facets[0].jones.sum_jones_chan))
    return fail(DDF_EINVAL, "SumJones buffers are null");

  ddf_plan* P = new ddf_plan();
  auto bail = [&](int code) {
    delete P;
    return code;
  };
  double fm = 0, wl = 0;
  rc = freq_means(vgrid, &fm, &wl, st);
  if (rc) return bail(rc);
  fill_batch(P->pg, g, &facets[0], smg, fm, wl);
  fill_batch(P->pd, g, &facets[0], smd, fm, wl);
  P->pg.dopsf = g->dopsf;
  // prepared Jones tables (FP64, facet-fastest), refreshed from the caller's tables at every
  // grid / degrid call so that new Jones values need no new plan
  for (int k = 0; k < 2; ++k) {
    const ddf_jones& j = facets[0].jones;
    const float* src = k ? j.beam : j.kms;
    const int32_t* dims = k ? j.beam_dims : j.kms_dims;
    if (!j.enabled || !src) continue;
    const size_t n = size_t(dims[0]) * dims[1] * dims[2] * dims[3] * 4;
    double2* dst = P->alloc<double2>(n);
    float2* dstf = P->alloc<float2>(n);
    if (!dst || !dstf) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    P->jsrc[k] = reinterpret_cast<const float2*>(src);
    P->jprep[k] = dst;
    P->jprepf[k] = dstf;
    for (BatchParams* bp : {&P->pg, &P->pd}) {
      if (k) {
        bp->jp.pb = dst;
        bp->jp.pbf = dstf;
      } else {
        bp->jp.pk = dst;
        bp->jp.pkf = dstf;
      }
    }
  }
  if (prepare_jones(P, st)) return bail(fail(DDF_EHIP, "Jones preparation failed"));
  ddf_plan_opts o{};
  if (opts) o = *opts;
  const int tile = o.tile > 0 ? o.tile : 16;
  const int chunk = o.chunk > 0 ? o.chunk : 2048;
  const int halo = (maxsup - 1) / 2;
  {
    // the tile passes hold one (T + 2 h)^2 complex64 image per polarisation in LDS
    int dev = 0, lds_max = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return bail(fail(DDF_EHIP, "device query failed"));
    const int W = tile + 2 * halo;
    const int np = std::max(std::min(smg.nstokes, g->npol), smd.nstokes);
    if (size_t(np) * W * W * 8 + 4096 > size_t(lds_max))
      return bail(fail(DDF_EINVAL, "uv tile too large for LDS at this support and polarisation count"));
    // stamp records hold the tile offset in 16 bits and the support in 5
    if (W > 255 || maxsup > 31) return bail(fail(DDF_EINVAL, "uv tile or CF support too large for the stamp records"));
    // ... and the oversampled sub-kernel index in 11 bits
    if (int64_t(c0.overs) * c0.overs > 0x7FF) return bail(fail(DDF_EINVAL, "CF oversampling too large for the stamp records"));
  }
  int G = o.group;
  if (G <= 0) {
    int ng = (nfacets + 31) / 32;
    G = (nfacets + ng - 1) / ng;
  }
  if (G > 64) G = 64;  // gridding pass A aligns its slots (G lanes each) to wavefronts
  // facet descriptors
  P->d_support = P->alloc<int32_t>(sup0.size());
  if (!P->d_support || hipMemcpyAsync(P->d_support, sup0.data(), sup0.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(DDF_EHIP, "plan upload failed"));
  std::vector<FacetDev> hf(static_cast<size_t>(nfacets));
  for (int i = 0; i < nfacets; ++i) {
    const ddf_facet& f = facets[i];
    FacetDev& d = hf[size_t(i)];
    d.grid = f.grid;
    d.model = models[i];
    d.cf = reinterpret_cast<const float2*>(f.cf.data);
    {
      // padded copy of the facet's CF planes for degrid pass A: rows of an even pitch, plane pl at
      // twice its offset
      const int nw = f.cf.nw, ov = f.cf.overs;
      std::vector<int64_t> off(size_t(2 * nw));
      if (hipMemcpyAsync(off.data(), f.cf.offset, off.size() * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return bail(fail(DDF_EHIP, "reading the CF plane offsets failed"));
      int64_t total = 0, total_pad = 0;
      for (int pl = 0; pl < 2 * nw; ++pl) {
        const int64_t S = sup0[size_t(pl < nw ? pl : pl - nw)];
        total = std::max(total, off[size_t(pl)] + int64_t(ov) * ov * S * S);
        total_pad = std::max(total_pad, 2 * off[size_t(pl)] + int64_t(ov) * ov * S * (S + (S & 1)));
      }
      // grid pass B addresses a facet's CF values by a 32-bit byte offset
      if (total * int64_t(sizeof(float2)) >= (int64_t(1) << 32))
        return bail(fail(DDF_EINVAL, "a facet's CF table reaches 4 GiB"));
      float2* t = P->alloc<float2>(size_t(total_pad));
      if (!t) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
      if (batched_cf_pad(d.cf, t, f.cf.offset, P->d_support, nw, ov, maxsup, st))
        return bail(fail(DDF_EHIP, "CF copy launch failed"));
      d.cft = t;
    }
    d.cf_offset = f.cf.offset;
    d.cu = f.cu;
    d.cv = f.cv;
    d.l0 = f.l0;
    d.m0 = f.m0;
    d.n0 = std::sqrt(1 - f.l0 * f.l0 - f.m0 * f.m0) - 1;
    d.dec_l0 = f.decorr.l0;
    d.dec_m0 = f.decorr.m0;
    d.kms_dir = f.jones.kms_dir;
    d.beam_dir = f.jones.beam_dir;
    d.sumwt = f.sumwt;
    d.sum_jones = f.jones.sum_jones;
    d.sum_jones_chan = f.jones.sum_jones_chan;
  }
  P->d_facets = P->alloc<FacetDev>(hf.size());
  P->queues = P->alloc<uint32_t>(32);
  if (!P->d_facets || !P->queues) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
  if (hipMemcpyAsync(P->d_facets, hf.data(), hf.size() * sizeof(FacetDev), hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(DDF_EHIP, "plan upload failed"));
  for (BatchParams* bp : {&P->pg, &P->pd}) {
    bp->facets = P->d_facets;
    bp->nf = nfacets;
    bp->support = P->d_support;
    bp->prec = o.precision;
    bp->tile = tile;
    bp->halo = halo;
    bp->ntiles_x = (g->nx + tile - 1) / tile;
    bp->ntiles_y = (g->ny + tile - 1) / tile;
    bp->group = G;
    bp->slots = 256 / G;
    bp->ngroups = (nfacets + G - 1) / G;
  }
  P->pd.jmode = vdegrid == vgrid ? P->pg.jmode : P->pd.jmode;
  // block indices of both maps
  const int64_t* sg = vgrid->bda_start;
  const int64_t* sd = vdegrid->bda_start;
  for (int k = 0; k < 2; ++k) {
    const ddf_vis* v = k ? vdegrid : vgrid;
    const int64_t*& s = k ? sd : sg;
    if (s || v->nblocks == 0) continue;
    int64_t* ns = P->alloc<int64_t>(size_t(v->nblocks));
    int64_t* tmp = P->alloc<int64_t>(size_t((v->nblocks + kScanChunk - 1) / kScanChunk));
    if (!ns || !tmp) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    rc = launch_bda_index(v->bda, v->nblocks, ns, tmp, st);
    if (rc) return bail(rc);
    s = ns;
  }
  fill_vis(P->vg, vgrid, sg);
  fill_vis(P->vd, vdegrid, sd);
  // per-block descriptors of the grid map (grid pass A reads them instead of the row lists)
  if (vgrid->bda_len >= (int64_t(1) << 32) || vgrid->nchan > 65535)
    return bail(fail(DDF_EINVAL, "BDA map too long or too many channels for the block descriptors"));
  {
    uint4* d = P->alloc<uint4>(size_t(vgrid->nblocks));
    if (!d) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    if (vgrid->nblocks && batched_block_desc(P->vg, d, st)) return bail(fail(DDF_EHIP, "plan_block_desc failed"));
    P->vg.bdesc = d;
  }
  P->pg.nblk = vgrid->nblocks;
  P->pd.nblk = vdegrid->nblocks;
  // first frequency and step of equidistant channels (the anchors of ChanPhasor)
  for (VisP* pv : {&P->vg, &P->vd}) {
    if (!pv->equi) continue;
    double f2[2] = {0, 0};
    if (hipMemcpyAsync(f2, pv->freqs, sizeof(f2), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bail(fail(DDF_EHIP, "reading the channel frequencies failed"));
    pv->f0 = f2[0];
    pv->df = f2[1] - f2[0];
  }
  // placements
  P->place_g = P->alloc<uint64_t>(size_t(vgrid->nblocks) * nfacets);
  P->place_d = P->alloc<uint64_t>(size_t(vdegrid->nblocks) * nfacets);
  if (!P->place_g || !P->place_d) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
  if (vgrid->nblocks && batched_place(P->pg, P->vg, false, P->place_g, st)) return bail(fail(DDF_EHIP, "plan_place failed"));
  if (vdegrid->nblocks && batched_place(P->pd, P->vd, true, P->place_d, st)) return bail(fail(DDF_EHIP, "plan_place failed"));
  // bins
  uint32_t ni = 0;
  int64_t nr = 0;
  rc = build_bins(P, P->pg, P->vg, P->place_g, chunk, &P->recs_g, &P->items_g, &ni, &nr, &P->pg.item_range,
                  &P->pg.nranges, true, st, &P->tile_first_g, &P->tile_count_g);
  if (rc) return bail(rc);
  P->pg.nitems_b = ni;
  P->stats[0] = nr;
  P->stats[2] = ni;
  P->bstamp_g = P->alloc<uint64_t>(size_t(nr));
  if (!P->bstamp_g || batched_stamp_records(P->pg, P->place_g, P->recs_g, nr, P->bstamp_g, st))
    return bail(fail(DDF_ENOMEM, "plan allocation failed"));
  // the grid map's placements reduced to pair codes (3.5 GB -> 0.44 GB at the 81-facet chunk)
  {
    const size_t np = size_t(vgrid->nblocks) * nfacets;
    P->pcode_g = P->alloc<uint8_t>(np);
    if (!P->pcode_g || batched_pair_codes(P->place_g, int64_t(np), P->pcode_g, st) ||
        hipStreamSynchronize(st) != hipSuccess)
      return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    P->release(P->place_g, np);
  }

  {
    // tile slots of the gridding items (plain stores + a deterministic reduction instead of
    // global atomics); without the memory for them the tiles are flushed with atomics
    const int W = tile + 2 * halo;
    const int npg = smg.nstokes < g->npol ? smg.nstokes : g->npol;
    const size_t n = size_t(ni) * size_t(npg) * 2 * size_t(W) * size_t(W);
    if (ni) {
      P->slots_g = P->alloc<float>(n);
      if (!P->slots_g) (void)hipGetLastError();
    }
  }
  rc = build_bins(P, P->pd, P->vd, P->place_d, chunk, &P->recs_d, &P->items_d, &ni, &nr, &P->pd.item_range,
                  &P->pd.nranges, false, st);
  if (rc) return bail(rc);
  P->pd.nitems_b = ni;
  P->stats[1] = nr;
  P->stats[3] = ni;
  P->bstamp_d = P->alloc<uint64_t>(size_t(nr));
  if (!P->bstamp_d || batched_stamp_records(P->pd, P->place_d, P->recs_d, nr, P->bstamp_d, st))
    return bail(fail(DDF_ENOMEM, "plan allocation failed"));
  {
    const size_t np = size_t(vdegrid->nblocks) * nfacets;
    P->pcode_d = P->alloc<uint8_t>(np);
    if (!P->pcode_d || batched_pair_codes(P->place_d, int64_t(np), P->pcode_d, st) ||
        hipStreamSynchronize(st) != hipSuccess)
      return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    P->release(P->place_d, np);
  }

  // pass A layout
  const int64_t bpi = o.blocks_per_item > 0 ? o.blocks_per_item : 128;
  P->pg.blocks_per_item = bpi;
  P->pg.nchunks = (vgrid->nblocks + bpi - 1) / bpi;
  if (int64_t(P->pg.ngroups) * P->pg.nchunks >= (int64_t(1) << 32)) return bail(fail(DDF_EINVAL, "too many work items"));
  P->pg.nitems_a = uint32_t(int64_t(P->pg.ngroups) * P->pg.nchunks);
  P->pd.blocks_per_item = bpi;
  P->pd.nchunks = (vdegrid->nblocks + bpi - 1) / bpi;
  if (int64_t(P->pd.ngroups) * P->pd.nchunks >= (int64_t(1) << 32)) return bail(fail(DDF_EINVAL, "too many work items"));
  P->pd.nitems_a = uint32_t(int64_t(P->pd.ngroups) * P->pd.nchunks);
  const size_t per_f = size_t(g->nband) * g->npol + 2 * size_t(g->nband);
  if (size_t(nfacets) * per_f * 8 > 40 * 1024) return bail(fail(DDF_EINVAL, "facets x bands too large for one plan"));
  // pass A LDS: per-facet sums, visibility staging, and (when they fit in 80 KB with the rest,
  // so that two workgroups share a compute unit) the group's per-channel Jones sums
  const size_t stage_bytes = batched_grid_a_stage_bytes(G, P->pg.read);
  P->pg.chan_lds = (P->pg.jmode != 0) &&
                   (size_t(nfacets) * per_f + size_t(G) * (2 * size_t(vgrid->nchan) + 1)) * 8 + stage_bytes <= 80 * 1024;
  if (size_t(nfacets) * per_f * 8 + stage_bytes > 80 * 1024)
    return bail(fail(DDF_EINVAL, "facets x bands too large for one plan"));
  // LDS tiles
  const int W = tile + 2 * halo;
  const int npg = smg.nstokes < g->npol ? smg.nstokes : g->npol;
  if (size_t(npg > smd.nstokes ? npg : smd.nstokes) * W * W * 8 > 64 * 1024)
    return bail(fail(DDF_EINVAL, "tile too large for LDS: lower opts.tile"));
  // record values
  P->V = P->alloc<float2>(size_t(vgrid->nblocks) * nfacets * smg.nstokes);
  P->M = P->alloc<float2>(size_t(vdegrid->nblocks) * nfacets * smd.nstokes);
  if (!P->V || !P->M) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
  if (hipMemsetAsync(P->M, 0, size_t(vdegrid->nblocks) * nfacets * smd.nstokes * 8, st) != hipSuccess)
    return bail(fail(DDF_EHIP, "memset failed"));
  // visibility list of the degrid map
  if (vdegrid->nblocks) {
    int32_t* cnt = P->alloc<int32_t>(size_t(vdegrid->nblocks));
    int64_t* vs = P->alloc<int64_t>(size_t(vdegrid->nblocks));
    int64_t* tmp = P->alloc<int64_t>(size_t((vdegrid->nblocks + kScanChunk - 1) / kScanChunk));
    if (!cnt || !vs || !tmp) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    if (batched_vis_counts(P->vd, cnt, st)) return bail(fail(DDF_EHIP, "launch failed"));
    rc = launch_scan(cnt, vdegrid->nblocks, 0, vs, tmp, st);
    if (rc) return bail(rc);
    int64_t last_start = 0;
    int32_t last_cnt = 0;
    if (hipMemcpyAsync(&last_start, vs + vdegrid->nblocks - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&last_cnt, cnt + vdegrid->nblocks - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bail(fail(DDF_EHIP, "copy failed"));
    P->nvl = last_start + last_cnt;
    P->vrc = P->alloc<uint64_t>(size_t(P->nvl));
    P->vblk = P->alloc<uint32_t>(size_t(P->nvl));
    if (!P->vrc || !P->vblk) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    if (batched_vlist(P->vd, vs, P->vrc, P->vblk, st)) return bail(fail(DDF_EHIP, "launch failed"));
    // runs of consecutive entries with one block and one Jones key: degrid pass B forms each
    // facet's Jones product once per run
    if (P->nvl >= (int64_t(1) << 32)) return bail(fail(DDF_EINVAL, "degrid visibility list too long for one plan"));
    Temp rt(st);
    int32_t* flag = static_cast<int32_t*>(rt.get(size_t(P->nvl) * sizeof(int32_t)));
    int64_t* excl = static_cast<int64_t*>(rt.get(size_t(P->nvl) * sizeof(int64_t)));
    int64_t* ctmp = static_cast<int64_t*>(rt.get(size_t((P->nvl + kScanChunk - 1) / kScanChunk) * sizeof(int64_t)));
    P->vrun = P->alloc<uint32_t>(size_t(P->nvl));
    if (!flag || !excl || !ctmp || !P->vrun) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    if (batched_run_ids(P->pd, P->vd, P->vrc, P->vblk, P->nvl, flag, nullptr, nullptr, nullptr, 0, st))
      return bail(fail(DDF_EHIP, "launch failed"));
    rc = launch_scan(flag, P->nvl, 0, excl, ctmp, st);
    if (rc) return bail(rc);
    int64_t last_excl = 0;
    int32_t last_flag = 0;
    if (hipMemcpyAsync(&last_excl, excl + P->nvl - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&last_flag, flag + P->nvl - 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bail(fail(DDF_EHIP, "copy failed"));
    P->nruns = last_excl + last_flag;
    P->run_first = P->alloc<uint32_t>(size_t(P->nruns));
    if (!P->run_first) return bail(fail(DDF_ENOMEM, "plan allocation failed"));
    if (batched_run_ids(P->pd, P->vd, P->vrc, P->vblk, P->nvl, flag, excl, P->vrun, P->run_first, 1, st))
      return bail(fail(DDF_EHIP, "launch failed"));
  }
  P->stats[4] = P->nvl;
  P->stats[6] = tile;
  P->stats[7] = G;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return bail(fail(DDF_EHIP, "device query failed"));
  P->nwg = unsigned(o.workgroups > 0 ? o.workgroups : prop.multiProcessorCount * 4);
  if (o.workgroups > 0) {
    for (unsigned& x : P->nwg_tiles) x = unsigned(o.workgroups);
  } else if (batched_tile_wgs(P->pg, P->pd, P->nwg_tiles)) {
    return bail(fail(DDF_EHIP, "occupancy query failed"));
  }
  if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(DDF_EHIP, "plan build failed"));
  *out = P;
  return DDF_OK;
}

extern "C" int ddf_plan_grid(ddf_plan* P, float* vis, int32_t dopsf, void* stream) {
  if (!P) return fail(DDF_EINVAL, "null plan");
  hipStream_t st = static_cast<hipStream_t>(stream);
  VisP v = P->vg;
  if (vis) v.vis = reinterpret_cast<float2*>(vis);
  BatchParams p = P->pg;
  p.dopsf = dopsf;
  if (prepare_jones(P, st)) return fail(DDF_EHIP, "Jones preparation failed");
  if (batched_reset(P->queues, 1 + kItemRanges, st)) return fail(DDF_EHIP, "launch failed");
  P->mark(DDF_PASS_GRID_A, 0, st);
  if (p.nitems_a && batched_grid_a(p, v, P->pcode_g, P->V, P->queues, P->nwg, st)) return fail(DDF_EHIP, "grid pass A launch failed");
  P->mark(DDF_PASS_GRID_A, 1, st);
  P->mark(DDF_PASS_GRID_B, 0, st);
  if (p.nitems_b && batched_grid_b(p, P->bstamp_g, P->V, P->items_g, P->recs_g, P->queues + 1, P->nwg_tiles[0], P->slots_g,
                                     P->tile_first_g, P->tile_count_g, st))
    return fail(DDF_EHIP, "grid pass B launch failed");
  P->mark(DDF_PASS_GRID_B, 1, st);
  return DDF_OK;
}

extern "C" int ddf_plan_degrid(ddf_plan* P, float* vis, void* stream) {
  if (!P) return fail(DDF_EINVAL, "null plan");
  hipStream_t st = static_cast<hipStream_t>(stream);
  VisP v = P->vd;
  if (vis) v.vis = reinterpret_cast<float2*>(vis);
  if (prepare_jones(P, st)) return fail(DDF_EHIP, "Jones preparation failed");
  if (batched_reset(P->queues + 1 + kItemRanges, kItemRanges, st)) return fail(DDF_EHIP, "launch failed");
  P->mark(DDF_PASS_DEGRID_A, 0, st);
  if (P->pd.nitems_b && batched_degrid_a(P->pd, P->bstamp_d, P->M, P->items_d, P->recs_d, P->queues + 1 + kItemRanges, P->nwg_tiles[1], st))
    return fail(DDF_EHIP, "degrid pass A launch failed");
  P->mark(DDF_PASS_DEGRID_A, 1, st);
  P->mark(DDF_PASS_DEGRID_B, 0, st);
  if (P->nvl && batched_degrid_b(P->pd, v, P->M, P->vrc, P->vblk, P->vrun, P->run_first, P->nvl, st))
    return fail(DDF_EHIP, "degrid pass B launch failed");
  P->mark(DDF_PASS_DEGRID_B, 1, st);
  return DDF_OK;
}

extern "C" int ddf_plan_degrid_adjoint(ddf_plan* P, const float* vis_bar, void* stream) {
  if (!P || !vis_bar) return fail(DDF_EINVAL, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prepare_jones(P, st)) return fail(DDF_EHIP, "Jones preparation failed");
  if (batched_reset(P->queues + 1 + 2 * kItemRanges, 1 + kItemRanges, st)) return fail(DDF_EHIP, "launch failed");
  const float2* vb = reinterpret_cast<const float2*>(vis_bar);
  P->mark(DDF_PASS_ADJ_B, 0, st);
  if (P->pd.nitems_a && batched_degrid_b_adjoint(P->pd, P->vd, P->pcode_d, vb, P->M, P->queues + 1 + 2 * kItemRanges, P->nwg, st))
    return fail(DDF_EHIP, "degrid adjoint pass B launch failed");
  P->mark(DDF_PASS_ADJ_B, 1, st);
  P->mark(DDF_PASS_ADJ_A, 0, st);
  if (P->pd.nitems_b && batched_degrid_a_adjoint(P->pd, P->bstamp_d, P->M, P->items_d, P->recs_d, P->queues + 2 + 2 * kItemRanges, P->nwg_tiles[2], st))
    return fail(DDF_EHIP, "degrid adjoint pass A launch failed");
  P->mark(DDF_PASS_ADJ_A, 1, st);
  return DDF_OK;
}

extern "C" int ddf_plan_info(const ddf_plan* P, int64_t* out, int32_t n) {
  if (!P || !out) return fail(DDF_EINVAL, "null argument");
  for (int i = 0; i < n && i < 8; ++i) out[i] = P->stats[i];
  for (int i = 8; i < n && i < 11; ++i) out[i] = int64_t(P->nwg_tiles[i - 8]);
  return DDF_OK;
}

extern "C" int ddf_plan_placements(const ddf_plan* P, int32_t degrid, uint64_t* out, int64_t n, void* stream) {
  if (!P || !out) return fail(DDF_EINVAL, "null argument");
  const BatchParams& p = degrid ? P->pd : P->pg;
  const int64_t nb = degrid ? P->vd.nblocks : P->vg.nblocks;
  const int64_t need = nb * int64_t(p.nf);
  if (n < need) return fail(DDF_EINVAL, "output holds fewer than nblocks * nfacets records");
  if (need == 0) return DDF_OK;
  // recomputed into out (the plan keeps only the pair codes): the same kernel on the same inputs
  if (batched_place(p, degrid ? P->vd : P->vg, degrid != 0, out, static_cast<hipStream_t>(stream)))
    return fail(DDF_EHIP, "plan_place failed");
  return DDF_OK;
}

extern "C" int ddf_plan_timing(ddf_plan* P, int32_t enable) {
  if (!P) return fail(DDF_EINVAL, "null plan");
  if (enable && !P->ev[0][0]) {
    for (auto& e : P->ev)
      for (hipEvent_t& x : e) HIP_TRY(hipEventCreate(&x));
  }
  P->timing = enable ? 1 : 0;
  return DDF_OK;
}

extern "C" int ddf_plan_pass_ms(ddf_plan* P, float* ms, int32_t n) {
  if (!P || !ms) return fail(DDF_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) {
    ms[i] = -1.f;
    if (i >= DDF_PLAN_NPASS || !P->ev_set[i]) continue;
    HIP_TRY(hipEventSynchronize(P->ev[i][1]));
    HIP_TRY(hipEventElapsedTime(&ms[i], P->ev[i][0], P->ev[i][1]));
  }
  return DDF_OK;
}

extern "C" int ddf_plan_destroy(ddf_plan* P) {
  delete P;
  return DDF_OK;
}
