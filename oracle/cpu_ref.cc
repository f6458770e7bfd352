// TEST INFRASTRUCTURE -- the CPU oracle of the facet gridder / degridder.
//
// A sequential C++ restatement of the reference native gridder semantics, taking the same
// argument structs as the HIP library (include/ddfgpu.h) but with HOST pointers.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker
// (or the timed CPU baseline); the product path never routes through it.
//
// Followed, function by function:
//   ref_grid     DDFacet/Gridder/gridder.h:78-377      (gridder::gridder, one facet)
//   ref_degrid   DDFacet/Gridder/degridder.h:55-266    (degridder::degridder, one facet)
//   stokes maps  DDFacet/Gridder/Stokes.h:18-226, dispatch GridderSmearPols.cc:128-206, 261-342
//   jones        DDFacet/Gridder/JonesServer.cc:25-305 (NormJones, GiveJones, updateJones)
//   decorr       DDFacet/Gridder/DecorrelationHelper.cc:42-63
//   phasor       DDFacet/Gridder/CorrelationCalculator.h:37-65
// Pinned by tests/test_oracle.py against golden vectors written by the reference module built
// from source (tests/golden/make_golden.py), plus closed forms (delta stamps, DFT, Stokes tables).
//
// Deliberate differences (documented, none reachable with valid inputs):
//   * rows >= nrow are skipped (the reference tests irow>nrows and reads one row past the end);
//   * an oversampling offset that lands outside [0, overs) (posx exactly on a .5 boundary) is
//     clamped into range instead of reading before/after the CF plane.
#include "../include/ddfgpu.h"
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
using cf = std::complex<float>;
constexpr double kC = 299792458.0;
constexpr double kPi = 3.141592653589793238462643383279502884197;

thread_local std::string g_err;

struct M2 {
  cd a[4];
};
inline M2 unity() { return M2{{cd(1), cd(0), cd(0), cd(1)}}; }
inline M2 mul(const M2& x, const M2& y) {
  return M2{{x.a[0] * y.a[0] + x.a[1] * y.a[2], x.a[0] * y.a[1] + x.a[1] * y.a[3],
             x.a[2] * y.a[0] + x.a[3] * y.a[2], x.a[2] * y.a[1] + x.a[3] * y.a[3]}};
}
inline M2 herm(const M2& x) {
  return M2{{std::conj(x.a[0]), std::conj(x.a[2]), std::conj(x.a[1]), std::conj(x.a[3])}};
}

// ---- correlation <-> Stokes policies as linear maps --------------------------------------
// Grid side: stokes[k] = sum_j G[k][j] * corr[j] (corr padded to 2x2 as the Read policies do);
// degrid side: corr[k] = sum_j D[k][j] * stokes[j].
struct Policy {
  int read;          // 4: Read_4Corr, 2: Read_2Corr_Pad, 1: Read_1Corr_Pad
  int nstokes;
  cd G[4][4];
  cd D[4][4];
};

bool codes_eq(const uint16_t* c, int n, std::initializer_list<int> want) {
  if (n != (int)want.size()) return false;
  int i = 0;
  for (int w : want)
    if (c[i++] != w) return false;
  return true;
}

const cd H(0.5, 0.0), mH(-0.5, 0.0), iH(0.0, 0.5), miH(0.0, -0.5), I1(0.0, 1.0);

// Rows of the grid-side maps (Stokes.h:18-34).
void grow(cd* r, char s, int lin) {  // lin: 1 linear (XX..), 0 circular (RR..), 2 Stokes-I input
  for (int j = 0; j < 4; ++j) r[j] = 0;
  if (lin == 2) {  // I_FROM_I
    r[0] = 1;
    return;
  }
  if (s == 'I') { r[0] = H; r[3] = H; return; }
  if (lin == 1) {
    if (s == 'Q') { r[0] = H; r[3] = mH; }
    if (s == 'U') { r[1] = H; r[2] = H; }
    if (s == 'V') { r[1] = miH; r[2] = iH; }   // -(V1-V2)*0.5i
  } else {
    if (s == 'Q') { r[1] = H; r[2] = H; }
    if (s == 'U') { r[1] = miH; r[2] = iH; }
    if (s == 'V') { r[0] = H; r[3] = mH; }
  }
}

// Dispatch of pyGridderWPol (GridderSmearPols.cc:128-206), including the {I,U} -> IQ quirk.
bool grid_policy(const ddf_vis* v, Policy* p) {
  const uint16_t* c = v->corr_codes;
  const uint16_t* s = v->stokes_codes;
  int nc = v->ncorr_codes, ns = v->nstokes;
  for (int i = 0; i < nc; ++i)
    if (!((c[i] >= 5 && c[i] <= 12) || c[i] == 1)) {
      g_err = "Only accepts I,RR,RL,LR,LL,XX,XY,YX,YY as correlation input type";
      return false;
    }
  for (int i = 0; i < ns; ++i)
    if (s[i] < 1 || s[i] > 4) {
      g_err = "Only accepts I,Q,U,V as polarization output type";
      return false;
    }
  std::memset((void*)p, 0, sizeof(*p));
  const char* names = "?IQUV";
  std::string want;
  for (int i = 0; i < ns; ++i) want += names[s[i]];
  int lin = -1;
  std::vector<std::string> ok;
  if (codes_eq(c, nc, {9, 10, 11, 12})) {
    p->read = 4; lin = 1;
    ok = {"I", "Q", "U", "V", "IQ", "IU", "IV", "QU", "IQU", "IQUV"};
  } else if (codes_eq(c, nc, {5, 6, 7, 8})) {
    p->read = 4; lin = 0;
    ok = {"I", "IQ", "IU", "IV", "QU", "IQU", "IQUV"};
  } else if (codes_eq(c, nc, {9, 12})) {
    p->read = 2; lin = 1;
    ok = {"I", "IQ"};
  } else if (codes_eq(c, nc, {5, 8})) {
    p->read = 2; lin = 0;
    ok = {"I", "IV"};
  } else if (codes_eq(c, nc, {1})) {
    p->read = 1; lin = 2;
    ok = {"I"};
  }
  bool found = false;
  for (auto& o : ok) found |= (o == want);
  if (!found) {
    g_err = "Cannot convert input correlations to desired output Stokes parameters.";
    return false;
  }
  std::string eff = want;
  if (lin == 1 && p->read == 4 && want == "IU") eff = "IQ";  // GridderSmearPols.cc:142-143
  p->nstokes = ns;
  for (int k = 0; k < ns; ++k) grow(p->G[k], eff[k], lin);
  return true;
}

// Dispatch of pyDeGridderWPol (GridderSmearPols.cc:261-342) with the Stokes.h:90-192 maps.
bool degrid_policy(const ddf_vis* v, Policy* p) {
  const uint16_t* c = v->corr_codes;
  const uint16_t* s = v->stokes_codes;
  int nc = v->ncorr_codes, ns = v->nstokes;
  for (int i = 0; i < nc; ++i)
    if (!((c[i] >= 5 && c[i] <= 12) || c[i] == 1)) {
      g_err = "Only accepts I,RR,RL,LR,LL,XX,XY,YX,YY as correlation output types";
      return false;
    }
  std::memset((void*)p, 0, sizeof(*p));
  std::string want;
  for (int i = 0; i < ns; ++i) {
    if (s[i] < 1 || s[i] > 4) {
      g_err = "Cannot convert input Stokes parameter to desired output correlations.";
      return false;
    }
    want += "?IQUV"[s[i]];
  }
  bool X4 = codes_eq(c, nc, {9, 10, 11, 12}), R4 = codes_eq(c, nc, {5, 6, 7, 8});
  bool X2 = codes_eq(c, nc, {9, 12}), R2 = codes_eq(c, nc, {5, 8}), I1c = codes_eq(c, nc, {1});
  auto& D = p->D;
  bool done = true;
  if (want == "I" && (X4 || X2 || R4 || R2 || I1c)) {
    D[0][0] = 1; D[3][0] = 1;
  } else if (want == "Q" && X4) {
    D[0][0] = 1; D[3][0] = -1;
  } else if (want == "U" && X4) {
    D[1][0] = 1; D[2][0] = 1;
  } else if (want == "V" && X4) {
    D[1][0] = I1; D[2][0] = -I1;
  } else if (want == "IQUV" && X4) {
    D[0][0] = 1; D[0][1] = 1; D[1][2] = 1; D[1][3] = I1; D[2][2] = 1; D[2][3] = -I1; D[3][0] = 1; D[3][1] = -1;
  } else if (want == "IQUV" && R4) {
    D[0][0] = 1; D[0][3] = 1; D[1][1] = 1; D[1][2] = I1; D[2][1] = 1; D[2][2] = -I1; D[3][0] = 1; D[3][3] = -1;
  } else if (want == "IQ" && (X4 || X2)) {
    D[0][0] = 1; D[0][1] = 1; D[3][0] = 1; D[3][1] = -1;
  } else if (want == "IQ" && R4) {
    D[0][0] = 1; D[1][1] = 1; D[2][1] = 1; D[3][0] = 1;
  } else if (want == "IU" && X4) {
    D[0][0] = 1; D[1][1] = 1; D[2][1] = 1; D[3][0] = 1;
  } else if (want == "IU" && R4) {
    D[0][0] = 1; D[1][1] = I1; D[2][1] = -I1; D[3][0] = 1;
  } else if (want == "IV" && (R4 || R2)) {
    D[0][0] = 1; D[0][1] = 1; D[3][0] = 1; D[3][1] = -1;
  } else if (want == "IV" && X4) {
    D[0][0] = 1; D[1][1] = I1; D[2][1] = -I1; D[3][0] = 1;
  } else if (want == "QU" && X4) {
    D[0][0] = 1; D[1][1] = 1; D[2][1] = 1; D[3][0] = -1;
  } else if (want == "IQU" && X4) {
    D[0][0] = 1; D[0][1] = 1; D[1][2] = 1; D[2][2] = 1; D[3][0] = 1; D[3][1] = -1;
  } else if (want == "IQU" && R4) {
    D[0][0] = 1; D[1][1] = 1; D[1][2] = I1; D[2][1] = 1; D[2][2] = -I1; D[3][0] = 1;
  } else {
    done = false;
  }
  if (!done) {
    g_err = "Cannot convert input Stokes parameter to desired output correlations.";
    return false;
  }
  p->read = X4 || R4 ? 4 : (X2 || R2 ? 2 : 1);
  p->nstokes = ns;
  return true;
}

// ---- Jones server (JonesServer.cc) ----------------------------------------------------------
struct Jones {
  const ddf_jones* j;
  double wl_mean;
  M2 J0, J1, J0H, J1H;
  double BB = 0;
  int cur_a0 = -1, cur_a1 = -1, kt = -1, kc = -1, bt = -1, bc = -1;
  M2 J0b = unity(), J1b = unity(), J0k = unity(), J1k = unity();

  M2 give(const float* tab, const int32_t* dims, int it, int ant, int dir, int ich) const {
    size_t nd = dims[1], na = dims[2], nch = dims[3];
    size_t off = ((size_t(it) * nd + size_t(dir)) * na + size_t(ant)) * nch * 4 + size_t(ich) * 4;
    M2 out;
    if (j->interp_mode == 0) {
      for (int p = 0; p < 4; ++p) out.a[p] = cd(tab[2 * (off + p)], tab[2 * (off + p) + 1]);
      return out;
    }
    // "Krigging" mode: the direction offset never varies (JonesServer.cc:71-93), so the result
    // is (sum of non-zero coefs)^2 times the Jones of direction `dir`.
    double jabs[4] = {0, 0, 0, 0};
    for (int p = 0; p < 4; ++p) out.a[p] = 0;
    for (int d = 0; d < j->ncoefs; ++d) {
      double c = j->coefs[d];
      if (c == 0) continue;
      for (int p = 0; p < 4; ++p) {
        cd val(tab[2 * (off + p)], tab[2 * (off + p) + 1]);
        double A = std::abs(val);
        out.a[p] += c / A * val;
        jabs[p] += c * A;
      }
    }
    for (int p = 0; p < 4; ++p) out.a[p] *= jabs[p];
    return out;
  }
  void norm(M2& J, const double* uvw) const {
    if (!j->apply_amp)
      for (int p = 0; p < 4; ++p) {
        double a = std::abs(J.a[p]);
        if (a != 0.) J.a[p] /= a;
      }
    if (!j->apply_phase)
      for (int p = 0; p < 4; ++p) J.a[p] = std::abs(J.a[p]);
    if (j->scale_jones) {
      double R2 = (uvw[0] * uvw[0] + uvw[1] * uvw[1]) / (wl_mean * wl_mean);
      double al = std::exp(-2. * kPi * j->calib_error * j->calib_error * R2);
      for (int p = 0; p < 4; ++p) {
        double a = std::abs(J.a[p]);
        if (a != 0.) J.a[p] *= (1. - al) / a + al;
      }
    }
  }
  void reset() { cur_a0 = cur_a1 = kt = kc = bt = bc = -1; }
  bool update(size_t row, size_t chan, const double* uvw) {
    int a0 = j->a0[row], a1 = j->a1[row];
    bool blchg = a0 != cur_a0 || a1 != cur_a1;
    bool chg = false;
    cur_a0 = a0;
    cur_a1 = a1;
    if (j->beam) {
      int t = j->beam_time[row], c = j->beam_chan[chan];
      if (blchg || bt != t || bc != c) {
        J0b = give(j->beam, j->beam_dims, t, a0, j->beam_dir, c);
        J1b = give(j->beam, j->beam_dims, t, a1, j->beam_dir, c);
        bt = t; bc = c; chg = true;
      }
    }
    if (j->kms) {
      int t = j->kms_time[row], c = j->kms_chan[chan];
      if (blchg || kt != t || kc != c) {
        J0k = give(j->kms, j->kms_dims, t, a0, j->kms_dir, c);
        J1k = give(j->kms, j->kms_dims, t, a1, j->kms_dir, c);
        norm(J0k, uvw);
        norm(J1k, uvw);
        kt = t; kc = c; chg = true;
      }
    }
    if (chg) {
      J0 = unity();
      J1 = unity();
      if (j->beam) { J0 = mul(J0b, J0); J1 = mul(J1b, J1); }
      if (j->kms) { J0 = mul(J0k, J0); J1 = mul(J1k, J1); }
      BB = (std::abs(J0.a[0]) * std::abs(J1.a[0]) + std::abs(J0.a[3]) * std::abs(J1.a[3])) / 2.;
      BB *= BB;
      J0H = herm(J0);
      J1H = herm(J1);
    }
    return chg;
  }
};

double decorr_get(const ddf_decorr& d, const double* uvw, double nu, size_t idx) {
  if (!d.enabled) return 1.;
  double n0 = std::sqrt(1. - d.l0 * d.l0 - d.m0 * d.m0) - 1.;
  double f = 1.;
  if (d.fsmear) {
    double ph = uvw[3 * idx] * d.l0 + uvw[3 * idx + 1] * d.m0 + uvw[3 * idx + 2] * n0;
    double phi = kPi * d.dnu / kC * ph;
    f *= (phi == 0.) ? 1.0 : std::sin(phi) / phi;
  }
  if (d.tsmear) {
    const double* q = d.uvw_dt + 3 * idx;
    double dph = (q[0] * d.l0 + q[1] * d.m0 + q[2] * n0) * d.dt;
    double phi = kPi * nu / kC * dph;
    f *= (phi == 0.) ? 1.0 : std::sin(phi) / phi;
  }
  return f;
}

// Per-row phasor with the equidistant-channel recurrence (CorrelationCalculator.h:37-65).
struct Phasor {
  bool equi;
  cd cur, step;
  int cur_ch = -1;
  void reset() { cur_ch = -1; }
  bool get(const double* f, size_t ch, double angle, cd* out) {
    if (!equi) {
      *out = std::polar(1., f[ch] * angle);
      return true;
    }
    if (cur_ch == -1) {
      cur = std::polar(1., f[ch] * angle);
      step = std::polar(1., (f[1] - f[0]) * angle);
      cur_ch = int(ch);
    } else if (cur_ch > int(ch)) {
      g_err = "Reverse channel ordering: the BDA mapping is invalid";
      return false;
    } else {
      while (size_t(cur_ch) < ch) {
        cur *= step;
        cur_ch++;
      }
    }
    *out = cur;
    return true;
  }
};

struct BlockIter {
  const int32_t* map;
  int64_t nblocks;
  const int32_t* sizes;
  const int32_t* rec;
  explicit BlockIter(const ddf_vis* v) {
    map = v->bda;
    nblocks = int64_t(uint32_t(map[0])) + (int64_t(map[1]) << 32);
    sizes = map + 2;
    rec = map + 2 + nblocks;
  }
};

// Oversampled CF offset of the reference (gridder.h:336-351) with the .5-boundary clamp.
struct Place {
  int locx, locy, sup, S, plane;
  int64_t cfoff;
};

bool place(const ddf_grid_desc* g, const ddf_cf& cf, double Umean, double Vmean, double Wmean,
           double recipWvl, Place* pl) {
  int iw = int(std::lrint((double(cf.nw) - 1) * std::fabs(Wmean) * (cf.ref_wave * recipWvl) / cf.wmax));
  if (iw >= cf.nw) return false;
  int plane = (Wmean > 0) ? iw : cf.nw + iw;
  int S = cf.support[iw];
  int O = cf.overs;
  int sup = (S - 1) / 2;
  double scale0 = g->nx * g->incr[0], scale1 = g->nx * g->incr[1];
  double posx = scale0 * Umean * recipWvl + double(g->nx / 2);
  double posy = scale1 * Vmean * recipWvl + double(g->ny / 2);
  int locx = int(std::lrint(posx)), locy = int(std::lrint(posy));
  if (locx - sup < 0 || locx + sup >= g->nx || locy - sup < 0 || locy + sup >= g->ny) return false;
  int offx = int(std::lrint((locx - posx) * O) + (S * O - 1) / 2);
  int offy = int(std::lrint((locy - posy) * O) + (S * O - 1) / 2);
  int io = offy - sup * O, jo = offx - sup * O;
  io = io < 0 ? 0 : (io >= O ? O - 1 : io);
  jo = jo < 0 ? 0 : (jo >= O ? O - 1 : jo);
  pl->locx = locx; pl->locy = locy; pl->sup = sup; pl->S = S; pl->plane = plane;
  pl->cfoff = cf.offset[plane] + int64_t(io * O + jo) * S * S;
  return true;
}

template <typename GT>
int grid_impl(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f) {
  Policy pol;
  if (!grid_policy(v, &pol)) return DDF_EINVAL;
  BlockIter bi(v);
  const double Cu = f->cu, Cv = f->cv, l0 = f->l0, m0 = f->m0;
  const double n0 = std::sqrt(1 - l0 * l0 - m0 * m0) - 1;
  const size_t nrow = size_t(v->nrow), nchan = size_t(v->nchan), ncorr = size_t(v->ncorr);
  GT* grid = reinterpret_cast<GT*>(f->grid);
  double wl_mean = 0, fmean0 = 0;
  for (size_t c = 0; c < nchan; ++c) {
    wl_mean += kC / v->freqs[c];
    fmean0 += v->freqs[c];
  }
  wl_mean /= double(nchan);
  fmean0 /= double(nchan);
  Jones JS{&f->jones, wl_mean};
  int jmode = f->jones.enabled ? f->jones.mode : 0;
  Phasor ph{v->chan_equidistant != 0};
  std::vector<double> sjc(nchan), swc(nchan);
  const int32_t* rec = bi.rec;
  for (int64_t b = 0; b < bi.nblocks; ++b) {
    JS.reset();
    const int nr = bi.sizes[b] - 2;
    const size_t ch0 = size_t(rec[0]), ch1 = size_t(rec[1]);
    const int32_t* rows = rec + 2;
    rec += bi.sizes[b];
    if (v->sparsification && !v->sparsification[b]) continue;
    cd Vis[4] = {0, 0, 0, 0};
    for (size_t c = 0; c < nchan; ++c) sjc[c] = swc[c] = 0;
    size_t midrow = size_t(rows[nr / 2]);
    double dec = midrow < nrow ? decorr_get(f->decorr, v->uvw, fmean0, midrow) : 1.;
    double fmean = 0, sw = 0, sj = 0, ssw = 0, um = 0, vm = 0, wm = 0;
    int nvis = 0;
    cd meas[4] = {dec, 0, 0, dec};
    bool have_psf = false;
    for (int r = 0; r < nr; ++r) {
      size_t row = size_t(rows[r]);
      if (row >= nrow) continue;
      const double* uvw = v->uvw + 3 * row;
      const double U = uvw[0], V = uvw[1], W = uvw[2];
      const double angle = -2. * kPi * (U * l0 + V * m0 + W * n0) / kC;
      ph.reset();
      for (size_t c = ch0; c < ch1; ++c) {
        size_t doff = (row * nchan + c) * ncorr;
        if (v->flags[doff]) continue;
        cd corr = 1;
        if (!g->dopsf && !ph.get(v->freqs, c, angle, &corr)) return DDF_ERUNTIME;
        if (jmode == 1) {
          bool chg = JS.update(row, c, uvw);
          if ((chg || !have_psf) && g->dopsf) {
            M2 m = mul(JS.J0, JS.J1H);
            m = mul(mul(JS.J0H, m), JS.J1);
            for (int p = 0; p < 4; ++p) meas[p] = m.a[p] * dec;
            have_psf = true;
          }
        }
        if (!g->dopsf) {
          const float* vp = v->vis + 2 * doff;
          if (pol.read == 4) {
            for (int p = 0; p < 4; ++p) meas[p] = cd(vp[2 * p], vp[2 * p + 1]);
          } else if (pol.read == 2) {
            meas[0] = cd(vp[0], vp[1]); meas[1] = meas[2] = 0; meas[3] = cd(vp[2], vp[3]);
          } else {
            meas[0] = cd(vp[0], vp[1]); meas[1] = meas[2] = 0; meas[3] = meas[0];
          }
        }
        const double fw = double(v->weights[row * nchan + c]);
        const cd wt = fw * corr;
        const double fwd = fw * dec * dec;
        ssw += fwd;
        swc[c] += fwd;
        if (jmode == 1) {
          if (!g->dopsf) {
            M2 m{{meas[0], meas[1], meas[2], meas[3]}};
            m = mul(mul(JS.J0H, m), JS.J1);
            for (int p = 0; p < 4; ++p) meas[p] = m.a[p];
          }
          sj += JS.BB * fwd;
          sjc[c] += JS.BB * fwd;
        }
        if (pol.read == 4) {
          for (int p = 0; p < 4; ++p) Vis[p] += meas[p] * wt;
        } else if (pol.read == 2) {
          Vis[0] += meas[0] * wt;
          Vis[3] += meas[3] * wt;
        } else {
          Vis[0] += meas[0] * wt;
        }
        um += U + W * Cu;
        vm += V + W * Cv;
        wm += W;
        fmean += v->freqs[c];
        sw += fw;
        ++nvis;
      }
    }
    if (nvis == 0) continue;
    um /= nvis; vm /= nvis; wm /= nvis; fmean /= nvis;
    if (jmode == 2) {
      double uvwm[3] = {um, vm, wm};
      JS.update(size_t(rows[nr / 2]), (ch0 + ch1) / 2, uvwm);
      M2 m{{Vis[0], Vis[1], Vis[2], Vis[3]}};
      if (g->dopsf) m = mul(mul(JS.J0, m), JS.J1H);
      m = mul(mul(JS.J0H, m), JS.J1);
      for (int p = 0; p < 4; ++p) Vis[p] = m.a[p];
      sj = ssw * JS.BB;
      for (size_t c = ch0; c < ch1; ++c) sjc[c] = swc[c] * JS.BB;
    }
    const int gchan = v->chan_map[ch0];
    cd stokes[4];
    for (int k = 0; k < 4; ++k) {
      stokes[k] = 0;
      for (int p = 0; p < 4; ++p) stokes[k] += pol.G[k][p] * Vis[p];
    }
    if (gchan < 0 || gchan >= g->nband) continue;
    const double recipWvl = fmean / kC;
    Place pl;
    if (!place(g, f->cf, um, vm, wm, recipWvl, &pl)) continue;
    const float* cfp = f->cf.data + 2 * pl.cfoff;
    for (int ip = 0; ip < pol.nstokes; ++ip) {
      if (ip >= g->npol) continue;
      size_t goff = (size_t(gchan) * g->npol + ip) * size_t(g->nx) * g->ny;
      const cd val = stokes[ip];
      for (int sy = -pl.sup; sy <= pl.sup; ++sy) {
        GT* gp = grid + 2 * (goff + size_t(pl.locy + sy) * g->nx + pl.locx);
        for (int sx = -pl.sup; sx <= pl.sup; ++sx) {
          int k = (sy + pl.sup) * pl.S + (sx + pl.sup);
          cd prod = val * cd(cfp[2 * k], cfp[2 * k + 1]);
          gp[2 * sx] = GT(double(gp[2 * sx]) + prod.real());
          gp[2 * sx + 1] = GT(double(gp[2 * sx + 1]) + prod.imag());
        }
      }
      f->sumwt[ip + gchan * g->npol] += sw;
      if (jmode) {
        f->jones.sum_jones[gchan] += sj;
        f->jones.sum_jones[gchan + g->nband] += ssw;
        for (size_t c = ch0; c < ch1; ++c) {
          f->jones.sum_jones_chan[c] += sjc[c];
          f->jones.sum_jones_chan[nchan + c] += swc[c];
        }
      }
    }
  }
  return DDF_OK;
}

std::mutex g_row_locks[3373];

int degrid_impl(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, bool locked) {
  Policy pol;
  if (!degrid_policy(v, &pol)) return DDF_EINVAL;
  if (pol.nstokes > g->npol) {
    g_err = "grid has fewer polarisations than the requested Stokes parameters";
    return DDF_EINVAL;
  }
  BlockIter bi(v);
  const double Cu = f->cu, Cv = f->cv, l0 = f->l0, m0 = f->m0;
  const double n0 = std::sqrt(1 - l0 * l0 - m0 * m0) - 1;
  const size_t nrow = size_t(v->nrow), nchan = size_t(v->nchan), ncorr = size_t(v->ncorr);
  const float* grid = reinterpret_cast<const float*>(f->grid);
  double wl_mean = 0;
  for (size_t c = 0; c < nchan; ++c) wl_mean += kC / v->freqs[c];
  wl_mean /= double(nchan);
  Jones JS{&f->jones, wl_mean};
  int jmode = f->jones.enabled ? f->jones.mode : 0;
  Phasor ph{v->chan_equidistant != 0};
  const int32_t* rec = bi.rec;
  for (int64_t b = 0; b < bi.nblocks; ++b) {
    const int nr = bi.sizes[b] - 2;
    const size_t ch0 = size_t(rec[0]), ch1 = size_t(rec[1]);
    const int32_t* rows = rec + 2;
    rec += bi.sizes[b];
    const int gchan = v->chan_map[ch0];
    if (gchan < 0 || gchan >= g->nband) continue;
    double fmean = 0;
    for (size_t c = ch0; c < ch1; ++c) fmean += v->freqs[c];
    fmean /= double(ch1 - ch0);
    int nv = 0;
    double um = 0, vm = 0, wm = 0;
    for (int r = 0; r < nr; ++r) {
      size_t row = size_t(rows[r]);
      if (row >= nrow) continue;
      const double* uvw = v->uvw + 3 * row;
      um += uvw[0] + uvw[2] * Cu;
      vm += uvw[1] + uvw[2] * Cv;
      wm += uvw[2];
      ++nv;
    }
    if (nv == 0) continue;
    um /= nv; vm /= nv; wm /= nv;
    const double recipWvl = fmean / kC;
    Place pl;
    if (!place(g, f->cf, um, vm, wm, recipWvl, &pl)) continue;
    const float* cfp = f->cf.data + 2 * pl.cfoff;
    cd st[4] = {0, 0, 0, 0};
    for (int ip = 0; ip < pol.nstokes; ++ip) {
      size_t goff = (size_t(gchan) * g->npol + ip) * size_t(g->nx) * g->ny;
      cd s = 0;
      for (int sy = -pl.sup; sy <= pl.sup; ++sy) {
        const float* gp = grid + 2 * (goff + size_t(pl.locy + sy) * g->nx + pl.locx);
        for (int sx = -pl.sup; sx <= pl.sup; ++sx) {
          int k = (sy + pl.sup) * pl.S + (sx + pl.sup);
          cf prod = cf(gp[2 * sx], gp[2 * sx + 1]) * cf(cfp[2 * k], cfp[2 * k + 1]);
          s += cd(prod);
        }
      }
      st[ip] = s;
    }
    M2 cv;
    for (int k = 0; k < 4; ++k) {
      cv.a[k] = 0;
      for (int p = 0; p < 4; ++p) cv.a[k] += pol.D[k][p] * st[p];
    }
    size_t midrow = size_t(rows[nr / 2]);
    if (jmode == 2 && midrow < nrow) {
      JS.update(midrow, (ch0 + ch1) / 2, v->uvw + 3 * midrow);
      cv = mul(mul(JS.J0, cv), JS.J1H);
    }
    double dec = midrow < nrow ? decorr_get(f->decorr, v->uvw, fmean, midrow) : 1.;
    for (int r = 0; r < nr; ++r) {
      size_t row = size_t(rows[r]);
      if (row >= nrow) continue;
      const double* uvw = v->uvw + 3 * row;
      const double angle = 2. * kPi * (uvw[0] * l0 + uvw[1] * m0 + uvw[2] * n0) / kC;
      ph.reset();
      std::unique_lock<std::mutex> lk;
      if (locked) lk = std::unique_lock<std::mutex>(g_row_locks[row % 3373]);
      for (size_t c = ch0; c < ch1; ++c) {
        size_t doff = (row * nchan + c) * ncorr;
        if (jmode == 1) JS.update(row, c, uvw);
        cd corr;
        if (!ph.get(v->freqs, c, angle, &corr)) return DDF_ERUNTIME;
        corr *= dec;
        M2 vb;
        if (jmode == 1) {
          vb = mul(mul(JS.J0, cv), JS.J1H);
          for (int p = 0; p < 4; ++p) vb.a[p] *= corr;
        } else {
          for (int p = 0; p < 4; ++p) vb.a[p] = cv.a[p] * corr;
        }
        float* vp = v->vis + 2 * doff;
        auto sub = [&](int slot, int p) {
          vp[2 * slot] = float(double(vp[2 * slot]) - vb.a[p].real());
          vp[2 * slot + 1] = float(double(vp[2 * slot + 1]) - vb.a[p].imag());
        };
        if (ncorr == 4) {
          for (int p = 0; p < 4; ++p) sub(p, p);
        } else if (ncorr == 2) {
          sub(0, 0);
          sub(1, 3);
        } else {
          sub(0, 0);
        }
      }
    }
  }
  return DDF_OK;
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

int ref_grid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f) {
  if (g->grid_f64) return grid_impl<double>(g, v, f);
  return grid_impl<float>(g, v, f);
}

int ref_degrid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f) {
  return degrid_impl(g, v, f, false);
}

// Integer placement maps of every BDA block for one facet: the block means (gridder.h:211-289 over
// unflagged visibilities for gridding, degridder.h:141-167 over all rows for degridding) and the
// placement of gridder.h:321-351 / degridder.h:170-200 with the .5-boundary clamp of `place`.
// out: int32 [nblocks][8] = valid, iw, conj (W <= 0), locx, locy, io, jo, band.
int ref_place(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, int degrid, int32_t* out) {
  BlockIter bi(v);
  const double Cu = f->cu, Cv = f->cv;
  const size_t nrow = size_t(v->nrow), nchan = size_t(v->nchan), ncorr = size_t(v->ncorr);
  const int32_t* rec = bi.rec;
  for (int64_t b = 0; b < bi.nblocks; ++b) {
    int32_t* o = out + 8 * b;
    for (int k = 0; k < 8; ++k) o[k] = 0;
    const int nr = bi.sizes[b] - 2;
    const size_t ch0 = size_t(rec[0]), ch1 = size_t(rec[1]);
    const int32_t* rows = rec + 2;
    rec += bi.sizes[b];
    if (!degrid && v->sparsification && !v->sparsification[b]) continue;
    double um = 0, vm = 0, wm = 0, fmean = 0;
    int nv = 0;
    if (!degrid) {
      for (int r = 0; r < nr; ++r) {
        size_t row = size_t(rows[r]);
        if (row >= nrow) continue;
        const double* uvw = v->uvw + 3 * row;
        const double U = uvw[0], V = uvw[1], W = uvw[2];
        for (size_t c = ch0; c < ch1; ++c) {
          if (v->flags[(row * nchan + c) * ncorr]) continue;
          um += U + W * Cu;
          vm += V + W * Cv;
          wm += W;
          fmean += v->freqs[c];
          ++nv;
        }
      }
      if (nv == 0) continue;
      um /= nv; vm /= nv; wm /= nv; fmean /= nv;
    } else {
      for (size_t c = ch0; c < ch1; ++c) fmean += v->freqs[c];
      fmean /= double(ch1 - ch0);
      for (int r = 0; r < nr; ++r) {
        size_t row = size_t(rows[r]);
        if (row >= nrow) continue;
        const double* uvw = v->uvw + 3 * row;
        um += uvw[0] + uvw[2] * Cu;
        vm += uvw[1] + uvw[2] * Cv;
        wm += uvw[2];
        ++nv;
      }
      if (nv == 0) continue;
      um /= nv; vm /= nv; wm /= nv;
    }
    const int gchan = v->chan_map[ch0];
    if (gchan < 0 || gchan >= g->nband) continue;
    Place pl;
    if (!place(g, f->cf, um, vm, wm, fmean / kC, &pl)) continue;
    const int nw = f->cf.nw, O = f->cf.overs;
    const int64_t sub = (pl.cfoff - f->cf.offset[pl.plane]) / (int64_t(pl.S) * pl.S);
    o[0] = 1;
    o[1] = pl.plane % nw;
    o[2] = (wm > 0) ? 0 : 1;
    o[3] = pl.locx;
    o[4] = pl.locy;
    o[5] = int32_t(sub / O);
    o[6] = int32_t(sub % O);
    o[7] = gchan;
  }
  return DDF_OK;
}

// All facets of one major cycle, one facet per thread, as the reference process pool runs them
// (DDFacet/Other/AsyncProcessPool.py:831-861; one APP job per facet, ClassFacetMachine.py:1589).
int ref_grid_facets(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, int nf, int nthreads) {
  int rc = DDF_OK;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(min : rc)
#endif
  for (int i = 0; i < nf; ++i) {
    int r = ref_grid(g, v, f + i);
    rc = r < rc ? r : rc;
  }
  (void)nthreads;
  return rc;
}

// Degridding of all facets into one shared vis, serialised per row as the reference's named
// semaphores do (degridder.h:236-263, Semaphores.cc:27-29, 3373 semaphores).
int ref_degrid_facets(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, int nf, int nthreads) {
  int rc = DDF_OK;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(min : rc)
#endif
  for (int i = 0; i < nf; ++i) {
    int r = degrid_impl(g, v, f + i, true);
    rc = r < rc ? r : rc;
  }
  (void)nthreads;
  return rc;
}

}  // extern "C"
