// Beam Jones producers on the device: the analytic primary-beam E-Jones tables of
// ClassJones.EstimateBeam for the GMRT and ATCA models, and the cell products of the Jones domain
// merge (ClassJonesDomains.MergeJones).  The tables they write have the layout of the ddf_jones
// inputs ([t][dir][ant][chan][2][2] complex64), so the gridders read them without a host trip.
//
//   beam     one thread per (direction, Jones channel), the phase centre as an extra direction:
//            the angular distance (ClassGMRTBeam.AngDist, DDFacet/Data/ClassGMRTBeam.py:29-40),
//            the model of every channel of the Jones channel (GMRT: a polynomial in d nu / 1e9
//            clipped past its first stationary point, ClassGMRTBeam.py:140-159; ATCA: the square
//            roots of two band polynomials mixed with the reference's weights, its value past
//            100 arcmin GHz, ClassATCABeam.py:191-233), their mean (GiveInstrumentBeam), in FP64
//            with the reference's operation order (Horner's rule as numpy's polyval, no
//            contraction); then one thread per output matrix writes the diagonal Jones, times the
//            inverse of the centre beam when CenterNorm is on (ClassJones.py:1045-1053).
//   merge    one thread per output matrix: the 2 x 2 complex64 product of ModLinAlg.BatchDot
//            (DDFacet/Array/ModLinAlg.py:94-134) of the two tables' cells, the identity where
//            either table has no cell for the merged channel.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ddfgpu.h"

namespace ddf {
void set_last_error(const std::string& msg);  // ddfgpu.hip: the message ddf_last_error() returns
}

namespace {

int jfail(int code, const std::string& m) {
  ddf::set_last_error(m);
  return code;
}

unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return unsigned(b < 1 ? 1 : b);
}

constexpr double kArcminPerRad = 180.0 / 3.141592653589793 * 60;

// numpy.polynomial.polynomial.polyval: c0 = c[n-1]; c0 = c[i] + c0 x for i = n-2 .. 0
__device__ inline double polyval_asc(const double* c, int n, double x) {
#pragma clang fp contract(off)
  double v = c[n - 1];
  for (int i = n - 2; i >= 0; --i) v = c[i] + v * x;
  return v;
}

// Beam value of channel ich at angular distance d (arcmin).
__device__ inline double beam_value(const ddf_beam_desc& b, int ich, double d) {
#pragma clang fp contract(off)
  const double f = b.freqs[ich];
  const double* c0 = b.coef0 + int64_t(ich) * b.ncoef;
  if (b.kind == DDF_BEAM_GMRT) {
    const double dnu = d * f / 1e9;
    double v = polyval_asc(c0, b.ncoef, dnu);
    if (d > double(b.x_null[ich])) v = double(b.y_null[ich]);
    return v;
  }
  // ATCA: bands at or below / at or above the channel
  const double* c1 = b.coef1 + int64_t(ich) * b.ncoef;
  const double f00 = b.band_f0[2 * ich], f01 = b.band_f0[2 * ich + 1];
  const double dnu = d * f * 1e-9;
  const double dnu0 = d * f00 * 1e-9, dnu1 = d * f01 * 1e-9;
  const double df0 = fabs(f - f00), df1 = fabs(f - f01);
  double v, ynull;
  if (df0 == 0. && df1 == 0.) {
    v = sqrt(fabs(polyval_asc(c0, b.ncoef, dnu)));
    ynull = sqrt(fabs(polyval_asc(c0, b.ncoef, 100.)));
  } else {
    const double b0 = sqrt(fabs(polyval_asc(c0, b.ncoef, dnu0)));
    const double b1 = sqrt(fabs(polyval_asc(c1, b.ncoef, dnu1)));
    v = (df0 * b0 + df1 * b1) / (df0 + df1);
    const double y0 = sqrt(fabs(polyval_asc(c0, b.ncoef, 100.)));
    const double y1 = sqrt(fabs(polyval_asc(c1, b.ncoef, 100.)));
    ynull = (df0 * y0 + df1 * y1) / (df0 + df1);
  }
  // the reference keeps the null value in a float32 array
  const double yn = double(float(ynull));
  if (dnu > double(b.x_null[ich])) v = yn / (dnu - yn);
  return v;
}

// mean[dir][j] for dir < ndir, and dir == ndir the phase centre (distance 0 by construction of
// the reference's call with the centre's own coordinates).
__global__ void beam_mean_k(const ddf_beam_desc b, const double* ra, const double* dec, int ndir, double ra_c,
                            double dec_c, double* mean) {
  const int64_t n = int64_t(ndir + 1) * b.nchan_jones;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int dir = int(i / b.nchan_jones), j = int(i - int64_t(dir) * b.nchan_jones);
    const double r = dir < ndir ? ra[dir] : ra_c, de = dir < ndir ? dec[dir] : dec_c;
    double dd;
    {
#pragma clang fp contract(off)
      double c = sin(de) * sin(dec_c) + cos(de) * cos(dec_c) * cos(r - ra_c);
      c = c > 1. ? 1. : (c < -1. ? -1. : c);
      dd = acos(c) * kArcminPerRad;
    }
    double s = 0;
    int cnt = 0;
    for (int ich = 0; ich < b.nchan; ++ich) {
      if (b.chan_to_jones[ich] != j) continue;
      s += beam_value(b, ich, dd);
      ++cnt;
    }
    // a Jones channel without channels: the mean of an empty set, NaN (numpy's mean)
    mean[i] = s / double(cnt);
  }
}

__global__ void beam_write_k(const double* mean, int ndir, int nchj, int ntime, int na, int center_norm,
                             float4* out) {
  const int64_t per_t = int64_t(ndir) * na * nchj;
  const int64_t n = per_t * ntime;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t k = i % per_t;
    const int j = int(k % nchj);
    const int dir = int(k / (int64_t(na) * nchj));
    double v = mean[int64_t(dir) * nchj + j];
    if (center_norm) {
      // BatchInverse of diag(B0, B0) then BatchDot with diag(B, B), the zero parts dropping out:
      // (B0 (1 / (B0 B0))) B
#pragma clang fp contract(off)
      const double b0 = mean[int64_t(ndir) * nchj + j];
      const double det = 1. / (b0 * b0);
      v = (b0 * det) * v;
    }
    const float x = float(v);
    if (x != x) {
      // an empty Jones channel: every element of the reference's mean matrix is NaN
      out[2 * i] = make_float4(x, x, x, x);
      out[2 * i + 1] = make_float4(x, x, x, x);
      continue;
    }
    out[2 * i] = make_float4(x, 0.f, 0.f, 0.f);
    out[2 * i + 1] = make_float4(0.f, 0.f, x, 0.f);
  }
}

struct C32 {
  float re, im;
};
__device__ __forceinline__ C32 cmul(C32 a, C32 b) {
#pragma clang fp contract(off)
  return C32{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ C32 cadd(C32 a, C32 b) { return C32{a.re + b.re, a.im + b.im}; }

__global__ void jones_merge_k(const ddf_jones_merge_desc m) {
  const int64_t cell = int64_t(m.nd) * m.na;
  const int64_t n = int64_t(m.nt) * cell * m.nch;
  const float2* j0 = static_cast<const float2*>(m.j0);
  const float2* j1 = static_cast<const float2*>(m.j1);
  float2* out = static_cast<float2*>(m.out);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int c = int(i % m.nch);
    const int64_t da = (i / m.nch) % cell;
    const int t = int(i / (int64_t(m.nch) * cell));
    const int c0 = m.ich0[c], c1 = m.ich1[c];
    const int t0 = m.it0[t], t1 = m.it1[t];
    float2* o = out + 4 * i;
    if (c0 >= m.nch0 || c1 >= m.nch1 || t0 < 0 || t0 >= m.nt0 || t1 < 0 || t1 >= m.nt1) {
      // an index map entry outside its table: NaN, not a read past the table
      const float q = __builtin_nanf("");
      for (int k = 0; k < 4; ++k) o[k] = make_float2(q, q);
      continue;
    }
    if (c0 < 0 || c1 < 0) {
      o[0] = make_float2(1.f, 0.f);
      o[1] = make_float2(0.f, 0.f);
      o[2] = make_float2(0.f, 0.f);
      o[3] = make_float2(1.f, 0.f);
      continue;
    }
    const float2* a = j0 + 4 * ((int64_t(t0) * cell + da) * m.nch0 + c0);
    const float2* b = j1 + 4 * ((int64_t(t1) * cell + da) * m.nch1 + c1);
    const C32 a00{a[0].x, a[0].y}, a01{a[1].x, a[1].y}, a10{a[2].x, a[2].y}, a11{a[3].x, a[3].y};
    const C32 b00{b[0].x, b[0].y}, b01{b[1].x, b[1].y}, b10{b[2].x, b[2].y}, b11{b[3].x, b[3].y};
    const C32 r00 = cadd(cmul(a00, b00), cmul(a01, b10));
    const C32 r01 = cadd(cmul(a00, b01), cmul(a01, b11));
    const C32 r10 = cadd(cmul(a10, b00), cmul(a11, b10));
    const C32 r11 = cadd(cmul(a10, b01), cmul(a11, b11));
    o[0] = make_float2(r00.re, r00.im);
    o[1] = make_float2(r01.re, r01.im);
    o[2] = make_float2(r10.re, r10.im);
    o[3] = make_float2(r11.re, r11.im);
  }
}

}  // namespace

extern "C" {

int ddf_beam_jones(const ddf_beam_desc* b, const double* ra, const double* dec, int32_t ndir, double ra_c,
                   double dec_c, int32_t ntime, int32_t na, int32_t center_norm, void* jones, void* stream) {
  if (!b) return jfail(DDF_EINVAL, "ddf_beam_jones: null descriptor");
  if (b->kind != DDF_BEAM_GMRT && b->kind != DDF_BEAM_ATCA) return jfail(DDF_EINVAL, "ddf_beam_jones: unknown beam model");
  if (b->nchan <= 0 || b->nchan_jones <= 0 || b->ncoef <= 0 || b->ncoef > 16)
    return jfail(DDF_EINVAL, "ddf_beam_jones: nchan, nchan_jones > 0 and 0 < ncoef <= 16 required");
  if (ndir < 0 || ntime < 0 || na < 0) return jfail(DDF_EINVAL, "ddf_beam_jones: negative table size");
  if (!b->freqs || !b->coef0 || !b->x_null || !b->y_null || !b->chan_to_jones ||
      (b->kind == DDF_BEAM_ATCA && (!b->coef1 || !b->band_f0)) || (ndir && (!ra || !dec)) ||
      (int64_t(ntime) * ndir * na && !jones))
    return jfail(DDF_EINVAL, "ddf_beam_jones: null array");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t nmean = int64_t(ndir + 1) * b->nchan_jones;
  double* mean = nullptr;
  if (hipMallocAsync(reinterpret_cast<void**>(&mean), size_t(nmean) * sizeof(double), st) != hipSuccess)
    return jfail(DDF_ENOMEM, "ddf_beam_jones: workspace allocation failed");
  hipLaunchKernelGGL(beam_mean_k, dim3(blocks_for(nmean)), dim3(256), 0, st, *b, ra, dec, ndir, ra_c, dec_c, mean);
  const int64_t nout = int64_t(ntime) * ndir * na * b->nchan_jones;
  if (nout)
    hipLaunchKernelGGL(beam_write_k, dim3(blocks_for(nout)), dim3(256), 0, st, mean, ndir, b->nchan_jones, ntime, na,
                       center_norm, static_cast<float4*>(jones));
  const hipError_t e = hipGetLastError();
  hipFreeAsync(mean, st);
  return e == hipSuccess ? DDF_OK : jfail(DDF_EHIP, std::string("ddf_beam_jones: ") + hipGetErrorString(e));
}

int ddf_jones_merge(const ddf_jones_merge_desc* m, void* stream) {
  if (!m) return jfail(DDF_EINVAL, "ddf_jones_merge: null descriptor");
  if (m->nt < 0 || m->nch < 0 || m->nd < 0 || m->na < 0 || m->nt0 <= 0 || m->nt1 <= 0 || m->nch0 <= 0 || m->nch1 <= 0)
    return jfail(DDF_EINVAL, "ddf_jones_merge: bad table sizes");
  const int64_t n = int64_t(m->nt) * m->nd * m->na * m->nch;
  if (!n) return DDF_OK;
  if (!m->j0 || !m->j1 || !m->out || !m->it0 || !m->it1 || !m->ich0 || !m->ich1)
    return jfail(DDF_EINVAL, "ddf_jones_merge: null array");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(jones_merge_k, dim3(blocks_for(n)), dim3(256), 0, st, *m);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DDF_OK : jfail(DDF_EHIP, std::string("ddf_jones_merge: ") + hipGetErrorString(e));
}

}  // extern "C"
