// Device-side building blocks of the gfx950 facet gridder / degridder.
//
// Complex and 2x2 Jones algebra, 64-lane wavefront reductions, the per-visibility phasor, the
// Jones-table lookup of DDEs::JonesServer (DDFacet/Gridder/JonesServer.cc:25-298), the
// decorrelation factor (DecorrelationHelper.cc:42-63) and the BDA-block placement arithmetic
// (gridder.h:319-351, degridder.h:170-200) evaluated without FMA contraction so that the integer
// maps it produces (w-plane, grid cell, oversampling offset) are those of the reference.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

namespace ddf {

// Global-memory views of pointers.  Pointers reached through a struct (kernel-argument structs,
// facet descriptors) are generic to the compiler, and generic loads are flat loads, which count
// on both the vector-memory and the LDS counters, so every LDS wait also waits for them (and the
// reverse).  Every buffer these kernels touch is device global memory; these views give the loads,
// stores and atomics their global forms (gptr for reads, gptr_w for writes and atomics).
#define DDF_GLOBAL __attribute__((address_space(1)))
// four floats read or written as one 16-byte access (a native vector type, which copies across
// address spaces)
typedef float v4f __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ const DDF_GLOBAL T* gptr(const T* p) {
  return (const DDF_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ DDF_GLOBAL T* gptr_w(T* p) {
  return (DDF_GLOBAL T*)p;
}
// A struct read from global memory (word by word: structs do not copy across address spaces).
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "gload reads whole 32-bit words");
  T out;
  const DDF_GLOBAL uint32_t* src = gptr(reinterpret_cast<const uint32_t*>(p));
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) dst[i] = src[i];
  return out;
}
// Complex values in global memory (the vector classes do not copy across address spaces).
__device__ __forceinline__ float2 ld2(const DDF_GLOBAL float2* p) {
  const DDF_GLOBAL float* q = reinterpret_cast<const DDF_GLOBAL float*>(p);
  return make_float2(q[0], q[1]);
}
__device__ __forceinline__ void st2(DDF_GLOBAL float2* p, float2 v) {
  DDF_GLOBAL float* q = reinterpret_cast<DDF_GLOBAL float*>(p);
  q[0] = v.x;
  q[1] = v.y;
}
// Relaxed device-scope atomic add on global memory (returns nothing).
template <typename T>
__device__ __forceinline__ void atomic_add_g(DDF_GLOBAL T* p, T v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


constexpr double kC = 299792458.0;
constexpr double kPi = 3.141592653589793238462643383279502884197;
// exp(i kPhaseScale (u l + v m + w n) nu) is the phasor of a visibility (nu in Hz)
constexpr double kPhaseScale = -2. * kPi / kC;
constexpr int kWave = 64;

// ---- complex -------------------------------------------------------------------------------
// cx<T>: complex in FP64 (cd, the reference's dcmplx) or FP32 (cf); M2T<T>: 2 x 2 complex matrix
// (DDF::Mat, JonesServer.h:38-83) with M2 = M2T<double>.
template <typename T>
struct cx {
  T re, im;
};
using cd = cx<double>;
using cf = cx<float>;
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>{a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>{a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return cx<T>{a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return cx<T>{a.re, -a.im}; }
__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ double cabs(cd a) { return hypot(a.re, a.im); }
template <typename T, typename U>
__device__ __forceinline__ cx<T> cvt(cx<U> a) { return cx<T>{T(a.re), T(a.im)}; }

template <typename T>
struct M2T {
  cx<T> a[4];
};
using M2 = M2T<double>;
__device__ __forceinline__ M2 unity() { return M2{{cd{1, 0}, cd{0, 0}, cd{0, 0}, cd{1, 0}}}; }
template <typename T>
__device__ __forceinline__ M2T<T> mul(const M2T<T>& x, const M2T<T>& y) {
  return M2T<T>{{x.a[0] * y.a[0] + x.a[1] * y.a[2], x.a[0] * y.a[1] + x.a[1] * y.a[3],
                 x.a[2] * y.a[0] + x.a[3] * y.a[2], x.a[2] * y.a[1] + x.a[3] * y.a[3]}};
}
template <typename T>
__device__ __forceinline__ M2T<T> herm(const M2T<T>& x) {
  return M2T<T>{{conj(x.a[0]), conj(x.a[2]), conj(x.a[1]), conj(x.a[3])}};
}
template <typename T, typename U>
__device__ __forceinline__ M2T<T> cvt(const M2T<U>& x) {
  return M2T<T>{{cvt<T>(x.a[0]), cvt<T>(x.a[1]), cvt<T>(x.a[2]), cvt<T>(x.a[3])}};
}

// ---- wavefront reductions (all lanes receive the total) ------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ---- phasor ---------------------------------------------------------------------------------
// exp(i*phase): the phase, formed in FP64, is reduced in FP64 to r = phase - q pi/2 with
// |r| <= pi/4 (pi/2 split in a head and a tail), then sin r and cos r are evaluated in FP32 by
// their Taylor polynomials (degrees 9 and 8: truncation below 3e-8 on |r| <= pi/4) and placed by
// the quadrant q mod 4.  The error is ~1e-7 for |phase| up to 1e15 rad (tested to 1e8; beyond,
// the FP64 phase itself is coarser than a radian), where an FP32 phase of 1e4 rad would already
// carry ~1e-3 rad.
// q is rounded by adding and subtracting 1.5 2^52 (exact for |phase 2/pi| < 2^51: the sum then
// lies in [2^52, 2^53), where the FP64 spacing is 1), whose low mantissa bits hold q mod 2^32.
__host__ __device__ __forceinline__ cd phasor(double phase) {
  const double two_over_pi = 0.63661977236758134308;
  const double pio2_hi = 1.57079632679489655800;
  const double pio2_lo = 6.12323399573676603587e-17;
  const double kRound = 6755399441055744.0;  // 1.5 2^52
  const double t = phase * two_over_pi + kRound;
  const double q = t - kRound;
  double r = fma(-q, pio2_hi, phase);
  r = fma(-q, pio2_lo, r);
  const float x = float(r), x2 = x * x;
  const float sp = x * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 1.f / 362880.f, -1.f / 5040.f), 1.f / 120.f), -1.f / 6.f), 1.f);
  const float cp = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 1.f / 40320.f, -1.f / 720.f), 1.f / 24.f), -0.5f), 1.f);
  const int k = int(static_cast<uint32_t>(__builtin_bit_cast(unsigned long long, t)) & 3u);
  const float s = (k & 1) ? cp : sp, c = (k & 1) ? sp : cp;
  const float ss = (k & 2) ? -s : s;
  const float cc = ((k + 1) & 2) ? -c : c;
  return cd{double(cc), double(ss)};
}

// exp(2 pi i rev) for a phase in revolutions (the mixed precision of the batched passes): the
// phase is reduced to r = rev - rint(rev) in [-1/2, 1/2] in FP64, then the hardware sine and
// cosine of gfx950 (v_sin_f32 / v_cos_f32, whose argument is in revolutions) give the phasor.
// Measured against FP64 sin / cos on gfx950: 1.8e-7 maximum absolute error over phases up to
// 5.5e3 rad (the error of r in FP32 is below 3e-8 revolutions at any phase).
__device__ __forceinline__ float2 phasor_rev(double rev) {
  const float x = float(rev - rint(rev));
  return make_float2(__builtin_amdgcn_cosf(x), __builtin_amdgcn_sinf(x));
}
constexpr double kInvTwoPi = 0.15915494309189533577;

// ---- Jones tables (JonesServer.cc) -----------------------------------------------------------
struct JonesP {
  int enabled, mode, interp;
  const float2* kms;
  int kdims[4];
  const int32_t *ktime, *kchan;
  int kdir;
  const float2* beam;
  int bdims[4];
  const int32_t *btime, *bchan;
  int bdir;
  const int32_t *a0, *a1;
  const float* coefs;
  int ncoefs;
  int apply_amp, apply_phase, scale;
  double calib_error;
  double* sum_jones;
  double* sum_jones_chan;
  // prepared tables of the batched plan (facet-fastest [t][ant][chan][dir][4], FP64): the DD
  // table with interpolation and NormJones amplitude / phase applied, the beam table with
  // interpolation applied; null when the raw tables above are read
  const double2* pk;
  const double2* pb;
  // the same prepared tables rounded to complex64 (read by the FP32 arithmetic of the mixed and
  // FP32 precisions)
  const float2* pkf;
  const float2* pbf;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// GiveJones (JonesServer.cc:54-95); indices clamped to the table so a bad map cannot fault.
__device__ inline M2 give_jones(const float2* tab, const int* dims, int it, int ant, int dir, int ich,
                                int interp, const float* coefs, int ncoefs) {
  it = clampi(it, 0, dims[0] - 1);
  dir = clampi(dir, 0, dims[1] - 1);
  ant = clampi(ant, 0, dims[2] - 1);
  ich = clampi(ich, 0, dims[3] - 1);
  size_t off = ((size_t(it) * dims[1] + dir) * size_t(dims[2]) + ant) * size_t(dims[3]) + ich;
  const DDF_GLOBAL float2* p = gptr(tab) + off * 4;
  M2 out;
  if (interp == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) out.a[q] = cd{p[q].x, p[q].y};
    return out;
  }
  // Interpolation mode 1 reads the same direction for every coefficient (JonesServer.cc:71-93):
  // the result is (sum of non-zero coefs)^2 times the Jones of `dir`, accumulated in the same
  // order as the reference.
  double jabs[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q) out.a[q] = cd{0, 0};
  for (int d = 0; d < ncoefs; ++d) {
    double c = gptr(coefs)[d];
    if (c == 0) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cd val{p[q].x, p[q].y};
      double A = cabs(val);
      out.a[q] = out.a[q] + val * (c / A);
      jabs[q] += c * A;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) out.a[q] = out.a[q] * jabs[q];
  return out;
}

// ScaleJones part of NormJones (JonesServer.cc:44-51): amplitudes pulled towards 1 by a factor
// that depends on the baseline length of the visibility.
__device__ inline void scale_jones(const JonesP& jp, M2& J, const double* uvw, double wl_mean) {
  double R2 = (uvw[0] * uvw[0] + uvw[1] * uvw[1]) / (wl_mean * wl_mean);
  double al = exp(-2. * kPi * jp.calib_error * jp.calib_error * R2);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double a = cabs(J.a[q]);
    if (a != 0.) J.a[q] = J.a[q] * ((1. - al) / a + al);
  }
}

// NormJones (JonesServer.cc:25-52), applied to the DD (killMS) Jones only.
__device__ inline void norm_jones(const JonesP& jp, M2& J, const double* uvw, double wl_mean) {
  if (!jp.apply_amp) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double a = cabs(J.a[q]);
      if (a != 0.) J.a[q] = cd{J.a[q].re / a, J.a[q].im / a};
    }
  }
  if (!jp.apply_phase) {
#pragma unroll
    for (int q = 0; q < 4; ++q) J.a[q] = cd{cabs(J.a[q]), 0.};
  }
  if (jp.scale) scale_jones(jp, J, uvw, wl_mean);
}

struct JonesPair {
  M2 J0, J1;
  double BB;
};

// updateJones (JonesServer.cc:193-298) evaluated afresh for (row, chan): J = JkMS * JBeam per
// antenna, BB = ((|J0_00||J1_00| + |J0_11||J1_11|)/2)^2.  `uvw` feeds ScaleJones only.
// Entry of a prepared table (JonesP::pk / pb), indices clamped like give_jones.
__device__ __forceinline__ M2 prepared_jones(const double2* tab, const int* dims, int it, int ant, int dir, int ich) {
  it = clampi(it, 0, dims[0] - 1);
  dir = clampi(dir, 0, dims[1] - 1);
  ant = clampi(ant, 0, dims[2] - 1);
  ich = clampi(ich, 0, dims[3] - 1);
  const DDF_GLOBAL double2* p = gptr(tab) + (((size_t(it) * dims[2] + ant) * size_t(dims[3]) + ich) * size_t(dims[1]) + dir) * 4;
  M2 out;
#pragma unroll
  for (int q = 0; q < 4; ++q) out.a[q] = cd{p[q].x, p[q].y};
  return out;
}

// PREP: read the prepared tables (JonesP::pk / pb, the batched plan), else the raw tables.  Both
// give the same values; the prepared path has the per-entry work done once.
template <bool PREP = false>
__device__ inline JonesPair jones_at(const JonesP& jp, int64_t row, int chan, const double* uvw,
                                     double wl_mean) {
  int a0 = gptr(jp.a0)[row], a1 = gptr(jp.a1)[row];
  if (PREP) {
    JonesPair r;
    if (jp.pk) {
      const int t = gptr(jp.ktime)[row], c = gptr(jp.kchan)[chan];
      r.J0 = prepared_jones(jp.pk, jp.kdims, t, a0, jp.kdir, c);
      r.J1 = prepared_jones(jp.pk, jp.kdims, t, a1, jp.kdir, c);
      if (jp.scale) {
        scale_jones(jp, r.J0, uvw, wl_mean);
        scale_jones(jp, r.J1, uvw, wl_mean);
      }
      if (jp.pb) {
        const int tb = gptr(jp.btime)[row], cb = gptr(jp.bchan)[chan];
        r.J0 = mul(r.J0, prepared_jones(jp.pb, jp.bdims, tb, a0, jp.bdir, cb));
        r.J1 = mul(r.J1, prepared_jones(jp.pb, jp.bdims, tb, a1, jp.bdir, cb));
      }
    } else {
      const int tb = gptr(jp.btime)[row], cb = gptr(jp.bchan)[chan];
      r.J0 = prepared_jones(jp.pb, jp.bdims, tb, a0, jp.bdir, cb);
      r.J1 = prepared_jones(jp.pb, jp.bdims, tb, a1, jp.bdir, cb);
    }
    const double bb = (cabs(r.J0.a[0]) * cabs(r.J1.a[0]) + cabs(r.J0.a[3]) * cabs(r.J1.a[3])) / 2.;
    r.BB = bb * bb;
    return r;
  }
  M2 J0 = unity(), J1 = unity();
  if (jp.beam) {
    int t = gptr(jp.btime)[row], c = gptr(jp.bchan)[chan];
    J0 = give_jones(jp.beam, jp.bdims, t, a0, jp.bdir, c, jp.interp, jp.coefs, jp.ncoefs);
    J1 = give_jones(jp.beam, jp.bdims, t, a1, jp.bdir, c, jp.interp, jp.coefs, jp.ncoefs);
  }
  if (jp.kms) {
    int t = gptr(jp.ktime)[row], c = gptr(jp.kchan)[chan];
    M2 K0 = give_jones(jp.kms, jp.kdims, t, a0, jp.kdir, c, jp.interp, jp.coefs, jp.ncoefs);
    M2 K1 = give_jones(jp.kms, jp.kdims, t, a1, jp.kdir, c, jp.interp, jp.coefs, jp.ncoefs);
    norm_jones(jp, K0, uvw, wl_mean);
    norm_jones(jp, K1, uvw, wl_mean);
    J0 = mul(K0, J0);
    J1 = mul(K1, J1);
  }
  JonesPair r;
  r.J0 = J0;
  r.J1 = J1;
  double bb = (cabs(J0.a[0]) * cabs(J1.a[0]) + cabs(J0.a[3]) * cabs(J1.a[3])) / 2.;
  r.BB = bb * bb;
  return r;
}

// jones_at<true> for a key read beforehand: antennas a0 / a1, DD Jones time and channel kt / kc,
// beam Jones time and channel bt / bc (the tables that are absent are not read).  No ScaleJones.
__device__ __forceinline__ JonesPair jones_keyed(const JonesP& jp, int a0, int a1, int kt, int kc, int bt, int bc) {
  JonesPair r;
  if (jp.pk) {
    r.J0 = prepared_jones(jp.pk, jp.kdims, kt, a0, jp.kdir, kc);
    r.J1 = prepared_jones(jp.pk, jp.kdims, kt, a1, jp.kdir, kc);
    if (jp.pb) {
      r.J0 = mul(r.J0, prepared_jones(jp.pb, jp.bdims, bt, a0, jp.bdir, bc));
      r.J1 = mul(r.J1, prepared_jones(jp.pb, jp.bdims, bt, a1, jp.bdir, bc));
    }
  } else {
    r.J0 = prepared_jones(jp.pb, jp.bdims, bt, a0, jp.bdir, bc);
    r.J1 = prepared_jones(jp.pb, jp.bdims, bt, a1, jp.bdir, bc);
  }
  // |x| |y| as sqrt(|x|^2 |y|^2): two square roots in place of four hypot calls, within an ulp or
  // two of the product of the moduli.  The product of the squared moduli leaves the normal FP64
  // range only where |x| |y| is above 1e154 or below 1e-154, where BB = bb^2 itself (nearly)
  // overflows or underflows.
  auto n2 = [](cd z) { return z.re * z.re + z.im * z.im; };
  const double bb = (sqrt(n2(r.J0.a[0]) * n2(r.J1.a[0])) + sqrt(n2(r.J0.a[3]) * n2(r.J1.a[3]))) / 2.;
  r.BB = bb * bb;
  return r;
}

// Entry of a prepared complex64 table (JonesP::pkf / pbf).
__device__ __forceinline__ M2T<float> prepared_jones_f(const float2* tab, const int* dims, int it, int ant, int dir,
                                                       int ich) {
  it = clampi(it, 0, dims[0] - 1);
  dir = clampi(dir, 0, dims[1] - 1);
  ant = clampi(ant, 0, dims[2] - 1);
  ich = clampi(ich, 0, dims[3] - 1);
  const DDF_GLOBAL float2* p = gptr(tab) + (((size_t(it) * dims[2] + ant) * size_t(dims[3]) + ich) * size_t(dims[1]) + dir) * 4;
  M2T<float> out;
#pragma unroll
  for (int q = 0; q < 4; ++q) out.a[q] = cf{p[q].x, p[q].y};
  return out;
}

// Whether jones_pair_t<T> reads the complex64 tables (jones_pair_f) rather than going through
// jones_at: FP32 arithmetic, no ScaleJones (which depends on the visibility), a complex64 table.
template <typename T>
__device__ __forceinline__ bool jones_f32_path(const JonesP& jp) {
  return !std::is_same<T, double>::value && !jp.scale && (jp.pkf || jp.pbf);
}

// The Jones key of (row, chan) for the complex64 tables: antennas, DD Jones time and channel, beam
// time and channel (the indices of an absent table are not read).  It does not depend on the
// direction, so a key read once serves every facet.
struct JonesKey {
  int a0, a1, kt, kc, bt, bc;
};
__device__ __forceinline__ JonesKey jones_key_f(const JonesP& jp, int64_t row, int chan) {
  JonesKey k;
  k.a0 = gptr(jp.a0)[row];
  k.a1 = gptr(jp.a1)[row];
  k.kt = k.kc = k.bt = k.bc = 0;
  if (jp.pkf) {
    k.kt = gptr(jp.ktime)[row];
    k.kc = gptr(jp.kchan)[chan];
  }
  if (jp.pbf) {
    k.bt = gptr(jp.btime)[row];
    k.bc = gptr(jp.bchan)[chan];
  }
  return k;
}

// The Jones pair of a key from the complex64 tables in the directions jp.kdir / jp.bdir: DD Jones
// times beam where both are given.
__device__ __forceinline__ void jones_pair_f(const JonesP& jp, const JonesKey& k, M2T<float>& J0, M2T<float>& J1) {
  if (jp.pkf) {
    J0 = prepared_jones_f(jp.pkf, jp.kdims, k.kt, k.a0, jp.kdir, k.kc);
    J1 = prepared_jones_f(jp.pkf, jp.kdims, k.kt, k.a1, jp.kdir, k.kc);
    if (jp.pbf) {
      J0 = mul(J0, prepared_jones_f(jp.pbf, jp.bdims, k.bt, k.a0, jp.bdir, k.bc));
      J1 = mul(J1, prepared_jones_f(jp.pbf, jp.bdims, k.bt, k.a1, jp.bdir, k.bc));
    }
  } else {
    J0 = prepared_jones_f(jp.pbf, jp.bdims, k.bt, k.a0, jp.bdir, k.bc);
    J1 = prepared_jones_f(jp.pbf, jp.bdims, k.bt, k.a1, jp.bdir, k.bc);
  }
}

// The Jones pair of (row, chan) in the arithmetic type T from the prepared tables: FP64 through
// jones_at; FP32 from the complex64 copies where jones_f32_path<T> holds.
template <typename T>
__device__ __forceinline__ void jones_pair_t(const JonesP& jp, int64_t row, int chan, const double* uvw,
                                             double wl_mean, M2T<T>& J0, M2T<T>& J1) {
  if (!jones_f32_path<T>(jp)) {
    const JonesPair jj = jones_at<true>(jp, row, chan, uvw, wl_mean);
    J0 = cvt<T>(jj.J0);
    J1 = cvt<T>(jj.J1);
    return;
  }
  M2T<float> F0, F1;
  jones_pair_f(jp, jones_key_f(jp, row, chan), F0, F1);
  J0 = cvt<T>(F0);
  J1 = cvt<T>(F1);
}

// Per-lane Jones cache with the refresh rule of JonesServer::updateJones (JonesServer.cc:193-298):
// the Jones pair is re-evaluated when the baseline, a time index or a Jones channel changes, with
// the uvw of the visibility that triggered the refresh (which only the ScaleJones term reads).
// Beam and DD parts are refreshed together (the reference refreshes them separately; the two
// differ only with ScaleJones on and a beam-only key change).
template <bool PREP = false>
struct JonesCache {
  int a0, a1, kt, kc, bt, bc;
  JonesPair cur;
  __device__ __forceinline__ void reset() { a0 = a1 = kt = kc = bt = bc = -1; }
  __device__ __forceinline__ const JonesPair& get(const JonesP& jp, int64_t row, int chan, const double* uvw,
                                                   double wl_mean) {
    const int r0 = gptr(jp.a0)[row], r1 = gptr(jp.a1)[row];
    const int k_t = jp.kms ? gptr(jp.ktime)[row] : -1, k_c = jp.kms ? gptr(jp.kchan)[chan] : -1;
    const int b_t = jp.beam ? gptr(jp.btime)[row] : -1, b_c = jp.beam ? gptr(jp.bchan)[chan] : -1;
    if (r0 != a0 || r1 != a1 || k_t != kt || k_c != kc || b_t != bt || b_c != bc) {
      cur = jones_at<PREP>(jp, row, chan, uvw, wl_mean);
      a0 = r0; a1 = r1; kt = k_t; kc = k_c; bt = b_t; bc = b_c;
    }
    return cur;
  }
};

// ---- decorrelation (DecorrelationHelper.cc:42-63) -------------------------------------------
struct DecorrP {
  int enabled, tsmear, fsmear;
  const double* uvw_dt;
  double dt, dnu, l0, m0;
  double fscale;  // pi dnu / c, formed on the host (decorr_fscale)
};
// The phase scales of the two terms, pi dnu / c and pi nu / c in the reference's operation order;
// the caller forms them where dnu and nu are known (the host for dnu and the mean frequency of a
// chunk, the block for a block's mean frequency), so that the passes hold no division.
__host__ __device__ inline double decorr_fscale(double dnu) { return kPi * dnu / kC; }
__host__ __device__ inline double decorr_tscale(double nu) { return kPi * nu / kC; }

// tscale = decorr_tscale(nu) for the frequency nu of the time-smearing term.
__device__ inline double decorr_factor(const DecorrP& d, const double* uvw, double tscale, int64_t idx) {
  if (!d.enabled) return 1.;
  double n0 = sqrt(1. - d.l0 * d.l0 - d.m0 * d.m0) - 1.;
  double f = 1.;
  if (d.fsmear) {
    const DDF_GLOBAL double* q = gptr(uvw) + 3 * idx;
    double ph = q[0] * d.l0 + q[1] * d.m0 + q[2] * n0;
    double phi = d.fscale * ph;
    f *= (phi == 0.) ? 1.0 : sin(phi) / phi;
  }
  if (d.tsmear) {
    const DDF_GLOBAL double* q = gptr(d.uvw_dt) + 3 * idx;
    double dph = (q[0] * d.l0 + q[1] * d.m0 + q[2] * n0) * d.dt;
    double phi = tscale * dph;
    f *= (phi == 0.) ? 1.0 : sin(phi) / phi;
  }
  return f;
}

// ---- CF tables and BDA-block placement -------------------------------------------------------
struct CFP {
  const float2* data;
  const int64_t* offset;  // [2*nw]
  const int32_t* support; // [nw]
  int nw, overs;
  double ref_wave, wmax;
};

struct Place {
  int locx, locy, sup, S;
  int64_t cfoff;  // complex-element offset of the oversampled sub-kernel
};

// w-plane, grid cell and oversampling offset of a block (gridder.h:321-351): evaluated in the
// reference operation order with contraction off.  Returns false when the block is dropped
// (iw >= Nw or support outside the grid), as the reference does.
__device__ inline bool place_block(const CFP& cf, int nx, int ny, double incr0, double incr1,
                                   double um, double vm, double wm, double recip_wvl, Place* pl) {
#pragma clang fp contract(off)
  double fiw = (double(cf.nw) - 1) * fabs(wm) * (cf.ref_wave * recip_wvl) / cf.wmax;
  if (!(fiw >= 0.0 && fiw < 2147483647.0)) return false;
  int iw = int(rint(fiw));
  if (iw >= cf.nw) return false;
  int plane = (wm > 0) ? iw : cf.nw + iw;
  int S = cf.support[iw];
  int O = cf.overs;
  int sup = (S - 1) / 2;
  double posx = double(nx) * incr0 * um * recip_wvl + double(nx / 2);
  double posy = double(nx) * incr1 * vm * recip_wvl + double(ny / 2);
  if (!(fabs(posx) < 1e9 && fabs(posy) < 1e9)) return false;
  int locx = int(rint(posx)), locy = int(rint(posy));
  if (locx - sup < 0 || locx + sup >= nx || locy - sup < 0 || locy + sup >= ny) return false;
  int offx = int(rint((double(locx) - posx) * double(O))) + (S * O - 1) / 2;
  int offy = int(rint((double(locy) - posy) * double(O))) + (S * O - 1) / 2;
  int io = clampi(offy - sup * O, 0, O - 1);
  int jo = clampi(offx - sup * O, 0, O - 1);
  pl->locx = locx;
  pl->locy = locy;
  pl->sup = sup;
  pl->S = S;
  pl->cfoff = cf.offset[plane] + int64_t(io * O + jo) * S * S;
  return true;
}

// Block-sum term U + W*Cu of the reference (gridder.h:272-274), no contraction.
__device__ __forceinline__ double shifted(double u, double w, double c) {
#pragma clang fp contract(off)
  return u + w * c;
}

}  // namespace ddf
