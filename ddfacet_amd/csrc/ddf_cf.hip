// On-device w-projection convolution functions of many facets (the w-plane loop of
// ModCF.ClassWTermModified, DDFacet/Imager/ModCF.py:444-597), for gfx950:
//
//   per (facet f, plane i of support S):
//     screen   w[a][b] = exp(-2 pi i (wlist_i / RefWave) n1_f(l_a, m_b)) |taper_S[a][b]|
//              on l_a, m_b = linspace(-lrad + dx / 2, lrad - dx / 2, S), dx = 2 lrad / S, n1_f the
//              facet's 2-D polynomial (ModFitPoly2D.polyval2d, coefficient k = i * 6 + j of
//              l^i m^j); with one w-plane the screen is |taper| (ModCF.py:479-491)
//     pad      zero-padded to N = S * OverS, centred as ModCF.ZeroPad (ModCF.py:68-84)
//     FFT      centred forward 2-D FFT / N^2 in FP64 (ModCF.fft2, ModCF.py:52-58; rocFFT double)
//     store    complex64, reorganised to [OverS][OverS][S][S] (ModCF.py:364-372); the conjugate
//              planes are the same steps on conj(screen) (ModCF.py:493-497)
//
// Output per facet: the planes in put() order (Wplanes 0..nw-1, then WplanesConj 0..nw-1)
// packed back to back, the layout the gridders take (ddf_cf.data / plane_offset).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/ddfgpu.h"

namespace ddf {
void set_last_error(const std::string& msg);
}

namespace {

int cfail(int code, const std::string& m) {
  ddf::set_last_error(m);
  return code;
}

constexpr int kPolyOrder = 5;
constexpr int kNCoef = (kPolyOrder + 1) * (kPolyOrder + 1);

// One support group: every (facet, plane) of support S, the plain screens then the conjugate
// ones.  t[(b * N + y) * N + x] is the ifftshift of the padded screen of batch entry b:
// t[j] = z[(j + N / 2) % N] per axis.
__global__ __launch_bounds__(256) void cf_screen(double2* t, int N, int S, int nb_half, const int* plane_of,
                                                 int nplanes_grp, const double* coefs, const double* wlist,
                                                 const double* taper, double ref_wave, double lrad, int nw) {
  const int64_t per = int64_t(N) * N;
  const int64_t total = per * 2 * nb_half;
  const int off = (N - S) / 2 + ((N % 2 == 0) ? 1 : 0);
  const double dx = 2. * lrad / double(S);
  const double start = -lrad + dx / 2., stop = lrad - dx / 2.;
  const double step = S > 1 ? (stop - start) / double(S - 1) : 0.;
  const double kPi = 3.141592653589793;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = i / per;
    const int r = int(i - b * per);
    const int jy = r / N, jx = r - (r / N) * N;
    int zy = jy + N / 2, zx = jx + N / 2;
    if (zy >= N) zy -= N;
    if (zx >= N) zx -= N;
    const int a = zy - off, c = zx - off;
    double2 v = make_double2(0., 0.);
    if (a >= 0 && a < S && c >= 0 && c < S) {
      const bool conj = b >= nb_half;
      const int64_t bb = conj ? b - nb_half : b;
      const int f = int(bb / nplanes_grp);
      const int plane = plane_of[bb - int64_t(f) * nplanes_grp];
      const double tap = taper[a * S + c];
      if (nw <= 1) {
        v = make_double2(tap, 0.);
      } else {
        const double l = start + a * step, m = start + c * step;
        const double* cf = coefs + int64_t(f) * kNCoef;
        // polyval2d: sum over (i, j) of coef[i * 6 + j] l^i m^j, terms in the reference's order
        double n1 = 0., li = 1.;
        for (int pi = 0; pi <= kPolyOrder; ++pi) {
          double mj = 1.;
          for (int pj = 0; pj <= kPolyOrder; ++pj) {
            n1 += cf[pi * (kPolyOrder + 1) + pj] * (li * mj);
            mj *= m;
          }
          li *= l;
        }
        double s, co;
        sincos(-2. * kPi * (wlist[plane] / ref_wave) * n1, &s, &co);
        v = make_double2(co * tap, (conj ? -s : s) * tap);
      }
    }
    t[i] = v;
  }
}

// fftshift, 1/N^2, complex64, [OverS][OverS][S][S]: out plane[(io * O + jo) * S^2 + p * S + q] =
// F_c[io + p O][jo + q O] with F_c[k] = F[(k + (N + 1) / 2) % N] per axis.
__global__ __launch_bounds__(256) void cf_store(const double2* F, int N, int S, int O, int nb_half,
                                                const int* plane_of, int nplanes_grp, float2* out,
                                                int64_t facet_stride, const int64_t* plane_off, int nw) {
  const int64_t per = int64_t(N) * N;
  const int64_t total = per * 2 * nb_half;
  const double norm = 1. / (double(N) * double(N));
  const int sh = (N + 1) / 2;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = i / per;
    const int r = int(i - b * per);
    const int ij = r / (S * S);
    const int pq = r - ij * S * S;
    const int io = ij / O, jo = ij - (ij / O) * O;
    const int p = pq / S, q = pq - (pq / S) * S;
    int ky = io + p * O + sh, kx = jo + q * O + sh;
    if (ky >= N) ky -= N;
    if (kx >= N) kx -= N;
    const double2 x = F[b * per + int64_t(ky) * N + kx];
    const bool conj = b >= nb_half;
    const int64_t bb = conj ? b - nb_half : b;
    const int f = int(bb / nplanes_grp);
    const int plane = plane_of[bb - int64_t(f) * nplanes_grp] + (conj ? nw : 0);
    out[int64_t(f) * facet_stride + plane_off[plane] + r] = make_float2(float(x.x * norm), float(x.y * norm));
  }
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65535) b = 65535;
  return unsigned(b < 1 ? 1 : b);
}

}  // namespace

extern "C" int ddf_cf_build(const ddf_cf_spec* s, float* out, int64_t* plane_offset, void* stream) {
  if (!s || !out || !plane_offset || !s->sups || !s->wlist || !s->coefs || !s->tapers)
    return cfail(DDF_EINVAL, "null argument");
  if (s->nw <= 0 || s->overs <= 0 || s->nfacets <= 0) return cfail(DDF_EINVAL, "bad CF sizes");
  const int nw = s->nw, O = s->overs, nf = s->nfacets;
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<int64_t> off(static_cast<size_t>(2 * nw));
  std::vector<int64_t> taper_off(static_cast<size_t>(nw));
  int64_t pos = 0, tpos = 0;
  for (int k = 0; k < 2 * nw; ++k) {
    const int S = s->sups[k % nw];
    if (S <= 0 || S % 2 == 0) return cfail(DDF_EINVAL, "CF supports must be odd and positive");
    off[size_t(k)] = pos;
    pos += int64_t(S) * O * S * O;
  }
  for (int i = 0; i < nw; ++i) {
    taper_off[size_t(i)] = tpos;
    tpos += int64_t(s->sups[i]) * s->sups[i];
  }
  const int64_t facet_stride = pos;
  for (int k = 0; k < 2 * nw; ++k) plane_offset[k] = off[size_t(k)];
  // device copies of the small tables
  std::map<int, std::vector<int>> groups;  // support -> planes
  for (int i = 0; i < nw; ++i) groups[s->sups[i]].push_back(i);
  std::vector<void*> tmp;
  auto cleanup = [&]() {
    for (void* p : tmp) (void)hipFree(p);
  };
  auto dalloc = [&](size_t bytes, void** p) -> int {
    if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) return cfail(DDF_ENOMEM, "CF build allocation failed");
    tmp.push_back(*p);
    return DDF_OK;
  };
  double *d_coefs = nullptr, *d_wlist = nullptr, *d_taper = nullptr;
  int64_t* d_off = nullptr;
  int rc = dalloc(size_t(nf) * kNCoef * 8, reinterpret_cast<void**>(&d_coefs));
  if (!rc) rc = dalloc(size_t(nw) * 8, reinterpret_cast<void**>(&d_wlist));
  if (!rc) rc = dalloc(size_t(tpos) * 8, reinterpret_cast<void**>(&d_taper));
  if (!rc) rc = dalloc(size_t(2 * nw) * 8, reinterpret_cast<void**>(&d_off));
  if (rc) {
    cleanup();
    return rc;
  }
  int err = DDF_OK;
  auto H = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && err == DDF_OK) err = cfail(DDF_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  H(hipMemcpyAsync(d_coefs, s->coefs, size_t(nf) * kNCoef * 8, hipMemcpyHostToDevice, st), "coefs");
  H(hipMemcpyAsync(d_wlist, s->wlist, size_t(nw) * 8, hipMemcpyHostToDevice, st), "wlist");
  H(hipMemcpyAsync(d_taper, s->tapers, size_t(tpos) * 8, hipMemcpyHostToDevice, st), "tapers");
  H(hipMemcpyAsync(d_off, off.data(), size_t(2 * nw) * 8, hipMemcpyHostToDevice, st), "offsets");
  if (rocfft_setup() != rocfft_status_success && err == DDF_OK) err = cfail(DDF_EHIP, "rocfft_setup failed");
  for (auto& g : groups) {
    if (err) break;
    const int S = g.first;
    const int N = S * O;
    const std::vector<int>& planes = g.second;
    const int npg = int(planes.size());
    const int nb_half = nf * npg;
    const int64_t elems = int64_t(N) * N * 2 * nb_half;
    double2* t = nullptr;
    int* d_planes = nullptr;
    if (dalloc(size_t(elems) * sizeof(double2), reinterpret_cast<void**>(&t)) ||
        dalloc(size_t(npg) * 4, reinterpret_cast<void**>(&d_planes))) {
      err = DDF_ENOMEM;
      break;
    }
    H(hipMemcpyAsync(d_planes, planes.data(), size_t(npg) * 4, hipMemcpyHostToDevice, st), "planes");
    const double* taper = d_taper + taper_off[size_t(planes[0])];
    hipLaunchKernelGGL(cf_screen, dim3(grid_for(elems)), dim3(256), 0, st, t, N, S, nb_half, d_planes, npg, d_coefs,
                       d_wlist, taper, s->ref_wave, s->lrad, nw);
    H(hipGetLastError(), "cf_screen");
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t lengths[2] = {size_t(N), size_t(N)};
    size_t wsize = 0;
    if (rocfft_plan_create(&plan, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                           rocfft_precision_double, 2, lengths, size_t(2 * nb_half), nullptr) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(plan, &wsize) != rocfft_status_success ||
        rocfft_execution_info_create(&info) != rocfft_status_success) {
      if (err == DDF_OK) err = cfail(DDF_EHIP, "rocfft plan failed");
    } else {
      if (wsize) {
        if (dalloc(wsize, &work) ||
            rocfft_execution_info_set_work_buffer(info, work, wsize) != rocfft_status_success)
          if (err == DDF_OK) err = cfail(DDF_EHIP, "rocfft work buffer failed");
      }
      if (!err && rocfft_execution_info_set_stream(info, st) != rocfft_status_success)
        err = cfail(DDF_EHIP, "rocfft stream failed");
      void* bufs[1] = {t};
      if (!err && rocfft_execute(plan, bufs, nullptr, info) != rocfft_status_success)
        err = cfail(DDF_EHIP, "rocfft_execute failed");
      if (!err) {
        hipLaunchKernelGGL(cf_store, dim3(grid_for(elems)), dim3(256), 0, st, t, N, S, O, nb_half, d_planes, npg,
                           reinterpret_cast<float2*>(out), facet_stride, d_off, nw);
        H(hipGetLastError(), "cf_store");
      }
      H(hipStreamSynchronize(st), "CF build");
    }
    if (info) rocfft_execution_info_destroy(info);
    if (plan) rocfft_plan_destroy(plan);
  }
  H(hipStreamSynchronize(st), "CF build");
  cleanup();
  return err;
}
