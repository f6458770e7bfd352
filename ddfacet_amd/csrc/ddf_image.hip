// Facet image side of the major cycle on gfx950: centred facet FFTs (rocFFT), the facet
// normalisation image, the facet stitch and the model-to-facet extraction.
//
//   ddf_fft2c          ModFFTW.FFTW_2Donly.fft / .ifft (DDFacet/ToolsDir/ModFFTW.py:557-600):
//                      forward  A <- fftshift(fft2(ifftshift(A))) / N^2,
//                      inverse  A <- fftshift(ifft2(ifftshift(A))) * N^2 (unnormalised),
//                      both times `scale` (GridToIm's OverS^2, ClassDDEGridMachine.py:1186-1190).
//   ddf_facet_norm     BuildFacetNormImage (ClassFacetMachine.py:1289-1325): sum of the facets'
//                      spatial weights on the main image, facets added in index order.
//   ddf_stitch         FacetsToIm_Channel, kind Dirty / PSF (ClassFacetMachine.py:1331-1416).
//   ddf_model_to_grid  GiveModelTessel(ToGrid=False, ApplyNorm=True) (ClassImToGrid.py:138-264)
//                      followed by the forward FFT of get(TranformModelInput="FT")
//                      (ClassDDEGridMachine.py:951-958).
// The stitch and the norm image are gathers over main-image pixels that add the covering facets
// in facet order: no atomics, and the FP32 sums are formed in the reference's order.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ddfgpu.h"

namespace ddf {
void set_last_error(const std::string& msg);  // ddfgpu.hip: the message ddf_last_error() returns
}

namespace {

int ifail(int code, const std::string& m) {
  ddf::set_last_error(m);
  return code;
}

#define HIP_OK(x)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) return ifail(DDF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- centred FFTs -------------------------------------------------------------------------------
// dst[p][y][x] = scale * src[p][(y + s) % n][(x + s) % n]
__global__ __launch_bounds__(256) void roll2(const float2* src, float2* dst, int n, int64_t nplanes, int s,
                                             float scale) {
  const int64_t total = int64_t(n) * n * nplanes;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = i / (int64_t(n) * n);
    const int r = int(i - p * int64_t(n) * n);
    const int y = r / n, x = r - (r / n) * n;
    int ys = y + s, xs = x + s;
    if (ys >= n) ys -= n;
    if (xs >= n) xs -= n;
    float2 v = src[(p * n + ys) * n + xs];
    dst[i] = make_float2(v.x * scale, v.y * scale);
  }
}

// A cached rocFFT plan is read-only once made, so calls on several streams or host threads may
// share it; each call brings its own execution info, work buffer and staging buffer.
struct FFTPlan {
  rocfft_plan plan = nullptr;
  size_t work_size = 0;
};

std::mutex g_fft_mu;
std::map<std::tuple<int, int64_t, int, int>, FFTPlan> g_fft_plans;  // (n, nplanes, dir, device)

int get_plan(int n, int64_t nplanes, int inverse, FFTPlan** out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_fft_mu);
  static bool setup = false;
  if (!setup) {
    if (rocfft_setup() != rocfft_status_success) return ifail(DDF_EHIP, "rocfft_setup failed");
    setup = true;
  }
  auto key = std::make_tuple(n, nplanes, inverse, dev);
  auto it = g_fft_plans.find(key);
  if (it != g_fft_plans.end()) {
    *out = &it->second;
    return DDF_OK;
  }
  FFTPlan P;
  size_t lengths[2] = {size_t(n), size_t(n)};
  if (rocfft_plan_create(&P.plan, rocfft_placement_inplace,
                         inverse ? rocfft_transform_type_complex_inverse : rocfft_transform_type_complex_forward,
                         rocfft_precision_single, 2, lengths, size_t(nplanes), nullptr) != rocfft_status_success)
    return ifail(DDF_EHIP, "rocfft_plan_create failed");
  if (rocfft_plan_get_work_buffer_size(P.plan, &P.work_size) != rocfft_status_success) {
    (void)rocfft_plan_destroy(P.plan);
    return ifail(DDF_EHIP, "rocfft work size query failed");
  }
  auto res = g_fft_plans.emplace(key, P);
  *out = &res.first->second;
  return DDF_OK;
}

// ---- facet geometry (GiveEdges, DDFacet/ToolsDir/GiveEdges.py:29-82, Parity=None) -------------
struct Edges {
  int x0m, x1m, y0m, y1m, x0f, x1f, y0f, y1f;
};

Edges give_edges(int xc, int yc, int m, int nfacet) {
  Edges e;
  const int h = nfacet / 2;
  int mx0 = xc - h;
  e.x0m = mx0 > 0 ? mx0 : 0;
  e.x0f = e.x0m - mx0;
  int mx1 = xc + h;
  int x1m = mx1 < m - 1 ? mx1 : m - 1;
  e.x1f = nfacet - (mx1 - x1m);
  e.x1m = x1m + 1;
  int my0 = yc - h;
  e.y0m = my0 > 0 ? my0 : 0;
  e.y0f = e.y0m - my0;
  int my1 = yc + h;
  int y1m = my1 < m - 1 ? my1 : m - 1;
  e.y1f = nfacet - (my1 - y1m);
  e.y1m = y1m + 1;
  return e;
}

struct ImgFacet {
  Edges e;
  const float* sphe;
  const float* inv_sphe;
  const float* sw;
  const float2* dirty;
  float2* model;
  const double* sumwt;
  const double* sum_jones;
};

// FacetNorm[X, Y] = sum over facets (in order) of SW_f[a, b]
__global__ __launch_bounds__(256) void facet_norm_kernel(const ImgFacet* F, int nf, int n, int m, float* norm) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < int64_t(m) * m;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int X = int(i / m), Y = int(i - int64_t(X) * m);
    float acc = 0.f;
    for (int f = 0; f < nf; ++f) {
      const Edges& e = F[f].e;
      if (X < e.x0m || X >= e.x1m || Y < e.y0m || Y >= e.y1m) continue;
      const int a = X - e.x0m + e.x0f, b = Y - e.y0m + e.y0f;
      if (a < e.x0f || a >= e.x1f || b < e.y0f || b >= e.y1f) continue;
      acc = acc + F[f].sw[size_t(a) * n + b];
    }
    norm[i] = acc;
  }
}

// Image[ch, pol, X, Y] = sum_f Im_f / FacetNorm, with Im_f[a, b] = Im[n-1-b, a] (the
// Im[::-1, :].T of the reference) and Im[i, j] = Re(dirty) * InvSphe * SW_grid / weights in FP64,
// stored FP32, zero where Sphe < 1e-3.  SW_grid[n-1-b, a] = SW[a, b].
__global__ __launch_bounds__(256) void stitch_kernel(const ImgFacet* F, int nf, int n, int nband, int npol, int m,
                                                     const float* norm, int use_norm, float* image) {
  const int64_t total = int64_t(nband) * npol * m * m;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t plane = i / (int64_t(m) * m);
    const int64_t pix = i - plane * int64_t(m) * m;
    const int ch = int(plane / npol), pol = int(plane - int64_t(ch) * npol);
    const int X = int(pix / m), Y = int(pix - int64_t(X) * m);
    float acc = 0.f;
    for (int f = 0; f < nf; ++f) {
      const ImgFacet& fa = F[f];
      const Edges& e = fa.e;
      if (X < e.x0m || X >= e.x1m || Y < e.y0m || Y >= e.y1m) continue;
      const int a = X - e.x0m + e.x0f, b = Y - e.y0m + e.y0f;
      const int gi = n - 1 - b, gj = a;
      const size_t g = size_t(gi) * n + gj;
      float v = 0.f;
      if (!(fa.sphe[g] < 1e-3f)) {
        double w = fa.sumwt[ch * npol + pol];
        double sjn = 1.0;
        if (fa.sum_jones) {
          double ssq = fa.sum_jones[nband + ch];
          if (ssq == 0.) ssq = 1.;
          sjn = fa.sum_jones[ch] / ssq;
          if (sjn == 0.) sjn = 1.;
        }
        w = w * sqrt(sjn);
        if (w == 0.) w = 1.0;
        const double re = double(fa.dirty[(size_t(plane) * n + gi) * n + gj].x);
        v = float(re * double(fa.inv_sphe[g]) * double(fa.sw[size_t(a) * n + b]) / w);
      }
      acc = acc + v;
    }
    if (use_norm) acc = acc / norm[pix];
    image[i] = acc;
  }
}

// Model grid (before the FFT) of facet f: B[i, j] = A[j, n-1-i] (A.T[::-1, :]) with
// A[a, b] = Image[ch, pol, x0d + a - x0p, y0d + b - y0p] / FacetNorm * SW[a, b] / Sphe[a, b]
// inside the facet box, 0 outside and where Sphe < 1e-3, then * (OverS * n)^2.  A plane whose
// cut is entirely zero is left zero without normalisation (ClassImToGrid.py:175-176).
__global__ __launch_bounds__(256) void model_cut_kernel(const ImgFacet* F, int f, int n, int nband, int npol, int m,
                                                        const float* image, const float* norm, const int* plane_nz,
                                                        float factor) {
  const ImgFacet& fa = F[f];
  const Edges& e = fa.e;
  const int64_t total = int64_t(nband) * npol * n * n;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t plane = i / (int64_t(n) * n);
    const int64_t r = i - plane * int64_t(n) * n;
    const int gi = int(r / n), gj = int(r - int64_t(gi) * n);
    const int a = gj, b = n - 1 - gi;
    float v = 0.f;
    if (plane_nz[plane] && a >= e.x0f && a < e.x1f && b >= e.y0f && b < e.y1f) {
      const int X = e.x0m + a - e.x0f, Y = e.y0m + b - e.y0f;
      const size_t ab = size_t(a) * n + b;
      v = image[(size_t(plane) * m + X) * m + Y];
      v = v / norm[size_t(X) * m + Y];
      v = v * fa.sw[ab];
      v = v / fa.sphe[ab];
      if (fa.sphe[ab] < 1e-3f) v = 0.f;
    } else if (plane_nz[plane]) {
      const size_t ab = size_t(a) * n + b;
      if (fa.sphe[ab] < 1e-3f) v = 0.f;
    }
    fa.model[i] = make_float2(v * factor, 0.f);
  }
}

__global__ __launch_bounds__(256) void model_nonzero_kernel(const ImgFacet* F, int f, int nband, int npol, int m,
                                                            const float* image, int* plane_nz) {
  const ImgFacet& fa = F[f];
  const Edges& e = fa.e;
  const int w = e.x1m - e.x0m, h = e.y1m - e.y0m;
  const int64_t total = int64_t(nband) * npol * w * h;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t plane = i / (int64_t(w) * h);
    const int64_t r = i - plane * int64_t(w) * h;
    const int X = e.x0m + int(r / h), Y = e.y0m + int(r - (r / h) * h);
    if (image[(size_t(plane) * m + X) * m + Y] != 0.f) plane_nz[plane] = 1;
  }
}

// Adjoint of the model cut: image_bar[ch, pol, X, Y] = sum over facets (in order) of
// factor / FacetNorm[X, Y] * SW[a, b] / Sphe[a, b] * Re(plane_bar[ch, pol, n-1-b, a]),
// zero where Sphe < 1e-3 (the forward's flips and box read backwards).
__global__ __launch_bounds__(256) void model_cut_adjoint_kernel(const ImgFacet* F, int nf, int n, int nband, int npol,
                                                                int m, const float* norm, float factor, float* image_bar) {
  const int64_t total = int64_t(nband) * npol * m * m;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t plane = i / (int64_t(m) * m);
    const int64_t pix = i - plane * int64_t(m) * m;
    const int X = int(pix / m), Y = int(pix - int64_t(X) * m);
    float acc = 0.f;
    for (int f = 0; f < nf; ++f) {
      const ImgFacet& fa = F[f];
      const Edges& e = fa.e;
      if (X < e.x0m || X >= e.x1m || Y < e.y0m || Y >= e.y1m) continue;
      const int a = X - e.x0m + e.x0f, b = Y - e.y0m + e.y0f;
      const size_t ab = size_t(a) * n + b;
      if (fa.sphe[ab] < 1e-3f) continue;
      const float g = fa.model[(size_t(plane) * n + (n - 1 - b)) * n + a].x;
      acc = acc + g * factor / norm[pix] * fa.sw[ab] / fa.sphe[ab];
    }
    image_bar[i] = acc;
  }
}

__global__ void any_nonzero_kernel(const float* x, int64_t n, int* flag) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    if (x[i] != 0.f) flag[0] = 1;
}

__global__ void zero_ints(int* x, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = 0;
}

unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return unsigned(b);
}

int upload_facets(const ddf_image_facet* facets, int nf, int n, int m, hipStream_t st, ImgFacet** dout,
                  std::vector<void*>& tmp) {
  std::vector<ImgFacet> h(static_cast<size_t>(nf));
  for (int i = 0; i < nf; ++i) {
    const ddf_image_facet& f = facets[i];
    if (!f.sphe || !f.sw) return ifail(DDF_EINVAL, "facet Sphe / SW pointer is null");
    h[size_t(i)].e = give_edges(f.xc, f.yc, m, n);
    h[size_t(i)].sphe = f.sphe;
    h[size_t(i)].inv_sphe = f.inv_sphe;
    h[size_t(i)].sw = f.sw;
    h[size_t(i)].dirty = reinterpret_cast<const float2*>(f.dirty);
    h[size_t(i)].model = reinterpret_cast<float2*>(f.model);
    h[size_t(i)].sumwt = f.sumwt;
    h[size_t(i)].sum_jones = f.sum_jones;
    const Edges& e = h[size_t(i)].e;
    if (e.x1m <= e.x0m || e.y1m <= e.y0m || e.x0f < 0 || e.y0f < 0 || e.x1f > n || e.y1f > n ||
        e.x1f - e.x0f != e.x1m - e.x0m || e.y1f - e.y0f != e.y1m - e.y0m)
      return ifail(DDF_EINVAL, "facet box does not intersect the main image consistently");
  }
  void* d = nullptr;
  HIP_OK(hipMallocAsync(&d, h.size() * sizeof(ImgFacet), st));
  tmp.push_back(d);
  HIP_OK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(ImgFacet), hipMemcpyHostToDevice, st));
  *dout = static_cast<ImgFacet*>(d);
  return DDF_OK;
}

struct TmpFree {
  std::vector<void*>& v;
  hipStream_t st;
  ~TmpFree() {
    for (void* p : v) (void)hipFreeAsync(p, st);
  }
};

}  // namespace

// Per-call resources of ddf_fft2c, released on every return path (stream-ordered frees).
struct FFTCall {
  hipStream_t st;
  void* tmp = nullptr;
  void* work = nullptr;
  rocfft_execution_info info = nullptr;
  ~FFTCall() {
    if (info) (void)rocfft_execution_info_destroy(info);
    if (work) (void)hipFreeAsync(work, st);
    if (tmp) (void)hipFreeAsync(tmp, st);
  }
};

extern "C" int ddf_fft2c(float* planes, int32_t n, int64_t nplanes, int32_t inverse, double scale, void* stream) {
  if (!planes || n <= 0 || nplanes <= 0) return ifail(DDF_EINVAL, "bad FFT arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  FFTPlan* P = nullptr;
  int rc = get_plan(n, nplanes, inverse, &P);
  if (rc) return rc;
  const int64_t total = int64_t(n) * n * nplanes;
  FFTCall call{st};
  HIP_OK(hipMallocAsync(&call.tmp, size_t(total) * sizeof(float2), st));
  if (rocfft_execution_info_create(&call.info) != rocfft_status_success)
    return ifail(DDF_EHIP, "rocfft_execution_info_create failed");
  if (P->work_size) {
    HIP_OK(hipMallocAsync(&call.work, P->work_size, st));
    if (rocfft_execution_info_set_work_buffer(call.info, call.work, P->work_size) != rocfft_status_success)
      return ifail(DDF_EHIP, "rocfft work buffer failed");
  }
  if (rocfft_execution_info_set_stream(call.info, st) != rocfft_status_success) return ifail(DDF_EHIP, "rocfft stream");
  float2* a = reinterpret_cast<float2*>(planes);
  float2* t = static_cast<float2*>(call.tmp);
  // ifftshift: t[j] = a[(j + n/2) % n]
  hipLaunchKernelGGL(roll2, dim3(blocks_for(total)), dim3(256), 0, st, a, t, n, nplanes, n / 2, 1.f);
  HIP_OK(hipGetLastError());
  void* bufs[1] = {call.tmp};
  if (rocfft_execute(P->plan, bufs, nullptr, call.info) != rocfft_status_success)
    return ifail(DDF_EHIP, "rocfft_execute failed");
  // fftshift: a[k] = t[(k + (n+1)/2) % n], with the normalisation
  double norm = inverse ? scale : scale / (double(n) * double(n));
  hipLaunchKernelGGL(roll2, dim3(blocks_for(total)), dim3(256), 0, st, t, a, n, nplanes, (n + 1) / 2, float(norm));
  HIP_OK(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_facet_norm(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t m,
                              float* facet_norm, void* stream) {
  if (!facets || nf <= 0 || npix_facet <= 0 || m <= 0 || !facet_norm) return ifail(DDF_EINVAL, "bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<void*> tmp;
  TmpFree guard{tmp, st};
  ImgFacet* d = nullptr;
  int rc = upload_facets(facets, nf, npix_facet, m, st, &d, tmp);
  if (rc) return rc;
  hipLaunchKernelGGL(facet_norm_kernel, dim3(blocks_for(int64_t(m) * m)), dim3(256), 0, st, d, nf, npix_facet, m,
                     facet_norm);
  HIP_OK(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_stitch(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband, int32_t npol,
                          int32_t m, const float* facet_norm, float* image, void* stream) {
  if (!facets || nf <= 0 || npix_facet <= 0 || m <= 0 || !facet_norm || !image || nband <= 0 || npol <= 0)
    return ifail(DDF_EINVAL, "bad arguments");
  for (int i = 0; i < nf; ++i)
    if (!facets[i].dirty || !facets[i].inv_sphe || !facets[i].sumwt) return ifail(DDF_EINVAL, "facet buffer is null");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<void*> tmp;
  TmpFree guard{tmp, st};
  ImgFacet* d = nullptr;
  int rc = upload_facets(facets, nf, npix_facet, m, st, &d, tmp);
  if (rc) return rc;
  // "if (FacetNorm != 0).any(): Image /= FacetNorm"
  int* flag = nullptr;
  HIP_OK(hipMallocAsync(reinterpret_cast<void**>(&flag), sizeof(int), st));
  tmp.push_back(flag);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, st, flag, 1);
  hipLaunchKernelGGL(any_nonzero_kernel, dim3(blocks_for(int64_t(m) * m)), dim3(256), 0, st, facet_norm,
                     int64_t(m) * m, flag);
  int hflag = 0;
  HIP_OK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  const int64_t total = int64_t(nband) * npol * m * m;
  hipLaunchKernelGGL(stitch_kernel, dim3(blocks_for(total)), dim3(256), 0, st, d, nf, npix_facet, nband, npol, m,
                     facet_norm, hflag, image);
  HIP_OK(hipGetLastError());
  return DDF_OK;
}

extern "C" int ddf_model_to_grid(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband,
                                 int32_t npol, int32_t m, const float* image, const float* facet_norm, int32_t overs,
                                 void* stream) {
  if (!facets || nf <= 0 || npix_facet <= 0 || m <= 0 || !image || !facet_norm || nband <= 0 || npol <= 0)
    return ifail(DDF_EINVAL, "bad arguments");
  for (int i = 0; i < nf; ++i)
    if (!facets[i].model) return ifail(DDF_EINVAL, "facet model pointer is null");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<void*> tmp;
  TmpFree guard{tmp, st};
  ImgFacet* d = nullptr;
  int rc = upload_facets(facets, nf, npix_facet, m, st, &d, tmp);
  if (rc) return rc;
  int* nz = nullptr;
  HIP_OK(hipMallocAsync(reinterpret_cast<void**>(&nz), sizeof(int) * size_t(nband) * npol, st));
  tmp.push_back(nz);
  const int n = npix_facet;
  const float factor = float(double(overs) * n) * float(double(overs) * n);
  const int64_t per = int64_t(nband) * npol * n * n;
  bool contiguous = true;  // models[f] == models[0] + f * per: one batched FFT for every facet
  for (int f = 1; f < nf; ++f)
    contiguous &= reinterpret_cast<float2*>(facets[f].model) == reinterpret_cast<float2*>(facets[0].model) + f * per;
  for (int f = 0; f < nf; ++f) {
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(256), 0, st, nz, nband * npol);
    hipLaunchKernelGGL(model_nonzero_kernel, dim3(blocks_for(per)), dim3(256), 0, st, d, f, nband, npol, m, image, nz);
    hipLaunchKernelGGL(model_cut_kernel, dim3(blocks_for(per)), dim3(256), 0, st, d, f, n, nband, npol, m, image,
                       facet_norm, nz, factor);
    HIP_OK(hipGetLastError());
    if (!contiguous) {
      rc = ddf_fft2c(facets[f].model, n, int64_t(nband) * npol, 0, 1.0, stream);
      if (rc) return rc;
    }
  }
  if (contiguous) return ddf_fft2c(facets[0].model, n, int64_t(nf) * nband * npol, 0, 1.0, stream);
  return DDF_OK;
}

extern "C" int ddf_model_to_grid_adjoint(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband,
                                         int32_t npol, int32_t m, float* image_bar, const float* facet_norm,
                                         int32_t overs, void* stream) {
  if (!facets || nf <= 0 || npix_facet <= 0 || m <= 0 || !image_bar || !facet_norm || nband <= 0 || npol <= 0)
    return ifail(DDF_EINVAL, "bad arguments");
  for (int i = 0; i < nf; ++i)
    if (!facets[i].model) return ifail(DDF_EINVAL, "facet model pointer is null");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = npix_facet;
  // adjoint of the forward centred FFT (1/n^2 fft): the centred inverse FFT scaled by 1/n^2
  const int64_t per = int64_t(nband) * npol * n * n;
  bool contiguous = true;
  for (int f = 1; f < nf; ++f)
    contiguous &= reinterpret_cast<float2*>(facets[f].model) == reinterpret_cast<float2*>(facets[0].model) + f * per;
  const double inv = 1.0 / (double(n) * double(n));
  if (contiguous) {
    int rc = ddf_fft2c(facets[0].model, n, int64_t(nf) * nband * npol, 1, inv, stream);
    if (rc) return rc;
  } else {
    for (int f = 0; f < nf; ++f) {
      int rc = ddf_fft2c(facets[f].model, n, int64_t(nband) * npol, 1, inv, stream);
      if (rc) return rc;
    }
  }
  std::vector<void*> tmp;
  TmpFree guard{tmp, st};
  ImgFacet* d = nullptr;
  int rc = upload_facets(facets, nf, npix_facet, m, st, &d, tmp);
  if (rc) return rc;
  const float factor = float(double(overs) * n) * float(double(overs) * n);
  const int64_t total = int64_t(nband) * npol * m * m;
  hipLaunchKernelGGL(model_cut_adjoint_kernel, dim3(blocks_for(total)), dim3(256), 0, st, d, nf, n, nband, npol, m,
                     facet_norm, factor, image_bar);
  HIP_OK(hipGetLastError());
  return DDF_OK;
}
