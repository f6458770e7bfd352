// Chunk staging on the device: the baseline-time sort of the chunk reader and the flag selections
// of ClassMS.UpdateFlags that follow the flag preparation (ddf_prepare_flags).
//
//   sort     ClassMS.ReadData (DDFacet/Data/ClassMS.py:575-586): the row order of
//            sorted(zip(A0, A1, TIME, range(nrow))), i.e. by antenna pair, then time, then row.
//            Two stable radix sorts (time, then the antenna pair), each keeping the previous
//            order among equal keys.
//   gather   rows of any per-row array (uvw, data, flags, weights) in that order.
//   select   ClassMS.UpdateFlags (:1169-1233): per-antenna fractions of flagged (row, channel)
//            cells counted on the equalised flags before the uv-range cut, the uv-range and time
//            range cuts, then every row of an antenna with a flagged fraction above the threshold
//            (or named by the caller: distance to the core, explicit names) flagged.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>

#include "../../include/ddfgpu.h"

namespace ddf {
void set_last_error(const std::string& msg);  // ddfgpu.hip: the message ddf_last_error() returns
}

namespace {

int cfail(int code, const std::string& m) {
  ddf::set_last_error(m);
  return code;
}

#define CHUNK_TRY(x)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) {                                                                   \
      rc = cfail(DDF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));                   \
      goto done;                                                                              \
    }                                                                                         \
  } while (0)

unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return unsigned(b);
}

__global__ void iota_k(int32_t* x, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = int32_t(i);
}

__global__ void copy_f64_k(const double* src, double* dst, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// key of the antenna pair of row perm[i]: A0 in the high half, A1 in the low half (both < 2^16)
__global__ void pair_keys_k(const int32_t* a0, const int32_t* a1, const int32_t* perm, int64_t n, uint32_t* keys) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t r = perm[i];
    keys[i] = (uint32_t(a0[r]) << 16) | (uint32_t(a1[r]) & 0xFFFFu);
  }
}

// rows of row_bytes bytes; 16-byte chunks where the row size and both buffers allow
__global__ void gather_rows_k(const uint8_t* src, uint8_t* dst, int64_t nrow, int64_t row_bytes, const int32_t* index,
                              int vec16) {
  const int64_t words = vec16 ? row_bytes / 16 : row_bytes;
  const int64_t total = nrow * words;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i / words, w = i - r * words;
    const int64_t s = index[r];
    if (vec16)
      reinterpret_cast<uint4*>(dst + r * row_bytes)[w] = reinterpret_cast<const uint4*>(src + s * row_bytes)[w];
    else
      dst[r * row_bytes + w] = src[s * row_bytes + w];
  }
}

// per antenna: flagged (row, channel) cells of its rows on correlation 0 of the equalised flags,
// and its rows (an autocorrelation row counts once)
__global__ void antenna_counts_k(const uint8_t* flags, const int32_t* a0, const int32_t* a1, int64_t nrow, int nchan,
                                 int ncorr, int na, unsigned long long* nflag, unsigned long long* nrows) {
  for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrow; r += int64_t(gridDim.x) * blockDim.x) {
    const uint8_t* f = flags + size_t(r) * nchan * ncorr;
    unsigned long long c = 0;
    for (int ch = 0; ch < nchan; ++ch) c += f[size_t(ch) * ncorr] ? 1u : 0u;
    const int p = a0[r], q = a1[r];
    if (p >= 0 && p < na) {
      atomicAdd(nflag + p, c);
      atomicAdd(nrows + p, 1ull);
    }
    if (q != p && q >= 0 && q < na) {
      atomicAdd(nflag + q, c);
      atomicAdd(nrows + q, 1ull);
    }
  }
}

// antenna flagged: named by the caller, or flagged fraction > threshold (the fraction formed as
// the reference forms it: flags1[rows].sum() / float(flags1[rows].size or 1), in FP64)
__global__ void antenna_decide_k(const unsigned long long* nflag, const unsigned long long* nrows, int na, int nchan,
                                 double threshold, const uint8_t* named, uint8_t* flag_ant, double* frac) {
  const int a = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const double cells = double(nrows[a] * (unsigned long long)(nchan));
  const double fr = double(nflag[a]) / (cells > 0.0 ? cells : 1.0);
  frac[a] = fr;
  flag_ant[a] = uint8_t((named && named[a]) || fr > threshold);
}

// row cuts: uv distance outside [d0, d1] (squared metres, u^2 + v^2), time outside [t0, t1],
// or either antenna flagged
__global__ void select_rows_k(uint8_t* flags, const double* uvw, const double* times, const int32_t* a0,
                              const int32_t* a1, int64_t nrow, int64_t row_cells, int use_uv, double d0, double d1,
                              int use_time, double t0, double t1, double time_offset, const uint8_t* flag_ant, int na) {
#pragma clang fp contract(off)
  for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrow; r += int64_t(gridDim.x) * blockDim.x) {
    bool cut = false;
    if (use_uv) {
      const double u = uvw[3 * r], v = uvw[3 * r + 1];
      const double duv = u * u + v * v;
      cut = cut || duv < d0 || duv > d1;
    }
    if (use_time) {
      const double t = times[r] - time_offset;
      cut = cut || t < t0 || t > t1;
    }
    const int p = a0[r], q = a1[r];
    if ((p >= 0 && p < na && flag_ant[p]) || (q >= 0 && q < na && flag_ant[q])) cut = true;
    if (!cut) continue;
    uint8_t* f = flags + r * row_cells;
    for (int64_t k = 0; k < row_cells; ++k) f[k] = 1;
  }
}

}  // namespace

extern "C" int ddf_sort_baseline_time(const int32_t* a0, const int32_t* a1, const double* times, int64_t nrow,
                                      int32_t* index, void* stream) {
  if (nrow < 0 || nrow >= (int64_t(1) << 31)) return cfail(DDF_EINVAL, "bad row count");
  if (nrow == 0) return DDF_OK;
  if (!a0 || !a1 || !times || !index) return cfail(DDF_EINVAL, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = DDF_OK;
  double *tk = nullptr, *tk2 = nullptr;
  uint32_t *pk = nullptr, *pk2 = nullptr;
  int32_t* perm2 = nullptr;
  void* tmp = nullptr;
  size_t tb1 = 0, tb2 = 0;
  const int n = int(nrow);
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&tk), size_t(n) * 8, st));
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&tk2), size_t(n) * 8, st));
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&pk), size_t(n) * 4, st));
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&pk2), size_t(n) * 4, st));
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&perm2), size_t(n) * 4, st));
  hipLaunchKernelGGL(copy_f64_k, dim3(grid_for(n)), dim3(256), 0, st, times, tk, int64_t(n));
  hipLaunchKernelGGL(iota_k, dim3(grid_for(n)), dim3(256), 0, st, index, int64_t(n));
  {
    // pass 1: by time (stable: equal times keep row order)
    hipcub::DoubleBuffer<double> k1(tk, tk2);
    hipcub::DoubleBuffer<int32_t> v1(index, perm2);
    CHUNK_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, k1, v1, n, 0, 64, st));
    hipcub::DoubleBuffer<uint32_t> k2(pk, pk2);
    hipcub::DoubleBuffer<int32_t> v2(index, perm2);
    CHUNK_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, k2, v2, n, 0, 32, st));
    CHUNK_TRY(hipMallocAsync(&tmp, tb1 > tb2 ? tb1 : tb2, st));
    CHUNK_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb1, k1, v1, n, 0, 64, st));
    // pass 2: by antenna pair over the time order (stable: time, then row, among equal pairs)
    int32_t* p1 = v1.Current();
    int32_t* p1o = (p1 == index) ? perm2 : index;
    hipLaunchKernelGGL(pair_keys_k, dim3(grid_for(n)), dim3(256), 0, st, a0, a1, p1, int64_t(n), pk);
    hipcub::DoubleBuffer<int32_t> v3(p1, p1o);
    CHUNK_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb2, k2, v3, n, 0, 32, st));
    if (v3.Current() != index)
      CHUNK_TRY(hipMemcpyAsync(index, v3.Current(), size_t(n) * 4, hipMemcpyDeviceToDevice, st));
  }
done:
  for (void* p : {static_cast<void*>(tk), static_cast<void*>(tk2), static_cast<void*>(pk), static_cast<void*>(pk2),
                  static_cast<void*>(perm2), tmp})
    if (p) (void)hipFreeAsync(p, st);
  return rc;
}

extern "C" int ddf_gather_rows(const void* src, void* dst, int64_t nrow, int64_t row_bytes, const int32_t* index,
                               void* stream) {
  if (nrow < 0 || row_bytes <= 0) return cfail(DDF_EINVAL, "bad sizes");
  if (nrow == 0) return DDF_OK;
  if (!src || !dst || !index) return cfail(DDF_EINVAL, "null argument");
  if (src == dst) return cfail(DDF_EINVAL, "the gather cannot run in place");
  const int vec16 = (row_bytes % 16 == 0) && (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                    (reinterpret_cast<uintptr_t>(dst) % 16 == 0);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t total = nrow * (vec16 ? row_bytes / 16 : row_bytes);
  hipLaunchKernelGGL(gather_rows_k, dim3(grid_for(total)), dim3(256), 0, st, static_cast<const uint8_t*>(src),
                     static_cast<uint8_t*>(dst), nrow, row_bytes, index, vec16);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DDF_OK : cfail(DDF_EHIP, hipGetErrorString(e));
}

extern "C" int ddf_flag_selection(uint8_t* flags, const double* uvw, const double* times, const int32_t* a0,
                                  const int32_t* a1, int64_t nrow, int32_t nchan, int32_t ncorr,
                                  const ddf_flag_select* sel, double* ant_frac, uint8_t* ant_flagged, void* stream) {
  if (!sel || nrow < 0 || nchan <= 0 || ncorr <= 0 || sel->na <= 0) return cfail(DDF_EINVAL, "bad arguments");
  if (nrow > 0 && (!flags || !a0 || !a1 || (sel->use_uv_range && !uvw) || (sel->use_time_range && !times)))
    return cfail(DDF_EINVAL, "null argument");
  if (!ant_frac || !ant_flagged) return cfail(DDF_EINVAL, "null output");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int na = sel->na;
  int rc = DDF_OK;
  unsigned long long* cnt = nullptr;
  uint8_t* named = nullptr;
  CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&cnt), size_t(2 * na) * 8, st));
  CHUNK_TRY(hipMemsetAsync(cnt, 0, size_t(2 * na) * 8, st));
  if (sel->named_antennas) {
    CHUNK_TRY(hipMallocAsync(reinterpret_cast<void**>(&named), size_t(na), st));
    CHUNK_TRY(hipMemcpyAsync(named, sel->named_antennas, size_t(na), hipMemcpyHostToDevice, st));
  }
  if (nrow > 0)
    hipLaunchKernelGGL(antenna_counts_k, dim3(grid_for(nrow)), dim3(256), 0, st, flags, a0, a1, nrow, nchan, ncorr, na,
                       cnt, cnt + na);
  hipLaunchKernelGGL(antenna_decide_k, dim3((na + 63) / 64), dim3(64), 0, st, cnt, cnt + na, na, nchan,
                     sel->ant_threshold, named, ant_flagged, ant_frac);
  if (nrow > 0)
    hipLaunchKernelGGL(select_rows_k, dim3(grid_for(nrow)), dim3(256), 0, st, flags, uvw, times, a0, a1, nrow,
                       int64_t(nchan) * ncorr, sel->use_uv_range, sel->uv_min_m2, sel->uv_max_m2, sel->use_time_range,
                       sel->time_min, sel->time_max, sel->time_offset, ant_flagged, na);
  CHUNK_TRY(hipGetLastError());
done:
  if (cnt) (void)hipFreeAsync(cnt, st);
  if (named) (void)hipFreeAsync(named, st);
  return rc;
}
