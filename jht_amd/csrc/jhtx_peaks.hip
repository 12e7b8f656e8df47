// Measured device rates for the roofline denominators: FP32 matrix (v_mfma_f32_16x16x4f32),
// FP64 vector FMA and device-to-device copy bandwidth, each timed with HIP events after a
// warm-up launch.  The compute loops keep 8 independent accumulation chains per wave and
// launch 2 waves per SIMD on every CU, so neither dependency latency nor occupancy limits
// them; every thread stores one value so the loops are not removed.

#include <hip/hip_runtime.h>

#include <string>

#include "jhtx_internal.h"

namespace jhtx {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kPeakBlock = 256;  // 4 waves

__global__ __launch_bounds__(kPeakBlock) void k_peak_mfma(float* out, int iters) {
  floatx4 d[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) d[c] = floatx4{0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * (float)(threadIdx.x & 63), b = 1.0001f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < 8; ++c) d[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d[c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) s += d[c][0] + d[c][1] + d[c][2] + d[c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(kPeakBlock) void k_peak_fma64(double* out, int iters) {
  double v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = 1e-3 * (double)threadIdx.x + (double)c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] = fma(v[c], 0.9999999, 1e-9);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s += v[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// dst[i] = src[i] for n 16-byte elements, grid-stride
__global__ __launch_bounds__(kPeakBlock) void k_peak_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ms of one launch of f after one warm-up launch
template <typename F>
hipError_t time_launch(hipStream_t s, F f, float& ms) {
  f();
  hipEvent_t a, b;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&b);
  if (e != hipSuccess) {
    (void)hipEventDestroy(a);
    return e;
  }
  (void)hipEventRecord(a, s);
  f();
  (void)hipEventRecord(b, s);
  e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (e == hipSuccess) e = hipGetLastError();
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return e;
}

}  // namespace
}  // namespace jhtx

using namespace jhtx;

extern "C" int jhtx_probe_peaks(void* buf, size_t bytes, double* out, void* stream) {
  if (!buf || !out) {
    set_error("jhtx_probe_peaks: null argument");
    return JHTX_EINVAL;
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
    set_error("jhtx_probe_peaks: no device");
    return JHTX_EDEVICE;
  }
  // 2 work-groups of 4 waves per CU: 2 waves per SIMD
  const int blocks = 2 * cus;
  const size_t need_out = (size_t)blocks * kPeakBlock * sizeof(double);
  constexpr size_t kCopyBytes = size_t(256) << 20;  // per direction
  if (bytes < 2 * kCopyBytes || bytes < need_out) {
    set_error("jhtx_probe_peaks: buffer smaller than " + std::to_string(2 * kCopyBytes) + " bytes");
    return JHTX_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int it_mfma = 20000, it_fma = 20000;
  float ms = 0.f;
  hipError_t e = time_launch(s, [&] {
    hipLaunchKernelGGL(k_peak_mfma, dim3(blocks), dim3(kPeakBlock), 0, s, reinterpret_cast<float*>(buf), it_mfma);
  }, ms);
  if (e != hipSuccess) {
    set_error(std::string("jhtx_probe_peaks: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  const double waves = (double)blocks * kPeakBlock / 64.0;
  out[0] = waves * it_mfma * 8 * 2.0 * 16 * 16 * 4 / (ms * 1e-3) / 1e12;  // TFLOP/s
  e = time_launch(s, [&] {
    hipLaunchKernelGGL(k_peak_fma64, dim3(blocks), dim3(kPeakBlock), 0, s, reinterpret_cast<double*>(buf), it_fma);
  }, ms);
  if (e != hipSuccess) {
    set_error(std::string("jhtx_probe_peaks: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  out[1] = (double)blocks * kPeakBlock * it_fma * 32 * 2.0 / (ms * 1e-3) / 1e12;  // TFLOP/s
  const long long n = (long long)(kCopyBytes / sizeof(float4));
  const float4* src = reinterpret_cast<const float4*>(buf);
  float4* dst = reinterpret_cast<float4*>(reinterpret_cast<char*>(buf) + kCopyBytes);
  e = time_launch(s, [&] {
    hipLaunchKernelGGL(k_peak_copy, dim3(blocks * 8), dim3(kPeakBlock), 0, s, src, dst, n);
  }, ms);
  if (e != hipSuccess) {
    set_error(std::string("jhtx_probe_peaks: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  out[2] = 2.0 * (double)kCopyBytes / (ms * 1e-3) / 1e9;  // GB/s, read + write
  return JHTX_OK;
}
