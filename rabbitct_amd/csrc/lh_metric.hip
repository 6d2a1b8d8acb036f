// lh_metric.hip -- the runner's -c verification on the GPU: HU conversion of
// the float volume and its error statistics against a uint16 reference
// (rabbitRunner.c:346-384).  Errors are integers, so the sums are exact
// (uint64) and the result does not depend on the reduction order.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "lolahip.h"
#include "lh_internal.h"

namespace {

// HU = floor(h0*v + h1 + 0.5) for v > 0 else 0, clamped to 4095; the product
// and sums rounded separately, as the C expression is evaluated without
// contraction.
__device__ __forceinline__ float hu_of(float v, float h0, float h1)
{
  float hu = 0.0f;
  if (v > 0.0f)
    hu = floorf(__fadd_rn(__fadd_rn(__fmul_rn(h0, v), h1), 0.5f));
  return fminf(hu, 4095.0f);
}

__global__ void hu_convert_kernel(const float *__restrict__ vol, uint64_t n, float h0, float h1,
    uint16_t *__restrict__ out)
{
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint16_t)hu_of(vol[i], h0, h1);
}

// acc[0] += sum e^2, acc[1] = max |e|, acc[2..12] += histogram of floor|e|
// (last bin: >= 10)
__global__ __launch_bounds__(256) void hu_metric_kernel(const float *__restrict__ vol, uint64_t n,
    const uint16_t *__restrict__ ref, float h0, float h1, unsigned long long *acc)
{
  __shared__ unsigned long long sh[13];
  if (threadIdx.x < 13)
    sh[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long se = 0, mx = 0;
  unsigned int hist[11] = { 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const long e = (long)hu_of(vol[i], h0, h1) - (long)ref[i];
    const unsigned long long ae = (unsigned long long)(e < 0 ? -e : e);
    se += ae * ae;
    mx = ae > mx ? ae : mx;
    ++hist[ae > 10 ? 10 : ae];
  }
  atomicAdd(&sh[0], se);
  atomicMax(&sh[1], mx);
  for (int b = 0; b < 11; ++b)
    if (hist[b])
      atomicAdd(&sh[2 + b], (unsigned long long)hist[b]);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], sh[0]);
    atomicMax(&acc[1], sh[1]);
    for (int b = 0; b < 11; ++b)
      atomicAdd(&acc[2 + b], sh[2 + b]);
  }
}

// HU slice z = L/2 - 1 of a device volume (the slice the runner's PGM and
// result.rctd hold, rabbitRunner.c:298,400) on the host.
int hu_slice(const float *dVol, uint32_t L, float h0, float h1, std::vector<uint16_t> &out)
{
  const uint64_t ss = (uint64_t)L * L;
  out.resize(ss);
  uint16_t *d   = nullptr;
  hipError_t e = hipMalloc(&d, ss * sizeof(uint16_t));
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "HU slice: %s", hipGetErrorString(e));
  int rc = lhHuConvert(dVol + (uint64_t)(L / 2 - 1) * ss, ss, h0, h1, d);
  if (rc == LH_OK) {
    e = hipMemcpy(out.data(), d, ss * sizeof(uint16_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
      rc = lh_fail(LH_E_HIP, "HU slice: %s", hipGetErrorString(e));
  }
  (void)hipFree(d);
  return rc;
}

// fwrite of n items or a failure status
bool put(FILE *f, const void *p, size_t size, size_t n)
{
  return fwrite(p, size, n, f) == n;
}

} // namespace

extern "C" {

int lhHuConvert(const float *dVol, uint64_t n, float hus0, float hus1, uint16_t *dOut)
{
  if (!dVol || !dOut || n == 0)
    return lh_fail(LH_E_ARG, "lhHuConvert: bad arguments");
  hu_convert_kernel<<<(unsigned)std::min<uint64_t>((n + 255) / 256, 16384), 256>>>(dVol, n, hus0,
      hus1, dOut);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhHuConvert: %s", hipGetErrorString(e));
  return LH_OK;
}

int lhHuMetric(const float *dVol, uint64_t n, const uint16_t *dRef, float hus0, float hus1,
    double *mse, double *maxAbs, uint32_t *hist)
{
  if (!dVol || !dRef || n == 0)
    return lh_fail(LH_E_ARG, "lhHuMetric: bad arguments");
  unsigned long long *dAcc = nullptr, h[13] = { 0 };
  hipError_t e = hipMalloc(&dAcc, sizeof h);
  if (e == hipSuccess)
    e = hipMemset(dAcc, 0, sizeof h);
  if (e == hipSuccess) {
    hu_metric_kernel<<<(unsigned)std::min<uint64_t>((n + 255) / 256, 4096), 256>>>(dVol, n, dRef,
        hus0, hus1, dAcc);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(h, dAcc, sizeof h, hipMemcpyDeviceToHost);
  (void)hipFree(dAcc);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhHuMetric: %s", hipGetErrorString(e));
  if (mse)
    *mse = (double)h[0] / (double)n;
  if (maxAbs)
    *maxAbs = (double)h[1];
  if (hist)
    for (int b = 0; b < 11; ++b)
      hist[b] = (uint32_t)h[2 + b];
  return LH_OK;
}

int lhWritePgmSlice(const char *path, uint32_t L, const float *dVol, float hus0, float hus1)
{
  if (!path || !dVol || L < 2)
    return lh_fail(LH_E_ARG, "lhWritePgmSlice: bad arguments");
  std::vector<uint16_t> hu;
  const int rc = hu_slice(dVol, L, hus0, hus1, hu);
  if (rc != LH_OK)
    return rc;
  // 16-bit binary PGM: samples big-endian
  for (uint16_t &v : hu)
    v = (uint16_t)((v >> 8) | (v << 8));
  FILE *f = fopen(path, "wb");
  if (!f)
    return lh_fail(LH_E_IO, "lhWritePgmSlice: cannot create %s", path);
  bool ok = fprintf(f, "P5\n%u %u\n4095\n", L, L) > 0 && put(f, hu.data(), 2, hu.size());
  ok = fclose(f) == 0 && ok;
  return ok ? LH_OK : lh_fail(LH_E_IO, "lhWritePgmSlice: write to %s failed", path);
}

int lhWriteResultFile(const char *path, uint32_t L, uint32_t numberOfImages,
    uint64_t meanTimeUsec, const float *dVol, const uint16_t *dRef, float hus0, float hus1,
    double *mse, double *maxAbs, uint32_t *hist)
{
  if (!path || !dVol || !dRef || L < 2)
    return lh_fail(LH_E_ARG, "lhWriteResultFile: bad arguments");
  const uint64_t n = (uint64_t)L * L * L;
  double err = 0.0, mx = 0.0;
  uint32_t h[11];
  int rc = lhHuMetric(dVol, n, dRef, hus0, hus1, &err, &mx, h);
  std::vector<uint16_t> hu;
  if (rc == LH_OK)
    rc = hu_slice(dVol, L, hus0, hus1, hu);
  if (rc != LH_OK)
    return rc;
  // the slice as the runner holds it after its -c loop: HU values as floats
  std::vector<float> slice(hu.begin(), hu.end());
  const std::vector<uint64_t> times(numberOfImages, meanTimeUsec);
  const int32_t magic = 591984, magicEnd = 489195, size = (int32_t)L;
  FILE *f = fopen(path, "wb");
  if (!f)
    return lh_fail(LH_E_IO, "lhWriteResultFile: cannot create %s", path);
  bool ok = put(f, &magic, 4, 1) && put(f, &size, 4, 1) && put(f, &numberOfImages, 4, 1) &&
            put(f, times.data(), 8, times.size()) && put(f, slice.data(), 4, slice.size()) &&
            put(f, &err, 8, 1) && put(f, &mx, 8, 1) && put(f, h, 4, 11) &&
            put(f, &magicEnd, 4, 1);
  ok = fclose(f) == 0 && ok;
  if (!ok)
    return lh_fail(LH_E_IO, "lhWriteResultFile: write to %s failed", path);
  if (mse)
    *mse = err;
  if (maxAbs)
    *maxAbs = mx;
  if (hist)
    for (int b = 0; b < 11; ++b)
      hist[b] = h[b];
  return LH_OK;
}

} // extern "C"
