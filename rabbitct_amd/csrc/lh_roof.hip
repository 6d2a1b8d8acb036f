// lh_roof.hip -- the practical gather roof of the backprojector: its inner
// step without the projection.  Per update one coordinate step along the
// lane's path, floor and fraction, the address of the 2x2 stencil, four taps
// and the bilinear interpolation (horizontal then vertical, as
// LolaOPT.c:162-169), accumulated in FP32 -- no reciprocal, no weight.
// The taps come either from an LDS window staged once per workgroup (the
// staged kernel's tap path) or from an image in L2 through buffer loads (its
// fallback path).  Lanes walk 2D-local paths like the projections of
// neighbouring voxel lines: lane (ly, lz) of a workgroup starts at
// (0.9 ly, 1.1 lz) px from the workgroup's corner and steps (0.6, 0.05) px
// per voxel, so a workgroup reads a window of about 40 x 10 px.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lolahip.h"
#include "lh_internal.h"

namespace {

constexpr int kRun = 16;       // updates per lane and pass, as the FP32 kernel's run
constexpr int kWinW = 128;     // LDS window: 128 x 16 px
constexpr int kWinH = 16;
constexpr int kRsrc3 = 0x00020000;

__global__ void fill_kernel(float *img, uint32_t n)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    img[i] = (float)((i * 2654435761u) >> 8) * 0x1p-24f;
}

template <bool LDS>
__global__ __launch_bounds__(256) void gather_roof_kernel(const float *__restrict__ img, int W,
    int H, int passes, float *__restrict__ out)
{
  __shared__ float win[LDS ? kWinW * kWinH : 1];
  // the workgroup's corner, spread over the image (the window stays inside)
  const int u0 = 2 + (int)((blockIdx.x * 37u) % (unsigned)(W - kWinW - 4));
  const int v0 = 2 + (int)((blockIdx.x * 11u) % (unsigned)(H - kWinH - 4));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float ul = 0.9f * (float)(lane & 31) + 0.25f;
  const float vl = 1.1f * (float)(2 * w + (lane >> 5)) + 0.25f;
  if constexpr (LDS) {
    for (int i = threadIdx.x; i < kWinW * kWinH; i += 256)
      win[i] = img[(size_t)(v0 + i / kWinW) * W + u0 + i % kWinW];
    __syncthreads();
  }
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)img, (short)0, (int)((uint32_t)W * (uint32_t)H * 4u), kRsrc3);
  float acc[kRun];
#pragma unroll
  for (int r = 0; r < kRun; ++r)
    acc[r] = 0.0f;
  for (int p = 0; p < passes; ++p) {
    // a new path each pass, in both coordinates (as every view moves the
    // projection), so nothing of the step is loop-invariant
    const float ub = ul + 0.003f * (float)p, vb = vl + 0.002f * (float)p;
#pragma unroll
    for (int r = 0; r < kRun; ++r) {
      const float u = fmaf((float)r, 0.6f, ub), v = fmaf((float)r, 0.05f, vb);
      const float ax = __builtin_amdgcn_fractf(u), ay = __builtin_amdgcn_fractf(v);
      const int ix = (int)u, iy = (int)v;
      float p00, p10, p01, p11;
      if constexpr (LDS) {
        const float *q = win + iy * kWinW + ix;
        p00 = q[0];
        p10 = q[1];
        p01 = q[kWinW];
        p11 = q[kWinW + 1];
      } else {
        const uint32_t off = (uint32_t)((v0 + iy) * W + u0 + ix) * 4u;
        p00 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
        p10 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u, 0, 0));
        p01 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u * W, 0, 0));
        p11 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u * W + 4u, 0, 0));
      }
      const float b = fmaf(ax, p10 - p00, p00);
      const float t = fmaf(ax, p11 - p01, p01);
      acc[r] += fmaf(ay, t - b, b);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int r = 0; r < kRun; ++r)
    s += acc[r];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

} // namespace

extern "C" {

int lhGatherRoof(int lds, int blocksPerCU, int passes, double *updatesPerSec, float *ms)
{
  // passes <= 1000 keeps every lane's path inside the 128 x 16 px window
  // (at most 40 px right of and 11 rows below the window's corner)
  if (blocksPerCU < 1 || blocksPerCU > 64 || passes < 1 || passes > 1000 || !updatesPerSec)
    return lh_fail(LH_E_ARG, "lhGatherRoof: bad arguments");
  const int W = 512, H = 512; // 1 MiB: resident in every XCD's L2
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhGatherRoof: %s", hipGetErrorString(e));
  const int blocks = cus * blocksPerCU;
  float *img = nullptr, *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float t = 0.0f;
  if ((e = hipMalloc(&img, (size_t)W * H * sizeof(float))) == hipSuccess &&
      (e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(float))) == hipSuccess &&
      (e = hipEventCreate(&e0)) == hipSuccess && (e = hipEventCreate(&e1)) == hipSuccess) {
    fill_kernel<<<256, 256>>>(img, (uint32_t)(W * H));
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) { // warm-up, then timed
      e = hipEventRecord(e0, 0);
      if (lds)
        gather_roof_kernel<true><<<blocks, 256>>>(img, W, H, passes, out);
      else
        gather_roof_kernel<false><<<blocks, 256>>>(img, W, H, passes, out);
      if (e == hipSuccess)
        e = hipGetLastError();
      if (e == hipSuccess)
        e = hipEventRecord(e1, 0);
      if (e == hipSuccess)
        e = hipEventSynchronize(e1);
    }
    if (e == hipSuccess)
      e = hipEventElapsedTime(&t, e0, e1);
  }
  for (void *p : { (void *)img, (void *)out })
    if (p)
      (void)hipFree(p);
  for (hipEvent_t ev : { e0, e1 })
    if (ev)
      (void)hipEventDestroy(ev);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhGatherRoof: %s", hipGetErrorString(e));
  *updatesPerSec = (double)blocks * 256.0 * passes * kRun / (t * 1e-3);
  if (ms)
    *ms = t;
  return LH_OK;
}

} // extern "C"
