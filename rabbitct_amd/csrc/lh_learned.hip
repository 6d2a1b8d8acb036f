// lh_learned.hip -- device helpers of the sparse-view + CNN path (no
// counterpart in the reference): the patch gather that feeds training, a
// ray-marching forward projector, and the held-out-view residual that decides
// whether the corrected volume is accepted or the full-view backprojection
// runs instead.
//
// Volumes are float32 [z][y][x] with voxel (x,y,z) at world
// (O + R x, O + R y, O + R z) (the layout and coordinates of the
// backprojected volume, rabbitRunner.c:189).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "lolahip.h"
#include "lh_internal.h"

namespace {

// out[p] = vols[spec.vol][z0+dz][y0+dy][x0+dx] for p = (dz, dy, dx) in P^3.
__global__ void patch_kernel(const float *__restrict__ vols, int L, const int *__restrict__ spec,
    int P, float *__restrict__ out)
{
  const int b = blockIdx.y;
  const int *s = spec + 4 * b;
  const size_t LLL = (size_t)L * L * L, PPP = (size_t)P * P * P;
  const float *src = vols + (size_t)s[0] * LLL;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < PPP;
       i += (size_t)gridDim.x * blockDim.x) {
    const int x = i % P, y = (i / P) % P, z = i / ((size_t)P * P);
    out[b * PPP + i] = src[((size_t)(s[1] + z) * L + (s[2] + y)) * L + (s[3] + x)];
  }
}

// Trilinear sample of the volume at voxel coordinates (zero outside).
__device__ __forceinline__ float trilinear(const float *__restrict__ vol, int L, float fx,
    float fy, float fz)
{
  const float x0f = floorf(fx), y0f = floorf(fy), z0f = floorf(fz);
  const int x0 = (int)x0f, y0 = (int)y0f, z0 = (int)z0f;
  const float ax = fx - x0f, ay = fy - y0f, az = fz - z0f;
  float acc = 0.0f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = x0 + (q & 1), y = y0 + ((q >> 1) & 1), z = z0 + (q >> 2);
    if ((unsigned)x < (unsigned)L && (unsigned)y < (unsigned)L && (unsigned)z < (unsigned)L) {
      const float w = ((q & 1) ? ax : 1.0f - ax) * (((q >> 1) & 1) ? ay : 1.0f - ay) *
                      ((q >> 2) ? az : 1.0f - az);
      acc = fmaf(w, vol[((size_t)z * L + y) * L + x], acc);
    }
  }
  return acc;
}

// Line integral of the volume along the ray of detector pixel
// (i*stride, j*stride): midpoint rule with step h (mm) between the entry and
// exit of the box that holds the trilinear support.  kSegLanes lanes share a
// ray, each taking every kSegLanes-th sample; their partial sums are reduced
// with cross-lane shuffles.  geo = S[3], Minv[9].
constexpr int kSegLanes = 16;

__global__ __launch_bounds__(256) void fp_kernel(const float *__restrict__ vol, int L, float R,
    float O, int Wd, int Hd, int stride, const double *__restrict__ geo, float h,
    float *__restrict__ out)
{
  const int ray = (blockIdx.x * blockDim.x + threadIdx.x) / kSegLanes;
  const int seg = threadIdx.x % kSegLanes;
  const int i = ray % Wd, j = ray / Wd;
  const bool live = j < Hd;
  const double *S = geo, *Mi = geo + 3;
  double d[3];
  for (int r = 0; r < 3; ++r)
    d[r] = Mi[3 * r] * (double)(i * stride) + Mi[3 * r + 1] * (double)(j * stride) + Mi[3 * r + 2];
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  // ray in voxel coordinates: p(t) = (S + t d - O) / R, t in mm
  double t0 = 0.0, t1 = 1e30;
  double p0[3], dp[3];
  const double lo = -1.0, hi = (double)L; // trilinear support
  for (int r = 0; r < 3; ++r) {
    d[r] /= nrm;
    p0[r] = (S[r] - (double)O) / (double)R;
    dp[r] = d[r] / (double)R;
    if (fabs(dp[r]) < 1e-30) {
      if (p0[r] <= lo || p0[r] >= hi)
        t1 = -1.0;
    } else {
      double ta = (lo - p0[r]) / dp[r], tb = (hi - p0[r]) / dp[r];
      if (ta > tb) {
        const double tt = ta;
        ta = tb;
        tb = tt;
      }
      t0 = fmax(t0, ta);
      t1 = fmin(t1, tb);
    }
  }
  float acc = 0.0f;
  if (live && t1 > t0) {
    // the chord of the box is at most sqrt(3) (L + 1) R
    const double nmax = 2.0 * (double)(L + 1) * (double)R / (double)h + 2.0;
    const int n = (int)fmin(ceil((t1 - t0) / (double)h), nmax);
    const double ts = t0 + 0.5 * (double)h;
    // FP64 ray origin at the first sample, FP32 steps from there
    const float px = (float)(p0[0] + ts * dp[0]), py = (float)(p0[1] + ts * dp[1]),
                pz = (float)(p0[2] + ts * dp[2]);
    const float sx = (float)(h * dp[0]), sy = (float)(h * dp[1]), sz = (float)(h * dp[2]);
    for (int k = seg; k < n; k += kSegLanes)
      acc += trilinear(vol, L, fmaf((float)k, sx, px), fmaf((float)k, sy, py),
          fmaf((float)k, sz, pz));
  }
  for (int o = kSegLanes / 2; o > 0; o >>= 1)
    acc += __shfl_xor(acc, o, kSegLanes);
  if (live && seg == 0)
    out[(size_t)j * Wd + i] = acc * h;
}

// sums[0] += sum (a - b)^2, sums[1] += sum b^2
__global__ void residual_kernel(const float *__restrict__ a, const float *__restrict__ b, size_t n,
    double *__restrict__ sums)
{
  double s0 = 0.0, s1 = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double d = (double)a[i] - (double)b[i];
    s0 += d * d;
    s1 += (double)b[i] * (double)b[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(sums, s0);
    atomicAdd(sums + 1, s1);
  }
}

} // namespace

extern "C" {

// Gathers n patches of P^3 voxels into dOut (n x P^3 floats, device) from
// the stack of L^3 volumes dVols (device); spec (host, n x 4 ints) holds
// (volume index, z0, y0, x0) per patch, each patch inside its volume.
int lhExtractPatches(const float *dVols, int nVols, int L, const int *spec, int n, int P,
    float *dOut)
{
  if (!dVols || !spec || !dOut || n < 1 || P < 1 || P > L)
    return lh_fail(LH_E_ARG, "lhExtractPatches: bad arguments");
  for (int b = 0; b < n; ++b) {
    const int *s = spec + 4 * b;
    if (s[0] < 0 || s[0] >= nVols || s[1] < 0 || s[2] < 0 || s[3] < 0 || s[1] + P > L ||
        s[2] + P > L || s[3] + P > L)
      return lh_fail(LH_E_ARG, "lhExtractPatches: patch %d (%d,%d,%d,%d) outside the volumes", b,
          s[0], s[1], s[2], s[3]);
  }
  int *dSpec = nullptr;
  hipError_t e = hipMalloc(&dSpec, (size_t)n * 4 * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(dSpec, spec, (size_t)n * 4 * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const size_t PPP = (size_t)P * P * P;
    patch_kernel<<<dim3((unsigned)std::min<size_t>((PPP + 255) / 256, 1024), n), 256>>>(dVols, L,
        dSpec, P, dOut);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  (void)hipFree(dSpec);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhExtractPatches: %s", hipGetErrorString(e));
  return LH_OK;
}

// Forward projection of the device volume dVol (L^3, voxel size R, origin O)
// for `count` views given by src (count x 3) and minv (count x 9, row-major;
// the ray through pixel (u,v) runs along Minv (u,v,1)), on the detector
// subsampled by `stride`: dOut receives count x ceil(H/stride) x
// ceil(W/stride) line integrals (device), marched with step h mm.
int lhForwardProject(const float *dVol, int L, float R, float O, int W, int H, int stride,
    uint32_t count, const double *src, const double *minv, float h, float *dOut)
{
  if (!dVol || !dOut || !src || !minv || L < 1 || W < 1 || H < 1 || stride < 1 || !(h > 0.0f))
    return lh_fail(LH_E_ARG, "lhForwardProject: bad arguments");
  const int Wd = (W + stride - 1) / stride, Hd = (H + stride - 1) / stride;
  std::vector<double> geo((size_t)count * 12);
  for (uint32_t v = 0; v < count; ++v) {
    for (int k = 0; k < 3; ++k)
      geo[12 * v + k] = src[3 * v + k];
    for (int k = 0; k < 9; ++k)
      geo[12 * v + 3 + k] = minv[9 * v + k];
  }
  double *dGeo = nullptr;
  hipError_t e = hipMalloc(&dGeo, geo.size() * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpy(dGeo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice);
  for (uint32_t v = 0; e == hipSuccess && v < count; ++v) {
    const size_t lanes = (size_t)Wd * Hd * kSegLanes;
    fp_kernel<<<(unsigned)((lanes + 255) / 256), 256>>>(dVol, L, R, O, Wd, Hd, stride,
        dGeo + 12 * v, h, dOut + (size_t)v * Wd * Hd);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  (void)hipFree(dGeo);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhForwardProject: %s", hipGetErrorString(e));
  return LH_OK;
}

// Relative residual sqrt(sum (a-b)^2 / sum b^2) of two device arrays.
int lhRelativeResidual(const float *dA, const float *dB, uint64_t n, double *out)
{
  if (!dA || !dB || !out || n == 0)
    return lh_fail(LH_E_ARG, "lhRelativeResidual: bad arguments");
  double *dS = nullptr;
  double h[2] = { 0.0, 0.0 };
  hipError_t e = hipMalloc(&dS, 2 * sizeof(double));
  if (e == hipSuccess)
    e = hipMemset(dS, 0, 2 * sizeof(double));
  if (e == hipSuccess) {
    residual_kernel<<<(unsigned)std::min<uint64_t>((n + 255) / 256, 2048), 256>>>(dA, dB, n, dS);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(h, dS, 2 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dS);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhRelativeResidual: %s", hipGetErrorString(e));
  *out = h[1] > 0.0 ? sqrt(h[0] / h[1]) : (h[0] > 0.0 ? INFINITY : 0.0);
  return LH_OK;
}

} // extern "C"
