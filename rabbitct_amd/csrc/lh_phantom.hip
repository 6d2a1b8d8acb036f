// lh_phantom.hip -- seeded synthetic projections on the GPU (no counterpart
// in the reference, which reads a recorded scan).
//
// Per view: analytic line integrals of an ellipsoid phantom for the ray of
// every detector pixel (FP64), times the FDK cosine weight and a short-scan
// redundancy weight per column, then a row-wise discrete ramp filter (Ram-Lak,
// spatial form) scaled by the angular step, a zero 1-pixel border, and the
// result as float32.  The same chain in FP64 C is the oracle
// (oracle/rct_oracle.c: orc_phantom_line_integrals, orc_ramp_filter_rows).

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "lolahip.h"
#include "lh_internal.h"

namespace {

// Chord of the ray s0 + t d (|d| = 1) through ellipsoid e: centre e[0..2],
// semi-axes e[3..5], rotation e[7..15] (row-major, columns = axes).
__device__ double chord(const double *e, const double *s0, const double *d)
{
  const double p[3] = { s0[0] - e[0], s0[1] - e[1], s0[2] - e[2] };
  const double *R   = e + 7;
  double A = 0.0, B = 0.0, C = -1.0;
  for (int k = 0; k < 3; ++k) {
    const double q0 = (R[0 * 3 + k] * p[0] + R[1 * 3 + k] * p[1] + R[2 * 3 + k] * p[2]) / e[3 + k];
    const double q1 = (R[0 * 3 + k] * d[0] + R[1 * 3 + k] * d[1] + R[2 * 3 + k] * d[2]) / e[3 + k];
    A += q1 * q1;
    B += q0 * q1;
    C += q0 * q0;
  }
  const double disc = B * B - A * C;
  return disc > 0.0 ? 2.0 * sqrt(disc) / A : 0.0;
}

__global__ void integrals_kernel(int W, int H, const double *__restrict__ src,
    const double *__restrict__ minv, const double *__restrict__ axis,
    const float *__restrict__ parker, int nEll, const double *__restrict__ ell,
    double *__restrict__ out)
{
  __shared__ double sEll[64 * 16];
  for (int i = threadIdx.x; i < nEll * 16; i += blockDim.x)
    sEll[i] = ell[i];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= W)
    return;
  double d[3];
  for (int r = 0; r < 3; ++r)
    d[r] = minv[3 * r + 0] * (double)i + minv[3 * r + 1] * (double)j + minv[3 * r + 2];
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  d[0] /= nrm;
  d[1] /= nrm;
  d[2] /= nrm;
  double acc = 0.0;
  for (int k = 0; k < nEll; ++k)
    acc += sEll[16 * k + 6] * chord(sEll + 16 * k, src, d);
  const double cosw = d[0] * axis[0] + d[1] * axis[1] + d[2] * axis[2];
  out[(size_t)j * W + i] = acc * cosw * (double)parker[i];
}

// One workgroup per row: the row in LDS, out[i] = scale * tau * sum_m h(i-m) x[m].
__global__ void ramp_kernel(int W, int H, double tau, double scale, int border,
    const double *__restrict__ in, float *__restrict__ out)
{
  extern __shared__ double row[];
  const int j = blockIdx.x;
  for (int i = threadIdx.x; i < W; i += blockDim.x)
    row[i] = in[(size_t)j * W + i];
  __syncthreads();
  const double h0 = 1.0 / (4.0 * tau * tau);
  const double c  = -1.0 / (M_PI * M_PI * tau * tau);
  for (int i = threadIdx.x; i < W; i += blockDim.x) {
    double acc = h0 * row[i];
    // odd distances only: h(k) = c / k^2
    for (int k = 1; k < W; k += 2) {
      const double hk = c / ((double)k * (double)k);
      const double a  = (i - k >= 0) ? row[i - k] : 0.0;
      const double b  = (i + k < W) ? row[i + k] : 0.0;
      acc += hk * (a + b);
    }
    float v = (float)(scale * tau * acc);
    if (border && (i == 0 || i == W - 1 || j == 0 || j == H - 1))
      v = 0.0f;
    out[(size_t)j * W + i] = v;
  }
}

// Raw line integrals of the pixels (i*stride, j*stride), no weights.
__global__ void raw_integrals_kernel(int Wd, int stride, const double *__restrict__ geo,
    int nEll, const double *__restrict__ ell, float *__restrict__ out)
{
  __shared__ double sEll[64 * 16];
  for (int i = threadIdx.x; i < nEll * 16; i += blockDim.x)
    sEll[i] = ell[i];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= Wd)
    return;
  const double *src = geo, *minv = geo + 3;
  double d[3];
  for (int r = 0; r < 3; ++r)
    d[r] = minv[3 * r + 0] * (double)(i * stride) + minv[3 * r + 1] * (double)(j * stride) +
           minv[3 * r + 2];
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  d[0] /= nrm;
  d[1] /= nrm;
  d[2] /= nrm;
  double acc = 0.0;
  for (int k = 0; k < nEll; ++k)
    acc += sEll[16 * k + 6] * chord(sEll + 16 * k, src, d);
  out[(size_t)j * Wd + i] = (float)acc;
}

// FDK weights of measured line integrals: out = raw * cos(angle to the
// principal axis) * parker[column], FP64 (the form integrals_kernel produces).
__global__ void fdk_weight_kernel(int W, const double *__restrict__ minv,
    const double *__restrict__ axis, const float *__restrict__ parker,
    const float *__restrict__ raw, double *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= W)
    return;
  double d[3];
  for (int r = 0; r < 3; ++r)
    d[r] = minv[3 * r + 0] * (double)i + minv[3 * r + 1] * (double)j + minv[3 * r + 2];
  const double nrm  = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double cosw = (d[0] * axis[0] + d[1] * axis[1] + d[2] * axis[2]) / nrm;
  out[(size_t)j * W + i] = (double)raw[(size_t)j * W + i] * cosw * (double)parker[i];
}

} // namespace

// Unweighted line integrals of an ellipsoid phantom on a detector subsampled
// by `stride` (pixels (i*stride, j*stride)), for `count` views; the result
// (count x ceil(H/stride) x ceil(W/stride) float32) goes to device memory
// dOut.  Ray geometry per view: source src[3], Minv[9] (row-major, the ray
// through pixel (u,v) runs along Minv (u,v,1)).
extern "C" int lhLineIntegrals(int W, int H, int stride, uint32_t count, const double *src,
    const double *minv, uint32_t nEll, const double *ell, float *dOut)
{
  if (W < 1 || H < 1 || stride < 1 || !src || !minv || !ell || !dOut || nEll == 0 || nEll > 64)
    return lh_fail(LH_E_ARG, "lhLineIntegrals: bad arguments");
  const int Wd = (W + stride - 1) / stride, Hd = (H + stride - 1) / stride;
  std::vector<double> geo((size_t)count * 12);
  for (uint32_t v = 0; v < count; ++v) {
    for (int k = 0; k < 3; ++k)
      geo[v * 12 + k] = src[3 * v + k];
    for (int k = 0; k < 9; ++k)
      geo[v * 12 + 3 + k] = minv[9 * v + k];
  }
  double *dGeo = nullptr, *dEll = nullptr;
  int rc = LH_OK;
  hipError_t e = hipMalloc(&dGeo, geo.size() * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&dEll, (size_t)nEll * 16 * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpy(dGeo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dEll, ell, (size_t)nEll * 16 * sizeof(double), hipMemcpyHostToDevice);
  for (uint32_t v = 0; e == hipSuccess && v < count; ++v) {
    raw_integrals_kernel<<<dim3((Wd + 255) / 256, Hd), 256>>>(Wd, stride, dGeo + 12 * v,
        (int)nEll, dEll, dOut + (size_t)v * Wd * Hd);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  if (e != hipSuccess)
    rc = lh_fail(LH_E_HIP, "lhLineIntegrals: %s", hipGetErrorString(e));
  (void)hipFree(dGeo);
  (void)hipFree(dEll);
  return rc;
}

// FDK pre-processing of measured projections (the step before the
// backprojection; the RabbitCT input carries its result): dRaw (count x H x W
// line integrals, device) -> dOut (device) = ramp_tau( raw * cos * parker ) *
// dbeta, optionally with a zero 1-pixel border.  minv / axis / parker as in
// lhGenerateViews (host).
extern "C" int lhFdkFilter(int W, int H, uint32_t count, const double *minv, const double *axis,
    const float *parker, double tau, double dbeta, int border, const float *dRaw, float *dOut)
{
  if (W < 2 || H < 1 || !minv || !axis || !parker || !dRaw || !dOut || !(tau > 0.0))
    return lh_fail(LH_E_ARG, "lhFdkFilter: bad arguments");
  double *dGeo = nullptr, *dW = nullptr;
  float *dPar = nullptr;
  std::vector<double> geo((size_t)count * 12);
  for (uint32_t v = 0; v < count; ++v) {
    for (int k = 0; k < 9; ++k)
      geo[12 * v + k] = minv[9 * v + k];
    for (int k = 0; k < 3; ++k)
      geo[12 * v + 9 + k] = axis[3 * v + k];
  }
  hipError_t e = hipMalloc(&dGeo, geo.size() * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&dW, (size_t)W * H * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&dPar, (size_t)count * W * sizeof(float));
  if (e == hipSuccess)
    e = hipMemcpy(dGeo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dPar, parker, (size_t)count * W * sizeof(float), hipMemcpyHostToDevice);
  for (uint32_t v = 0; e == hipSuccess && v < count; ++v) {
    const size_t img = (size_t)v * W * H;
    fdk_weight_kernel<<<dim3((W + 255) / 256, H), 256>>>(W, dGeo + 12 * v, dGeo + 12 * v + 9,
        dPar + (size_t)v * W, dRaw + img, dW);
    ramp_kernel<<<H, 256, W * sizeof(double)>>>(W, H, tau, dbeta, border, dW, dOut + img);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  (void)hipFree(dGeo);
  (void)hipFree(dW);
  (void)hipFree(dPar);
  if (e != hipSuccess)
    return lh_fail(LH_E_HIP, "lhFdkFilter: %s", hipGetErrorString(e));
  return LH_OK;
}

extern "C" int lhGenerateViews(LhContext *c, uint32_t firstSlot, uint32_t count, const int *ids,
    const double *matrices, const double *src, const double *minv, const double *axis,
    const float *parker, uint32_t nEll, const double *ell, double tau, double dbeta,
    int border, float *hostImages)
{
  if (!c || !ids || !matrices || !src || !minv || !axis || !parker || !ell || nEll == 0 ||
      nEll > 64)
    return lh_fail(LH_E_ARG, "lhGenerateViews: bad arguments");
  const LhConfig &g = lh_config(c);
  if (firstSlot + count > g.nSlots)
    return lh_fail(LH_E_ARG, "lhGenerateViews: slots out of range");
  for (uint32_t v = 0; v < count; ++v)
    if (ids[v] < 1 || (uint32_t)ids[v] > g.nViews)
      return lh_fail(LH_E_ARG, "lhGenerateViews: view id %d out of range", ids[v]);
  const int W = (int)g.W, H = (int)g.H;
  hipStream_t s = lh_stream(c);
  double *dGeo = nullptr, *dEll = nullptr, *dInt = nullptr;
  float *dPar = nullptr, *dImg = nullptr;
  const size_t per = 3 + 9 + 3;
  std::vector<double> geo((size_t)count * per);
  for (uint32_t v = 0; v < count; ++v) {
    for (int k = 0; k < 3; ++k)
      geo[v * per + k] = src[3 * v + k];
    for (int k = 0; k < 9; ++k)
      geo[v * per + 3 + k] = minv[9 * v + k];
    for (int k = 0; k < 3; ++k)
      geo[v * per + 12 + k] = axis[3 * v + k];
  }
#define LH_TRY(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      rc = lh_fail(LH_E_HIP, "lhGenerateViews: %s: %s", #call, hipGetErrorString(e_));   \
      goto done;                                                                         \
    }                                                                                    \
  } while (0)
  int rc = LH_OK;
  LH_TRY(hipSetDevice(g.device));
  LH_TRY(hipMalloc(&dGeo, geo.size() * sizeof(double)));
  LH_TRY(hipMalloc(&dEll, (size_t)nEll * 16 * sizeof(double)));
  LH_TRY(hipMalloc(&dPar, (size_t)count * W * sizeof(float)));
  LH_TRY(hipMalloc(&dInt, (size_t)W * H * sizeof(double)));
  LH_TRY(hipMalloc(&dImg, (size_t)W * H * sizeof(float)));
  LH_TRY(hipMemcpyAsync(dGeo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice, s));
  LH_TRY(hipMemcpyAsync(dEll, ell, (size_t)nEll * 16 * sizeof(double), hipMemcpyHostToDevice, s));
  LH_TRY(hipMemcpyAsync(dPar, parker, (size_t)count * W * sizeof(float), hipMemcpyHostToDevice, s));
  for (uint32_t v = 0; v < count; ++v) {
    const double *gv = dGeo + v * per;
    integrals_kernel<<<dim3((W + 255) / 256, H), 256, 0, s>>>(W, H, gv, gv + 3, gv + 12,
        dPar + (size_t)v * W, (int)nEll, dEll, dInt);
    ramp_kernel<<<H, 256, W * sizeof(double), s>>>(W, H, tau, dbeta, border, dInt, dImg);
    LH_TRY(hipGetLastError());
    rc = lhLoadViewDevice(c, firstSlot + v, ids[v], matrices + 12 * v, dImg);
    if (rc != LH_OK)
      goto done;
    if (hostImages)
      LH_TRY(hipMemcpyAsync(hostImages + (size_t)v * W * H, dImg, (size_t)W * H * sizeof(float),
          hipMemcpyDeviceToHost, s));
  }
  LH_TRY(hipStreamSynchronize(s));
done:
  (void)hipStreamSynchronize(s);
  (void)hipFree(dGeo);
  (void)hipFree(dEll);
  (void)hipFree(dPar);
  (void)hipFree(dInt);
  (void)hipFree(dImg);
  return rc;
#undef LH_TRY
}
