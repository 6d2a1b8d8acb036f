/* rct_oracle.c -- CPU oracle for the RabbitCT backprojection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke check
 * and the CPU-baseline leg of the benchmark may call it; the product library
 * (rabbitct_amd/liblolahip.so) never links or loads it.
 *
 * Plain C, FP64 throughout.  Each function names the reference file:line it
 * restates (paths relative to the RabbitCT source tree).  Pinning (the checks
 * that make these restatements trustworthy) lives in tests/test_oracle.py:
 *   orc_backproject_bunny      golden vectors of the compiled LolaBunny, closed
 *                              forms (constant image, single pixel, 1/s^2 scaling)
 *   orc_backproject_clipped    golden vectors of the compiled LolaOPT/LolaOMP,
 *                              agreement with bunny inside the clip ranges
 *   orc_backproject_lines      equality with orc_backproject_clipped on the
 *                              same lines (it is that loop on a line subset)
 *   orc_backproject_opt_f32    golden vectors of the compiled LolaOPT (within
 *                              5e-7 of the range, closer than the FP64 form)
 *   orc_backproject_asm        golden vectors of the compiled LolaASM (the
 *                              do-while "at least one vector" semantics)
 *   orc_line_ranges            byte equality with clip files written by the
 *                              compiled reference (analyseGeometry.c)
 *   orc_shadow                 the padding the compiled LolaASM prints
 *   orc_hu_metric              the numbers the compiled runner prints for -c
 *   orc_phantom_line_integrals closed-form sphere chords, brute-force numeric
 *                              ray marching on a tiny detector
 *   orc_ramp_filter_rows       FFT convolution with the same kernel (numpy)
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Projection of a world point with a column-major 3x4 matrix a[c*3+r]
 * (LolaBunny.c:64-66, analyseGeometry.c:95-100). */
static inline void project(const double *a, double x, double y, double z, double *u,
    double *v, double *w)
{
  double wn = a[2] * x + a[5] * y + a[8] * z + a[11];
  *w        = wn;
  *u        = (a[0] * x + a[3] * y + a[6] * z + a[9]) / wn;
  *v        = (a[1] * x + a[4] * y + a[7] * z + a[10]) / wn;
}

/* C's (int) conversion of a double (truncation toward zero), with values far
 * outside the int range mapped to a sentinel that is off every detector. */
static inline long trunc_index(double t)
{
  if (!(t > -1.0e9 && t < 1.0e9))
    return -2000000000L;
  return (long)t;
}

/* pN: bounds-checked pixel fetch, zero outside the detector (LolaBunny.c:11-17). */
static inline double pixel(const float *img, long i, long j, long W, long H)
{
  if (i >= 0 && i < W && j >= 0 && j < H)
    return (double)img[j * W + i];
  return 0.0;
}

/* pHatN: bilinear interpolation with truncation toward zero; the fractional
 * weights are negative for coordinates in (-1,0) (LolaBunny.c:19-30). */
static inline double bilinear(const float *img, double u, double v, long W, long H)
{
  long i   = trunc_index(u);
  long j   = trunc_index(v);
  if (i < -2 || j < -2 || i > W || j > H)
    return 0.0;
  double a = u - (double)i;
  double b = v - (double)j;
  return (1.0 - a) * (1.0 - b) * pixel(img, i, j, W, H) +
         a * (1.0 - b) * pixel(img, i + 1, j, W, H) +
         (1.0 - a) * b * pixel(img, i, j + 1, W, H) + a * b * pixel(img, i + 1, j + 1, W, H);
}

/* Unclipped FP64 backprojection of one view, all L^3 voxels
 * (LolaBunny.c:44-75).  Accumulates in double (the reference accumulates the
 * FP64 value rounded to float, LolaBunny.c:68). */
void orc_backproject_bunny(uint32_t L, float rL, float oL, uint32_t W, uint32_t H,
    const double *a, const float *img, double *vol)
{
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)L; k++) {
    double z = (double)oL + (double)k * (double)rL;
    for (long j = 0; j < (long)L; j++) {
      double y = (double)oL + (double)j * (double)rL;
      for (long i = 0; i < (long)L; i++) {
        double x = (double)oL + (double)i * (double)rL;
        double u, v, w;
        project(a, x, y, z, &u, &v, &w);
        vol[(k * L + j) * L + i] += 1.0 / (w * w) * bilinear(img, u, v, W, H);
      }
    }
  }
}

/* Clipped FP64 backprojection of one view: only voxels x in [start,stop) of
 * their (z,y) line are updated; taps outside the detector read zero, as the
 * zero-padded image does (LolaOPT.c:91-179, LolaOMP.c:65-135).  ranges holds
 * L*L int16 pairs (start, stop) for this view, index z*L+y
 * (LolaOPT.c:125-126). */
void orc_backproject_clipped(uint32_t L, float rL, float oL, uint32_t W, uint32_t H,
    const double *a, const float *img, const int16_t *ranges, double *vol)
{
#pragma omp parallel for schedule(static)
  for (long z = 0; z < (long)L; z++) {
    double wz = (double)oL + (double)z * (double)rL;
    for (long y = 0; y < (long)L; y++) {
      long start = ranges[2 * (z * L + y)];
      long stop  = ranges[2 * (z * L + y) + 1];
      double wy  = (double)oL + (double)y * (double)rL;
      for (long x = start; x < stop; x++) {
        double wx = (double)oL + (double)x * (double)rL;
        double u, v, w;
        project(a, wx, wy, wz, &u, &v, &w);
        vol[(z * L + y) * L + x] += bilinear(img, u, v, W, H) / (w * w);
      }
    }
  }
}

/* The clipped FP64 form restricted to a list of (z, y) lines: out[k*L + x]
 * += the view's contribution to voxel x of line (zs[k], ys[k]); for sampled
 * checks at sizes where the whole volume is out of reach (L=1024).  ranges
 * holds the view's L*L (start, stop) pairs. */
void orc_backproject_lines(uint32_t L, float rL, float oL, uint32_t W, uint32_t H,
    const double *a, const float *img, const int16_t *ranges, uint32_t nLines,
    const int32_t *zs, const int32_t *ys, double *out)
{
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)nLines; k++) {
    const long z = zs[k], y = ys[k];
    const double wz = (double)oL + (double)z * (double)rL;
    const double wy = (double)oL + (double)y * (double)rL;
    const long start = ranges[2 * (z * L + y)];
    const long stop  = ranges[2 * (z * L + y) + 1];
    for (long x = start; x < stop; x++) {
      const double wx = (double)oL + (double)x * (double)rL;
      double u, v, w;
      project(a, wx, wy, wz, &u, &v, &w);
      out[k * L + x] += bilinear(img, u, v, W, H) / (w * w);
    }
  }
}

/* LolaASM semantics: like the clipped form, but every line of every view
 * processes at least one vector of vs voxels starting at start, because each
 * fastrabbit kernel is a do-while loop (fastRabbitAVX.S:128-131,
 * fastRabbitAVX512.S:150-152 and its tail at :228-256); lines with a negative
 * start are skipped (LolaASM.c:152-154). */
void orc_backproject_asm(uint32_t L, float rL, float oL, uint32_t W, uint32_t H,
    const double *a, const float *img, const int16_t *ranges, int vs, double *vol)
{
#pragma omp parallel for schedule(static)
  for (long z = 0; z < (long)L; z++) {
    double wz = (double)oL + (double)z * (double)rL;
    for (long y = 0; y < (long)L; y++) {
      long start = ranges[2 * (z * L + y)];
      long stop  = ranges[2 * (z * L + y) + 1];
      if (start < 0)
        continue;
      long count = stop - start;
      if (count < vs)
        count = vs;
      double wy = (double)oL + (double)y * (double)rL;
      for (long x = start; x < start + count && x < (long)L; x++) {
        double wx = (double)oL + (double)x * (double)rL;
        double u, v, w;
        project(a, wx, wy, wz, &u, &v, &w);
        vol[(z * L + y) * L + x] += bilinear(img, u, v, W, H) / (w * w);
      }
    }
  }
}

/* Hit test of the clip scan: (int)u in [0, W-1] and (int)v in [0, H-1].  The
 * reference writes "iix > isx - 1" with isx unsigned, so iix = -1 is a miss as
 * well (analyseGeometry.c:104-106,235,261). */
static inline int clip_hit(const double *a, float wx, float wy, float wz, uint32_t W,
    uint32_t H)
{
  double u, v, w;
  project(a, (double)wx, (double)wy, (double)wz, &u, &v, &w);
  long iix = trunc_index(u);
  long iiy = trunc_index(v);
  return iix >= 0 && iix <= (long)W - 1 && iiy >= 0 && iiy <= (long)H - 1;
}

/* Line ranges of one view (analyseGeometry.c:207-281): for every (z,y) line
 * the first voxel whose ray hits the detector, scanning left to right, and the
 * last one, scanning right to left, as [start, stop); start is rounded down
 * and stop up to a multiple of vs; a line that never hits is [0,0).  World z
 * and y are float and advance by "+= mm" (analyseGeometry.c:214-220), world x
 * is the float x*mm+O (analyseGeometry.c:231,257). */
void orc_line_ranges(uint32_t L, float rL, float oL, uint32_t W, uint32_t H,
    const double *a, int vs, int16_t *out)
{
  /* Sequential float accumulation of wz and wy, exactly as the scan does. */
  float *wzs = (float *)malloc(L * sizeof(float));
  float *wys = (float *)malloc(L * sizeof(float));
  float acc  = oL;
  for (uint32_t i = 0; i < L; i++, acc += rL)
    wzs[i] = acc;
  acc = oL;
  for (uint32_t i = 0; i < L; i++, acc += rL)
    wys[i] = acc;

#pragma omp parallel for schedule(static)
  for (long z = 0; z < (long)L; z++) {
    for (long y = 0; y < (long)L; y++) {
      long xstart = -1, xstop = -1;
      for (long x = 0; x < (long)L; x++) {
        float wx = (float)x * rL + oL;
        if (clip_hit(a, wx, wys[y], wzs[z], W, H)) {
          xstart = x;
          break;
        }
      }
      int16_t *r = out + 2 * (z * L + y);
      if (xstart == -1) {
        r[0] = 0;
        r[1] = 0;
        continue;
      }
      for (long x = (long)L - 1; x >= 0; x--) {
        float wx = (float)x * rL + oL;
        if (clip_hit(a, wx, wys[y], wzs[z], W, H)) {
          xstop = x + 1;
          break;
        }
      }
      long al;
      if ((al = xstop % vs))
        xstop += vs - al;
      if ((al = xstart % vs))
        xstart -= al;
      r[0] = (int16_t)xstart;
      r[1] = (int16_t)xstop;
    }
  }
  free(wzs);
  free(wys);
}

/* Shadow of the volume on the detector: floor/ceil of the bounding box of the
 * eight corner-voxel projections over all views (analyseGeometry.c:50-93).
 * out = {Umin, Umax, Vmin, Vmax}. */
void orc_shadow(uint32_t L, float rL, float oL, uint32_t W, uint32_t H, uint32_t N,
    const double *geom, int32_t *out)
{
  double minU = (double)W + 1, minV = (double)H + 1, maxU = -2, maxV = -2;
  for (uint32_t n = 0; n < N; n++) {
    const double *a = geom + 12 * n;
    for (int c = 0; c < 8; c++) {
      double x = (double)oL + (double)((c & 2) ? L - 1 : 0) * rL;
      double y = (double)oL + (double)((c & 1) ? L - 1 : 0) * rL;
      double z = (double)oL + (double)((c & 4) ? L - 1 : 0) * rL;
      double u, v, w;
      project(a, x, y, z, &u, &v, &w);
      if (u < minU)
        minU = u;
      if (v < minV)
        minV = v;
      if (u > maxU)
        maxU = u;
      if (v > maxV)
        maxV = v;
    }
  }
  out[0] = (int32_t)floor(minU);
  out[1] = (int32_t)ceil(maxU);
  out[2] = (int32_t)floor(minV);
  out[3] = (int32_t)ceil(maxV);
}

/* The -c quality measure (rabbitRunner.c:23,338-373): HU = floor(h0*v+h1+0.5)
 * for v > 0, else 0, clamped to 4095; squared and absolute HU error against a
 * uint16 reference, an 11-bin histogram of floor(|err|).  res = {MSE, max}. */
void orc_hu_metric(const float *vol, uint64_t n, const uint16_t *ref, float h0, float h1,
    double *res, uint32_t *hist)
{
  double mse = 0.0, mx = 0.0;
  memset(hist, 0, 11 * sizeof(uint32_t));
  for (uint64_t i = 0; i < n; i++) {
    float hu = 0.0f;
    if (vol[i] > 0.0f)
      hu = floorf(h0 * vol[i] + h1 + 0.5f);
    if (hu > 4095.0f)
      hu = 4095.0f;
    float e  = hu - (float)ref[i];
    float ae = fabsf(e);
    if ((double)ae > mx)
      mx = (double)ae;
    mse += (double)(e * e);
    uint32_t b = (uint32_t)floorf(ae);
    if (b > 10)
      b = 10;
    hist[b]++;
  }
  res[0] = mse / (double)n;
  res[1] = mx;
}

/* HU conversion of a float volume into uint16 (the same rounding and clamp as
 * above); used to write Reference<L>.vol files for -c. */
void orc_hu_convert(const float *vol, uint64_t n, float h0, float h1, uint16_t *out)
{
  for (uint64_t i = 0; i < n; i++) {
    float hu = 0.0f;
    if (vol[i] > 0.0f)
      hu = floorf(h0 * vol[i] + h1 + 0.5f);
    if (hu > 4095.0f)
      hu = 4095.0f;
    out[i] = (uint16_t)hu;
  }
}

/* ---------------------------------------------------------------------------
 * Synthetic data: the FP64 form of the phantom generator (no counterpart in the
 * reference).  Ellipsoid e: centre c[3], rotation R (row-major 3x3, columns are
 * the ellipsoid's axes in world coordinates), semi-axes s[3], density rho.
 * ell = n_ell * 16 doubles: c0 c1 c2  s0 s1 s2  rho  R00..R22.
 * A ray through pixel (i,j) of view n starts at the source src[3n..] and runs
 * along Minv[9n..] * (i, j, 1) (row-major), normalised.  Line integral =
 * sum over ellipsoids of rho * chord length.
 * ------------------------------------------------------------------------- */
static double chord(const double *e, const double *s0, const double *d)
{
  /* Ellipsoid frame: q = diag(1/s) R^T (p - c). */
  double p[3] = { s0[0] - e[0], s0[1] - e[1], s0[2] - e[2] };
  const double *R = e + 7;
  double q0[3], q1[3];
  for (int k = 0; k < 3; k++) {
    double pk = R[0 * 3 + k] * p[0] + R[1 * 3 + k] * p[1] + R[2 * 3 + k] * p[2];
    double dk = R[0 * 3 + k] * d[0] + R[1 * 3 + k] * d[1] + R[2 * 3 + k] * d[2];
    q0[k]     = pk / e[3 + k];
    q1[k]     = dk / e[3 + k];
  }
  double A    = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
  double B    = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2];
  double C    = q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] - 1.0;
  double disc = B * B - A * C;
  if (disc <= 0.0)
    return 0.0;
  return 2.0 * sqrt(disc) / A;
}

void orc_phantom_line_integrals(uint32_t W, uint32_t H, uint32_t N, const double *src,
    const double *minv, uint32_t n_ell, const double *ell, double *out)
{
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (long n = 0; n < (long)N; n++) {
    for (long j = 0; j < (long)H; j++) {
      const double *M = minv + 9 * n;
      for (long i = 0; i < (long)W; i++) {
        double d[3];
        for (int r = 0; r < 3; r++)
          d[r] = M[3 * r + 0] * (double)i + M[3 * r + 1] * (double)j + M[3 * r + 2];
        double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        d[0] /= nrm;
        d[1] /= nrm;
        d[2] /= nrm;
        double acc = 0.0;
        for (uint32_t k = 0; k < n_ell; k++)
          acc += ell[16 * k + 6] * chord(ell + 16 * k, src + 3 * n, d);
        out[(n * H + j) * W + i] = acc;
      }
    }
  }
}

/* Row-wise discrete ramp filter (Ram-Lak, spatial form): h[0] = 1/(4 t^2),
 * h[k odd] = -1/(pi^2 k^2 t^2), h[k even] = 0; out[i] = t * sum_k h[k] in[i-k]
 * over the row (zero outside).  rows = number of rows, each of length W. */
void orc_ramp_filter_rows(uint64_t rows, uint32_t W, double t, const double *in,
    double *out)
{
  double *h = (double *)malloc(W * sizeof(double));
  h[0]      = 1.0 / (4.0 * t * t);
  for (uint32_t k = 1; k < W; k++)
    h[k] = (k & 1) ? -1.0 / (M_PI * M_PI * (double)k * (double)k * t * t) : 0.0;
#pragma omp parallel for schedule(static)
  for (uint64_t r = 0; r < rows; r++) {
    const double *x = in + r * W;
    double *y       = out + r * W;
    for (long i = 0; i < (long)W; i++) {
      double acc = 0.0;
      for (long m = 0; m < (long)W; m++) {
        long k = i - m;
        if (k < 0)
          k = -k;
        acc += h[k] * x[m];
      }
      y[i] = t * acc;
    }
  }
  free(h);
}

/* LolaOPT's own FP32 arithmetic, statement by statement (LolaOPT.c:131-173):
 * float world coordinates with wx advanced by "+= mm", the line-invariant part
 * tmp[] evaluated in double and rounded to float, u/v/w = tmp + wx*a[k] in
 * double rounded to float, rcpW = 1.0f/w, truncation, float lerp and a float
 * volume.  Taps outside the detector read zero (the zero padding).  Used to
 * measure how far the reference's own rounding lies from the FP64 result. */
void orc_backproject_opt_f32(uint32_t L, float mm, float o, uint32_t W, uint32_t H,
    const double *a, const float *img, const int16_t *ranges, float *vol)
{
#pragma omp parallel for schedule(static)
  for (long z = 0; z < (long)L; z++) {
    for (long y = 0; y < (long)L; y++) {
      int start = ranges[2 * (z * L + y)];
      int stop  = ranges[2 * (z * L + y) + 1];
      if (stop - start == 0)
        continue;
      float wz = (float)z * mm + o;
      float wy = (float)y * mm + o;
      float wx = (float)start * mm + o;
      float tmp0 = (float)(a[3] * wy + a[6] * wz + a[9]);
      float tmp1 = (float)(a[4] * wy + a[7] * wz + a[10]);
      float tmp2 = (float)(a[5] * wy + a[8] * wz + a[11]);
      for (int x = start; x < stop; ++x) {
        float u     = (float)(tmp0 + wx * a[0]);
        float v     = (float)(tmp1 + wx * a[1]);
        float w     = (float)(tmp2 + wx * a[2]);
        float rcpW  = 1.0f / w;
        float rcp2W = rcpW * rcpW;
        float ix    = u * rcpW;
        float iy    = v * rcpW;
        long iix    = trunc_index(ix);
        long iiy    = trunc_index(iy);
        float scalx = ix - (float)iix;
        float scaly = iy - (float)iiy;
        wx += mm;
        float bl = (float)pixel(img, iix, iiy, W, H);
        float br = (float)pixel(img, iix + 1, iiy, W, H);
        float tl = (float)pixel(img, iix, iiy + 1, W, H);
        float tr = (float)pixel(img, iix + 1, iiy + 1, W, H);
        float valb = (1 - scalx) * bl + scalx * br;
        float valt = (1 - scalx) * tl + scalx * tr;
        float val  = (1 - scaly) * valb + scaly * valt;
        vol[(z * L + y) * L + x] += val * rcp2W;
      }
    }
  }
}
