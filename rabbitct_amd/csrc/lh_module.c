/* lh_module.c -- LolaHIP as a RabbitCT algorithm module (C host side).
 *
 * prepare / iterate / finish of type int (*)(RabbitCtGlobalData *)
 * (moduleLoader_types.h:10-12) over the device layer of lh_device.hip.
 * Module state lives in statics, as in the reference modules
 * (LolaOPT.c:18-21).  Errors print a message and exit(EXIT_FAILURE)
 * (error.h:21-31).
 *
 * Tunables (environment; the runner's command line stays unchanged):
 *   LOLAHIP_PRECISION       f32 (default) | bf16 | f16   projection storage
 *   LOLAHIP_VIEWS_PER_PASS  views per register pass (default 16; 1 = plain)
 *   LOLAHIP_GRAPHS          1: replay each batch as a captured HIP graph
 *                           (default 0: in one runner process the capture
 *                           costs more than the replays save)
 *   LOLAHIP_RUN             voxels per thread along x: 8 | 16 (default) | 32
 *   LOLAHIP_VECTORSIZE      clip alignment for computed ranges: 4 | 8 | 16 (default)
 *   LOLAHIP_CLIP_DEVICE     gpu (default) | cpu: where missing ranges are computed
 *   LOLAHIP_DEVICE          HIP device ordinal (default 0)
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "lolahip.h"

/* ---- geometry analysis (CPU) ------------------------------------------- */

/* (int) of a double as C truncates it; values far outside int map off-detector. */
static inline long lh_trunc(double t)
{
  return (t > -1.0e9 && t < 1.0e9) ? (long)t : -2000000000L;
}

/* Hit test of the clip scan (analyseGeometry.c:95-100,235,261): the ray of the
 * voxel at float world (wx,wy,wz) hits the detector iff (int)u in [0,W-1] and
 * (int)v in [0,H-1]; "iix > isx-1" compares unsigned, so -1 misses too. */
static inline int lh_hit(const double *a, float wx, float wy, float wz, uint32_t W,
    uint32_t H)
{
  const double x = wx, y = wy, z = wz;
  const double w = a[2] * x + a[5] * y + a[8] * z + a[11];
  const double u = (a[0] * x + a[3] * y + a[6] * z + a[9]) / w;
  const double v = (a[1] * x + a[4] * y + a[7] * z + a[10]) / w;
  const long iix = lh_trunc(u), iiy = lh_trunc(v);
  return iix >= 0 && iix <= (long)W - 1 && iiy >= 0 && iiy <= (long)H - 1;
}

int lhComputeLineRanges(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *geometry, int vs, LhLineRange *out)
{
  if (!geometry || !out || L == 0 || (vs != 4 && vs != 8 && vs != 16))
    return LH_E_ARG;
  /* world z and y advance by "+= mm" in float (analyseGeometry.c:214-220) */
  float *wacc = (float *)malloc(L * sizeof(float));
  if (!wacc)
    return LH_E_NOMEM;
  float acc = O_L;
  for (uint32_t i = 0; i < L; i++, acc += R_L)
    wacc[i] = acc;

#pragma omp parallel for schedule(dynamic, 16) collapse(2)
  for (long n = 0; n < (long)nViews; n++) {
    for (long z = 0; z < (long)L; z++) {
      const double *a = geometry + 12 * n;
      for (long y = 0; y < (long)L; y++) {
        LhLineRange *r = out + ((size_t)n * L + z) * L + y;
        long xs = -1, xe = -1;
        for (long x = 0; x < (long)L; x++)
          if (lh_hit(a, (float)x * R_L + O_L, wacc[y], wacc[z], W, H)) {
            xs = x;
            break;
          }
        if (xs < 0) {
          r->start = 0;
          r->stop  = 0;
          continue;
        }
        for (long x = (long)L - 1; x >= 0; x--)
          if (lh_hit(a, (float)x * R_L + O_L, wacc[y], wacc[z], W, H)) {
            xe = x + 1;
            break;
          }
        long al;
        if ((al = xe % vs))
          xe += vs - al;
        if ((al = xs % vs))
          xs -= al;
        r->start = (int16_t)xs;
        r->stop  = (int16_t)xe;
      }
    }
  }
  free(wacc);
  return LH_OK;
}

void lhComputeShadow(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *g, LhShadow *out)
{
  double minU = (double)W + 1, minV = (double)H + 1, maxU = -2, maxV = -2;
  for (uint32_t n = 0; n < nViews; n++) {
    const double *a = g + 12 * n;
    for (int c = 0; c < 8; c++) {
      const double x = (double)O_L + (double)((c & 1) ? L - 1 : 0) * R_L;
      const double y = (double)O_L + (double)((c & 2) ? L - 1 : 0) * R_L;
      const double z = (double)O_L + (double)((c & 4) ? L - 1 : 0) * R_L;
      const double w = a[2] * x + a[5] * y + a[8] * z + a[11];
      const double u = (a[0] * x + a[3] * y + a[6] * z + a[9]) / w;
      const double v = (a[1] * x + a[4] * y + a[7] * z + a[10]) / w;
      minU = u < minU ? u : minU;
      maxU = u > maxU ? u : maxU;
      minV = v < minV ? v : minV;
      maxV = v > maxV ? v : maxV;
    }
  }
  out->Umin = (int32_t)floor(minU);
  out->Umax = (int32_t)ceil(maxU);
  out->Vmin = (int32_t)floor(minV);
  out->Vmax = (int32_t)ceil(maxV);
}

static int env_int(const char *name, int dflt)
{
  const char *s = getenv(name);
  return (s && *s) ? atoi(s) : dflt;
}

static int use_device_clip(void)
{
  const char *s = getenv("LOLAHIP_CLIP_DEVICE");
  return !(s && strcmp(s, "cpu") == 0);
}

int lhLineRangesFile(const char *clipFilename, uint32_t L, float R_L, float O_L,
    uint32_t W, uint32_t H, uint32_t N, const double *geometry, int vs, LhLineRange *out)
{
  char fallback[64];
  const char *path = clipFilename;
  if (path == NULL) {
    if (L != 128 && L != 256 && L != 512 && L != 1024) {
      printf("Unsupported problem size: %u\n", L);
      return LH_E_ARG;
    }
    snprintf(fallback, sizeof fallback, "RabbitInput/LineRange%u-%d.rct", L, vs);
    path = fallback;
    printf("No clipping file specified, falling back to %s\n", path);
  }
  const size_t want = (size_t)N * L * L * sizeof(LhLineRange);
  FILE *f           = fopen(path, "rb");
  if (f != NULL) {
    fseek(f, 0L, SEEK_END);
    const long fsize = ftell(f);
    if (fsize < 0 || (size_t)fsize != want) {
      printf("corrupted file %s: size is %lu (should be %lu)\n", path, (unsigned long)fsize,
          (unsigned long)want);
      fclose(f);
      return LH_E_IO;
    }
    fseek(f, 0L, SEEK_SET);
    const size_t got = fread(out, sizeof(LhLineRange), (size_t)N * L * L, f);
    fclose(f);
    if (got != (size_t)N * L * L) {
      fprintf(stderr, "error reading from %s\n", path);
      return LH_E_IO;
    }
    printf("Read %lu bytes of clipping data from %s\n", (unsigned long)want, path);
    return LH_OK;
  }
  if (errno != ENOENT) {
    fprintf(stderr, "fopen %s: %s\n", path, strerror(errno));
    return LH_E_IO;
  }
  if ((f = fopen(path, "wb")) == NULL) {
    fprintf(stderr, "fopen %s: %s\n", path, strerror(errno));
    return LH_E_IO;
  }
  int rc = LH_E_NODEV;
  if (use_device_clip())
    rc = lhComputeLineRangesDevice(L, R_L, O_L, W, H, N, geometry, vs, out);
  if (rc != LH_OK)
    rc = lhComputeLineRanges(L, R_L, O_L, W, H, N, geometry, vs, out);
  if (rc != LH_OK) {
    fclose(f);
    return rc;
  }
  const size_t put = fwrite(out, sizeof(LhLineRange), (size_t)N * L * L, f);
  fclose(f);
  if (put != (size_t)N * L * L) {
    fprintf(stderr, "error writing to %s\n", path);
    return LH_E_IO;
  }
  printf("Wrote %lu bytes of clipping data to %s\n", (unsigned long)want, path);
  return LH_OK;
}

/* ---- module ------------------------------------------------------------- */

static LhContext *Ctx;
static uint32_t TotalViews;   /* N from prepare */
static uint32_t PerPass;      /* views per register pass */
static uint32_t BatchBase;    /* first slot of the batch being filled */
static uint32_t Pending;      /* views loaded into the current batch */

static void die(const char *what, int rc)
{
  fprintf(stderr, "LolaHIP: %s failed (%d): %s\n", what, rc, lhLastError());
  exit(EXIT_FAILURE);
}

static int parse_precision(void)
{
  const char *s = getenv("LOLAHIP_PRECISION");
  if (!s || !*s || strcmp(s, "f32") == 0)
    return LH_PREC_F32;
  if (strcmp(s, "bf16") == 0)
    return LH_PREC_BF16;
  if (strcmp(s, "f16") == 0)
    return LH_PREC_F16;
  fprintf(stderr, "LolaHIP: unknown LOLAHIP_PRECISION '%s' (f32, bf16, f16)\n", s);
  exit(EXIT_FAILURE);
}

int lolaHipPrepare(RabbitCtGlobalData *rcgd)
{
  /* numberOfProjections is N only when the runner got -a (rabbitRunner.c:209-217) */
  if (rcgd->numberOfProjections == 0) {
    fprintf(stderr,
        "This module needs global geometry information.\n"
        "Please add -a switch!\n");
    exit(EXIT_FAILURE);
  }
  const uint32_t L = rcgd->problemSize, N = rcgd->numberOfProjections;
  const int vs     = env_int("LOLAHIP_VECTORSIZE", 16);
  int per          = env_int("LOLAHIP_VIEWS_PER_PASS", 16);
  if (per < 1)
    per = 1;
  if ((uint32_t)per > N)
    per = (int)N;

  LhConfig cfg;
  lhDefaultConfig(&cfg, L, rcgd->imageWidth, rcgd->imageHeight, N);
  cfg.R_L          = rcgd->voxelSize;
  cfg.O_L          = rcgd->O_Index;
  cfg.viewsPerPass = (uint32_t)per;
  cfg.nSlots       = 2u * (uint32_t)per;
  cfg.kernel       = per > 1 ? LH_KERNEL_FUSED : LH_KERNEL_PLAIN;
  cfg.precision    = parse_precision();
  cfg.runLength    = (uint32_t)env_int("LOLAHIP_RUN", 16);
  cfg.clip         = 1;
  cfg.device       = env_int("LOLAHIP_DEVICE", 0);

  int rc = lhCreate(&cfg, &Ctx);
  if (rc != LH_OK)
    die("device set-up", rc);
  /* batches replay captured HIP graphs with LOLAHIP_GRAPHS=1 */
  if ((rc = lhSetGraphs(Ctx, env_int("LOLAHIP_GRAPHS", 0))) != LH_OK)
    die("graph set-up", rc);

  LhLineRange *range = (LhLineRange *)malloc((size_t)N * L * L * sizeof(LhLineRange));
  if (range == NULL) {
    perror("malloc");
    exit(EXIT_FAILURE);
  }
  rc = lhLineRangesFile(rcgd->clipFilename, L, rcgd->voxelSize, rcgd->O_Index,
      rcgd->imageWidth, rcgd->imageHeight, N, rcgd->globalGeometry, vs, range);
  if (rc != LH_OK)
    exit(EXIT_FAILURE);
  if ((rc = lhSetRanges(Ctx, range)) != LH_OK)
    die("range upload", rc);
  free(range);
  /* the module accumulates (+=) into the caller's volume */
  if ((rc = lhSetVolume(Ctx, rcgd->volumeData)) != LH_OK)
    die("volume upload", rc);

  TotalViews = N;
  PerPass    = (uint32_t)per;
  BatchBase  = 0;
  Pending    = 0;
  return 1;
}

static void parallel_copy(float *dst, const float *src, size_t n)
{
  const long parts = 64;
  const size_t chunk = (n + (size_t)parts - 1) / (size_t)parts;
#pragma omp parallel for schedule(static)
  for (long k = 0; k < parts; k++) {
    const size_t a = (size_t)k * chunk;
    if (a < n)
      memcpy(dst + a, src + a, ((a + chunk < n) ? chunk : n - a) * sizeof(float));
  }
}

/* Launches the batch being filled; returns 1 if it launched work. */
static int flush_batch(void)
{
  if (Pending == 0)
    return 0;
  int rc = lhBackproject(Ctx, BatchBase, Pending);
  if (rc != LH_OK)
    die("backprojection", rc);
  BatchBase = (BatchBase == 0) ? PerPass : 0;
  Pending   = 0;
  return 1;
}

int lolaHipBackprojection(RabbitCtGlobalData *rcgd)
{
  if (Ctx == NULL) {
    fprintf(stderr, "LolaHIP: backprojection called before prepare\n");
    exit(EXIT_FAILURE);
  }
  int launched = 0;
  for (uint32_t p = 0; p < rcgd->numberOfProjections; p++) {
    const Projection *pr = &rcgd->projectionBuffer[p];
    /* the image goes through the context's pinned ring (copied by all host
     * threads); its H2D copy and staging run behind the kernels in flight */
    float *buf = NULL;
    int rc     = lhAcquireUpload(Ctx, &buf);
    if (rc != LH_OK)
      die("view upload", rc);
    parallel_copy(buf, pr->image, (size_t)rcgd->imageWidth * rcgd->imageHeight);
    if ((rc = lhCommitUpload(Ctx, BatchBase + Pending, pr->id, pr->matrix)) != LH_OK)
      die("view upload", rc);
    Pending++;
    if (Pending == PerPass || (uint32_t)pr->id == TotalViews)
      launched |= flush_batch();
  }
  /* A call that launched a backprojection returns only when that work (and
   * the uploads enqueued before it) has finished, so the runner's timer
   * around this call (rabbitRunner.c:55-58,279) measures it.  Views that are
   * only buffered for the next batch may still be crossing PCIe. */
  if (launched) {
    int rc = lhSynchronize(Ctx);
    if (rc != LH_OK)
      die("synchronize", rc);
  }
  return 1;
}

int lolaHipQueryIdle(void)
{
  if (Ctx == NULL)
    return -1;
  return lhQueryIdle(Ctx);
}

int lolaHipFinish(RabbitCtGlobalData *rcgd)
{
  if (Ctx == NULL)
    return 1;
  flush_batch();
  int rc = lhGetVolume(Ctx, rcgd->volumeData);
  if (rc != LH_OK)
    die("volume download", rc);
  lhDestroy(Ctx);
  Ctx = NULL;
  return 1;
}

/* ---------------------------------------------------------------------------
 * .rct ingest straight into the device (no counterpart in the reference, whose
 * runner reads each record into its own buffers, ctFileReader.c:47-61): views
 * [firstView, firstView+count) are read with pread by all host threads into
 * the context's pinned upload ring and committed to slots firstSlot.., so the
 * file reads overlap the H2D copies and staging of earlier views; the kernel
 * is asked to read the next kReadAhead records ahead (POSIX_FADV_WILLNEED),
 * so a file not in the page cache streams in while earlier views are copied.
 * The record layout is ctFileReader.c:51-58.  matrices (optional,
 * count x 12) receives the views' matrices.
 * ------------------------------------------------------------------------- */
static int pread_all(int fd, char *dst, size_t len, off_t off)
{
  while (len > 0) {
    const ssize_t r = pread(fd, dst, len, off);
    if (r <= 0) {
      if (r < 0 && errno == EINTR)
        continue;
      return -1;
    }
    dst += r;
    len -= (size_t)r;
    off += r;
  }
  return 0;
}

#ifndef LH_READ_AHEAD
#define LH_READ_AHEAD 32
#endif
enum { kReadAhead = LH_READ_AHEAD };

int lhLoadRctFile(LhContext *ctx, const char *path, uint32_t firstView, uint32_t count,
    uint32_t firstSlot, double *matrices)
{
  LhConfig cfg;
  if (!ctx || !path || lhGetConfig(ctx, &cfg) != LH_OK)
    return LH_E_ARG;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) {
    fprintf(stderr, "LolaHIP: cannot open %s: %s\n", path, strerror(errno));
    return LH_E_IO;
  }
  uint32_t hdr[5];
  int rc = LH_OK;
  if (pread_all(fd, (char *)hdr, sizeof hdr, 0) != 0) {
    rc = LH_E_IO;
  } else if (hdr[0] != cfg.W || hdr[1] != cfg.H || firstView + count > hdr[2]) {
    fprintf(stderr, "LolaHIP: %s holds %u views of %ux%u, asked for views [%u,%u) of %ux%u\n",
        path, hdr[2], hdr[0], hdr[1], firstView, firstView + count, cfg.W, cfg.H);
    rc = LH_E_ARG;
  }
  const size_t imgBytes = (size_t)cfg.W * cfg.H * sizeof(float);
  const size_t record   = 12 * sizeof(double) + imgBytes;
  uint32_t advised      = 0; /* records [0, advised) asked for */
  for (uint32_t v = 0; rc == LH_OK && v < count; v++) {
    const off_t off = (off_t)20 + (off_t)(firstView + v) * (off_t)record;
    if (advised < count && advised < v + kReadAhead / 2) {
      /* refilled to kReadAhead records ahead each time fewer than half remain */
      const uint32_t end = v + kReadAhead < count ? v + kReadAhead : count;
      (void)posix_fadvise(fd, (off_t)20 + (off_t)(firstView + advised) * (off_t)record,
          (off_t)(end - advised) * (off_t)record, POSIX_FADV_WILLNEED);
      advised = end;
    }
    double mat[12];
    float *buf = NULL;
    if (pread_all(fd, (char *)mat, sizeof mat, off) != 0) {
      rc = LH_E_IO;
      break;
    }
    if ((rc = lhAcquireUpload(ctx, &buf)) != LH_OK)
      break;
    const long parts   = 16;
    const size_t chunk = (imgBytes + (size_t)parts - 1) / (size_t)parts;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (long k = 0; k < parts; k++) {
      const size_t a = (size_t)k * chunk;
      if (a < imgBytes) {
        const size_t len = (a + chunk < imgBytes) ? chunk : imgBytes - a;
        bad |= pread_all(fd, (char *)buf + a, len, off + (off_t)sizeof mat + (off_t)a) != 0;
      }
    }
    if (bad) {
      (void)lhCancelUpload(ctx);
      rc = LH_E_IO;
      break;
    }
    if ((rc = lhCommitUpload(ctx, firstSlot + v, (int)(firstView + v + 1), mat)) != LH_OK)
      break;
    if (matrices)
      memcpy(matrices + 12 * (size_t)v, mat, sizeof mat);
  }
  close(fd);
  if (rc == LH_E_IO)
    fprintf(stderr, "LolaHIP: short read from %s\n", path);
  return rc;
}
