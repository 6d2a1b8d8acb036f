/* lolahip.h -- C-ABI of LolaHIP, an MI355X (gfx950) cone-beam backprojector
 * that plugs into RabbitCT's algorithm registry.
 *
 * Two layers, both plain C linkage with plain pointers and sizes:
 *
 *  1. The RabbitCT module interface: three functions of type
 *     int (*)(RabbitCtGlobalData *), registered as
 *     { "LolaHIP", lolaHipPrepare, lolaHipBackprojection, lolaHipFinish }.
 *  2. The device layer under it (lh*): a reconstruction context that owns the
 *     volume, the line ranges and a set of projection slots in HBM, for
 *     callers that keep their inputs resident on the GPU.
 *
 * Reference interfaces named below are files of the RabbitCT sources.
 */
#ifndef LOLAHIP_H
#define LOLAHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The module's data carrier.  Layout-identical to the reference's
 * Projection / RabbitCtGlobalData (rabbitCt.h:9-27); when the reference header
 * has been included first its definitions are used instead (same include
 * guard), so a registry file can include both headers. */
#ifndef RABBITRUNNER_TYPES_H
#define RABBITRUNNER_TYPES_H
typedef struct {
  int id;         /* 1-based view number (ctFileReader.c:47-61) */
  double *matrix; /* 12 doubles, column-major 3x4: a[c*3+r] */
  float *image;   /* imageHeight x imageWidth floats, row-major */
} Projection;

typedef struct {
  uint32_t problemSize;   /* L */
  uint32_t imageWidth;    /* W */
  uint32_t imageHeight;   /* H */
  float voxelSize;        /* R_L, mm */
  float O_Index;          /* O_L = -R_L*(L-1)/2, mm */
  float *volumeData;      /* L^3 floats, x fastest: f[(z*L+y)*L+x] */
  uint32_t numberOfProjections; /* N at prepare (with -a), b at iterate */
  double *globalGeometry; /* N x 12 doubles (valid only with -a) */
  Projection *projectionBuffer; /* b buffered views */
  char *clipFilename;     /* -C file or NULL */
} RabbitCtGlobalData;
#endif

/* One line range: voxels x in [start, stop) of a (z,y) line are updated
 * (analyseGeometry_types.h:23-26). */
typedef struct {
  int16_t start;
  int16_t stop;
} LhLineRange;

/* Bounding box of the volume's shadow on the detector
 * (OutShadowType, analyseGeometry_types.h:16-21). */
typedef struct {
  int32_t Umin, Umax, Vmin, Vmax;
} LhShadow;

/* ---------------------------------------------------------------------------
 * 1. RabbitCT module interface (moduleLoader_types.h:10-12).
 * Errors follow the reference's convention: a message on stderr and
 * exit(EXIT_FAILURE) (error.h:21-31, LolaOPT.c:26-31).  Return value 1.
 * ------------------------------------------------------------------------- */

/* Replaces lolaOptPrepare (LolaOPT.c:23-84): requires the global geometry
 * (-a), builds or loads the line ranges through the reference's clip-file
 * contract (analyseGeometry.c:102-298), allocates the volume, the ranges and
 * the projection slots on the GPU and uploads the (zeroed) volume. */
int lolaHipPrepare(RabbitCtGlobalData *rcgd);

/* Replaces lolaOptBackprojection (LolaOPT.c:91-179): backprojects the
 * rcgd->numberOfProjections buffered views into the volume (+=).  Views are
 * collected into batches of LOLAHIP_VIEWS_PER_PASS (a batch may span calls:
 * 496 views admit only -b values that divide 496); a batch launches when it
 * is full or holds view id == N, and a call that launched a batch returns
 * only after that work has finished, so the runner's timer around the call
 * (rabbitRunner.c:55-58,279) covers it. */
int lolaHipBackprojection(RabbitCtGlobalData *rcgd);

/* Replaces lolaOptFinish (LolaOPT.c:86-89): flushes pending views, copies the
 * volume into rcgd->volumeData and releases the GPU resources. */
int lolaHipFinish(RabbitCtGlobalData *rcgd);

/* 1 when the module has no GPU work in flight, 0 while some runs, -1 before
 * prepare (lhQueryIdle on the module's context).  A check of the timing
 * contract above; no counterpart in the reference. */
int lolaHipQueryIdle(void);

/* ---------------------------------------------------------------------------
 * Geometry analysis with plain pointers (analyseGeometry.h:12-13).
 * ------------------------------------------------------------------------- */

/* computeLineRanges (analyseGeometry.c:207-281), CPU form: ranges for views
 * [0, nViews) of geometry (nViews x 12 doubles) into out[nViews*L*L],
 * aligned to vectorSize (4, 8 or 16).  Returns 0 or a negative LH_E* code. */
int lhComputeLineRanges(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *geometry, int vectorSize, LhLineRange *out);

/* The same on the GPU (bit-identical); out is host memory. */
int lhComputeLineRangesDevice(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *geometry, int vectorSize, LhLineRange *out);

/* The clip-file contract of computeLineRanges (analyseGeometry.c:110-197,
 * 283-296): path = clipFilename, or "RabbitInput/LineRange<L>-<VS>.rct" when
 * NULL; an existing file of N*L*L*4 bytes is read, a missing one is computed
 * and written.  Returns 0, or a negative LH_E* code (size mismatch, I/O). */
int lhLineRangesFile(const char *clipFilename, uint32_t L, float R_L, float O_L,
    uint32_t W, uint32_t H, uint32_t nViews, const double *geometry, int vectorSize,
    LhLineRange *out);

/* computeShadowOfProjection (analyseGeometry.c:50-93). */
void lhComputeShadow(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *geometry, LhShadow *out);

/* ---------------------------------------------------------------------------
 * 2. Device layer.  No counterpart in the reference: it is the GPU form of
 * the state LolaOPT keeps in statics (LolaOPT.c:18-21) plus the volume.
 * ------------------------------------------------------------------------- */

enum { LH_PREC_F32 = 0, LH_PREC_BF16 = 1, LH_PREC_F16 = 2 };
enum { LH_KERNEL_PLAIN = 0, LH_KERNEL_FUSED = 1 };

enum {
  LH_OK = 0,
  LH_E_ARG = -1,      /* invalid argument / unsupported size */
  LH_E_NODEV = -2,    /* no visible GPU */
  LH_E_NOMEM = -3,    /* device or pinned allocation failed */
  LH_E_HIP = -4,      /* other HIP runtime error */
  LH_E_IO = -5,       /* file read/write/size error */
  LH_E_STATE = -6     /* call out of order */
};

typedef struct {
  uint32_t L;          /* volume edge, multiple of 64 */
  float R_L;           /* voxel size, mm */
  float O_L;           /* volume origin, mm */
  uint32_t W, H;       /* detector size, pixels */
  uint32_t nViews;     /* N: number of views (ids 1..N) */
  uint32_t nSlots;     /* projection slots resident in HBM */
  int precision;       /* LH_PREC_*: storage type of the projections */
  int kernel;          /* LH_KERNEL_*: plain launch or persistent fused */
  uint32_t viewsPerPass; /* fused: views accumulated per register pass */
  uint32_t runLength;  /* voxels along x per thread: 8, 16 or 32 (0 = auto) */
  int clip;            /* 1: line ranges (LolaOPT semantics), 0: all voxels */
  int staging;         /* 1: per-view footprints staged in LDS (FP32), 0: L2 gathers */
  int device;          /* HIP device ordinal */
} LhConfig;

typedef struct LhContext LhContext;

/* Fills cfg with defaults for (L, W, H, nViews). */
void lhDefaultConfig(LhConfig *cfg, uint32_t L, uint32_t W, uint32_t H, uint32_t nViews);

/* Creates a context on cfg->device.  On failure *out is null and nothing
 * stays allocated. */
int lhCreate(const LhConfig *cfg, LhContext **out);
int lhDestroy(LhContext *ctx);

/* Line ranges of all N views (N*L*L entries, host memory); required when
 * cfg.clip == 1. */
int lhSetRanges(LhContext *ctx, const LhLineRange *ranges);
/* Computes the ranges on the GPU from geometry (N x 12) instead. */
int lhSetRangesFromGeometry(LhContext *ctx, const double *geometry, int vectorSize);

/* Volume upload / zero / download (host memory, L^3 floats). */
int lhSetVolume(LhContext *ctx, const float *host);
int lhZeroVolume(LhContext *ctx);
int lhGetVolume(LhContext *ctx, float *host);

/* Uploads view `id` (1-based) into projection slot `slot`: the matrix
 * (12 doubles) and the W x H float image (host memory), converted to the
 * context's storage precision. */
int lhLoadView(LhContext *ctx, uint32_t slot, int id, const double *matrix,
    const float *image);
/* Same, with the image already in device memory (W x H floats).  The image
 * is read by kernels on the context's stream: it must stay unchanged until
 * the stream has passed this call (work enqueued on the same stream after it
 * may overwrite it). */
int lhLoadViewDevice(LhContext *ctx, uint32_t slot, int id, const double *matrix,
    const float *dImage);

/* lhLoadView in two halves, so the caller can fill the pinned buffer itself
 * (e.g. with several threads).  Acquire returns the next of the context's
 * pinned W x H float buffers, waiting only until its previous H2D copy has
 * read it; commit enqueues its copy on the context's upload stream and the
 * staging into `slot` on the compute stream, ordered by events. */
int lhAcquireUpload(LhContext *ctx, float **hostBuffer);
int lhCommitUpload(LhContext *ctx, uint32_t slot, int id, const double *matrix);
/* Gives an acquired buffer back without uploading it. */
int lhCancelUpload(LhContext *ctx);

/* Enqueues the backprojection of slots [firstSlot, firstSlot+count) on the
 * context's stream (asynchronous). */
int lhBackproject(LhContext *ctx, uint32_t firstSlot, uint32_t count);

/* lhBackproject without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory while they run (code under HIP
 * graph capture, tight serving loops).  The footprint boxes of each register
 * pass go to a fixed 32 MiB table that the library holds per GPU instead of
 * the context's table; the caller's slots, ranges and volume are used as by
 * lhBackproject, with the same result bit for bit.  Launches are eager (no
 * graph replay).  Returns LH_E_ARG when one pass needs more than the table's
 * 4 Mi boxes (views per pass x tiles, tiles = (L/runLength) x (L/32) x
 * (L/8): at most 128 views per pass at L=512 with 16-voxel runs).  Calls
 * that overlap on one GPU share that table and race.  The persistent tile
 * walk uses the list of tiles holding work that an earlier lhBackproject of
 * the same launch shape built; without one it walks every tile (the same
 * result, slower where the volume's ends are clipped). */
int lhBackprojectNoAlloc(LhContext *ctx, uint32_t firstSlot, uint32_t count);

/* Runs `reps` times the backprojection of all slots [0, count) in passes of
 * cfg.viewsPerPass and returns, measured with HIP events on the context's
 * stream, the mean milliseconds per rep and per-launch statistics. */
int lhTimedBackproject(LhContext *ctx, uint32_t count, int reps, float *msPerRep,
    float *msPerLaunch, uint32_t *launchesPerRep);

int lhSynchronize(LhContext *ctx);

/* 1 when all work enqueued on the context's streams (compute, upload,
 * staging) has completed, 0 while some is still running, a negative LH_E*
 * code on error.  Does not wait.  No counterpart in the reference. */
int lhQueryIdle(LhContext *ctx);

/* Compute units and peak engine clock (kHz) of a device, for rate
 * ceilings.  No counterpart in the reference. */
int lhDeviceInfo(int device, int *computeUnits, int *clockKHz);

/* Pinned (page-locked) host memory, so H2D copies from it run asynchronously
 * at the DMA rate.  The platform refuses to pin memory that already exists,
 * so callers that upload from host memory allocate it here.  No counterpart
 * in the reference (memoryUtilsAllocate, memoryUtils.c:43-67, allocates
 * pageable memory). */
int lhHostAlloc(void **p, uint64_t bytes);
int lhHostFree(void *p);

/* The whole reconstruction from host memory, as a caller that holds the views
 * in RAM runs it (the per-view copy inside LolaOPT's timed iterate,
 * LolaOPT.c:111-117, plus the backprojection): views [0, count) of images
 * (count x H x W floats, host; pinned memory from lhHostAlloc makes the
 * copies asynchronous), matrices (count x 12) and ids (count, 1-based) are
 * copied into a device ring of two batches on the context's upload stream and
 * staged into the slots (one launch per batch) and backprojected (+=) on the
 * compute stream, in batches of viewsPerPass views (fused; 16 for plain
 * launches; the first and the last two batches a quarter as long, so the
 * pipeline fills and drains quickly) over nSlots / batch slot sets: batch k+1
 * crosses PCIe while batch k is backprojected.  With lhSetGraphs on, each
 * batch's backprojection replays a captured HIP graph.  Asynchronous:
 * lhSynchronize waits for the end.  The host arrays must stay unchanged until
 * then.  No counterpart in the reference. */
int lhReconstructHost(LhContext *ctx, const float *images, const double *matrices,
    const int *ids, uint32_t count);

/* Half-precision host copies of views for lhReconstructHostHalf: count views
 * of W x H floats (images) become count x W x H halves (FP16 or BF16 bit
 * patterns, precision LH_PREC_F16 / LH_PREC_BF16) and one storage exponent m
 * per view: FP16 holds p * 4^-m with m such that max|p| 4^-m lies in
 * (4096, 16384], BF16 holds p (m = 0); both round to nearest even, as the
 * device staging of float uploads rounds.  Host code (all host threads).  No
 * counterpart in the reference, whose projections cross into the module as
 * float32 (rabbitCt.h:20-27). */
int lhHostHalfViews(int precision, const float *images, uint32_t count, uint32_t W, uint32_t H,
    uint16_t *halves, int32_t *exponents);

/* lhReconstructHost from half-precision host views (lhHostHalfViews output:
 * count x H x W halves and their storage exponents, for the context's
 * precision, BF16 or FP16): half the bytes cross PCIe.  The staging forms
 * each slot's pairs from the uploaded halves: h(i,j) and the forward
 * difference h(i+1,j) - h(i,j) rounded to the storage precision (float uploads
 * take the difference from the unrounded p(i+1,j)).  No counterpart in the
 * reference. */
int lhReconstructHostHalf(LhContext *ctx, const uint16_t *halves, const int32_t *exponents,
    const double *matrices, const int *ids, uint32_t count);

/* The batch sizes lhReconstructHost uses for `count` views in batches of
 * `batch` (the first and the last two shortened): returns the number of
 * batches and writes up to maxSizes of them to sizes (may be null). */
uint32_t lhHostBatchPlan(uint32_t count, uint32_t batch, uint32_t *sizes, uint32_t maxSizes);

/* lhReconstructHost `reps` times, each from idle streams, timed with HIP
 * events from before the first H2D copy to after the last backprojection:
 * mean milliseconds per repetition. */
int lhTimedReconstructHost(LhContext *ctx, const float *images, const double *matrices,
    const int *ids, uint32_t count, int reps, float *msPerRep);

/* lhTimedReconstructHost for lhReconstructHostHalf. */
int lhTimedReconstructHostHalf(LhContext *ctx, const uint16_t *halves, const int32_t *exponents,
    const double *matrices, const int *ids, uint32_t count, int reps, float *msPerRep);

/* The PCIe leg of lhReconstructHost alone: the H2D copies of views
 * [0, count) of images (host, count x H x W floats) into its device ring, one
 * copy per batch as lhReconstructHost makes them, back to back on the upload
 * stream, `reps` times; mean milliseconds per repetition (HIP events).  No
 * counterpart in the reference. */
int lhTimedUploadHost(LhContext *ctx, const float *images, uint32_t count, int reps,
    float *msPerRep);

/* lhTimedUploadHost for half-precision views (count x H x W halves). */
int lhTimedUploadHostHalf(LhContext *ctx, const uint16_t *halves, uint32_t count, int reps,
    float *msPerRep);

/* enable != 0: lhBackproject replays one captured HIP graph per batch (slot
 * range and launch shape), capturing it on first use. */
int lhSetGraphs(LhContext *ctx, int enable);

/* Changes the launch shape of later backprojections (kernel, views per
 * register pass, voxels per thread along x, LDS staging) without touching
 * resident data; the autotuner's knob. */
int lhSetLaunch(LhContext *ctx, int kernel, uint32_t viewsPerPass, uint32_t runLength,
    int staging);

/* Tile shape of later backprojections.  16-voxel runs: tileY = 32 (or 0,
 * the default) for 16 x 32 x 8 voxels (x, y, z) staged through a 128 x
 * 39-pixel LDS window, 16 for 16 x 16 x 16 (an 80 x 62 window: narrow, tall
 * footprints), 64 for 16 x 64 x 4 (224 x 22: wide, short ones).  32-voxel
 * runs: 32 (or 0) for 32 x 32 x 8, 64 for 32 x 64 x 4.  8-voxel runs keep
 * 8 x 32 x 8.  A footprint larger than its window is
 * gathered from L2, so every shape gives the volume LolaOPT's arithmetic
 * defines; results may differ between shapes in the last bits (each tile's
 * local coordinate origin is its footprint's).  No counterpart in the
 * reference (LolaOPT.c:131-173 walks voxels one line at a time); the
 * autotuner's second knob. */
int lhSetTile(LhContext *ctx, uint32_t tileY);

/* Restricts later backprojections to the z-slab [z0, z1) of the volume
 * (multiples of 8; [0, L) after lhCreate): one GPU's share of a volume
 * sharded into z-slabs.  Voxels outside the slab are left untouched, and
 * lhEffectiveUpdates still counts the whole volume.  No counterpart in the
 * reference, whose loops always cover the whole volume. */
int lhSetSlab(LhContext *ctx, uint32_t z0, uint32_t z1);

/* Gather roof: updates/s of the backprojector's inner step without the
 * projection (coordinate step, floor/fraction, four taps, bilinear lerp,
 * accumulate) over 2D-local lane paths, taps from LDS (lds = 1) or from an
 * L2-resident image; blocksPerCU workgroups of 256 lanes per CU, `passes`
 * runs of 16 updates per lane (<= 1000).  No counterpart in the reference. */
int lhGatherRoof(int lds, int blocksPerCU, int passes, double *updatesPerSec, float *ms);

/* Device-to-device copy of `bytes` on the context's stream, timed with HIP
 * events: the reference copy against which HBM counters are scaled. */
int lhCopyCalibration(LhContext *ctx, uint64_t bytes, float *ms);

/* Effective voxel-view updates of views [0,count) of the slots: the number of
 * voxels inside their line ranges (L^3 per view without clipping). */
uint64_t lhEffectiveUpdates(LhContext *ctx, uint32_t count);

/* Synthetic input (no counterpart in the reference): for views
 * ids[0..count) (1-based; they select the line ranges) fills slots
 * [firstSlot, firstSlot+count) with ramp-filtered,
 * cosine- and redundancy-weighted analytic line integrals of an ellipsoid
 * phantom.  Per view v: src[3v..] source position, minv[9v..] row-major
 * inverse of the matrix's 3x3 part (ray through pixel (i,j) along
 * minv*(i,j,1)), axis[3v..] unit principal axis, parker[W*v..] column
 * weights; ell = nEll x 16 doubles (centre 3, semi-axes 3, density, rotation
 * 9 row-major); tau = ramp-filter sample spacing (mm), dbeta = angular step
 * (rad), border = zero the outermost pixel ring.  hostImages (optional,
 * count*H*W floats) receives the float32 projections. */
int lhGenerateViews(LhContext *ctx, uint32_t firstSlot, uint32_t count, const int *ids,
    const double *matrices, const double *src, const double *minv, const double *axis,
    const float *parker, uint32_t nEll, const double *ell, double tau, double dbeta,
    int border, float *hostImages);

/* vol *= s on the context's stream (the weight of a view subset: a sparse
 * scan of every k-th view is scaled by k). */
int lhScaleVolume(LhContext *ctx, float s);

/* Device pointer of the context's resident volume (L^3 floats, [z][y][x]);
 * borrowed, valid until lhDestroy. */
int lhVolumeDevice(LhContext *ctx, float **dVol);

/* ---------------------------------------------------------------------
 * Sparse-view + learned correction (no counterpart in the reference).
 * All pointers named d* are device memory; calls are synchronous.
 * ------------------------------------------------------------------- */

/* Unweighted analytic line integrals of an ellipsoid phantom (the raw data
 * of held-out views) on the detector subsampled by `stride`:
 * dOut = count x ceil(H/stride) x ceil(W/stride) floats. */
int lhLineIntegrals(int W, int H, int stride, uint32_t count, const double *src,
    const double *minv, uint32_t nEll, const double *ell, float *dOut);

/* Ray-marching forward projector of a volume (L^3, voxel size R mm, origin
 * O mm) at `count` views (src count x 3, minv count x 9), detector
 * subsampled by `stride`, step h mm, trilinear samples, midpoint rule. */
int lhForwardProject(const float *dVol, int L, float R, float O, int W, int H, int stride,
    uint32_t count, const double *src, const double *minv, float h, float *dOut);

/* out = sqrt(sum (a-b)^2 / sum b^2) over n floats. */
int lhRelativeResidual(const float *dA, const float *dB, uint64_t n, double *out);

/* Gathers n patches of P^3 voxels from a stack of nVols volumes of L^3:
 * spec (host) = n x (volume, z0, y0, x0). */
int lhExtractPatches(const float *dVols, int nVols, int L, const int *spec, int n, int P,
    float *dOut);

/* 3-level 3D U-Net: 3x3x3 convolutions (BF16 weights and activations, FP32
 * accumulation on MFMA), ReLU, 2x2x2 average pooling, nearest upsampling,
 * channel concatenation, a 1x1x1 head and a residual connection:
 *   out = in + head(dec1(cat(up(dec2(cat(up(bottom(pool(enc2(pool(enc1(in))))))),
 *                                  enc2))), enc1))
 * channels enc1 16, enc2 32, bottom 64, dec2 32, dec1 16.  Parameters are
 * FP32 masters updated by Adam (beta 0.9/0.999, eps 1e-8). */
typedef struct LhUnet LhUnet;
int lhUnetCreate(LhUnet **out, uint64_t seed);
int lhUnetDestroy(LhUnet *net);
uint64_t lhUnetParamCount(void);
/* Flat FP32 parameters: per conv layer W[cout][27][cin] then b[cout] (layers
 * enc1a enc1b enc2a enc2b bottom_a bottom_b dec2a dec2b dec1a dec1b; enc1a
 * has 16 input channels of which only channel 0 is fed), then head w[16], b. */
int lhUnetGetParams(LhUnet *net, float *host);
int lhUnetSetParams(LhUnet *net, const float *host);
int lhUnetGetGrads(LhUnet *net, float *host);
/* dOut = net(dIn) for n cubes of D^3 floats (D a multiple of 64). */
int lhUnetForward(LhUnet *net, const float *dIn, float *dOut, int n, int D);
/* firstBlock != 0 (the default): lhUnetForward evaluates the first two
 * convolutions in one kernel (the first layer's activations recomputed on
 * the second's halo, never stored); 0: one kernel per layer, as training
 * runs them.  The outputs are the same bit for bit. */
int lhUnetSetFusion(LhUnet *net, int firstBlock);
/* MSE loss of net(dIn) against dTarget, its gradients, and with lr > 0 one
 * Adam step. */
int lhUnetTrainStep(LhUnet *net, const float *dIn, const float *dTarget, int n, int D, float lr,
    float *loss);
/* Single-convolution entries (tests): BF16 channels-last tensors, W FP32
 * [cout][27][cin] (host).  backward = 1: backward-data pass (dx holds cout
 * channels, dy receives cin). */
int lhConv3dTest(const uint16_t *dx, int cin, int cout, const float *W, const float *b,
    uint16_t *dy, int n, int D, int relu, int backward);
int lhConv3dWgradTest(const uint16_t *dx, int cin, const uint16_t *dg, int cout, float *dW,
    int n, int D);
/* Decoder form (tests): channels [0, c0) from dx0 ([n*Ds^3][c0], Ds = D/2 when
 * half0 = 1, nearest upsampling read in place), channels [c0, cin) from dx1
 * ([n*D^3][cin-c0]); dy receives [n*D^3][cout]. */
int lhConv3dSplitTest(const uint16_t *dx0, int c0, int half0, const uint16_t *dx1, int cin,
    int cout, const float *W, const float *b, uint16_t *dy, int n, int D, int relu);

/* Device memory for callers without another allocator. */
int lhDeviceAlloc(void **p, uint64_t bytes);
int lhDeviceFree(void *p);
int lhCopyToDevice(void *dDst, const void *src, uint64_t bytes);
int lhCopyToHost(void *dst, const void *dSrc, uint64_t bytes);

/* ---------------------------------------------------------------------
 * Before and after the path.
 * ------------------------------------------------------------------- */

/* The context's configuration (as created, with later lhSetLaunch changes). */
int lhGetConfig(const LhContext *ctx, LhConfig *out);

/* Reads views [firstView, firstView+count) of a .rct file (the record layout
 * of ctFileReader.c:51-58) straight into slots firstSlot.. (ids firstView+1..)
 * through the pinned upload ring, with all host threads issuing pread, so
 * reads overlap the copies and staging of earlier views; matrices (optional,
 * count x 12) receives their matrices. */
int lhLoadRctFile(LhContext *ctx, const char *path, uint32_t firstView, uint32_t count,
    uint32_t firstSlot, double *matrices);

/* FDK pre-processing of measured line integrals (count x H x W floats,
 * device) into the filtered projections the RabbitCT input carries: cosine
 * weight from the ray direction (minv, count x 9) against the principal axis
 * (axis, count x 3), column weights parker (count x W, e.g. short-scan
 * redundancy), row ramp filter of sample spacing tau (mm), times dbeta (rad),
 * optional zero 1-pixel border.  dOut: count x H x W floats, device. */
int lhFdkFilter(int W, int H, uint32_t count, const double *minv, const double *axis,
    const float *parker, double tau, double dbeta, int border, const float *dRaw, float *dOut);

/* The runner's -c verification (rabbitRunner.c:346-384) on the GPU:
 * HU = floor(hus0*v + hus1 + 0.5) for v > 0 (else 0), clamped to 4095, against
 * a uint16 reference (device); mse, max |error| and the 11-bin histogram of
 * floor|error| (last bin >= 10).  Exact (integer sums). */
int lhHuMetric(const float *dVol, uint64_t n, const uint16_t *dRef, float hus0, float hus1,
    double *mse, double *maxAbs, uint32_t *hist);
/* The same HU conversion into a uint16 volume (device). */
int lhHuConvert(const float *dVol, uint64_t n, float hus0, float hus1, uint16_t *dOut);

/* The runner's -p slice image (rabbitRunner.c:292-315): a 16-bit binary PGM,
 * header "P5\n<L> <L>\n4095\n", then the HU values (lhHuConvert) of slice
 * z = L/2 - 1 of the L^3 device volume as big-endian uint16. */
int lhWritePgmSlice(const char *path, uint32_t L, const float *dVol, float hus0, float hus1);

/* The runner's -c result file (rabbitRunner.c:318-407, result.rctd): int32
 * 591984, int32 L, uint32 numberOfImages, numberOfImages x uint64
 * meanTimeUsec (the runner stores its mean time per image for every image),
 * the HU slice z = L/2 - 1 as L x L floats, double mse, double max |error|,
 * uint32 hist[11] (lhHuMetric of the L^3 device volume against the device
 * uint16 reference), int32 489195; little-endian, as the runner writes on
 * x86.  mse / maxAbs / hist (optional) receive the metric. */
int lhWriteResultFile(const char *path, uint32_t L, uint32_t numberOfImages,
    uint64_t meanTimeUsec, const float *dVol, const uint16_t *dRef, float hus0, float hus1,
    double *mse, double *maxAbs, uint32_t *hist);

/* Last error message of the calling thread. */
const char *lhLastError(void);

/* Byte offsets of the RabbitCtGlobalData fields, for ABI checks:
 * out[0..9] = offsetof each field in declaration order, out[10] = sizeof. */
void lhAbiLayout(uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* LOLAHIP_H */
