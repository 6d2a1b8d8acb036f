// lh_device.hip -- LolaHIP device layer: CDNA4 (gfx950) backprojection kernels
// and the extern "C" context API declared in include/lolahip.h.
//
// The path is a gather-interpolate-accumulate, not a contraction: no MFMA.
// Work decomposition of the backprojection kernel (bp_kernel):
//   * a workgroup is 4 waves; lane l of wave w owns the voxels
//     (x0 .. x0+R-1, y0+l, z0+w): one x-run of R voxels per lane, 64 adjacent
//     lines per wave, so one instruction's 64 lanes step along y and their
//     detector taps stay within a few cache lines;
//   * the R accumulators live in VGPRs across all views of a launch (a "pass"),
//     so the volume is read and written once per pass, not once per view;
//   * per view and lane, the homogeneous (u,v,w) at the run start is computed
//     in FP64 (LolaOPT does the same per line, LolaOPT.c:140-142) and rounded
//     to FP32; voxel r of the run then uses fma(r, a*R_L, origin) in FP32, one
//     rounding from the origin, so nothing drifts along x;
//   * the view's matrix and increments are wave-uniform (scalar loads); the
//     lane's line range (int16 start, stop) is one coalesced dword per lane;
//   * taps come from a zero-bordered copy of the projection through
//     buffer loads (the bounds of the buffer resource turn any stray offset
//     into a zero instead of a fault); coordinates are clamped to [-2, W] x
//     [-2, H] so every tap lands in the 2-pixel zero border or the image.
//
// Interpolation semantics follow LolaOPT.c:144-173: truncation toward zero
// (negative fractions for coordinates in (-1,0)), horizontal then vertical
// lerp, weight 1/w^2, zero outside the detector, update only x in
// [start, stop) of the line's range (LolaOPT.c:125-128).

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <cstdarg>
#include <cstddef>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lolahip.h"
#include "lh_internal.h"

namespace {
thread_local std::string g_err;
}

int lh_fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {
#define fail lh_fail

#define LH_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(e_ == hipErrorOutOfMemory ? LH_E_NOMEM : LH_E_HIP, "%s:%d %s: %s",      \
          __FILE__, __LINE__, #call, hipGetErrorString(e_));                             \
  } while (0)

constexpr uint32_t kBorder   = 2;          // zero columns/rows on each side
constexpr uint32_t kRsrcWord3 = 0x00020000; // raw buffer descriptor, gfx9 family
constexpr int kWaves         = 4;          // waves per workgroup
constexpr int kWave          = 64;

// Per-slot view parameters, wave-uniform in the kernel (scalar loads).
struct alignas(16) ViewParams {
  double a[12];    // FP64 matrix, column-major a[c*3+r], times k = 2^-m
  float du, dv, dw; // a[0..2] * R_L in FP32: per-voxel step along x
  int id;          // 1-based view id (selects the line ranges)
  float scale;     // storage scale s = 4^m of the projection (1 for FP32)
  int pad[3];
};
static_assert(sizeof(ViewParams) == 128, "ViewParams layout");

struct BpArgs {
  float *vol;
  const int *ranges;       // LhLineRange as packed int32: low half start, high half stop
  const ViewParams *views;
  const uint8_t *slots;
  uint64_t slotBytes;      // stride between slots
  uint32_t rowBytes;       // stride between padded rows
  uint32_t slotRecords;    // bytes of one slot addressed by the buffer resource
  uint32_t baseOffset;     // byte offset of pixel (0,0) relative to element (-2,-2)
  int L;
  float R_L, O_L;
  float Wf, Hf;            // clamp bounds: W and H as float
  int firstSlot, nViews;
  int clip;
  int tilesX, tilesY, tilesZ; // tile grid (x-runs, 64-line groups, 4-slice groups)
  int tileZ0;              // first z tile of the launch (a z-slab: tiles [tileZ0, +tilesZ))
  int persistent;
  const int2 *boxes;       // staged kernel: footprint box per (tile, slot), [tile][boxSlots]
  int boxSlots;            // slots per tile row of boxes
  uint32_t boxBytes;       // bytes of the boxes array (< 4 GiB)
  // persistent form: XCD x walks entries [xcdTile[x], xcdTile[x + 1]) of
  // the tile list `tiles` (tile indices of the launch, z-major, relative to
  // tileZ0), or of the tiles themselves where `tiles` is null
  const int *tiles;
  int xcdTile[9];
  int *taken;              // persistent form: entries taken per XCD range (zero at launch)
};

// Per-view parameters are read through the constant address space: uniform
// addresses then become scalar loads.
typedef const __attribute__((address_space(4))) ViewParams CViewParams;
__device__ __forceinline__ CViewParams &view_params(const BpArgs &A, int slot)
{
  return ((CViewParams *)(uintptr_t)A.views)[slot];
}

// The kernel's argument block read again through the kernarg segment pointer,
// which an empty asm hides from the compiler: fields used through it in a
// loop are then reloaded with scalar loads where they are needed instead of
// being held live (and spilled to VGPR lanes) across the whole loop.
typedef const __attribute__((address_space(4))) BpArgs KernelArgs;
__device__ __forceinline__ KernelArgs &kernel_args()
{
  KernelArgs *k = (KernelArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *k;
}

template <class Args>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(Args &A, int slot)
{
  return __builtin_amdgcn_make_buffer_rsrc(
      (void *)(A.slots + (uint64_t)slot * A.slotBytes), (short)0, (int)A.slotRecords,
      (int)kRsrcWord3);
}

__device__ __forceinline__ float bf16_lo(uint32_t q) { return __uint_as_float(q << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t q) { return __uint_as_float(q & 0xffff0000u); }
__device__ __forceinline__ float f16_lo(uint32_t q)
{
  return __half2float(__ushort_as_half((unsigned short)(q & 0xffffu)));
}
__device__ __forceinline__ float f16_hi(uint32_t q)
{
  return __half2float(__ushort_as_half((unsigned short)(q >> 16)));
}

// Horizontal lerp of one "pair" dword (p0, d) = (p(i,j), p(i+1,j) - p(i,j))
// at fraction ax: p0 + ax d, one FMA with both halves widened (FP16: inside
// the FMA; BF16: shift and mask).  The same operations on both tap paths (LDS
// and L2).
template <int PREC>
__device__ __forceinline__ float lerp_pair(uint32_t q, float ax)
{
  if constexpr (PREC == LH_PREC_F16)
    return fmaf(ax, f16_hi(q), f16_lo(q));
  else
    return fmaf(ax, bf16_hi(q), bf16_lo(q));
}

// Coordinates.  Each tile-view has a footprint box (box_kernel) whose origin
// (iu0, iv0) is an integer pixel; voxel coordinates are carried relative to
// it, u' = u - iu0 w and v' = v - iv0 w, from an FP64 origin per run (the
// subtraction is done in FP64, so the integer offset costs nothing in
// accuracy).  The local detector coordinates ix' = u'/w, iy' = v'/w are at
// most the box size (< 128 px where the footprint is staged) instead of up to
// the detector size, so their FP32 rounding is 8-16x smaller than that of
// global coordinates near pixel 1000.  Three wave-uniform modes per tile-view:
//  kInterior  the box lies inside the detector and every coordinate is >= 1
//             px: no clamp; the local coordinates are >= 1, so truncation is
//             the floor and x - trunc(x) is fract(x) exactly;
//  kClamp     coordinates clamped to the box [0, fw-2] x [0, fh-2]; the box
//             origin is >= 0, so the global coordinate is >= 0 and its
//             truncation is the local floor;
//  kEdge      clamped, and the box origin is left of or above pixel 0: the
//             global coordinate ix' + iu0 (exact: |iu0| <= 2) is truncated
//             toward zero, as LolaOPT.c:144-173 truncates (negative fractions
//             extrapolate in (-1, 0)).
enum { kInterior = 0, kClamp = 1, kEdge = 2 };

// Local cell of a coordinate pair: fractions (ax, ay) and the integer corner
// (cx, cy), relative to the box origin (kInterior, kClamp) or global (kEdge).
template <int MODE>
__device__ __forceinline__ void cell_of(float ix, float iy, float hiu, float hiv, float ou,
    float ov, float &ax, float &ay, float &fx, float &fy)
{
  if constexpr (MODE == kInterior) {
    ax = __builtin_amdgcn_fractf(ix);
    ay = __builtin_amdgcn_fractf(iy);
    fx = ix;   // converted by truncation by the caller
    fy = iy;
  } else {
    ix = __builtin_amdgcn_fmed3f(ix, 0.0f, hiu);
    iy = __builtin_amdgcn_fmed3f(iy, 0.0f, hiv);
    if constexpr (MODE == kEdge) {
      ix += ou;
      iy += ov;
    }
    fx = __builtin_truncf(ix);
    fy = __builtin_truncf(iy);
    ax = ix - fx;
    ay = iy - fy;
  }
}

// The four taps of the bilinear stencil from the slot image in L2, at byte
// offset `off` of the upper-left tap.
//  FP32: row-major bordered image, four 4-byte loads (rows j and j+1).
//  BF16/F16: "pair" layout, element (i,j) holds p(i,j) and p(i+1,j) - p(i,j)
//  as 2 halves in one dword: two 4-byte loads (rows j and j+1).
template <int PREC>
__device__ __forceinline__ float l2_lerp(__amdgpu_buffer_rsrc_t rs, uint32_t off, uint32_t row,
    float ax, float ay)
{
  float b, t;
  if constexpr (PREC == LH_PREC_F32) {
    const float p00 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
    const float p10 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u, 0, 0));
    const float p01 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, (int)row, 0));
    const float p11 =
        __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u, (int)row, 0));
    b = fmaf(ax, p10 - p00, p00);
    t = fmaf(ax, p11 - p01, p01);
  } else {
    const uint32_t q0 = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
    const uint32_t q1 = __builtin_amdgcn_raw_buffer_load_b32(rs, off, (int)row, 0);
    b = lerp_pair<PREC>(q0, ax);
    t = lerp_pair<PREC>(q1, ax);
  }
  return fmaf(ay, t - b, b);
}

// One voxel-view update with taps from L2 (LolaOPT.c:145-170): returns
// val / w^2 for the voxel at local homogeneous (u', v', w).  `base` is the
// byte offset of the box origin pixel in the slot (kEdge: of pixel (0, 0)),
// `pitch` the row pitch in elements.  The coordinate arithmetic is that of
// lds_update, so both tap paths give the same bits.
template <int PREC, int MODE>
__device__ __forceinline__ float l2_update(__amdgpu_buffer_rsrc_t rs, uint32_t base, int pitch,
    uint32_t row, float u, float v, float w, float hiu, float hiv, float ou, float ov)
{
  const float rw = __builtin_amdgcn_rcpf(w);
  float ax, ay, fx, fy;
  cell_of<MODE>(u * rw, v * rw, hiu, hiv, ou, ov, ax, ay, fx, fy);
  const uint32_t off = (uint32_t)(__mul24((int)fy, pitch) + (int)fx) * 4u + base;
  return l2_lerp<PREC>(rs, off, row, ax, ay) * (rw * rw);
}

// ---- LDS-staged tile ------------------------------------------------------
// Tile of R (x) x LY (y) x LZ (z) voxels (the Shape below): lane l of wave w
// owns (x0..x0+R-1, y0 + l%LY, z0 + w*64/LY + l/LY).  Per (view, tile)
// box_kernel projects the tile's 8 corners and keeps the bounding box of
// their detector positions (plus a 1-pixel margin and the right/lower tap);
// the kernel copies that footprint of the bordered projection into LDS with
// direct-to-LDS buffer loads (16 bytes per lane, no VGPR round trip).  The
// footprints are double-buffered: view v+1 streams in while view v computes,
// one barrier per view.  Voxel coordinates are clamped to the box, or, for
// boxes the detector edge does not clip, lie in it by construction, so every
// LDS index is in bounds.  A footprint larger than the window falls back to
// L2 gathers for that view.
// updates per scheduling group in the staged loop, by precision (measured
// at L=512 with R=16: FP32 fastest with groups of 16, that is one fence at
// the end of the run, 47.4 against 48.5 ms per step with groups of 2 or 4;
// the half-precision kernels fastest with 4)
#ifndef LH_FENCE32
#define LH_FENCE32 16
#endif
template <int PREC>
constexpr int kFence = PREC == LH_PREC_F32 ? LH_FENCE32 : 4;
// updates per scheduling group of the L2-gather fallback: four voxels' tap
// loads in flight per lane (measured at L=256, where most footprints take it:
// 124 views on 16-voxel runs 3.65 ms with groups of one, 3.27 with two, 3.16
// with four, 3.43 with eight; 32-voxel runs 3.64 / 3.31 / 4.24 ms with two /
// four / eight)
template <int R>
constexpr int kFenceL2 = 4;
// voxels per parking segment of a run (the line ranges' alignment)
template <int R>
constexpr int kSeg = R < 16 ? R : 16;
#ifndef LH_OCC16
#define LH_OCC16 4
#endif

// Shape of a staged tile and of its LDS footprint window.  A workgroup of 4
// waves covers R (x, a template parameter of its own) x LY (y) x LZ (z) voxels,
// LZ = 256 / LY: lane l of wave w owns the x-run at y0 + l % LY,
// z0 + w * (64 / LY) + l / LY.  Per view the tile's detector footprint is
// staged into a window of FH rows of PITCH dwords, double-buffered; a
// footprint up to PITCH x FH pixels takes the staged path, a larger one L2
// gathers.  The copy moves 1 KB chunks of the window (64 lanes x 16 bytes,
// direct-to-LDS), wave w chunks w, w + 4, ...; lane l's chunk dword
// 256 c + 4 l falls at row (256 c + 4 l) / PITCH, column (256 c + 4 l) % PITCH
// (PITCH is a multiple of 4, so a lane's 16 bytes never straddle a row).
// A power-of-two pitch window may end in a partial chunk (copied by the
// lanes inside it); other windows are staged in whole chunks only, so a
// footprint fits when its rows lie within them (FH rows, at most ROWS).
// OCC: workgroups per CU the register and LDS budget is sized for.
template <int LY_, int PITCH_, int ROWS_, int OCC_>
struct Shape {
  static constexpr int LY = LY_, LZ = 256 / LY_, ZW = kWave / LY_; // z per wave
  static constexpr bool POW2  = (PITCH_ & (PITCH_ - 1)) == 0;
  static constexpr int PITCH = PITCH_, FW = PITCH_, OCC = OCC_;
  static constexpr int BUF    = PITCH_ * ROWS_;                   // dwords per buffer
  static constexpr int CHUNKS = (BUF + 255) / 256;                // 1 KB chunks per buffer
  static constexpr int KMAX   = (CHUNKS + kWaves - 1) / kWaves;   // chunks per wave
  static constexpr int FH     = POW2 ? ROWS_ : (BUF / 256) * 256 / PITCH_;
  static constexpr int LOG2   = PITCH_ == 64 ? 6 : PITCH_ == 128 ? 7 : PITCH_ == 256 ? 8 : 0;
  static_assert(PITCH_ % 4 == 0 && PITCH_ + 1 <= 255, "ds_read2 offsets are 8-bit dwords");
  static_assert(!POW2 || LOG2 > 0, "power-of-two pitch without its shift");
  static_assert(LY_ * (256 / LY_) == 256 && kWave % LY_ == 0, "lane layout");
  static_assert(OCC_ * (2 * BUF * 4 + 2 * kWaves * 4) <= 160 * 1024, "LDS of OCC workgroups");
  static_assert(FH >= 2, "window rows");
};
// 32 (y) x 8 (z) tiles, 128 x 39 window, 4 workgroups per CU: every footprint
// of a 16-voxel run fits at L=512 (and of a 32-voxel run at L=1024)
typedef Shape<32, 128, 39, LH_OCC16> ShapeY32;
// 32 x 8 tiles of 32-voxel runs where they span 16 mm (L <= 512): a 160 x 40
// window (at L=512 99.7% of the footprints fit 160 x 39, against 90.6% at
// 128), three workgroups per CU
typedef Shape<32, 160, 40, 3> ShapeY32W;
// 16 x 16 tiles: narrow, tall footprints (80 x 60: at L=512 every 16-voxel
// run's fits)
typedef Shape<16, 80, 62, 4> ShapeY16;
// 64 x 4 tiles: wide, short footprints (224 x 21); for 16- and 32-voxel runs
typedef Shape<64, 224, 22, 4> ShapeY64;

struct StageBox {
  int iu0, iv0;   // footprint origin (pixels)
  int fw, fh;     // footprint size
  bool fits;      // staged in LDS (else L2 gathers)
  bool interior;  // the detector edge does not clip the footprint and every
                  // coordinate of the tile is >= 1 pixel: no clamp is needed
};

// Footprint boxes are computed once per (slot, tile) by box_kernel and packed
// as int16 {iu0, iv0, fw | interior << 15, fh}; fw = 0 marks "no box" (the
// L2 fallback).  They are stored [tile][slot], so a tile's boxes for
// consecutive views share cache lines, and are read with vector loads issued
// two views ahead (they land under the same vmcnt wait as the footprints):
// a scalar load would be waited for by every LDS access (lgkmcnt).
template <class Args> __device__ __forceinline__ __amdgpu_buffer_rsrc_t box_rsrc(Args &A)
{
  return __builtin_amdgcn_make_buffer_rsrc((void *)A.boxes, (short)0, (int)A.boxBytes,
      (int)kRsrcWord3);
}

template <class Args>
__device__ __forceinline__ uint2 fetch_box(__amdgpu_buffer_rsrc_t rs, Args &A, int slot,
    int tile)
{
  const uint32_t off = ((uint32_t)tile * (uint32_t)A.boxSlots + (uint32_t)slot) * 8u;
  const auto q       = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
  return make_uint2(q[0], q[1]);
}

// STAGE = 0: every footprint is gathered from L2 (the staged tile shape and
// coordinate arithmetic without the LDS copies).
template <int STAGE, class SH>
__device__ __forceinline__ StageBox decode_box(uint2 raw)
{
  const int qx = __builtin_amdgcn_readfirstlane((int)raw.x);
  const int qy = __builtin_amdgcn_readfirstlane((int)raw.y);
  StageBox b;
  b.iu0      = (int)(short)(qx & 0xffff);
  b.iv0      = qx >> 16;
  b.fw       = qy & 0x7fff;
  b.interior = (qy & 0x8000) != 0;
  b.fh       = qy >> 16;
  b.fits     = STAGE && b.fw >= 2 && b.fh >= 2 && b.fw <= SH::FW && b.fh <= SH::FH;
  return b;
}

// One thread per (slot, tile): project the tile's 8 corner voxels with the
// slot's FP64 matrix, clamp to the bordered detector [-2, W] x [-2, H], and
// keep [floor(min) - 1, floor(max) + 3): the taps of a coordinate t lie in
// [trunc(t), trunc(t) + 1] within [floor(t), floor(t) + 2], plus one pixel of
// margin on each side for the FP32 rounding of the per-voxel coordinates.
// The projection of the tile (a box in front of the source) is the convex hull
// of its corners' projections, so every voxel's coordinate lies in the box.
// "interior": no corner was clamped, all lie at >= 1 pixel and below W (H),
// so the clamp is the identity for the whole tile and the coordinates are
// positive (truncation is the floor).
__global__ void box_kernel(const ViewParams *__restrict__ views, int firstSlot, int nTiles,
    int tilesX, int tilesY, int R, int ly, int lz, float O_L, float R_L, int W, int H,
    int nSlots, int2 *boxes)
{
  const int t    = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = firstSlot + blockIdx.y;
  if (t >= nTiles)
    return;
  const double *a = views[slot].a;
  const int tx = t % tilesX, ty = (t / tilesX) % tilesY, tz = t / (tilesX * tilesY);
  const double O = O_L, RL = R_L;
  double umin = 1e30, umax = -1e30, vmin = 1e30, vmax = -1e30;
  bool interior = true;
  for (int c = 0; c < 8; ++c) {
    const double x = O + (double)(tx * R + ((c & 1) ? R - 1 : 0)) * RL;
    const double y = O + (double)(ty * ly + ((c & 2) ? ly - 1 : 0)) * RL;
    const double z = O + (double)(tz * lz + ((c & 4) ? lz - 1 : 0)) * RL;
    const double hw = fma(a[2], x, fma(a[5], y, fma(a[8], z, a[11])));
    const double u  = fma(a[0], x, fma(a[3], y, fma(a[6], z, a[9]))) / hw;
    const double v  = fma(a[1], x, fma(a[4], y, fma(a[7], z, a[10]))) / hw;
    interior = interior && hw > 0.0 && u >= 1.0 && v >= 1.0 && u < (double)W && v < (double)H;
    const double uc = fmin(fmax(u, -2.0), (double)W), vc = fmin(fmax(v, -2.0), (double)H);
    umin = fmin(umin, uc);
    umax = fmax(umax, uc);
    vmin = fmin(vmin, vc);
    vmax = fmax(vmax, vc);
  }
  const int iu0 = max((int)floor(umin) - 1, -2);
  const int iv0 = max((int)floor(vmin) - 1, -2);
  const int iu1 = min((int)floor(umax) + 3, W + 2);
  const int iv1 = min((int)floor(vmax) + 3, H + 2);
  int2 q;
  q.x = (int)(uint16_t)(int16_t)iu0 | (int)((uint32_t)(uint16_t)(int16_t)iv0 << 16);
  q.y = (int)((uint32_t)(iu1 - iu0) & 0x7fffu) | (interior ? 0x8000 : 0) |
        (int)((uint32_t)(uint16_t)(iv1 - iv0) << 16);
  boxes[(uint64_t)t * (uint64_t)nSlots + slot] = q;
}

// Issues the direct-to-LDS loads of footprint `b` of the slot image into buf
// (one dword per element: an FP32 pixel, or for half storage p(i,j) and the
// forward difference p(i+1,j) - p(i,j)).  The window is copied in 1 KB
// chunks, each one buffer_load_dwordx4 ... lds (lane l lands 16 bytes at
// 16 l); wave w copies chunks w, w + 4, ... of those that hold rows < fh.
// rel(k) is the lane's source byte offset, relative to the box origin, in its
// k-th chunk: for a 128-dword pitch the chunks are row pairs and rel(k) is
// linear in k; otherwise a per-lane table made once per tile.  Columns at or
// beyond fw and rows at or beyond fh within a copied chunk are copied too and
// never read; no lane writes past the end of the window; past the end of the
// slot the buffer bounds return zero.  Completion is awaited by the caller
// (s_waitcnt vmcnt(0) and a barrier).
// Lanes whose 16 bytes lie at or beyond the footprint width fw in their row
// are masked off, so the copies fetch about fw, not PITCH, pixels per row.
// a source offset beyond every slot (slots are < 2 GiB): the buffer bounds
// check returns zeros without a memory access
constexpr uint32_t kNoSource = 0x80000000u;

template <class SH>
struct ChunkOffsets {
  static_assert(!SH::POW2 || 1024 % SH::PITCH == 0, "whole rows per 4 chunks");
  static_assert(SH::PITCH <= 255, "a column fits 8 bits");
  uint32_t rel[SH::POW2 ? 1 : SH::KMAX];
  uint32_t cols[SH::POW2 ? 1 : (SH::KMAX + 3) / 4]; // column of chunk k: byte k % 4 of word k / 4
  uint32_t lane0; // 128-dword pitch: rel(0)
  uint32_t step;  // 128-dword pitch: rel(k+1) - rel(k)
  __device__ __forceinline__ void init(int wv, int lane, uint32_t rowBytes)
  {
    if constexpr (SH::POW2) {
      const uint32_t d = (uint32_t)(256 * wv + 4 * lane);
      lane0   = (d >> SH::LOG2) * rowBytes + (d & (SH::PITCH - 1)) * 4u;
      step    = (uint32_t)(1024 / SH::PITCH) * rowBytes;
      cols[0] = d & (SH::PITCH - 1); // the same in every chunk
    } else {
#pragma unroll
      for (int k = 0; k < (SH::KMAX + 3) / 4; ++k)
        cols[k] = 0;
#pragma unroll
      for (int k = 0; k < SH::KMAX; ++k) {
        const uint32_t d = (uint32_t)(256 * (wv + kWaves * k) + 4 * lane);
        rel[k] = (d / SH::PITCH) * rowBytes + (d % SH::PITCH) * 4u;
        cols[k / 4] |= (d % SH::PITCH) << (8 * (k % 4));
      }
    }
  }
  __device__ __forceinline__ uint32_t at(int k) const
  {
    if constexpr (SH::POW2)
      return lane0 + (uint32_t)k * step;
    else
      return rel[k];
  }
  __device__ __forceinline__ int col(int k) const
  {
    if constexpr (SH::POW2)
      return (int)cols[0];
    else
      return (int)__builtin_amdgcn_ubfe(cols[k / 4], 8 * (k % 4), 8);
  }
};

template <class SH>
__device__ __forceinline__ void stage_issue(__amdgpu_buffer_rsrc_t rs, const StageBox &b,
    int pitch, float *buf, int wv, int lane, const ChunkOffsets<SH> &co)
{
  const uint32_t base = (uint32_t)((b.iv0 + (int)kBorder) * pitch + b.iu0 + (int)kBorder) * 4u;
  const int need      = (b.fh * SH::PITCH + 255) >> 8; // chunks holding rows < fh
  constexpr int kTail = SH::BUF & 255;                   // dwords of a partial last chunk
  constexpr int kWhole = kTail ? SH::CHUNKS - 1 : SH::CHUNKS;
  if constexpr (SH::POW2) {
    // 128-dword pitch: a chunk is a row pair, wave w copies pairs w, w + 4,
    // ... in a rolled loop, and the lower half-wave alone a last odd row (the
    // window's last row too when FH is odd).  Measured 2.5% faster per step
    // than a loop over chunk counts at L=512.
    // The lane masks are source offsets past the slot (the buffer bounds
    // return zeros, fetched from nowhere) rather than branches: the view
    // loop then has only wave-uniform branches.  A last odd row is copied
    // by all 64 lanes in two 4-byte loads (a 16-byte load's upper half-wave
    // would land in the row past the window).
    static_assert(SH::PITCH == 128, "row pairs");
    uint32_t off = co.col(0) < b.fw ? base + co.lane0 : kNoSource;
    int r0       = 2 * wv;
    for (; r0 + 1 < b.fh; r0 += 2 * kWaves, off += co.step)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(buf + r0 * SH::PITCH), 16, (int)off, 0, 0,
          0);
    if (r0 < b.fh) {
      const uint32_t row = base + (uint32_t)(r0 * pitch) * 4u + 4u * (uint32_t)lane;
      const uint32_t lo  = lane < b.fw ? row : kNoSource;
      const uint32_t hi  = 64 + lane < b.fw ? row + 256u : kNoSource;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(buf + r0 * SH::PITCH), 4, (int)lo, 0, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(buf + r0 * SH::PITCH + 64), 4, (int)hi, 0,
          0, 0);
    }
  } else {
    // whole chunks only (FH keeps need <= CHUNKS); the opaque copy of the
    // wave index keeps the per-chunk bounds and LDS offsets from being
    // hoisted out of the view loop into registers of their own
    static_assert(kTail == 0 || SH::FH * SH::PITCH <= 256 * kWhole, "tail chunk");
    int w = wv;
    __asm__("" : "+s"(w));
#pragma unroll
    for (int k = 0; k < SH::KMAX; ++k) {
      const int ch = w + kWaves * k;
      if (ch >= need)
        break;
      const uint32_t src = co.col(k) < b.fw ? base + co.at(k) : kNoSource;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(buf + 256 * ch), 16, (int)src, 0, 0, 0);
    }
  }
}

// One voxel-view update from the LDS footprint (LolaOPT.c:145-170) at local
// homogeneous (u', v', w), in two halves: lds_fetch computes the cell and
// issues the tap reads, lds_finish interpolates and adds val / w^2 to the
// accumulator, so a run can have the taps of later voxels in flight while it
// interpolates earlier ones.  sBase[v * PITCH + u] is footprint pixel (u, v)
// (kEdge: sBase is offset by the box origin, so it is indexed by global
// pixels).  The coordinate arithmetic is cell_of's, shared with l2_update.
// LDS address of the upper-left tap of cell (fx, fy) (integer-valued floats).
template <int MODE, class SH>
__device__ __forceinline__ const float *lds_tap_address(const float *sBase, float fx, float fy)
{
  if constexpr (MODE == kInterior) {
    // byte address base + 4 * column + row * 4 * PITCH as a shift-add and a
    // shift-add (power-of-two pitch) or a 24-bit multiply-add: the empty asm
    // hides the partial sum, which the compiler would otherwise re-associate
    // into two shifts and a three-way add
    typedef const __attribute__((address_space(3))) float LdsFloat;
    uint32_t col = (uint32_t)(uintptr_t)(LdsFloat *)sBase + ((uint32_t)(int)fx << 2);
    __asm__("" : "+v"(col));
    if constexpr (SH::POW2)
      return (const float *)(LdsFloat *)(uintptr_t)(col +
          ((uint32_t)(int)fy << (SH::LOG2 + 2)));
    else
      return (const float *)(LdsFloat *)(uintptr_t)(
          __umul24((uint32_t)(int)fy, 4u * SH::PITCH) + col);
  } else {
    // fy*PITCH + fx is an exact integer in FP32 (|.| < 2^24): one conversion
    return sBase + (int)fmaf(fy, (float)SH::PITCH, fx);
  }
}

struct LdsTaps {
  float p0, p1, p2, p3; // FP32: p00, p10, p01, p11; half: pair dwords of rows j, j+1 in p0, p1
  float ax, ay, w2;     // fractions and the weight (1/w)^2
};

template <int PREC, int MODE, class SH>
__device__ __forceinline__ LdsTaps lds_fetch(const float *sBase, float u, float v, float w,
    float hiu, float hiv, float ou, float ov)
{
  const float rw = __builtin_amdgcn_rcpf(w);
  LdsTaps t;
  float fx, fy;
  cell_of<MODE>(u * rw, v * rw, hiu, hiv, ou, ov, t.ax, t.ay, fx, fy);
  t.w2 = rw * rw;
  const float *q = lds_tap_address<MODE, SH>(sBase, fx, fy);
  if constexpr (PREC == LH_PREC_F32) {
    t.p0 = q[0];
    t.p1 = q[1];
    t.p2 = q[SH::PITCH];
    t.p3 = q[SH::PITCH + 1];
  } else {
    // one dword per pixel: p(i,j) and the forward difference to p(i+1,j)
    t.p0 = q[0];
    t.p1 = q[SH::PITCH];
  }
  return t;
}

template <int PREC>
__device__ __forceinline__ void lds_finish(float &acc, const LdsTaps &t)
{
  float b, c;
  if constexpr (PREC == LH_PREC_F32) {
    b = fmaf(t.ax, t.p1 - t.p0, t.p0);
    c = fmaf(t.ax, t.p3 - t.p2, t.p2);
  } else {
    b = lerp_pair<PREC>(__float_as_uint(t.p0), t.ax);
    c = lerp_pair<PREC>(__float_as_uint(t.p1), t.ax);
  }
  acc += fmaf(t.ay, c - b, b) * t.w2;
}

// The whole update in one expression (returns val / w^2): the form the
// unpipelined runs use.  Written as lds_finish(lds_fetch(...)), with the
// weight formed before the tap reads, the 16-voxel FP32 kernel took 127
// instead of 101 VGPRs and ran 15% slower.
template <int PREC, int MODE, class SH>
__device__ __forceinline__ float lds_update(const float *sBase, float u, float v, float w,
    float hiu, float hiv, float ou, float ov)
{
  const float rw = __builtin_amdgcn_rcpf(w);
  float ax, ay, fx, fy;
  cell_of<MODE>(u * rw, v * rw, hiu, hiv, ou, ov, ax, ay, fx, fy);
  const float *q = lds_tap_address<MODE, SH>(sBase, fx, fy);
  float b, t;
  if constexpr (PREC == LH_PREC_F32) {
    const float p00 = q[0], p10 = q[1];
    const float p01 = q[SH::PITCH], p11 = q[SH::PITCH + 1];
    b = fmaf(ax, p10 - p00, p00);
    t = fmaf(ax, p11 - p01, p01);
  } else {
    // one dword per pixel: p(i,j) and the forward difference to p(i+1,j)
    const uint32_t *qq = reinterpret_cast<const uint32_t *>(q);
    b = lerp_pair<PREC>(qq[0], ax);
    t = lerp_pair<PREC>(qq[SH::PITCH], ax);
  }
  return fmaf(ay, t - b, b) * (rw * rw);
}

// f(std::integral_constant<int, I>) for I in [I0, I1): loop indices that
// are constants from the start, so arrays indexed by them stay in registers
template <int I0, int I1, class F>
__device__ __forceinline__ void static_for(F &&f)
{
  if constexpr (I0 < I1) {
    f(std::integral_constant<int, I0>{});
    static_for<I0 + 1, I1>(f);
  }
}

// Voxels whose taps are in flight while a voxel is interpolated, by run
// length: 32-voxel runs (3 workgroups per CU, registers to spare) read the
// taps of voxel r + 1 while interpolating voxel r (measured at L=512: FP32
// 45.6 -> 44.8 ms per step, FP16 44.5 -> 43.8, BF16 47.3 -> 46.4; 2 to 4
// voxels ahead ran 0.5-1% slower than one).  Shorter runs (4 workgroups
// per CU) keep one scheduling group per kFence voxels: pipelined, the
// compiler hoisted their coordinate arithmetic and spilled (16-voxel FP32
// 46.3 -> 50.8 ms per step).
template <int R>
constexpr int kAhead = R >= 32 ? 1 : 0;

// Voxels [r0, r0 + N) of a run at coordinates fma(r, d, o) (r = 0: o), from
// the LDS footprint.  kAhead > 0: the taps of voxel r + kAhead are read while
// voxel r is interpolated, one scheduling group per voxel, so the LDS latency
// is covered by the arithmetic of kAhead voxels; otherwise one scheduling
// group per kFence voxels.  The same operations as lds_update per voxel, in
// another order: the same bits.
template <int PREC, int MODE, class SH, int N, int r0, int R>
__device__ __forceinline__ void lds_run(float (&acc)[R], const float *sBase, float uo, float vo,
    float wo, float du, float dv, float dw, float hiu, float hiv, float ou = 0.0f,
    float ov = 0.0f)
{
  if constexpr (kAhead<R> == 0) {
#pragma unroll
    for (int r = r0; r < r0 + N; ++r) {
      const float fr = (float)r;
      acc[r] += lds_update<PREC, MODE, SH>(sBase, r ? fmaf(fr, du, uo) : uo,
          r ? fmaf(fr, dv, vo) : vo, r ? fmaf(fr, dw, wo) : wo, hiu, hiv, ou, ov);
      if ((r % kFence<PREC>) == kFence<PREC> - 1)
        __builtin_amdgcn_sched_barrier(0);
    }
  } else {
  constexpr int D = kAhead<R> < N ? kAhead<R> : N;
  auto fetch = [&](int r) {
    const float fr = (float)r;
    return lds_fetch<PREC, MODE, SH>(sBase, r ? fmaf(fr, du, uo) : uo,
        r ? fmaf(fr, dv, vo) : vo, r ? fmaf(fr, dw, wo) : wo, hiu, hiv, ou, ov);
  };
  LdsTaps t[D];
#pragma unroll
  for (int k = 0; k < D; ++k)
    t[k] = fetch(r0 + k);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    LdsTaps next;
    if (k + D < N)
      next = fetch(r0 + k + D);
    lds_finish<PREC>(acc[r0 + k], t[k % D]);
    if (k + D < N)
      t[k % D] = next;
    __builtin_amdgcn_sched_barrier(0);
  }
  }
}

// Voxels [r0, r0 + N) of a run at coordinates fma(r, d, o), taps from L2,
// in scheduling groups of kFenceL2 voxels: l2_update per voxel, the same bits.
template <int PREC, int MODE, int N, int r0, int R>
__device__ __forceinline__ void l2_run(float (&acc)[R], __amdgpu_buffer_rsrc_t rs, uint32_t base,
    int pitch, uint32_t row, float uo, float vo, float wo, float du, float dv, float dw, float hiu,
    float hiv)
{
#pragma unroll
  for (int r = r0; r < r0 + N; ++r) {
    const float fr = (float)r;
    acc[r] += l2_update<PREC, MODE>(rs, base, pitch, row, r ? fmaf(fr, du, uo) : uo,
        r ? fmaf(fr, dv, vo) : vo, r ? fmaf(fr, dw, wo) : wo, hiu, hiv, 0.0f, 0.0f);
    if ((r % kFenceL2<R>) == kFenceL2<R> - 1)
      __builtin_amdgcn_sched_barrier(0);
  }
}

template <int R, int PREC, int STAGE, class SH>
__device__ __forceinline__ void bp_tile_lds(const BpArgs &A, int tx, int ty, int tz, float *sImg,
    int *sAct)
{
  constexpr int ELEM    = 4;         // FP32 pixel or half pair: one dword per element
  constexpr float kPark   = 0x1p100f; // w of a parked lane: (1/w)^2 underflows to +0
  constexpr float kParkUV = 0x1p99f;  // its u' and v': coordinate (1/2, 1/2) in the box
  const int tile = tx + A.tilesX * (ty + A.tilesY * tz);
  const int lane = threadIdx.x & (kWave - 1);
  const int wv   = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int x0   = tx * R;
  const int y0   = ty * SH::LY;
  const int z0   = tz * SH::LZ;
  const int y    = y0 + (lane & (SH::LY - 1));
  const int z    = z0 + SH::ZW * wv + lane / SH::LY;
  const int L    = A.L;
  const double O = A.O_L, RL = A.R_L;

  const double wx0 = O + (double)x0 * RL;
  const double wy  = O + (double)y * RL;
  const double wz  = O + (double)z * RL;
  const uint64_t line = (uint64_t)z * (uint64_t)L + (uint64_t)y;
  const uint64_t LL   = (uint64_t)L * (uint64_t)L;
  const int pitch     = (int)(A.rowBytes / ELEM);
  const int n         = A.nViews;
  // the lane's source offsets in the footprint copies
  ChunkOffsets<SH> co;
  co.init(wv, lane, A.rowBytes);

  auto range_of = [&](int v) -> int {
    return A.clip ? A.ranges[(uint64_t)(view_params(A, A.firstSlot + v).id - 1) * LL + line]
                  : (L << 16);
  };
  // the lane's run [lo, hi) within its line's range [start, stop)
  auto decode = [&](int rr, int &lo, int &hi) {
    lo = max((int)(short)(rr & 0xffff) - x0, 0);
    hi = min((rr >> 16) - x0, R);
  };

  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    acc[r] = 0.0f;
  bool touched = false;

  // prologue: view 0 activity and footprint.  Each range is decoded once,
  // for the activity vote one view ahead, and kept for its own view.
  int lo_c, hi_c;
  decode(range_of(0), lo_c, hi_c);
  int rr_nxt = n > 1 ? range_of(1) : 0;
  // the vote runs on the whole wave, only the store is lane 0's
  const bool act0 = __any(lo_c < hi_c);
  if (lane == 0)
    sAct[wv] = act0 ? 1 : 0;
  __syncthreads();
  bool blockAct = __builtin_amdgcn_readfirstlane(sAct[0] | sAct[1] | sAct[2] | sAct[3]) != 0;
  const __amdgpu_buffer_rsrc_t boxRs = box_rsrc(A);
  StageBox box = decode_box<STAGE, SH>(fetch_box(boxRs, A, A.firstSlot, tile));
  // box of view v+1, in flight during view v-1
  uint2 boxRaw = n > 1 ? fetch_box(boxRs, A, A.firstSlot + 1, tile) : make_uint2(0u, 0u);
  if (blockAct && box.fits)
    stage_issue<SH>(slot_rsrc(A, A.firstSlot), box, pitch, sImg, wv, lane, co);

  for (int v = 0; v < n; ++v) {
    const int slot      = A.firstSlot + v;
    CViewParams &P      = view_params(A, slot);
    const bool hasNext  = v + 1 < n;
    // footprint(v) and the range loads have landed; publish activity of v+1
    __builtin_amdgcn_s_waitcnt(0);
    int lo_n = 0, hi_n = 0;
    decode(rr_nxt, lo_n, hi_n);
    const bool actNext = hasNext && __any(lo_n < hi_n);
    sAct[((v + 1) & 1) * kWaves + wv] = actNext ? 1 : 0; // the same value from every lane
    __syncthreads();
    const int *an = sAct + ((v + 1) & 1) * kWaves;
    const bool blockActNext = __builtin_amdgcn_readfirstlane(an[0] | an[1] | an[2] | an[3]) != 0;
    if (v + 2 < n)
      rr_nxt = range_of(v + 2);
    StageBox boxNext = box;
    if (hasNext) {
      boxNext = decode_box<STAGE, SH>(boxRaw);
      KernelArgs &K = kernel_args();
      if (v + 2 < n)
        boxRaw = fetch_box(box_rsrc(K), K, slot + 2, tile);
      if (blockActNext && boxNext.fits)
        stage_issue<SH>(slot_rsrc(K, slot + 1), boxNext, pitch,
            sImg + ((v + 1) & 1) * SH::BUF, wv, lane, co);
    }

    const int rlo     = lo_c;
    const int rhi     = hi_c;
    lo_c              = lo_n;
    hi_c              = hi_n;
    const bool active = rlo < rhi;
    if (blockAct && __builtin_amdgcn_readfirstlane(__any(active)) != 0) {
      touched |= active;
      // FP64 per-line origin A (x0, y, z, 1) relative to the box origin:
      // u' = u - iu0 w, v' = v - iv0 w in FP64, then rounded once to FP32; the
      // run's voxels step from it in FP32, one rounding each, so nothing
      // drifts along x at any L
      // (a6 wz first as a product: each instruction then reads one scalar
      // operand, which the vector ALU allows without copying the matrix into
      // vector registers)
      const double U0 = fma(P.a[0], wx0, fma(P.a[3], wy, P.a[6] * wz)) + P.a[9];
      const double V0 = fma(P.a[1], wx0, fma(P.a[4], wy, P.a[7] * wz)) + P.a[10];
      const double W0 = fma(P.a[2], wx0, fma(P.a[5], wy, P.a[8] * wz)) + P.a[11];
      float u0 = (float)fma(-(double)box.iu0, W0, U0);
      float v0 = (float)fma(-(double)box.iv0, W0, V0);
      float w0 = (float)W0;
      // dw in a vector register (the empty asm): a voxel's w step r*dw + w'
      // is then one v_fmamk_f32 with r as its literal, where an SGPR dw would
      // need r in a VGPR of its own for each r beyond the inline constants
      float dw = P.dw;
      __asm__("" : "+v"(dw));
      // the x-steps of u' and v': a*R_L - iu0 * a[2]*R_L
      const float du = fmaf(-(float)box.iu0, dw, P.du), dv = fmaf(-(float)box.iv0, dw, P.dv);
      // Voxels outside their line's range are "parked": (u', v', w) =
      // (1/2, 1/2, 1) * 2^100 with a zero step puts the coordinate half a
      // pixel inside the box and the weight (1/w)^2 underflows to +0, so the
      // update adds exactly zero and the accumulators take one straight path
      // through every branch below.  Ranges aligned to 16 voxels (the clip
      // files' VECTORSIZE 16) leave every 16-voxel segment of a run all in or
      // all out: such waves park whole segments (six selects per segment) on
      // the interior and clamp paths; other waves take the edge path, which
      // parks voxel by voxel.
      const bool aligned = __builtin_amdgcn_readfirstlane(
          __all(!active || ((rlo % kSeg<R>) == 0 && (rhi % kSeg<R>) == 0))) != 0;
      // Each update loop starts from its own opaque copy of the run origin
      // (an empty asm naming the loop): the loops then share no leading
      // instructions, which the compiler would otherwise hoist (with their
      // reciprocals, clamps and bound conversions) into the block every path
      // runs.  The clamp bounds are converted inside the loops that use them.
#define LH_LOOP_ORIGIN(tag)                                                              \
  float uo = u0, vo = v0, wo = w0;                                                       \
  __asm__("; " tag : "+v"(uo), "+v"(vo), "+v"(wo));                                      \
  const float hiu = (float)(box.fw - 2), hiv = (float)(box.fh - 2)
      // origin and steps of segment h (voxels [h S, h S + S), S = kSeg<R>):
      // the run's own, or parked with zero steps where the segment lies
      // outside the lane's range; a voxel's coordinate is fma(r, step,
      // origin) either way, so in-range voxels get the bits of the unsegmented
      // form
#define LH_SEGMENT(h)                                                                    \
  const bool on_ = rlo <= (h) * kSeg<R> && (h + 1) * kSeg<R> <= rhi;                      \
  const float us = on_ ? uo : kParkUV, vs = on_ ? vo : kParkUV, ws = on_ ? wo : kPark;   \
  const float dus = on_ ? du : 0.0f, dvs = on_ ? dv : 0.0f, dws = on_ ? dw : 0.0f
      // Boxes whose origin is left of or above pixel 0 (the detector edge)
      // truncate the global coordinate ix' + iu0; the others add 0, which
      // leaves the local coordinate as it is (kEdge arithmetic with a zero
      // offset is kClamp's).
      const bool edge = box.iu0 < 0 || box.iv0 < 0;
      if (box.fits) {
        // sImg buffer of this view
        const float *sBuf = sImg + (v & 1) * SH::BUF;
        // groups of kFence updates are fenced so the scheduler keeps at most
        // that many updates' taps in flight per lane
        if (aligned && box.interior) {
          LH_LOOP_ORIGIN("lds interior");
          static_for<0, R / kSeg<R>>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
            LH_SEGMENT(h);
            lds_run<PREC, kInterior, SH, kSeg<R>, h * kSeg<R>>(acc, sBuf, us, vs, ws, dus, dvs, dws,
                hiu, hiv);
          });
        } else if (aligned && !edge) {
          LH_LOOP_ORIGIN("lds clamp");
          static_for<0, R / kSeg<R>>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
            LH_SEGMENT(h);
            lds_run<PREC, kClamp, SH, kSeg<R>, h * kSeg<R>>(acc, sBuf, us, vs, ws, dus, dvs, dws,
                hiu, hiv);
          });
        } else if (aligned) {
          // detector-edge boxes (global truncation) with whole segments parked
          LH_LOOP_ORIGIN("lds edge");
          const float ou = (float)box.iu0, ov = (float)box.iv0;
          const float *sB = sBuf - (box.iv0 * SH::PITCH + box.iu0);
          static_for<0, R / kSeg<R>>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
            LH_SEGMENT(h);
            lds_run<PREC, kEdge, SH, kSeg<R>, h * kSeg<R>>(acc, sB, us, vs, ws, dus, dvs, dws, hiu,
                hiv, ou, ov);
          });
        } else {
          LH_LOOP_ORIGIN("lds edge or partial");
          // detector-edge boxes and partial runs, with
          // per-voxel parking (parked coordinates lie inside every box); for
          // an interior box the clamp is the identity and the bits are those
          // of the interior loop (for coordinates >= 0, trunc is the floor
          // and x - trunc(x) is fract(x))
          const float ou = edge ? (float)box.iu0 : 0.0f, ov = edge ? (float)box.iv0 : 0.0f;
          // indexed by global pixels for edge boxes
          const float *sB = edge ? sBuf - (box.iv0 * SH::PITCH + box.iu0) : sBuf;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const float fr = (float)r;
            const bool on  = r >= rlo && r < rhi;
            acc[r] += lds_update<PREC, kEdge, SH>(sB, on ? fmaf(fr, du, uo) : kParkUV,
                on ? fmaf(fr, dv, vo) : kParkUV, on ? fmaf(fr, dw, wo) : kPark, hiu, hiv, ou,
                ov);
            if ((r % kFence<PREC>) == kFence<PREC> - 1)
              __builtin_amdgcn_sched_barrier(0);
          }
        }
      } else {
        // footprint larger than the LDS window (or taps from L2 by
        // configuration): L2 gathers in groups of kFenceL2 voxels with the
        // clamp loop's arithmetic; segment parking where the box is off
        // the detector edge and the ranges are aligned, as on the LDS paths
        KernelArgs &K = kernel_args();
        const __amdgpu_buffer_rsrc_t rs = slot_rsrc(K, slot);
        const int pitchE   = (int)(K.rowBytes >> 2);
        const uint32_t row = K.rowBytes;
        if (aligned && !edge) {
          LH_LOOP_ORIGIN("l2 clamp");
          const uint32_t base = K.baseOffset + (uint32_t)(box.iv0 * pitchE + box.iu0) * 4u;
          static_for<0, R / kSeg<R>>([&](auto hc) {
            constexpr int h = decltype(hc)::value;
            LH_SEGMENT(h);
            l2_run<PREC, kClamp, kSeg<R>, h * kSeg<R>>(acc, rs, base, pitchE, row, us, vs, ws,
                dus, dvs, dws, hiu, hiv);
          });
        } else {
          LH_LOOP_ORIGIN("l2 edge or partial");
          const float ou = edge ? (float)box.iu0 : 0.0f, ov = edge ? (float)box.iv0 : 0.0f;
          // byte offset of the box origin pixel (edge boxes: of pixel (0, 0))
          const uint32_t base =
              K.baseOffset + (edge ? 0u : (uint32_t)(box.iv0 * pitchE + box.iu0) * 4u);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const float fr = (float)r;
            const bool on  = r >= rlo && r < rhi;
            acc[r] += l2_update<PREC, kEdge>(rs, base, pitchE, row,
                on ? fmaf(fr, du, uo) : kParkUV, on ? fmaf(fr, dv, vo) : kParkUV,
                on ? fmaf(fr, dw, wo) : kPark, hiu, hiv, ou, ov);
            if ((r % kFenceL2<R>) == kFenceL2<R> - 1)
              __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
#undef LH_LOOP_ORIGIN
#undef LH_SEGMENT
    }
    box      = boxNext;
    blockAct = blockActNext;
  }

  if (touched) {
    float4 *p = reinterpret_cast<float4 *>(A.vol + line * (uint64_t)L + (uint64_t)x0);
#pragma unroll
    for (int q = 0; q < R / 4; ++q) {
      float4 o = p[q];
      o.x += acc[4 * q + 0];
      o.y += acc[4 * q + 1];
      o.z += acc[4 * q + 2];
      o.w += acc[4 * q + 3];
      p[q] = o;
    }
  }
}

// Grid form: one workgroup per tile.  Persistent form: a grid sized to the
// chip walks the tiles that hold work (tile_list).  Workgroups run on the 8
// XCDs in turn in dispatch order, so blockIdx mod 8 names one XCD's
// workgroups; each XCD's range is one contiguous part of the z-major list
// (one z-slab, so its L2 serves one band of detector rows), the ranges cut
// to equal work, and a workgroup takes tiles from another XCD's range only
// when its own is used up.
// A must stay the kernel's only argument: kernel_args() reads it back from
// offset 0 of the kernarg segment.
// SH::OCC: workgroups per CU the register and LDS budget is sized for
// (4: 128 VGPRs, 3: 168).
template <int R, int PREC, int STAGE, class SH>
__global__ __launch_bounds__(kWaves *kWave, SH::OCC) void bp_kernel(BpArgs A)
{
  __shared__ float sImg[STAGE ? 2 * SH::BUF : 1];
  __shared__ int sAct[2 * kWaves];
  if (!A.persistent) {
    bp_tile_lds<R, PREC, STAGE, SH>(A, blockIdx.x, blockIdx.y, A.tileZ0 + blockIdx.z, sImg, sAct);
    return;
  }
  // Each workgroup takes the next entry of its XCD's range (one atomic per
  // tile, lane 0, handed to the other waves through LDS), then, once that
  // range is used up, of the other XCDs' ranges: tiles of unequal work
  // (clipped lines, the cylinder's edge) no longer leave workgroups idle
  // while others still hold a static share.  (Taking the next entry a tile
  // ahead, to hide the atomic, ran 10% slower: a workgroup then holds a
  // tile that an idle one could run.  Ordering each range heaviest first
  // tightened the end of the launch but lost more to L2 locality than the
  // z-major order: 28.7 -> 29.4 ms per step at L=512.)
  __shared__ int sNext;
  const int xcd = blockIdx.x & 7;
  for (;;) {
    __syncthreads(); // the previous tile's last LDS reads (footprints, votes, sNext) are done
    if (threadIdx.x == 0) {
      int e = -1;
      for (int k = 0; k < 8 && e < 0; ++k) {
        const int x = (xcd + k) & 7;
        const int n = A.xcdTile[x + 1] - A.xcdTile[x];
        if (n > 0) {
          const int i = atomicAdd(A.taken + x, 1);
          if (i < n)
            e = A.xcdTile[x] + i;
        }
      }
      sNext = e;
    }
    __syncthreads();
    const int e = __builtin_amdgcn_readfirstlane(sNext);
    if (e < 0)
      break;
    const int t  = A.tiles ? A.tiles[e] : e;
    const int tx = t % A.tilesX;
    const int ty = (t / A.tilesX) % A.tilesY;
    const int tz = A.tileZ0 + t / (A.tilesX * A.tilesY);
    bp_tile_lds<R, PREC, STAGE, SH>(A, tx, ty, tz, sImg, sAct);
  }
}

// zeroes the persistent walk's per-XCD counters before a launch (a kernel,
// not a memset, so that captured batches replay it as a kernel node)
__global__ void zero_taken_kernel(int *taken)
{
  if (threadIdx.x < 8)
    taken[threadIdx.x] = 0;
}

// ---- projection staging -------------------------------------------------

// max |p| over an image, as the bit pattern of a non-negative float (the
// order of non-negative floats is that of their bit patterns).  16-byte loads
// when the image is 16-byte aligned; one atomic per workgroup, since atomics
// on one address serialise in L2 (one per wave took ~49 us per 1248 x 960
// image, one per workgroup of a 256-workgroup grid ~5 us).
constexpr int kMaxAbsThreads = 256;
__global__ __launch_bounds__(kMaxAbsThreads) void maxabs_kernel(const float *__restrict__ img,
    uint32_t n, uint32_t *out, uint64_t imgStride = 0)
{
  // a batch of images: image blockIdx.y at img + blockIdx.y * imgStride
  img += (uint64_t)blockIdx.y * imgStride;
  out += blockIdx.y;
  __shared__ uint32_t sMax[kMaxAbsThreads / kWave];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  const uint32_t n4  = ((uintptr_t)img & 15) == 0 ? n / 4 : 0;
  const float4 *i4   = reinterpret_cast<const float4 *>(img);
  uint32_t m = 0;
  for (uint32_t i = tid; i < n4; i += nt) {
    const float4 q = i4[i];
    m = max(m, max(max(__float_as_uint(fabsf(q.x)), __float_as_uint(fabsf(q.y))),
                   max(__float_as_uint(fabsf(q.z)), __float_as_uint(fabsf(q.w)))));
  }
  for (uint32_t i = 4 * n4 + tid; i < n; i += nt)
    m = max(m, __float_as_uint(fabsf(img[i])));
  for (int o = 32; o > 0; o >>= 1)
    m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if ((threadIdx.x & (kWave - 1)) == 0)
    sMax[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMaxAbsThreads / kWave; ++w)
      m = max(m, sMax[w]);
    atomicMax(out, m);
  }
}

// Storage scale s = 4^m: FP16 keeps max|p|/s in (4096, 16384], so forward
// differences (at most 2 max|p|/s) stay below FP16's largest finite 65504;
// BF16 and FP32 use s = 1.  A power of four keeps k = s^-1/2 a power of two, so scaling the
// matrix by k is exact and 1/w^2 picks up exactly the factor s.
__device__ __forceinline__ int storage_exponent(int prec, uint32_t maxbits)
{
  if (prec != LH_PREC_F16 || maxbits == 0)
    return 0;
  const float mx = __uint_as_float(maxbits);
  int m = 0;
  float s = 1.0f;
  while (mx / s > 16384.0f && m < 60) { s *= 4.0f; ++m; }
  while (mx / s <= 4096.0f && m > -60) { s *= 0.25f; --m; }
  return m;
}

__device__ __forceinline__ uint16_t to_bf16(float f)
{
  uint32_t b = __float_as_uint(f);
  if ((b & 0x7fffffffu) > 0x7f800000u)
    return (uint16_t)((b >> 16) | 0x40u); // quiet NaN
  b += 0x7fffu + ((b >> 16) & 1u);        // round to nearest even
  return (uint16_t)(b >> 16);
}

__device__ __forceinline__ uint16_t to_f16(float f)
{
  return __half_as_ushort(__float2half_rn(f));
}

// Writes the bordered slot image from a W x H float image.  FP32: element
// (i,j), i in [-2, W+1], j in [-2, H+1]; halves: pair element (i,j), i in
// [-2, W], j in [-2, H+1], holding p(i,j) rounded and the difference
// p(i+1,j) - (the rounded p(i,j)) rounded, so p0 + d reproduces p(i+1,j) to
// the difference's own rounding.  Block (0,0) thread 0 also applies the storage scale
// to the slot's view parameters.
template <int PREC>
__device__ __forceinline__ void stage_element(const float *__restrict__ img, int W, int H,
    uint8_t *slot, uint32_t rowBytes, const uint32_t *maxbits, ViewParams vp, ViewParams *out,
    int ei, int ej, bool leader)
{
  // only FP16 storage is scaled (maxabs_kernel filled *maxbits)
  const int m    = PREC == LH_PREC_F16 ? storage_exponent(PREC, *maxbits) : 0;
  const float is = ldexpf(1.0f, -2 * m);
  if (leader) {
    if (PREC != LH_PREC_F32) {
      const double k = ldexp(1.0, -m);
      for (int c = 0; c < 12; ++c)
        vp.a[c] *= k;
      vp.du *= (float)k;
      vp.dv *= (float)k;
      vp.dw *= (float)k;
      vp.scale = ldexpf(1.0f, 2 * m);
    }
    *out = vp;
  }
  auto pix = [&](int i, int j) -> float {
    return (i >= 0 && i < W && j >= 0 && j < H) ? img[(size_t)j * W + i] : 0.0f;
  };
  const int i = ei - (int)kBorder, j = ej - (int)kBorder;
  if constexpr (PREC == LH_PREC_F32) {
    if (ei >= W + 4)
      return;
    reinterpret_cast<float *>(slot + (size_t)ej * rowBytes)[ei] = pix(i, j);
  } else {
    if (ei >= W + 3)
      return;
    const float v0 = pix(i, j) * is, v1 = pix(i + 1, j) * is;
    const uint16_t h0 = PREC == LH_PREC_BF16 ? to_bf16(v0) : to_f16(v0);
    const float p0 = PREC == LH_PREC_BF16 ? bf16_lo(h0) : f16_lo(h0);
    const uint16_t h1 = PREC == LH_PREC_BF16 ? to_bf16(v1 - p0) : to_f16(v1 - p0);
    const uint32_t q = (uint32_t)h0 | ((uint32_t)h1 << 16);
    reinterpret_cast<uint32_t *>(slot + (size_t)ej * rowBytes)[ei] = q;
  }
}

template <int PREC>
__global__ void stage_kernel(const float *__restrict__ img, int W, int H, uint8_t *slot,
    uint32_t rowBytes, const uint32_t *maxbits, ViewParams vp, ViewParams *out)
{
  stage_element<PREC>(img, W, H, slot, rowBytes, maxbits, vp, out,
      blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y,
      blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0);
}

// A batch of views in one launch: view z = blockIdx.z from img + z * imgStride
// into slot z of `slots`, with view parameters vpIn[z] (the host-computed
// FP64 matrix and steps) and max-abs cell maxbits[z].
template <int PREC>
__global__ void stage_batch_kernel(const float *__restrict__ img, uint64_t imgStride, int W,
    int H, uint8_t *slots, uint64_t slotBytes, uint32_t rowBytes, const uint32_t *maxbits,
    const ViewParams *__restrict__ vpIn, ViewParams *out)
{
  const int z = blockIdx.z;
  stage_element<PREC>(img + (uint64_t)z * imgStride, W, H, slots + (uint64_t)z * slotBytes,
      rowBytes, maxbits + z, vpIn[z], out + z, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y,
      blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0);
}

// A batch of views uploaded in half precision (lhReconstructHostHalf): view
// z = blockIdx.z, W x H halves at img + z * imgStride, already scaled by the
// host (lhHostHalfViews), into the pair layout of slot z: element (i,j)
// holds h(i,j) and the forward difference h(i+1,j) - h(i,j) of the uploaded
// halves, rounded to the storage precision (zero outside the image).  The
// view parameters vpIn[z] carry the scale already.
template <int PREC>
__global__ void stage_half_batch_kernel(const uint16_t *__restrict__ img, uint64_t imgStride,
    int W, int H, uint8_t *slots, uint64_t slotBytes, uint32_t rowBytes,
    const ViewParams *__restrict__ vpIn, ViewParams *out)
{
  const int z  = blockIdx.z;
  const int ei = blockIdx.x * blockDim.x + threadIdx.x, ej = blockIdx.y;
  img += (uint64_t)z * imgStride;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
    out[z] = vpIn[z];
  if (ei >= W + 3)
    return;
  const int i = ei - (int)kBorder, j = ej - (int)kBorder;
  auto hpix = [&](int ii) -> uint32_t {
    return (ii >= 0 && ii < W && j >= 0 && j < H) ? img[(size_t)j * W + ii] : 0u;
  };
  const uint32_t h0 = hpix(i), h1 = hpix(i + 1);
  const float p0 = PREC == LH_PREC_BF16 ? bf16_lo(h0) : f16_lo(h0);
  const float p1 = PREC == LH_PREC_BF16 ? bf16_lo(h1) : f16_lo(h1);
  const uint16_t d = PREC == LH_PREC_BF16 ? to_bf16(p1 - p0) : to_f16(p1 - p0);
  reinterpret_cast<uint32_t *>(slots + (uint64_t)z * slotBytes + (size_t)ej * rowBytes)[ei] =
      h0 | ((uint32_t)d << 16);
}

// vol *= s (float4 body, scalar tail)
__global__ void scale_kernel(float *__restrict__ vol, uint64_t n, float s)
{
  const uint64_t n4 = n / 4;
  float4 *v4 = reinterpret_cast<float4 *>(vol);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4;
       i += (uint64_t)gridDim.x * blockDim.x) {
    float4 q = v4[i];
    q.x *= s;
    q.y *= s;
    q.z *= s;
    q.w *= s;
    v4[i] = q;
  }
  for (uint64_t i = 4 * n4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    vol[i] *= s;
}

// ---- line ranges on the GPU ----------------------------------------------
// Bit-identical form of analyseGeometry.c:207-281: one thread per (z,y) line
// scans x from the left for the first hit and from the right for the last;
// the hit test is (int)u in [0,W-1] and (int)v in [0,H-1] in FP64 on float
// world coordinates (wz, wy accumulated by "+= mm" in float beforehand).
__device__ __forceinline__ bool clip_hit(const double *a, float wx, float wy, float wz,
    int W, int H)
{
  const double x = wx, y = wy, z = wz;
  // the same association as the C expression a2*x + a5*y + a8*z + a11
  const double wn = __dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(a[2], x), __dmul_rn(a[5], y)),
                                  __dmul_rn(a[8], z)), a[11]);
  const double un = __ddiv_rn(__dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(a[0], x),
                        __dmul_rn(a[3], y)), __dmul_rn(a[6], z)), a[9]), wn);
  const double vn = __ddiv_rn(__dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(a[1], x),
                        __dmul_rn(a[4], y)), __dmul_rn(a[7], z)), a[10]), wn);
  if (!(un > -1.0e9 && un < 1.0e9 && vn > -1.0e9 && vn < 1.0e9))
    return false;
  const long long iix = (long long)un, iiy = (long long)vn;
  return iix >= 0 && iix <= W - 1 && iiy >= 0 && iiy <= H - 1;
}

__global__ void line_ranges_kernel(int L, float rL, float oL, int W, int H,
    const double *__restrict__ geom, int vs, const float *__restrict__ wacc, int *out)
{
  const int view = blockIdx.y;
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= L * L)
    return;
  const int z = line / L, y = line % L;
  double a[12];
  for (int c = 0; c < 12; ++c)
    a[c] = geom[view * 12 + c];
  const float wz = wacc[z], wy = wacc[y];
  int xstart = -1, xstop = -1;
  for (int x = 0; x < L; ++x) {
    const float wx = __fadd_rn(__fmul_rn((float)x, rL), oL);
    if (clip_hit(a, wx, wy, wz, W, H)) { xstart = x; break; }
  }
  int packed = 0;
  if (xstart >= 0) {
    for (int x = L - 1; x >= 0; --x) {
      const float wx = __fadd_rn(__fmul_rn((float)x, rL), oL);
      if (clip_hit(a, wx, wy, wz, W, H)) { xstop = x + 1; break; }
    }
    int al;
    if ((al = xstop % vs))
      xstop += vs - al;
    if ((al = xstart % vs))
      xstart -= al;
    packed = (int)(uint16_t)(int16_t)xstart | (int)((uint32_t)(uint16_t)(int16_t)xstop << 16);
  }
  out[(size_t)view * L * L + line] = packed;
}

} // namespace

// ---- scratch of the allocation-free entry point -----------------------------
// lhBackprojectNoAlloc keeps the footprint boxes of one register pass in this
// file-scope array instead of the context's table (which lhBackproject
// allocates on first use of a launch shape): 4 Mi boxes of 8 bytes (32 MiB),
// one copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, resolved once
// per device at context creation (scratch_boxes).  Calls on one GPU share the
// array: calls that overlap on one GPU race.
constexpr uint64_t kScratchBoxes = 4ull << 20;
static __device__ __align__(256) int2 scratch[kScratchBoxes];

namespace {
std::mutex g_scratchMu;
std::vector<int2 *> g_scratchPtr; // device address of `scratch`, by device ordinal

// The device address of `scratch` on the current device `dev`.
int scratch_boxes(int dev, int2 **out)
{
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if ((int)g_scratchPtr.size() <= dev)
    g_scratchPtr.resize(dev + 1, nullptr);
  if (!g_scratchPtr[dev]) {
    void *p             = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return lh_fail(LH_E_HIP, "scratch table of device %d: hipGetSymbolAddress: %s", dev,
          hipGetErrorString(e));
    g_scratchPtr[dev] = static_cast<int2 *>(p);
  }
  *out = g_scratchPtr[dev];
  return LH_OK;
}
} // namespace

// ---- context ---------------------------------------------------------------

// The tiles of one launch shape that hold a voxel inside some view's line
// ranges (persistent walk order, on the device) and the per-XCD cuts.
struct TileList {
  int *dTiles = nullptr;
  int cut[9]  = {};
};

struct LhContext {
  LhConfig cfg{};
  hipStream_t stream = nullptr;
  float *dVol        = nullptr;
  int *dRanges       = nullptr;
  ViewParams *dViews = nullptr;
  uint8_t *dSlots    = nullptr;
  uint32_t *dMax     = nullptr;     // one per slot
  int *dTaken        = nullptr;     // the persistent walk's per-XCD counters (8)
  // host-upload ring: pinned host buffers (allocated on first use) and device
  // staging buffers of W*H floats; the H2D copies run on upStream
  hipStream_t upStream = nullptr;
  int nRing = 0, ringNext = 0, ringAcquired = -1;
  std::vector<float *> hRing, dRing;
  std::vector<hipEvent_t> copyDone, stageDone;
  // host-memory reconstruction (lhReconstructHost): a device ring of two
  // halves of `hostBatch` W*H images and view parameters, filled by H2D copies
  // on upStream and staged into the slots by one launch per batch on the
  // compute stream; per half, an event after its copies and one after the
  // staging launch that read it
  uint32_t hostBatch = 0;
  float *dHostRing   = nullptr;
  ViewParams *dHostParams = nullptr;
  ViewParams *hHostParams = nullptr; // pinned, one per view of the call
  uint32_t hHostParamsCap = 0;
  hipEvent_t halfCopied[2] = { nullptr, nullptr }, halfFree[2] = { nullptr, nullptr };
  // captured backprojection batches, by (first slot, count, kernel, run, staging)
  bool graphs = false;
  std::map<std::tuple<uint32_t, uint32_t, int, int, int>, hipGraphExec_t> graphCache;
  uint64_t slotBytes = 0;
  uint32_t rowBytes  = 0;
  uint32_t slotRecords = 0;
  uint32_t baseOffset  = 0;
  int numCU = 0;
  std::vector<uint64_t> effPerView; // by id-1
  // per line (z L + y): the union over views of the line ranges, [lineLo,
  // lineHi) (empty when lineLo >= lineHi; no vectors: ranges not counted),
  // and the voxels inside the ranges summed over views
  std::vector<int> lineLo, lineHi;
  std::vector<uint32_t> lineWork;
  // tiles holding work per (run, tile y, tile z, slab z0, slab z1)
  std::map<std::tuple<int, int, int, uint32_t, uint32_t>, TileList> tileLists;
  std::vector<int> slotId;          // id held by each slot (0 = empty)
  uint32_t slabZ0 = 0, slabZ1 = 0;  // z-slab of the volume backprojected ([0, L) by default)
  int2 *dBoxes  = nullptr;          // staged kernel: footprint box per (slot, tile)
  int2 *scratchBoxes = nullptr;     // device address of `scratch` on this context's device
  long boxTiles = 0;                // tiles per slot in dBoxes
  std::vector<int> boxR;            // tile (run length, y extent) the slot's boxes were built
                                    // for, as R | ly << 8 (0 = none)
  uint32_t tileY = 0;               // tile y extent set by lhSetTile (0: the run length's default)
};

// Grid of one launch: one workgroup per tile, or for the persistent form as
// many workgroups as the chip keeps resident (occupancy x CUs).
template <int R, int PREC, int STAGE, class SH>
static void launch_shape(LhContext *c, const BpArgs &A)
{
  static int perCU = 0;
  dim3 grid(A.tilesX, A.tilesY, A.tilesZ);
  if (A.persistent) {
    if (perCU == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU,
                           bp_kernel<R, PREC, STAGE, SH>, kWaves * kWave, 0) != hipSuccess ||
                          perCU < 1))
      perCU = SH::OCC;
    grid = dim3((unsigned)std::max<long>(1, std::min<long>((long)c->numCU * perCU,
        (long)A.tilesX * A.tilesY * A.tilesZ)));
  }
  bp_kernel<R, PREC, STAGE, SH><<<grid, dim3(kWaves * kWave), 0, c->stream>>>(A);
}

// Staged tile shape of a launch: (y, z) extents of the tile and the window
// kernel instance it runs.  The context's tile choice (lhSetTile) picks the
// y extent for 16-voxel runs; 8- and 32-voxel runs take 32 x 8 tiles.
enum TileKind { kTileY32, kTileY32W, kTileY16, kTileY64 };
struct TileDims {
  int ly, lz, kind;
};
static TileDims tile_dims(const LhContext *c);
static void drop_tile_lists(LhContext *c);

const LhConfig &lh_config(const LhContext *c) { return c->cfg; }
hipStream_t lh_stream(const LhContext *c) { return c->stream; }

extern "C" {

const char *lhLastError(void) { return g_err.c_str(); }

void lhAbiLayout(uint64_t *out)
{
  out[0]  = offsetof(RabbitCtGlobalData, problemSize);
  out[1]  = offsetof(RabbitCtGlobalData, imageWidth);
  out[2]  = offsetof(RabbitCtGlobalData, imageHeight);
  out[3]  = offsetof(RabbitCtGlobalData, voxelSize);
  out[4]  = offsetof(RabbitCtGlobalData, O_Index);
  out[5]  = offsetof(RabbitCtGlobalData, volumeData);
  out[6]  = offsetof(RabbitCtGlobalData, numberOfProjections);
  out[7]  = offsetof(RabbitCtGlobalData, globalGeometry);
  out[8]  = offsetof(RabbitCtGlobalData, projectionBuffer);
  out[9]  = offsetof(RabbitCtGlobalData, clipFilename);
  out[10] = sizeof(RabbitCtGlobalData);
}

void lhDefaultConfig(LhConfig *cfg, uint32_t L, uint32_t W, uint32_t H, uint32_t nViews)
{
  std::memset(cfg, 0, sizeof *cfg);
  cfg->L            = L;
  cfg->R_L          = L ? 256.0f / (float)L : 0.0f;
  cfg->O_L          = -0.5f * cfg->R_L * ((float)L - 1.0f);
  cfg->W            = W;
  cfg->H            = H;
  cfg->nViews       = nViews;
  cfg->viewsPerPass = 32;
  cfg->nSlots       = 2 * cfg->viewsPerPass;
  cfg->precision    = LH_PREC_F32;
  cfg->kernel       = LH_KERNEL_FUSED;
  cfg->runLength    = 16;
  cfg->clip         = 1;
  cfg->staging      = 1;
  cfg->device       = 0;
}

static int create_buffers(LhContext *c);

int lhCreate(const LhConfig *cfg, LhContext **out)
{
  *out = nullptr;
  if (!cfg || cfg->L == 0 || cfg->L % 64 || cfg->L > 4096 || cfg->W < 2 || cfg->H < 2 ||
      cfg->W > 16384 || cfg->H > 16384 || cfg->nViews == 0 || cfg->nSlots == 0)
    return fail(LH_E_ARG, "lhCreate: unsupported sizes (L must be a multiple of 64)");
  if (cfg->precision < LH_PREC_F32 || cfg->precision > LH_PREC_F16)
    return fail(LH_E_ARG, "lhCreate: bad precision %d", cfg->precision);
  const uint32_t R = cfg->runLength ? cfg->runLength : 16;
  if ((R != 8 && R != 16 && R != 32) || cfg->L % R)
    return fail(LH_E_ARG, "lhCreate: runLength must be 8, 16 or 32 and divide L");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(LH_E_NODEV, "no HIP device visible");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(LH_E_ARG, "device %d out of range (%d visible)", cfg->device, ndev);

  LhContext *c = new (std::nothrow) LhContext;
  if (!c)
    return fail(LH_E_NOMEM, "host allocation failed");
  c->cfg           = *cfg;
  c->cfg.runLength = R;
  const int rc     = create_buffers(c);
  if (rc != LH_OK) {
    lhDestroy(c); // releases whatever was created; the error message stays
    return rc;
  }
  *out = c;
  return LH_OK;
}

// Device buffers, streams and the upload ring of a new context (c->cfg set).
static int create_buffers(LhContext *c)
{
  const LhConfig *cfg = &c->cfg;
  LH_HIP(hipSetDevice(cfg->device));
  {
    const int rc = scratch_boxes(cfg->device, &c->scratchBoxes);
    if (rc != LH_OK)
      return rc;
  }
  LH_HIP(hipDeviceGetAttribute(&c->numCU, hipDeviceAttributeMultiprocessorCount, cfg->device));
  LH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

  const uint64_t L = cfg->L;
  LH_HIP(hipMalloc(&c->dVol, L * L * L * sizeof(float)));
  LH_HIP(hipMemsetAsync(c->dVol, 0, L * L * L * sizeof(float), c->stream));
  if (cfg->clip)
    LH_HIP(hipMalloc(&c->dRanges, (uint64_t)cfg->nViews * L * L * sizeof(int)));
  LH_HIP(hipMalloc(&c->dViews, cfg->nSlots * sizeof(ViewParams)));
  LH_HIP(hipMemsetAsync(c->dViews, 0, cfg->nSlots * sizeof(ViewParams), c->stream));

  // slot layout: FP32 rows of W+4 floats; halves: rows of W+3 pairs (one
  // dword: two halves); H+4 rows.  Rows padded to 256 bytes.
  const uint64_t elem  = 4;
  const uint64_t cols  = cfg->precision == LH_PREC_F32 ? cfg->W + 4 : cfg->W + 3;
  const uint64_t rows  = cfg->H + 4;
  c->rowBytes          = (uint32_t)((cols * elem + 255) / 256 * 256);
  c->slotRecords       = (uint32_t)(c->rowBytes * rows);
  c->slotBytes         = ((uint64_t)c->slotRecords + 4095) / 4096 * 4096;
  c->baseOffset        = (uint32_t)(kBorder * c->rowBytes + kBorder * elem);
  LH_HIP(hipMalloc(&c->dSlots, c->slotBytes * cfg->nSlots));
  LH_HIP(hipMemsetAsync(c->dSlots, 0, c->slotBytes * cfg->nSlots, c->stream));
  LH_HIP(hipMalloc(&c->dMax, cfg->nSlots * sizeof(uint32_t)));
  LH_HIP(hipMalloc(&c->dTaken, 8 * sizeof(int)));
  // one ring entry per view of a register pass (2..32), so a whole batch can
  // be in flight while the previous one is backprojected
  LH_HIP(hipStreamCreateWithFlags(&c->upStream, hipStreamNonBlocking));
  c->nRing = (int)std::min<uint32_t>(32u, std::max<uint32_t>(2u, cfg->viewsPerPass));
  c->hRing.assign(c->nRing, nullptr);
  c->dRing.assign(c->nRing, nullptr);
  c->copyDone.assign(c->nRing, nullptr);
  c->stageDone.assign(c->nRing, nullptr);
  for (int k = 0; k < c->nRing; ++k) {
    LH_HIP(hipMalloc(&c->dRing[k], (uint64_t)cfg->W * cfg->H * sizeof(float)));
    LH_HIP(hipEventCreateWithFlags(&c->copyDone[k], hipEventDisableTiming));
    LH_HIP(hipEventCreateWithFlags(&c->stageDone[k], hipEventDisableTiming));
  }
  c->slabZ0 = 0;
  c->slabZ1 = cfg->L;
  c->effPerView.assign(cfg->nViews, cfg->clip ? 0 : L * L * L);
  c->slotId.assign(cfg->nSlots, 0);
  c->boxR.assign(cfg->nSlots, 0);
  LH_HIP(hipStreamSynchronize(c->stream));
  return LH_OK;
}

int lhDestroy(LhContext *c)
{
  if (!c)
    return LH_OK;
  // every stream drains before the buffers its work reads or writes go away
  for (hipStream_t st : { c->stream, c->upStream })
    if (st)
      (void)hipStreamSynchronize(st);
  (void)hipFree(c->dVol);
  (void)hipFree(c->dRanges);
  (void)hipFree(c->dViews);
  (void)hipFree(c->dSlots);
  for (int h = 0; h < 2; ++h) {
    if (c->halfCopied[h])
      (void)hipEventDestroy(c->halfCopied[h]);
    if (c->halfFree[h])
      (void)hipEventDestroy(c->halfFree[h]);
  }
  (void)hipFree(c->dHostRing);
  (void)hipFree(c->dHostParams);
  if (c->hHostParams)
    (void)hipHostFree(c->hHostParams);
  for (auto &kv : c->graphCache)
    (void)hipGraphExecDestroy(kv.second);
  (void)hipFree(c->dMax);
  (void)hipFree(c->dTaken);
  (void)hipFree(c->dBoxes);
  drop_tile_lists(c);
  for (int k = 0; k < c->nRing; ++k) {
    (void)hipFree(c->dRing[k]);
    if (c->hRing[k])
      (void)hipHostFree(c->hRing[k]);
    if (c->copyDone[k])
      (void)hipEventDestroy(c->copyDone[k]);
    if (c->stageDone[k])
      (void)hipEventDestroy(c->stageDone[k]);
  }
  if (c->upStream)
    (void)hipStreamDestroy(c->upStream);
  if (c->stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
  return LH_OK;
}

static void drop_graphs(LhContext *c);

// Voxels inside the line ranges per view, and per line the union of the
// views' ranges.  The persistent walk's tile lists follow the unions, so
// lists and the batches captured with them are dropped.
static void count_effective(LhContext *c, const int *packed)
{
  const uint32_t L  = c->cfg.L;
  const uint64_t LL = (uint64_t)L * L;
  drop_graphs(c);
  drop_tile_lists(c);
  c->lineLo.assign(LL, (int)L);
  c->lineHi.assign(LL, 0);
  c->lineWork.assign(LL, 0);
  for (uint32_t n = 0; n < c->cfg.nViews; ++n) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < LL; ++i) {
      const int v     = packed[n * LL + i];
      const int start = std::max((int)(short)(v & 0xffff), 0);
      const int stop  = std::min(v >> 16, (int)L);
      if (stop > start) {
        s += (uint64_t)(stop - start);
        c->lineLo[i] = std::min(c->lineLo[i], start);
        c->lineHi[i] = std::max(c->lineHi[i], stop);
        c->lineWork[i] += (uint32_t)(stop - start);
      }
    }
    c->effPerView[n] = s;
  }
}

// The persistent walk of a launch shape (BpArgs::tiles, xcdTile): the tiles
// whose lines' range unions meet the tile's x extent, z-major, cut into 8
// contiguous ranges of equal work (a tile's work: its lines' voxels inside
// the ranges over all views, each line's share by the overlap of the tile
// with the line's union, plus 1/16 of a full tile per view for the view
// loop an idle tile still runs).  Without it, a volume whose top and bottom slabs project past
// the detector (L=512: work in the middle ~70% of z) left two XCDs walking
// mostly idle tiles and four at 1.41x the mean work; walking the idle tiles
// costs a view loop each (range loads and a barrier per view).  Built on the
// first launch of a shape after the ranges are set (mayAllocate); otherwise,
// or without clipping, every tile in 8 even ranges.
static void drop_tile_lists(LhContext *c)
{
  for (auto &kv : c->tileLists)
    (void)hipFree(kv.second.dTiles);
  c->tileLists.clear();
}

static int tile_list(LhContext *c, int R, const TileDims &td, int tilesX, int tilesY,
    int tileZ0, int tilesZ, bool mayAllocate, const int **tiles, int *cut)
{
  *tiles = nullptr;
  const long nAll = (long)tilesX * tilesY * tilesZ;
  for (int x = 0; x <= 8; ++x)
    cut[x] = (int)(nAll * x / 8);
  if (!c->cfg.clip || c->lineLo.empty())
    return LH_OK;
  const auto key = std::make_tuple(R, td.ly, td.lz, c->slabZ0, c->slabZ1);
  auto it = c->tileLists.find(key);
  if (it == c->tileLists.end()) {
    if (!mayAllocate)
      return LH_OK;
    const int L = (int)c->cfg.L;
    std::vector<int> list;
    std::vector<double> work;
    const double idle = (double)R * td.ly * td.lz / 16.0 * c->cfg.nViews;
    for (int r = 0; r < tilesZ; ++r)
      for (int ty = 0; ty < tilesY; ++ty)
        for (int tx = 0; tx < tilesX; ++tx) {
          bool inside = false;
          double w    = 0.0;
          for (int z = (tileZ0 + r) * td.lz; z < (tileZ0 + r + 1) * td.lz; ++z)
            for (int y = ty * td.ly; y < (ty + 1) * td.ly; ++y) {
              const size_t line = (size_t)z * L + y;
              const int lo = std::max(c->lineLo[line], tx * R);
              const int hi = std::min(c->lineHi[line], tx * R + R);
              if (hi > lo) {
                inside = true;
                w += (double)c->lineWork[line] * (hi - lo) /
                     (double)(c->lineHi[line] - c->lineLo[line]);
              }
            }
          if (inside) {
            list.push_back(tx + tilesX * (ty + tilesY * r));
            work.push_back(w + idle);
          }
        }
    TileList tl;
    double total = 0.0;
    for (double w : work)
      total += w;
    const int n = (int)list.size();
    tl.cut[0] = 0;
    tl.cut[8] = n;
    double acc = 0.0;
    int i      = 0;
    for (int x = 1; x < 8; ++x) {
      while (i < n && acc + 0.5 * work[(size_t)i] < total * x / 8.0)
        acc += work[(size_t)i++];
      tl.cut[x] = i;
    }
    if (n > 0) {
      LH_HIP(hipMalloc(&tl.dTiles, (size_t)n * sizeof(int)));
      const hipError_t e =
          hipMemcpy(tl.dTiles, list.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(tl.dTiles);
        return fail(LH_E_HIP, "tile list: %s", hipGetErrorString(e));
      }
    }
    it = c->tileLists.emplace(key, tl).first;
  }
  // an empty list (no work in any view) walks nothing: every cut is 0
  *tiles = it->second.dTiles;
  for (int x = 0; x <= 8; ++x)
    cut[x] = it->second.cut[x];
  return LH_OK;
}

int lhSetRanges(LhContext *c, const LhLineRange *ranges)
{
  if (!c || !c->cfg.clip || !ranges)
    return fail(LH_E_STATE, "lhSetRanges: context has no clipping");
  const uint64_t n = (uint64_t)c->cfg.nViews * c->cfg.L * c->cfg.L;
  static_assert(sizeof(LhLineRange) == sizeof(int), "packed range");
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipMemcpyAsync(c->dRanges, ranges, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  LH_HIP(hipStreamSynchronize(c->stream));
  count_effective(c, reinterpret_cast<const int *>(ranges));
  return LH_OK;
}

int lhSetRangesFromGeometry(LhContext *c, const double *geometry, int vs)
{
  if (!c || !c->cfg.clip || !geometry || (vs != 4 && vs != 8 && vs != 16))
    return fail(LH_E_ARG, "lhSetRangesFromGeometry: bad arguments");
  const LhConfig &g = c->cfg;
  std::vector<float> wacc(g.L);
  float acc = g.O_L;
  for (uint32_t i = 0; i < g.L; ++i, acc += g.R_L)
    wacc[i] = acc;
  double *dGeom = nullptr;
  float *dW     = nullptr;
  // freed on every exit, also when a step below fails
  struct Scratch {
    double *&g;
    float *&w;
    ~Scratch()
    {
      if (g)
        (void)hipFree(g);
      if (w)
        (void)hipFree(w);
    }
  } scratch{dGeom, dW};
  LH_HIP(hipSetDevice(g.device));
  LH_HIP(hipMalloc(&dGeom, (size_t)g.nViews * 12 * sizeof(double)));
  LH_HIP(hipMalloc(&dW, g.L * sizeof(float)));
  LH_HIP(hipMemcpyAsync(dGeom, geometry, (size_t)g.nViews * 12 * sizeof(double),
      hipMemcpyHostToDevice, c->stream));
  LH_HIP(hipMemcpyAsync(dW, wacc.data(), g.L * sizeof(float), hipMemcpyHostToDevice, c->stream));
  dim3 grid((g.L * g.L + 255) / 256, g.nViews);
  line_ranges_kernel<<<grid, 256, 0, c->stream>>>((int)g.L, g.R_L, g.O_L, (int)g.W, (int)g.H,
      dGeom, vs, dW, c->dRanges);
  LH_HIP(hipGetLastError());
  std::vector<int> host((size_t)g.nViews * g.L * g.L);
  LH_HIP(hipMemcpyAsync(host.data(), c->dRanges, host.size() * sizeof(int),
      hipMemcpyDeviceToHost, c->stream));
  LH_HIP(hipStreamSynchronize(c->stream));
  count_effective(c, host.data());
  return LH_OK;
}

int lhComputeLineRangesDevice(uint32_t L, float R_L, float O_L, uint32_t W, uint32_t H,
    uint32_t nViews, const double *geometry, int vectorSize, LhLineRange *out)
{
  LhConfig cfg;
  lhDefaultConfig(&cfg, L, W, H, nViews);
  cfg.R_L    = R_L;
  cfg.O_L    = O_L;
  cfg.nSlots = 1;
  LhContext *c = nullptr;
  int rc       = lhCreate(&cfg, &c);
  if (rc == LH_OK)
    rc = lhSetRangesFromGeometry(c, geometry, vectorSize);
  if (rc == LH_OK) {
    hipError_t e = hipMemcpy(out, c->dRanges, (size_t)nViews * L * L * sizeof(int),
        hipMemcpyDeviceToHost);
    if (e != hipSuccess)
      rc = fail(LH_E_HIP, "copy of ranges: %s", hipGetErrorString(e));
  }
  lhDestroy(c);
  return rc;
}

int lhSetVolume(LhContext *c, const float *host)
{
  if (!c || !host)
    return fail(LH_E_ARG, "lhSetVolume: null argument");
  const uint64_t L = c->cfg.L;
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipMemcpyAsync(c->dVol, host, L * L * L * sizeof(float), hipMemcpyHostToDevice, c->stream));
  LH_HIP(hipStreamSynchronize(c->stream));
  return LH_OK;
}

int lhZeroVolume(LhContext *c)
{
  if (!c)
    return fail(LH_E_ARG, "lhZeroVolume: null context");
  const uint64_t L = c->cfg.L;
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipMemsetAsync(c->dVol, 0, L * L * L * sizeof(float), c->stream));
  return LH_OK;
}

int lhGetVolume(LhContext *c, float *host)
{
  if (!c || !host)
    return fail(LH_E_ARG, "lhGetVolume: null argument");
  const uint64_t L = c->cfg.L;
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipMemcpyAsync(host, c->dVol, L * L * L * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  LH_HIP(hipStreamSynchronize(c->stream));
  return LH_OK;
}

int lhScaleVolume(LhContext *c, float s)
{
  if (!c)
    return fail(LH_E_ARG, "lhScaleVolume: null context");
  const uint64_t n = (uint64_t)c->cfg.L * c->cfg.L * c->cfg.L;
  LH_HIP(hipSetDevice(c->cfg.device));
  scale_kernel<<<(unsigned)std::min<uint64_t>((n / 4 + 255) / 256, 8192), 256, 0, c->stream>>>(
      c->dVol, n, s);
  LH_HIP(hipGetLastError());
  return LH_OK;
}

int lhGetConfig(const LhContext *c, LhConfig *out)
{
  if (!c || !out)
    return fail(LH_E_ARG, "lhGetConfig: null argument");
  *out = c->cfg;
  return LH_OK;
}

int lhVolumeDevice(LhContext *c, float **dVol)
{
  if (!c || !dVol)
    return fail(LH_E_ARG, "lhVolumeDevice: null argument");
  *dVol = c->dVol;
  return LH_OK;
}

// Enqueues on stream `st`: for FP16 storage the max-abs of the W x H image
// src (device), which sets the slot's storage scale, then the conversion into
// slot `slot` with the view parameters of `matrix` (passed by value to the
// staging kernel, which writes them to the slot).  FP32 and BF16 are stored
// unscaled and skip the max-abs pass.
static int load_view_common(LhContext *c, uint32_t slot, int id, const double *matrix,
    const float *src, hipStream_t st)
{
  const LhConfig &g = c->cfg;
  if (slot >= g.nSlots || id < 1 || (uint32_t)id > g.nViews || !matrix)
    return fail(LH_E_ARG, "lhLoadView: slot %u / id %d out of range", slot, id);
  ViewParams vp;
  std::memset(&vp, 0, sizeof vp);
  for (int k = 0; k < 12; ++k)
    vp.a[k] = matrix[k];
  vp.du    = (float)(matrix[0] * (double)g.R_L);
  vp.dv    = (float)(matrix[1] * (double)g.R_L);
  vp.dw    = (float)(matrix[2] * (double)g.R_L);
  vp.id    = id;
  vp.scale = 1.0f;
  if (g.precision == LH_PREC_F16) {
    LH_HIP(hipMemsetAsync(c->dMax + slot, 0, sizeof(uint32_t), st));
    const uint32_t n = g.W * g.H;
    // about four 16-byte loads per thread, at most 256 workgroups (atomics)
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(
        (n / 4 + 4 * kMaxAbsThreads - 1) / (4 * kMaxAbsThreads), 256u));
    maxabs_kernel<<<blocks, kMaxAbsThreads, 0, st>>>(src, n, c->dMax + slot);
  }
  const int cols = g.precision == LH_PREC_F32 ? (int)g.W + 4 : (int)g.W + 3;
  const int rows = (int)g.H + 4;
  dim3 grid((cols + 255) / 256, rows);
  uint8_t *dst = c->dSlots + (uint64_t)slot * c->slotBytes;
  switch (g.precision) {
  case LH_PREC_F32:
    stage_kernel<LH_PREC_F32><<<grid, 256, 0, st>>>(src, g.W, g.H, dst, c->rowBytes,
        c->dMax + slot, vp, c->dViews + slot);
    break;
  case LH_PREC_BF16:
    stage_kernel<LH_PREC_BF16><<<grid, 256, 0, st>>>(src, g.W, g.H, dst, c->rowBytes,
        c->dMax + slot, vp, c->dViews + slot);
    break;
  default:
    stage_kernel<LH_PREC_F16><<<grid, 256, 0, st>>>(src, g.W, g.H, dst, c->rowBytes,
        c->dMax + slot, vp, c->dViews + slot);
    break;
  }
  LH_HIP(hipGetLastError());
  c->slotId[slot] = id;
  c->boxR[slot]   = 0;
  return LH_OK;
}

int lhAcquireUpload(LhContext *c, float **hostBuffer)
{
  if (!c || !hostBuffer)
    return fail(LH_E_ARG, "lhAcquireUpload: null argument");
  if (c->ringAcquired >= 0)
    return fail(LH_E_STATE, "lhAcquireUpload: previous buffer not committed");
  LH_HIP(hipSetDevice(c->cfg.device));
  const int k = c->ringNext;
  if (!c->hRing[k])
    LH_HIP(hipHostMalloc((void **)&c->hRing[k], (uint64_t)c->cfg.W * c->cfg.H * sizeof(float),
        hipHostMallocDefault));
  // the buffer's previous H2D copy must have read it
  LH_HIP(hipEventSynchronize(c->copyDone[k]));
  c->ringAcquired = k;
  *hostBuffer     = c->hRing[k];
  return LH_OK;
}

int lhCancelUpload(LhContext *c)
{
  if (!c)
    return fail(LH_E_ARG, "lhCancelUpload: null context");
  c->ringAcquired = -1;  // the buffer stays next in the ring
  return LH_OK;
}

int lhCommitUpload(LhContext *c, uint32_t slot, int id, const double *matrix)
{
  if (!c)
    return fail(LH_E_ARG, "lhCommitUpload: null context");
  const int k = c->ringAcquired;
  if (k < 0)
    return fail(LH_E_STATE, "lhCommitUpload: no buffer acquired");
  c->ringAcquired = -1;
  LH_HIP(hipSetDevice(c->cfg.device));
  const size_t bytes = (size_t)c->cfg.W * c->cfg.H * sizeof(float);
  // copy stream: wait until the staging kernel that last read dRing[k] ran
  LH_HIP(hipStreamWaitEvent(c->upStream, c->stageDone[k], 0));
  LH_HIP(hipMemcpyAsync(c->dRing[k], c->hRing[k], bytes, hipMemcpyHostToDevice, c->upStream));
  LH_HIP(hipEventRecord(c->copyDone[k], c->upStream));
  // compute stream: stage once the copy has landed
  LH_HIP(hipStreamWaitEvent(c->stream, c->copyDone[k], 0));
  const int rc = load_view_common(c, slot, id, matrix, c->dRing[k], c->stream);
  if (rc != LH_OK)
    return rc;
  LH_HIP(hipEventRecord(c->stageDone[k], c->stream));
  c->ringNext = (k + 1) % c->nRing;
  return LH_OK;
}

int lhLoadView(LhContext *c, uint32_t slot, int id, const double *matrix, const float *image)
{
  if (!c || !image)
    return fail(LH_E_ARG, "lhLoadView: null argument");
  float *buf = nullptr;
  int rc     = lhAcquireUpload(c, &buf);
  if (rc != LH_OK)
    return rc;
  std::memcpy(buf, image, (size_t)c->cfg.W * c->cfg.H * sizeof(float));
  return lhCommitUpload(c, slot, id, matrix);
}

int lhLoadViewDevice(LhContext *c, uint32_t slot, int id, const double *matrix,
    const float *dImage)
{
  if (!c || !dImage)
    return fail(LH_E_ARG, "lhLoadViewDevice: null argument");
  LH_HIP(hipSetDevice(c->cfg.device));
  return load_view_common(c, slot, id, matrix, dImage, c->stream);
}

int lhQueryIdle(LhContext *c)
{
  if (!c)
    return fail(LH_E_ARG, "lhQueryIdle: null context");
  LH_HIP(hipSetDevice(c->cfg.device));
  for (hipStream_t st : { c->stream, c->upStream }) {
    const hipError_t e = hipStreamQuery(st);
    if (e == hipErrorNotReady)
      return 0;
    if (e != hipSuccess)
      return fail(LH_E_HIP, "lhQueryIdle: %s", hipGetErrorString(e));
  }
  return 1;
}

int lhDeviceInfo(int device, int *computeUnits, int *clockKHz)
{
  LH_HIP(hipDeviceGetAttribute(computeUnits, hipDeviceAttributeMultiprocessorCount, device));
  LH_HIP(hipDeviceGetAttribute(clockKHz, hipDeviceAttributeClockRate, device));
  return LH_OK;
}

int lhHostAlloc(void **p, uint64_t bytes)
{
  if (!p || bytes == 0)
    return fail(LH_E_ARG, "lhHostAlloc: bad arguments");
  *p = nullptr;
  LH_HIP(hipHostMalloc(p, (size_t)bytes, hipHostMallocDefault));
  return LH_OK;
}

int lhHostFree(void *p)
{
  if (p)
    LH_HIP(hipHostFree(p));
  return LH_OK;
}

static TileDims tile_dims(const LhContext *c)
{
  const LhConfig &g = c->cfg;
  if (g.runLength == 16 && c->tileY == 16)
    return {16, 16, kTileY16};
  if (g.runLength == 16 && c->tileY == 64)
    return {64, 4, kTileY64};
  if (g.runLength == 32 && c->tileY == 64)
    return {64, 4, kTileY64};
  // A staged 32-voxel run spanning 16 mm along x (L = 512) has footprints
  // wider than 128 pixels at some views: the 160-pixel window, 3 workgroups
  // per CU.  At 8 mm (L = 1024) the 128-pixel window fits them at 4; at 32 mm
  // (L = 256) most exceed both windows and take L2 gathers, which run fastest
  // at 4 workgroups per CU, as does the unstaged form at every size.
  if (g.runLength == 32 && g.staging && 32.0f * g.R_L > 8.0f && 32.0f * g.R_L <= 16.0f)
    return {32, 8, kTileY32W};
  return {32, 8, kTileY32};
}

static void drop_graphs(LhContext *c)
{
  for (auto &kv : c->graphCache)
    (void)hipGraphExecDestroy(kv.second);
  c->graphCache.clear();
}

// Box table storage for nTiles tiles per slot (reallocation invalidates the
// captured batches, which hold its address).
static int box_storage(LhContext *c, long nTiles)
{
  if (c->boxTiles == nTiles)
    return LH_OK;
  drop_graphs(c);
  (void)hipFree(c->dBoxes);
  c->dBoxes   = nullptr;
  c->boxTiles = 0;
  // the kernels address the boxes through one buffer resource (32-bit range)
  const uint64_t bytes = (uint64_t)c->cfg.nSlots * (uint64_t)nTiles * sizeof(int2);
  if (bytes >= (1ull << 32))
    return fail(LH_E_ARG, "footprint boxes of %u slots x %ld tiles exceed 4 GiB", c->cfg.nSlots,
        nTiles);
  LH_HIP(hipMalloc(&c->dBoxes, (size_t)bytes));
  c->boxTiles = nTiles;
  std::fill(c->boxR.begin(), c->boxR.end(), 0);
  return LH_OK;
}

// Footprint boxes of slots [first, first+count) for the staged tile grid of
// run length R; built once per (slot, R) and kept until the slot is reloaded
// (with `force`, rebuilt for every slot: the form a captured batch replays).
static int ensure_boxes(LhContext *c, uint32_t first, uint32_t count, int R, const TileDims &td,
    int tilesX, int tilesY, int tilesZ, bool force)
{
  const LhConfig &g = c->cfg;
  const long nTiles = (long)tilesX * tilesY * tilesZ;
  const int key     = R | td.ly << 8;
  const int rc      = box_storage(c, nTiles);
  if (rc != LH_OK)
    return rc;
  uint32_t s = first;
  while (s < first + count) {
    if (!force && c->boxR[s] == key) {
      ++s;
      continue;
    }
    uint32_t e = s;
    while (e < first + count && (force || c->boxR[e] != key) && e - s < 65535)
      ++e;
    box_kernel<<<dim3((unsigned)((nTiles + 255) / 256), e - s), 256, 0, c->stream>>>(c->dViews,
        (int)s, (int)nTiles, tilesX, tilesY, R, td.ly, td.lz, g.O_L, g.R_L, (int)g.W, (int)g.H,
        (int)g.nSlots, c->dBoxes);
    LH_HIP(hipGetLastError());
    for (uint32_t k = s; k < e; ++k)
      c->boxR[k] = key;
    s = e;
  }
  return LH_OK;
}

// Launches the backprojection of slots [first, first+count).  extBoxes:
// footprint boxes of exactly these slots, [tile][count], built by the caller
// (the kernel then addresses views, slots and boxes relative to `first`);
// nullptr: the context's box table, built here where needed.
static int launch_bp(LhContext *c, uint32_t first, uint32_t count, bool forceBoxes = false,
    const int2 *extBoxes = nullptr)
{
  const LhConfig &g = c->cfg;
  BpArgs A;
  A.vol         = c->dVol;
  A.ranges      = c->dRanges;
  A.views       = c->dViews;
  A.slots       = c->dSlots;
  A.slotBytes   = c->slotBytes;
  A.rowBytes    = c->rowBytes;
  A.slotRecords = c->slotRecords;
  A.baseOffset  = c->baseOffset;
  A.L           = (int)g.L;
  A.R_L         = g.R_L;
  A.O_L         = g.O_L;
  A.Wf          = (float)g.W;
  A.Hf          = (float)g.H;
  A.firstSlot   = (int)first;
  A.nViews      = (int)count;
  A.clip        = g.clip;
  const int R   = (int)g.runLength;
  const int stage = g.staging ? 1 : 0;
  const TileDims td = tile_dims(c);
  if (c->slabZ0 % td.lz || c->slabZ1 % td.lz)
    return fail(LH_E_ARG, "slab [%u, %u) is not whole %d-voxel tiles along z", c->slabZ0,
        c->slabZ1, td.lz);
  // both tap paths walk the same tiles and footprint boxes (the boxes fix
  // the local coordinate origin)
  A.tilesX      = (int)g.L / R;
  A.tilesY      = (int)g.L / td.ly;
  A.tilesZ      = (int)g.L / td.lz;
  // the box table covers every z tile; a slab launch walks its own
  const int tilesZAll = A.tilesZ;
  A.tileZ0      = (int)c->slabZ0 / td.lz;
  A.tilesZ      = (int)(c->slabZ1 - c->slabZ0) / td.lz;
  A.persistent  = g.kernel == LH_KERNEL_FUSED;
  A.tiles       = nullptr;
  for (int x = 0; x <= 8; ++x)
    A.xcdTile[x] = 0;
  if (A.persistent) {
    // the allocation-free entry point (extBoxes) walks a list only where an
    // earlier launch of this shape built it
    const int rc = tile_list(c, R, td, A.tilesX, A.tilesY, A.tileZ0, A.tilesZ,
        extBoxes == nullptr, &A.tiles, A.xcdTile);
    if (rc != LH_OK)
      return rc;
  }
  A.boxSlots    = (int)g.nSlots;
  if (extBoxes) {
    A.views     = c->dViews + first;
    A.slots     = c->dSlots + (uint64_t)first * c->slotBytes;
    A.firstSlot = 0;
    A.boxes     = extBoxes;
    A.boxSlots  = (int)count;
    A.boxBytes  = (uint32_t)((uint64_t)A.tilesX * A.tilesY * tilesZAll * count * sizeof(int2));
  } else {
    const int rc =
        ensure_boxes(c, first, count, R, td, A.tilesX, A.tilesY, tilesZAll, forceBoxes);
    if (rc != LH_OK)
      return rc;
    A.boxes    = c->dBoxes;
    A.boxBytes = (uint32_t)((uint64_t)g.nSlots * (uint64_t)c->boxTiles * sizeof(int2));
  }
#define LH_BY_PREC(RR, SH)                                                               \
  switch (g.precision) {                                                                 \
  case LH_PREC_F32:                                                                      \
    if (stage) launch_shape<RR, LH_PREC_F32, 1, SH>(c, A);                                \
    else launch_shape<RR, LH_PREC_F32, 0, SH>(c, A);                                      \
    break;                                                                               \
  case LH_PREC_BF16:                                                                     \
    if (stage) launch_shape<RR, LH_PREC_BF16, 1, SH>(c, A);                               \
    else launch_shape<RR, LH_PREC_BF16, 0, SH>(c, A);                                     \
    break;                                                                               \
  default:                                                                               \
    if (stage) launch_shape<RR, LH_PREC_F16, 1, SH>(c, A);                                \
    else launch_shape<RR, LH_PREC_F16, 0, SH>(c, A);                                      \
    break;                                                                               \
  }
  A.taken = c->dTaken;
  if (A.persistent) {
    zero_taken_kernel<<<1, 64, 0, c->stream>>>(c->dTaken);
    LH_HIP(hipGetLastError());
  }
  switch (R * 16 + td.kind) {
  case 8 * 16 + kTileY32: LH_BY_PREC(8, ShapeY32); break;
  case 16 * 16 + kTileY32: LH_BY_PREC(16, ShapeY32); break;
  case 16 * 16 + kTileY16: LH_BY_PREC(16, ShapeY16); break;
  case 16 * 16 + kTileY64: LH_BY_PREC(16, ShapeY64); break;
  case 32 * 16 + kTileY64: LH_BY_PREC(32, ShapeY64); break;
  case 32 * 16 + kTileY32: LH_BY_PREC(32, ShapeY32); break;
  case 32 * 16 + kTileY32W: LH_BY_PREC(32, ShapeY32W); break;
  default: return fail(LH_E_ARG, "no %d x %d x %d staged tile", R, td.ly, td.lz);
  }
#undef LH_BY_PREC
  LH_HIP(hipGetLastError());
  return LH_OK;
}

// The batch [first, first+count) as one HIP graph per (slots, launch shape):
// captured on first use (footprint boxes of every slot + the backprojection
// launches), replayed afterwards; the slots' contents may change between
// replays, their addresses do not.
static int graph_backproject(LhContext *c, uint32_t first, uint32_t count, uint32_t per)
{
  const LhConfig &g = c->cfg;
  const TileDims td = tile_dims(c);
  const auto key = std::make_tuple(first, count, g.kernel * 4 + g.precision,
      (int)g.runLength | td.ly << 8,
      g.kernel == LH_KERNEL_FUSED ? (int)(per << 1 | (uint32_t)g.staging) : g.staging);
  auto it = c->graphCache.find(key);
  if (it == c->graphCache.end()) {
    {
      // allocate the box table and the tile list outside the capture
      const int R = (int)g.runLength;
      int rc      = box_storage(c, (long)(g.L / R) * (g.L / td.ly) * (g.L / td.lz));
      if (rc == LH_OK && g.kernel == LH_KERNEL_FUSED) {
        const int *tiles = nullptr;
        int cut[9];
        rc = tile_list(c, R, td, (int)g.L / R, (int)g.L / td.ly, (int)c->slabZ0 / td.lz,
            (int)(c->slabZ1 - c->slabZ0) / td.lz, true, &tiles, cut);
      }
      if (rc != LH_OK)
        return rc;
    }
    hipGraph_t graph = nullptr;
    LH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int rc = LH_OK;
    for (uint32_t s = first; s < first + count && rc == LH_OK; s += per)
      rc = launch_bp(c, s, std::min(per, first + count - s), true);
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != LH_OK || e != hipSuccess) {
      if (graph)
        (void)hipGraphDestroy(graph);
      return rc != LH_OK ? rc : fail(LH_E_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
      return fail(LH_E_HIP, "graph instantiate: %s", hipGetErrorString(ei));
    it = c->graphCache.emplace(key, exec).first;
  }
  LH_HIP(hipGraphLaunch(it->second, c->stream));
  for (uint32_t s = first; s < first + count; ++s)
    c->boxR[s] = (int)g.runLength | td.ly << 8;
  return LH_OK;
}

int lhBackproject(LhContext *c, uint32_t first, uint32_t count)
{
  if (!c)
    return fail(LH_E_ARG, "lhBackproject: null context");
  const LhConfig &g = c->cfg;
  if (count == 0)
    return LH_OK;
  if (first + count > g.nSlots)
    return fail(LH_E_ARG, "lhBackproject: slots [%u,%u) beyond %u", first, first + count, g.nSlots);
  for (uint32_t s = first; s < first + count; ++s)
    if (c->slotId[s] == 0)
      return fail(LH_E_STATE, "lhBackproject: slot %u holds no view", s);
  LH_HIP(hipSetDevice(g.device));
  const uint32_t per = g.kernel == LH_KERNEL_FUSED ? std::max(1u, g.viewsPerPass) : 1u;
  if (c->graphs)
    return graph_backproject(c, first, count, per);
  for (uint32_t s = first; s < first + count; s += per) {
    const int rc = launch_bp(c, s, std::min(per, first + count - s));
    if (rc != LH_OK)
      return rc;
  }
  return LH_OK;
}

// View parameters of `matrix` for view `id` (unscaled: the staging applies
// the FP16 storage scale).
static ViewParams view_params_of(const LhConfig &g, int id, const double *matrix)
{
  ViewParams vp;
  std::memset(&vp, 0, sizeof vp);
  for (int k = 0; k < 12; ++k)
    vp.a[k] = matrix[k];
  vp.du    = (float)(matrix[0] * (double)g.R_L);
  vp.dv    = (float)(matrix[1] * (double)g.R_L);
  vp.dw    = (float)(matrix[2] * (double)g.R_L);
  vp.id    = id;
  vp.scale = 1.0f;
  return vp;
}

// Batch sizes of lhReconstructHost: batches of `batch` views, the first and
// the last two `ramp` views long, so the pipeline fills after one short upload
// and drains after one short backprojection (the copies and the kernels run
// at about the same rate).
static std::vector<uint32_t> host_batches(uint32_t count, uint32_t batch)
{
  std::vector<uint32_t> out;
  const uint32_t ramp = std::max(1u, batch / 4);
  uint32_t rem        = count;
  if (count <= 2 * batch || ramp == batch) {
    for (; rem > 0; rem -= std::min(batch, rem))
      out.push_back(std::min(batch, rem));
    return out;
  }
  out.push_back(ramp);
  rem -= ramp;
  while (rem > batch + ramp) {
    out.push_back(batch);
    rem -= batch;
  }
  if (rem > ramp)
    out.push_back(rem - ramp);
  out.push_back(std::min(ramp, rem));
  return out;
}

// The host pipeline's device ring (two halves of `batch` images) and its
// parameter buffers, kept for later calls of the same batch size.
static int host_ring(LhContext *c, uint32_t batch)
{
  // sized for float images; half-precision uploads use the first half of
  // each half of the ring
  if (c->hostBatch == batch)
    return LH_OK;
  const size_t px = (size_t)c->cfg.W * c->cfg.H;
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->upStream);
  (void)hipFree(c->dHostRing);
  (void)hipFree(c->dHostParams);
  c->dHostRing   = nullptr;
  c->dHostParams = nullptr;
  c->hostBatch   = 0;
  LH_HIP(hipMalloc(&c->dHostRing, 2 * (size_t)batch * px * sizeof(float)));
  LH_HIP(hipMalloc(&c->dHostParams, 2 * (size_t)batch * sizeof(ViewParams)));
  for (int h = 0; h < 2; ++h) {
    if (!c->halfCopied[h])
      LH_HIP(hipEventCreateWithFlags(&c->halfCopied[h], hipEventDisableTiming));
    if (!c->halfFree[h])
      LH_HIP(hipEventCreateWithFlags(&c->halfFree[h], hipEventDisableTiming));
  }
  c->hostBatch = batch;
  return LH_OK;
}

uint32_t lhHostBatchPlan(uint32_t count, uint32_t batch, uint32_t *sizes, uint32_t maxSizes)
{
  if (count == 0 || batch == 0)
    return 0;
  const std::vector<uint32_t> p = host_batches(count, batch);
  for (size_t k = 0; sizes && k < p.size() && k < maxSizes; ++k)
    sizes[k] = p[k];
  return (uint32_t)p.size();
}

// lhReconstructHost (halves == nullptr: float images) and
// lhReconstructHostHalf (halves and their storage exponents).
static int reconstruct_host(LhContext *c, const float *images, const uint16_t *halves,
    const int32_t *exps, const double *matrices, const int *ids, uint32_t count)
{
  const LhConfig &g    = c->cfg;
  const bool half      = halves != nullptr;
  const size_t elem    = half ? sizeof(uint16_t) : sizeof(float);
  const uint8_t *src   = half ? (const uint8_t *)halves : (const uint8_t *)images;
  const uint32_t batch = std::min(g.nSlots,
      g.kernel == LH_KERNEL_FUSED ? std::max(1u, g.viewsPerPass) : 16u);
  const uint32_t nsets = g.nSlots / batch;
  const size_t px      = (size_t)g.W * g.H;
  LH_HIP(hipSetDevice(g.device));
  {
    const int rc = host_ring(c, batch);
    if (rc != LH_OK)
      return rc;
  }
  if (c->hHostParamsCap < count) {
    // the previous call's parameter copies must have read the old buffer
    LH_HIP(hipStreamSynchronize(c->upStream));
    if (c->hHostParams)
      (void)hipHostFree(c->hHostParams);
    c->hHostParams    = nullptr;
    c->hHostParamsCap = 0;
    LH_HIP(hipHostMalloc((void **)&c->hHostParams, count * sizeof(ViewParams),
        hipHostMallocDefault));
    c->hHostParamsCap = count;
  } else {
    LH_HIP(hipStreamSynchronize(c->upStream)); // the buffer is rewritten below
  }
  for (uint32_t v = 0; v < count; ++v) {
    ViewParams &vp = c->hHostParams[v];
    vp             = view_params_of(g, ids[v], matrices + 12 * (size_t)v);
    if (half && exps[v] != 0) {
      // the storage scale 4^m folded into the matrix exactly, as the
      // staging of float uploads does (stage_element)
      const double k = std::ldexp(1.0, -exps[v]);
      for (int q = 0; q < 12; ++q)
        vp.a[q] *= k;
      vp.du *= (float)k;
      vp.dv *= (float)k;
      vp.dw *= (float)k;
      vp.scale = std::ldexp(1.0f, 2 * exps[v]);
    }
  }
  // the first copies wait for the compute stream's earlier work (a previous
  // call's staging may still read the ring)
  LH_HIP(hipEventRecord(c->halfFree[0], c->stream));
  LH_HIP(hipEventRecord(c->halfFree[1], c->stream));
  const std::vector<uint32_t> sizes = host_batches(count, batch);
  const int cols = g.precision == LH_PREC_F32 ? (int)g.W + 4 : (int)g.W + 3;
  uint32_t v0 = 0;
  for (size_t k = 0; k < sizes.size(); ++k) {
    const uint32_t n = sizes[k], hr = (uint32_t)(k & 1);
    const uint32_t base = (uint32_t)(k % nsets) * batch;
    float *ring         = c->dHostRing + (size_t)hr * batch * px;
    ViewParams *params  = c->dHostParams + (size_t)hr * batch;
    // upload stream: the ring half's previous staging has read it
    LH_HIP(hipStreamWaitEvent(c->upStream, c->halfFree[hr], 0));
    // the batch's images are contiguous on both sides: one copy
    LH_HIP(hipMemcpyAsync(ring, src + (size_t)v0 * px * elem, (size_t)n * px * elem,
        hipMemcpyHostToDevice, c->upStream));
    LH_HIP(hipMemcpyAsync(params, c->hHostParams + v0, n * sizeof(ViewParams),
        hipMemcpyHostToDevice, c->upStream));
    LH_HIP(hipEventRecord(c->halfCopied[hr], c->upStream));
    // compute stream: stage the batch (stream order puts it after the
    // backprojection that last read these slots), then backproject it
    LH_HIP(hipStreamWaitEvent(c->stream, c->halfCopied[hr], 0));
    uint8_t *slots = c->dSlots + (uint64_t)base * c->slotBytes;
    if (half) {
      const dim3 grid(((int)g.W + 3 + 255) / 256, g.H + 4, n);
      const uint16_t *ring16 = reinterpret_cast<const uint16_t *>(ring);
      if (g.precision == LH_PREC_BF16)
        stage_half_batch_kernel<LH_PREC_BF16><<<grid, 256, 0, c->stream>>>(ring16, px, g.W, g.H,
            slots, c->slotBytes, c->rowBytes, params, c->dViews + base);
      else
        stage_half_batch_kernel<LH_PREC_F16><<<grid, 256, 0, c->stream>>>(ring16, px, g.W, g.H,
            slots, c->slotBytes, c->rowBytes, params, c->dViews + base);
      LH_HIP(hipGetLastError());
    } else if (g.precision == LH_PREC_F16) {
      LH_HIP(hipMemsetAsync(c->dMax + base, 0, n * sizeof(uint32_t), c->stream));
      const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(
          (uint32_t)((px / 4 + 4 * kMaxAbsThreads - 1) / (4 * kMaxAbsThreads)), 256u));
      maxabs_kernel<<<dim3(blocks, n), kMaxAbsThreads, 0, c->stream>>>(ring, (uint32_t)px,
          c->dMax + base, px);
    }
    const dim3 grid((cols + 255) / 256, g.H + 4, n);
    switch (half ? -1 : (int)g.precision) {
    case -1:
      break;
    case LH_PREC_F32:
      stage_batch_kernel<LH_PREC_F32><<<grid, 256, 0, c->stream>>>(ring, px, g.W, g.H, slots,
          c->slotBytes, c->rowBytes, c->dMax + base, params, c->dViews + base);
      break;
    case LH_PREC_BF16:
      stage_batch_kernel<LH_PREC_BF16><<<grid, 256, 0, c->stream>>>(ring, px, g.W, g.H, slots,
          c->slotBytes, c->rowBytes, c->dMax + base, params, c->dViews + base);
      break;
    default:
      stage_batch_kernel<LH_PREC_F16><<<grid, 256, 0, c->stream>>>(ring, px, g.W, g.H, slots,
          c->slotBytes, c->rowBytes, c->dMax + base, params, c->dViews + base);
      break;
    }
    LH_HIP(hipGetLastError());
    LH_HIP(hipEventRecord(c->halfFree[hr], c->stream));
    for (uint32_t j = 0; j < n; ++j) {
      c->slotId[base + j] = ids[v0 + j];
      c->boxR[base + j]   = 0;
    }
    const int rc = lhBackproject(c, base, n);
    if (rc != LH_OK)
      return rc;
    v0 += n;
  }
  return LH_OK;
}

static int host_args_check(LhContext *c, const void *images, const double *matrices,
    const int *ids, uint32_t count, const char *who)
{
  if (!c || !images || !matrices || !ids || count == 0)
    return fail(LH_E_ARG, "%s: bad arguments", who);
  for (uint32_t v = 0; v < count; ++v)
    if (ids[v] < 1 || (uint32_t)ids[v] > c->cfg.nViews)
      return fail(LH_E_ARG, "%s: view id %d out of range", who, ids[v]);
  return LH_OK;
}

int lhReconstructHost(LhContext *c, const float *images, const double *matrices, const int *ids,
    uint32_t count)
{
  const int rc = host_args_check(c, images, matrices, ids, count, "lhReconstructHost");
  return rc != LH_OK ? rc : reconstruct_host(c, images, nullptr, nullptr, matrices, ids, count);
}

static int half_args_check(LhContext *c, const uint16_t *halves, const int32_t *exps,
    const double *matrices, const int *ids, uint32_t count, const char *who)
{
  const int rc = host_args_check(c, halves, matrices, ids, count, who);
  if (rc != LH_OK)
    return rc;
  if (!exps)
    return fail(LH_E_ARG, "%s: no storage exponents", who);
  if (c->cfg.precision == LH_PREC_F32)
    return fail(LH_E_ARG, "%s: the context stores FP32 projections", who);
  for (uint32_t v = 0; v < count; ++v)
    if (exps[v] < -60 || exps[v] > 60 || (c->cfg.precision == LH_PREC_BF16 && exps[v] != 0))
      return fail(LH_E_ARG, "%s: storage exponent %d of view %u out of range", who, exps[v], v);
  return LH_OK;
}

int lhReconstructHostHalf(LhContext *c, const uint16_t *halves, const int32_t *exps,
    const double *matrices, const int *ids, uint32_t count)
{
  const int rc = half_args_check(c, halves, exps, matrices, ids, count, "lhReconstructHostHalf");
  return rc != LH_OK ? rc : reconstruct_host(c, nullptr, halves, exps, matrices, ids, count);
}

// Repeats of reconstruct_host, each from idle streams, timed with HIP events
// from before the first H2D copy to after the last backprojection.
static int timed_reconstruct(LhContext *c, const float *images, const uint16_t *halves,
    const int32_t *exps, const double *matrices, const int *ids, uint32_t count, int reps,
    float *msPerRep)
{
  LH_HIP(hipSetDevice(c->cfg.device));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LH_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(LH_E_HIP, "host reconstruction timing: event creation failed");
  }
  double total = 0.0;
  int rc       = LH_OK;
  for (int rep = 0; rep < reps && rc == LH_OK; ++rep) {
    // each repetition starts on idle streams, at its first copy
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c->upStream);
    if (e == hipSuccess)
      e = hipEventRecord(e0, c->upStream);
    if (e != hipSuccess) {
      rc = fail(LH_E_HIP, "host reconstruction timing: %s", hipGetErrorString(e));
      break;
    }
    rc = reconstruct_host(c, images, halves, exps, matrices, ids, count);
    if (rc != LH_OK)
      break;
    float ms = 0.0f;
    e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess)
      e = hipEventSynchronize(e1);
    if (e == hipSuccess)
      e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      rc = fail(LH_E_HIP, "host reconstruction timing: %s", hipGetErrorString(e));
    total += ms;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == LH_OK && msPerRep)
    *msPerRep = (float)(total / reps);
  return rc;
}

int lhTimedReconstructHost(LhContext *c, const float *images, const double *matrices,
    const int *ids, uint32_t count, int reps, float *msPerRep)
{
  int rc = host_args_check(c, images, matrices, ids, count, "lhTimedReconstructHost");
  if (rc == LH_OK && reps < 1)
    rc = fail(LH_E_ARG, "lhTimedReconstructHost: bad arguments");
  return rc != LH_OK ? rc
                     : timed_reconstruct(c, images, nullptr, nullptr, matrices, ids, count, reps,
                           msPerRep);
}

int lhTimedReconstructHostHalf(LhContext *c, const uint16_t *halves, const int32_t *exps,
    const double *matrices, const int *ids, uint32_t count, int reps, float *msPerRep)
{
  int rc = half_args_check(c, halves, exps, matrices, ids, count, "lhTimedReconstructHostHalf");
  if (rc == LH_OK && reps < 1)
    rc = fail(LH_E_ARG, "lhTimedReconstructHostHalf: bad arguments");
  return rc != LH_OK ? rc
                     : timed_reconstruct(c, nullptr, halves, exps, matrices, ids, count, reps,
                           msPerRep);
}

// The H2D copies of reconstruct_host alone (elem bytes per pixel).
static int timed_upload(LhContext *c, const void *images, size_t elem, uint32_t count, int reps,
    float *msPerRep)
{
  const LhConfig &g    = c->cfg;
  const uint32_t batch = std::min(g.nSlots,
      g.kernel == LH_KERNEL_FUSED ? std::max(1u, g.viewsPerPass) : 16u);
  const size_t px = (size_t)g.W * g.H;
  LH_HIP(hipSetDevice(g.device));
  {
    const int rc = host_ring(c, batch);
    if (rc != LH_OK)
      return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LH_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(LH_E_HIP, "host upload timing: event creation failed");
  }
  hipError_t e = hipSuccess;
  double total = 0.0;
  const std::vector<uint32_t> sizes = host_batches(count, batch);
  for (int rep = 0; rep < reps && e == hipSuccess; ++rep) {
    // the ring is overwritten: the compute stream, which may read it, drains first
    for (hipStream_t st : { c->stream, c->upStream })
      if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e == hipSuccess)
      e = hipEventRecord(e0, c->upStream);
    uint32_t v0 = 0;
    for (size_t k = 0; k < sizes.size() && e == hipSuccess; v0 += sizes[k], ++k)
      e = hipMemcpyAsync(c->dHostRing + (k & 1) * (size_t)batch * px,
          (const uint8_t *)images + (size_t)v0 * px * elem, (size_t)sizes[k] * px * elem,
          hipMemcpyHostToDevice, c->upStream);
    if (e == hipSuccess)
      e = hipEventRecord(e1, c->upStream);
    if (e == hipSuccess)
      e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess)
      e = hipEventElapsedTime(&ms, e0, e1);
    total += ms;
  }
  (void)hipStreamSynchronize(c->upStream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess)
    return fail(LH_E_HIP, "host upload timing: %s", hipGetErrorString(e));
  if (msPerRep)
    *msPerRep = (float)(total / reps);
  return LH_OK;
}

int lhTimedUploadHost(LhContext *c, const float *images, uint32_t count, int reps,
    float *msPerRep)
{
  if (!c || !images || count == 0 || reps < 1)
    return fail(LH_E_ARG, "lhTimedUploadHost: bad arguments");
  return timed_upload(c, images, sizeof(float), count, reps, msPerRep);
}

int lhTimedUploadHostHalf(LhContext *c, const uint16_t *halves, uint32_t count, int reps,
    float *msPerRep)
{
  if (!c || !halves || count == 0 || reps < 1)
    return fail(LH_E_ARG, "lhTimedUploadHostHalf: bad arguments");
  return timed_upload(c, halves, sizeof(uint16_t), count, reps, msPerRep);
}

int lhBackprojectNoAlloc(LhContext *c, uint32_t first, uint32_t count)
{
  if (!c)
    return fail(LH_E_ARG, "lhBackprojectNoAlloc: null context");
  const LhConfig &g = c->cfg;
  if (count == 0)
    return LH_OK;
  if (first + count > g.nSlots)
    return fail(LH_E_ARG, "lhBackprojectNoAlloc: slots [%u,%u) beyond %u", first, first + count,
        g.nSlots);
  for (uint32_t s = first; s < first + count; ++s)
    if (c->slotId[s] == 0)
      return fail(LH_E_STATE, "lhBackprojectNoAlloc: slot %u holds no view", s);
  if (!c->scratchBoxes)
    return fail(LH_E_STATE, "lhBackprojectNoAlloc: scratch table not resolved");
  const uint32_t per   = g.kernel == LH_KERNEL_FUSED ? std::max(1u, g.viewsPerPass) : 1u;
  const int R          = (int)g.runLength;
  const TileDims td    = tile_dims(c);
  const int tilesX     = (int)g.L / R, tilesY = (int)g.L / td.ly, tilesZ = (int)g.L / td.lz;
  const uint64_t tiles = (uint64_t)tilesX * tilesY * tilesZ;
  const uint64_t need  = tiles * std::min(per, count);
  if (need > kScratchBoxes)
    return fail(LH_E_ARG,
        "lhBackprojectNoAlloc: %u views per pass x %llu tiles need %llu footprint boxes; the "
        "scratch table holds %llu (32 MiB): use fewer views per pass",
        std::min(per, count), (unsigned long long)tiles, (unsigned long long)need,
        (unsigned long long)kScratchBoxes);
  LH_HIP(hipSetDevice(g.device));
  // per pass: its boxes into the scratch table, then its launch (stream
  // order keeps the next pass's boxes behind this pass's reads)
  for (uint32_t s = first; s < first + count; s += per) {
    const uint32_t n = std::min(per, first + count - s);
    box_kernel<<<dim3((unsigned)((tiles + 255) / 256), n), 256, 0, c->stream>>>(c->dViews + s, 0,
        (int)tiles, tilesX, tilesY, R, td.ly, td.lz, g.O_L, g.R_L, (int)g.W, (int)g.H, (int)n,
        c->scratchBoxes);
    LH_HIP(hipGetLastError());
    const int rc = launch_bp(c, s, n, false, c->scratchBoxes);
    if (rc != LH_OK)
      return rc;
  }
  return LH_OK;
}

int lhSetGraphs(LhContext *c, int enable)
{
  if (!c)
    return fail(LH_E_ARG, "lhSetGraphs: null context");
  c->graphs = enable != 0;
  return LH_OK;
}

int lhTimedBackproject(LhContext *c, uint32_t count, int reps, float *msPerRep,
    float *msPerLaunch, uint32_t *launchesPerRep)
{
  if (!c || reps < 1 || count == 0 || count > c->cfg.nSlots)
    return fail(LH_E_ARG, "lhTimedBackproject: bad arguments");
  const LhConfig &g = c->cfg;
  for (uint32_t s = 0; s < count; ++s)
    if (c->slotId[s] == 0)
      return fail(LH_E_STATE, "lhTimedBackproject: slot %u holds no view", s);
  const uint32_t per = g.kernel == LH_KERNEL_FUSED ? std::max(1u, g.viewsPerPass) : 1u;
  const uint32_t nl  = (count + per - 1) / per;
  LH_HIP(hipSetDevice(g.device));
  std::vector<hipEvent_t> ev(nl + 1);
  for (auto &e : ev)
    LH_HIP(hipEventCreate(&e));
  double total = 0.0, launchSum = 0.0;
  for (int rep = 0; rep < reps; ++rep) {
    LH_HIP(hipEventRecord(ev[0], c->stream));
    for (uint32_t k = 0; k < nl; ++k) {
      const uint32_t s = k * per;
      const int rc     = launch_bp(c, s, std::min(per, count - s));
      if (rc != LH_OK)
        return rc;
      LH_HIP(hipEventRecord(ev[k + 1], c->stream));
    }
    LH_HIP(hipEventSynchronize(ev[nl]));
    float ms = 0.0f;
    LH_HIP(hipEventElapsedTime(&ms, ev[0], ev[nl]));
    total += ms;
    for (uint32_t k = 0; k < nl; ++k) {
      float m = 0.0f;
      LH_HIP(hipEventElapsedTime(&m, ev[k], ev[k + 1]));
      launchSum += m;
    }
  }
  for (auto &e : ev)
    (void)hipEventDestroy(e);
  if (msPerRep)
    *msPerRep = (float)(total / reps);
  if (msPerLaunch)
    *msPerLaunch = (float)(launchSum / ((double)reps * nl));
  if (launchesPerRep)
    *launchesPerRep = nl;
  return LH_OK;
}

int lhSetLaunch(LhContext *c, int kernel, uint32_t viewsPerPass, uint32_t runLength,
    int staging)
{
  if (!c || (kernel != LH_KERNEL_PLAIN && kernel != LH_KERNEL_FUSED) || viewsPerPass == 0 ||
      (runLength != 8 && runLength != 16 && runLength != 32) || c->cfg.L % runLength)
    return fail(LH_E_ARG, "lhSetLaunch: bad launch shape");
  c->cfg.kernel       = kernel;
  c->cfg.viewsPerPass = viewsPerPass;
  c->cfg.runLength    = runLength;
  c->cfg.staging      = staging ? 1 : 0;
  return LH_OK;
}

int lhSetTile(LhContext *c, uint32_t tileY)
{
  if (!c || (tileY != 0 && tileY != 16 && tileY != 32 && tileY != 64))
    return fail(LH_E_ARG, "lhSetTile: tile y extent must be 16, 32 or 64 (0: automatic)");
  if (tileY == 16 && (c->slabZ0 % 16 || c->slabZ1 % 16))
    return fail(LH_E_ARG, "lhSetTile: 16 x 16 tiles need a slab of multiples of 16");
  c->tileY = tileY;
  return LH_OK;
}

int lhSetSlab(LhContext *c, uint32_t z0, uint32_t z1)
{
  if (!c || z0 >= z1 || z1 > c->cfg.L || z0 % 8 || z1 % 8)
    return fail(LH_E_ARG, "lhSetSlab: need 0 <= z0 < z1 <= L, multiples of 8 (got %u, %u)", z0,
        z1);
  if (z0 != c->slabZ0 || z1 != c->slabZ1)
    drop_graphs(c); // captured launches hold the old tile range
  c->slabZ0 = z0;
  c->slabZ1 = z1;
  return LH_OK;
}

int lhCopyCalibration(LhContext *c, uint64_t bytes, float *ms)
{
  if (!c || bytes == 0)
    return fail(LH_E_ARG, "lhCopyCalibration: bad arguments");
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipMalloc(&a, bytes));
  LH_HIP(hipMalloc(&b, bytes));
  LH_HIP(hipMemsetAsync(a, 1, bytes, c->stream));
  LH_HIP(hipEventCreate(&e0));
  LH_HIP(hipEventCreate(&e1));
  LH_HIP(hipEventRecord(e0, c->stream));
  LH_HIP(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, c->stream));
  LH_HIP(hipEventRecord(e1, c->stream));
  LH_HIP(hipEventSynchronize(e1));
  float t = 0.0f;
  LH_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms)
    *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return LH_OK;
}

int lhSynchronize(LhContext *c)
{
  if (!c)
    return fail(LH_E_ARG, "lhSynchronize: null context");
  LH_HIP(hipSetDevice(c->cfg.device));
  LH_HIP(hipStreamSynchronize(c->stream));
  return LH_OK;
}

uint64_t lhEffectiveUpdates(LhContext *c, uint32_t count)
{
  if (!c)
    return 0;
  uint64_t s = 0;
  for (uint32_t k = 0; k < count && k < c->cfg.nSlots; ++k) {
    const int id = c->slotId[k];
    if (id >= 1)
      s += c->effPerView[id - 1];
  }
  return s;
}

} // extern "C"
