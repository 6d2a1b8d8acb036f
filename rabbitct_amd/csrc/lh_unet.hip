// lh_unet.hip -- 3-level 3D U-Net for the sparse-view correction (no
// counterpart in the reference; it learns the map from a 124-view to a
// 496-view backprojection).
//
// Tensors are channels-last BF16, [voxel][ld] with a channel slice at an
// offset, so the two concatenations of the decoder are channel slices of one
// buffer and cost no copy.  Every 3x3x3 convolution (stride 1, zero padding 1)
// is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   M = 16 output voxels per wave (consecutive x), N = output channels (16 per
//   MFMA tile), K = 27 taps x input channels.
// K is packed in 32-wide chunks so that each chunk reads one wave-uniform tap
// (or one of two taps, chosen by a lane bit):
//   section 1: for each tap, for each full block of 32 input channels;
//   section 2 (input channels = 16 mod 32): the last 16 channels of two
//   taps of one z offset, (dz, e) and (dz, e+1) with e = 3 dy + dx even,
//   share one chunk: 5 chunks per dz (the tenth tap of each is zero).
// Lane l holds A rows (voxel l%16) and B columns (channel l%16) for the 8
// K entries 8*(l/16)..+7 of a chunk: 8 consecutive channels of one tap, one
// 16-byte load each.  The wide-output forms (more than 32 output channels)
// pass the same two operands in the other order, so that their results
// come out voxel by voxel with four consecutive channels per lane.  The backward-data pass is the same kernel with the taps
// mirrored and input/output channels swapped in the weight packing.  The
// weight gradient is a separate MFMA kernel that reduces over voxels (K)
// from transposed rows staged in LDS.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "lolahip.h"
#include "lh_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// raw buffer resource word 3 for gfx950 (32-bit untyped access); a load
// whose offset lies past num_records returns zero
constexpr int kRsrcWord3 = 0x00020000;
constexpr uint32_t kMissOffset = 0x80000000u;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// FP32 -> bfloat16, round to nearest even (one v_cvt_pk_bf16_f32; two
// values share one when the compiler pairs them)
__device__ __forceinline__ uint16_t f2bf(float f)
{
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

// Number of 32-wide K chunks and the chunk/entry of (tap, ci) for C input
// channels (C a multiple of 16).
__host__ __device__ __forceinline__ int kchunks(int C) { return 27 * (C / 32) + ((C % 32) ? 15 : 0); }
__host__ __device__ __forceinline__ int kpos(int C, int tap, int ci)
{
  const int F = C / 32;
  if (ci < 32 * F)
    return (tap * F + ci / 32) * 32 + ci % 32;
  const int dz = tap / 9, e = tap % 9;  // e = 3 dy + dx
  return (27 * F + 5 * dz + e / 2) * 32 + (e & 1) * 16 + (ci - 32 * F);
}

// ---------------------------------------------------------------------------
// LDS-staged form.  A block computes a 16 (x) x 4 (y) x MT (z) tile of output
// voxels; wave w owns the MT x-rows at y = w.  The input halo
// 18 x 6 x (MT+2) voxels x CC channels is staged in LDS once per pass of CC
// input channels (16-byte loads, zeros outside the volume), and every A
// operand is one ds_read_b128 of 16 consecutive voxels x 8 channels.  A
// weight operand (global, L1/L2-resident) feeds MT MFMAs.
//
// Input channels [0, c0) may come from a tensor at half resolution (the
// nearest-neighbour upsampling of the decoder, read in place), channels
// [c0, cin) from a full-resolution tensor: the concatenation is never
// materialised.
struct ConvSrc {
  const uint16_t *p;  // tensor base
  int ld, ofs;        // voxel stride and channel offset (elements)
  int half;           // 1: the tensor has edge D/2 and is read at (x/2, y/2, z/2)
  int bytes;          // addressable bytes from p (buffer range, < 2^31)
};

struct ConvLdsArgs {
  ConvSrc s0, s1;     // channels [0, c0) from s0, [c0, cin) from s1
  int c0;
  const uint16_t *w;  // packed weights [cout][kp]
  const float *bias;
  uint16_t *y;
  int ldy, yofs;
  int D, n, cin, kp;
  int relu, accumulate;
  int tilesPerXcd;    // > 0: tiles split into 8 contiguous ranges, one per XCD
  // fused 1x1x1 head (16 output channels only): hout[v] = hin[v] + hb +
  // sum_c hw[c] y[v][c]; y itself is written only when non-null
  const float *hin, *hw, *hb;
  float *hout;
  // fused 2x2x2 average pooling of the output (ReLU'd, bfloat16) into a
  // half-resolution tensor [voxel][ldp] at channel offset pofs; null: none
  uint16_t *pool;
  int ldp, pofs;
};

constexpr int kHX = 18, kHY = 6;

// Stages the halo of input channels [cbase, cbase + CC) (CC = 32 or 16) as
// [hz][hy][hx][CC] bfloat16.  Each thread keeps kStageBatch 16-byte loads in
// flight before writing them to LDS.
#ifndef LH_STAGE_BATCH
#define LH_STAGE_BATCH 8
#endif
// 1: passes whose channels come from one tensor are staged by direct-to-LDS
// buffer loads
#ifndef LH_STAGE_DMA
#define LH_STAGE_DMA 1
#endif
constexpr int kStageBatch = LH_STAGE_BATCH;

template <int CC, int MT>
__device__ __forceinline__ void stage_halo(const ConvLdsArgs &a, uint16_t *lds, int cbase, int n,
    int x0, int y0, int z0)
{
  constexpr int Q     = CC / 8;                // 16-byte units per voxel
  constexpr int HV    = kHX * kHY * (MT + 2);  // halo voxels
  constexpr int TOTAL = HV * Q;
  constexpr int ITER  = (TOTAL + 255) / 256;
  const int D = a.D;
  // A pass whose channels all come from one tensor (the usual case: the
  // decoder's splits fall on pass boundaries) selects that tensor once, so
  // the buffer descriptor is wave-uniform; a pass that straddles the split
  // selects per 8-channel unit.
  const bool straddle = cbase < a.c0 && cbase + CC > a.c0;
  if (!straddle) {
    const bool first  = cbase < a.c0;
    const ConvSrc &sr = first ? a.s0 : a.s1;
    const int half    = sr.half;
    const int Ds      = half ? (D >> 1) : D;
    const uint32_t ld = (uint32_t)sr.ld;
    const uint32_t c0 = (uint32_t)(sr.ofs + (first ? cbase : cbase - a.c0));
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)sr.p, (short)0, sr.bytes, kRsrcWord3);
#if LH_STAGE_DMA
    // direct-to-LDS: unit i (voxel i / Q, channels 8 (i % Q) ..) lands 16
    // bytes at lds + 8 i, every unit of the thread in flight at once and no
    // registers held; units outside the volume read past the buffer range
    // and land zeros.  Returns once this wave's copies have landed.
    const int wbase = threadIdx.x & ~63;
#pragma unroll 1
    for (int k = 0; k < ITER; ++k) {
      const int i = threadIdx.x + k * 256;
      if (i < TOTAL) {
        const int q = i % Q, v = i / Q;
        const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        const bool in = (unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D &&
                        (unsigned)gz < (unsigned)D;
        const int sx = half ? (gx >> 1) : gx, sy = half ? (gy >> 1) : gy,
                  sz = half ? (gz >> 1) : gz;
        const uint32_t off =
            in ? ((((uint32_t)n * Ds + sz) * Ds + sy) * Ds + sx) * ld * 2u + (c0 + 8 * q) * 2u
               : kMissOffset;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
            (__attribute__((address_space(3))) void *)(lds + 8 * (wbase + k * 256)), 16,
            (int)off, 0, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    return;
#endif
#pragma unroll 1
    for (int b0 = 0; b0 < ITER; b0 += kStageBatch) {
      u32x4 val[kStageBatch];
#pragma unroll
      for (int k = 0; k < kStageBatch; ++k) {
        const int i = threadIdx.x + (b0 + k) * 256;
        if (b0 + k < ITER && i < TOTAL) {
          const int q = i % Q, v = i / Q;
          const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
          const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
          const bool in = (unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D &&
                          (unsigned)gz < (unsigned)D;
          const int sx = half ? (gx >> 1) : gx, sy = half ? (gy >> 1) : gy,
                    sz = half ? (gz >> 1) : gz;
          const uint32_t off =
              in ? ((((uint32_t)n * Ds + sz) * Ds + sy) * Ds + sx) * ld * 2u + (c0 + 8 * q) * 2u
                 : kMissOffset;
          val[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        }
      }
#pragma unroll
      for (int k = 0; k < kStageBatch; ++k) {
        const int i = threadIdx.x + (b0 + k) * 256;
        if (b0 + k < ITER && i < TOTAL)
          *reinterpret_cast<u32x4 *>(lds + (i / Q) * CC + 8 * (i % Q)) = val[k];
      }
    }
    return;
  }
#pragma unroll 1
  for (int b0 = 0; b0 < ITER; b0 += kStageBatch) {
    u32x4 val[kStageBatch];
#pragma unroll
    for (int k = 0; k < kStageBatch; ++k) {
      const int i = threadIdx.x + (b0 + k) * 256;
      if (b0 + k < ITER && i < TOTAL) {
        const int q = i % Q, v = i / Q;
        const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        const bool in = (unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D &&
                        (unsigned)gz < (unsigned)D;
        const int c = cbase + 8 * q;
        const bool first = c < a.c0;
        const ConvSrc &sr = first ? a.s0 : a.s1;
        const int cc = first ? c : c - a.c0;
        const int Ds = sr.half ? (D >> 1) : D;
        const int sx = sr.half ? (gx >> 1) : gx, sy = sr.half ? (gy >> 1) : gy,
                  sz = sr.half ? (gz >> 1) : gz;
        const uint32_t off = in ? (uint32_t)(((((uint32_t)n * Ds + sz) * Ds + sy) * Ds + sx) *
                                             (uint32_t)sr.ld + sr.ofs + cc) * 2u
                                : kMissOffset;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)sr.p, (short)0, sr.bytes, kRsrcWord3);
        val[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
      }
    }
#pragma unroll
    for (int k = 0; k < kStageBatch; ++k) {
      const int i = threadIdx.x + (b0 + k) * 256;
      if (b0 + k < ITER && i < TOTAL)
        *reinterpret_cast<u32x4 *>(lds + (i / Q) * CC + 8 * (i % Q)) = val[k];
    }
  }
}

// Coarse form for the upsampled channels [cbase, cbase + 32) of a decoder
// convolution (a.s0 at half resolution): the fine halo [x0-1, x0+16] x
// [y0-1, y0+4] x [z0-1, z0+MT] reads coarse voxels [x0/2-1, x0/2+8] x
// [y0/2-1, y0/2+2] x [z0/2-1, z0/2+MT/2], staged once as
// [MT/2 + 2][kCHY][kCHX][32] (zeros outside the coarse volume, which is where
// the fine halo leaves the volume) instead of 8 replicated copies.
#ifndef LH_UPC
#define LH_UPC 1
#endif
constexpr int kCHX = 10, kCHY = 4;
template <int MT>
__device__ __forceinline__ void stage_coarse(const ConvLdsArgs &a, uint16_t *lds, int cbase, int n,
    int x0, int y0, int z0)
{
  constexpr int CHZ   = MT / 2 + 2;
  constexpr int TOTAL = kCHX * kCHY * CHZ * 4;  // 16-byte units (32 channels per voxel)
  constexpr int ITER  = (TOTAL + 255) / 256;
  const int Dc = a.D >> 1;
  const ConvSrc &sr = a.s0;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)sr.p, (short)0, sr.bytes, kRsrcWord3);
#if LH_STAGE_DMA
  const int wbase = threadIdx.x & ~63;   // as in stage_halo: unit i lands at lds + 8 i
#endif
  u32x4 val[ITER];
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int i = threadIdx.x + k * 256;
    if (i < TOTAL) {
      const int q = i % 4, v = i / 4;
      const int cx = (x0 >> 1) - 1 + v % kCHX, cy = (y0 >> 1) - 1 + (v / kCHX) % kCHY,
                cz = (z0 >> 1) - 1 + v / (kCHX * kCHY);
      const bool in = (unsigned)cx < (unsigned)Dc && (unsigned)cy < (unsigned)Dc &&
                      (unsigned)cz < (unsigned)Dc;
      const uint32_t off = in ? (uint32_t)(((((uint32_t)n * Dc + cz) * Dc + cy) * Dc + cx) *
                                           (uint32_t)sr.ld + sr.ofs + cbase + 8 * q) * 2u
                              : kMissOffset;
#if LH_STAGE_DMA
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(lds + 8 * (wbase + k * 256)), 16,
          (int)off, 0, 0, 0);
#else
      val[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
#endif
    }
  }
#if LH_STAGE_DMA
  __builtin_amdgcn_s_waitcnt(0);
  return;
#endif
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int i = threadIdx.x + k * 256;
    if (i < TOTAL)
      *reinterpret_cast<u32x4 *>(lds + (i / 4) * 32 + 8 * (i % 4)) = val[k];
  }
}

template <int NT>
__device__ __forceinline__ void load_b(bf16x8 (&B)[NT], const uint16_t *wrow, size_t wtile,
    int chunk)
{
#pragma unroll
  for (int t = 0; t < NT; ++t)
    B[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(wrow + t * wtile + chunk * 32));
}

// LDS bytes of a launch: the larger of the 32-channel halo and the output
// transpose buffer (4 waves x MT rows x 16 voxels x NT*16 channels).
template <int NT, int MT> constexpr size_t conv_lds_bytes()
{
  const size_t halo = (size_t)(MT + 2) * kHY * kHX * 32 * 2;
  const size_t outb = (size_t)4 * MT * 16 * NT * 16 * 2;
  return halo > outb ? halo : outb;
}

// The 16-channel pass on a staged halo [MT+2][kHY][kHX][16] (K chunks
// k0 + 5 dz + j: lanes kg < 2 hold tap (dz, 2j), kg >= 2 tap (dz, 2j+1), each
// lane 8 of the 16 channels), input planes outer: the A operand of halo plane
// p at (dy, dx) serves output rows p - dz for dz = 0, 1, 2.
// PRE: the first chunks' weight operands (k0 + 5 dz) come loaded in B0.
template <int NT, int MT, bool PRE = false>
__device__ __forceinline__ void tail16_mfma(f32x4 (&acc)[MT][NT], const uint16_t *lds,
    const uint16_t *wrow, size_t wtile, int k0, int w, int m, int kg,
    const bf16x8 (&B0)[PRE ? 3 : 1][NT])
{
  constexpr int RS16 = kHX * kHY * 16;
  const int half = kg >> 1, cofs = 8 * (kg & 1);
  bf16x8 Bn[3][NT];
#pragma unroll
  for (int dz = 0; dz < 3; ++dz)
#pragma unroll
    for (int t = 0; t < NT; ++t)
      if constexpr (PRE)
        Bn[dz][t] = B0[dz][t];
  if constexpr (!PRE)
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
      load_b(Bn[dz], wrow, wtile, k0 + 5 * dz);
  for (int j = 0; j < 5; ++j) {
    bf16x8 B[3][NT];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        B[dz][t] = Bn[dz][t];
    if (j + 1 < 5)
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
        load_b(Bn[dz], wrow, wtile, k0 + 5 * dz + j + 1);
    // e = 9 is the pad: it reads tap 8's operand, whose products meet the
    // pad's weights, zero in the packing (finite activations)
    const int e  = 2 * j + half;
    const int ee = e >= 9 ? 8 : e;
    const uint16_t *base = lds + ((w + ee / 3) * kHX + m + ee % 3) * 16 + cofs;
#pragma unroll
    for (int p = 0; p < MT + 2; ++p) {
      const bf16x8 A =
          __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + p * RS16));
#pragma unroll
      for (int dz = 0; dz < 3; ++dz) {
        const int r = p - dz;
        if (r >= 0 && r < MT)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[dz][t], acc[r][t], 0, 0, 0);
      }
    }
  }
}

// tail16_mfma with the 15 K chunks' weight operands held in registers
// (W[5 dz + j]), loaded once by a persistent kernel for all its tiles: no
// weight load is issued between a tile's prefetch of the next halo and its
// MFMAs, so waiting for a weight operand never waits for that prefetch.  The
// same MFMAs in the same order as tail16_mfma.
template <int NT, int MT, int NW>
__device__ __forceinline__ void tail16_mfma_reg(f32x4 (&acc)[MT][NT], const uint16_t *lds,
    const bf16x8 (&W)[NW][NT], int w, int m, int kg)
{
  constexpr int RS16 = kHX * kHY * 16;
  const int half = kg >> 1, cofs = 8 * (kg & 1);
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int e  = 2 * j + half;
    const int ee = e >= 9 ? 8 : e;   // the pad: zero weights, as in tail16_mfma
    const uint16_t *base = lds + ((w + ee / 3) * kHX + m + ee % 3) * 16 + cofs;
#pragma unroll
    for (int p = 0; p < MT + 2; ++p) {
      const bf16x8 A =
          __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + p * RS16));
#pragma unroll
      for (int dz = 0; dz < 3; ++dz) {
        const int r = p - dz;
        if (r >= 0 && r < MT)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, W[5 * dz + j][t], acc[r][t],
                0, 0, 0);
      }
    }
  }
}

// Epilogue of a convolution tile: bias, ReLU, bfloat16 into an LDS transpose
// buffer (this wave's rows: [MT][16 voxels][COUT]), then 16-byte stores of 8
// channels; the fused pooling and 1x1x1 head read the same buffer.  Begins
// with a barrier: every wave's reads of the tile's input from lds are done.
// The epilogue's own loads, which a persistent kernel issues before its
// prefetch of the next tile (loads return in order: one issued after the
// prefetch would make its use wait for it); the one-tile kernels load them
// in the epilogue (PRE = false).  bias[t]: this lane's output
// channel 16 t + m; hin[k]: the fused head's input at this wave's voxel
// 64 k + lane of the tile.
template <int MT> constexpr int head_per_lane() { return (MT * 16 + 63) / 64; }

template <int NT>
__device__ __forceinline__ void load_bias(const ConvLdsArgs &a, int m, float (&bias)[NT])
{
#pragma unroll
  for (int t = 0; t < NT; ++t)
    bias[t] = a.bias ? a.bias[16 * t + m] : 0.0f;
}

template <int NT, int MT>
__device__ __forceinline__ void load_hin(const ConvLdsArgs &a, int n, int x0, int y0, int z0,
    int w, int lane, float (&hin)[head_per_lane<MT>()])
{
#pragma unroll
  for (int k = 0; k < head_per_lane<MT>(); ++k)
    hin[k] = 0.0f;
  if constexpr (NT == 1) {
    if (a.hout) {
      const int D = a.D;
#pragma unroll
      for (int k = 0; k < head_per_lane<MT>(); ++k) {
        const int vv = 64 * k + lane;
        if ((MT * 16) % 64 == 0 || vv < MT * 16) {
          const size_t vo = (((size_t)n * D + z0 + vv / 16) * D + y0 + w) * D + x0 + vv % 16;
          hin[k] = a.hin[vo];
        }
      }
    }
  }
}

template <int NT, int MT, bool PRE = false>
__device__ __forceinline__ void conv_epilogue(const ConvLdsArgs &a, f32x4 (&acc)[MT][NT],
    uint16_t *lds, int n, int x0, int y0, int z0, int w, int lane, int m, int kg,
    const float (&biasv)[NT], const float (&hin)[head_per_lane<MT>()])
{
  const int D = a.D;
  constexpr int COUT = NT * 16;
  __syncthreads();
  uint16_t *ob = lds + (size_t)w * MT * 16 * COUT;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = 16 * t + m;
    const float bias = PRE ? biasv[t] : (a.bias ? a.bias[co] : 0.0f);
#pragma unroll
    for (int r = 0; r < MT; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float val = acc[r][t][i] + bias;
        if (a.relu)
          val = fmaxf(val, 0.0f);
        ob[(r * 16 + 4 * kg + i) * COUT + co] = f2bf(val);
      }
  }
  __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0): this wave's LDS writes are done
  __builtin_amdgcn_wave_barrier();
  constexpr int UNITS = MT * 16 * COUT / 8;  // 16-byte units of this wave
  if (a.y) {
#pragma unroll
    for (int u0 = 0; u0 < UNITS; u0 += 64) {
      const int u = u0 + lane;
      if (UNITS % 64 == 0 || u < UNITS) {
        const int c8 = u % (COUT / 8), vx = (u / (COUT / 8)) % 16, r = u / (COUT / 8 * 16);
        uint4 val = *reinterpret_cast<const uint4 *>(ob + (r * 16 + vx) * COUT + 8 * c8);
        const size_t vo = (((size_t)n * D + z0 + r) * D + y0 + w) * D + x0 + vx;
        uint4 *dst = reinterpret_cast<uint4 *>(a.y + vo * a.ldy + a.yofs + 8 * c8);
        if (a.accumulate) {
          const uint4 old = *dst;
          const uint32_t nv[4] = { val.x, val.y, val.z, val.w };
          const uint32_t ov[4] = { old.x, old.y, old.z, old.w };
          uint32_t rv[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            rv[k] = (uint32_t)f2bf(__uint_as_float(nv[k] << 16) + __uint_as_float(ov[k] << 16)) |
                    ((uint32_t)f2bf(__uint_as_float(nv[k] & 0xffff0000u) +
                                    __uint_as_float(ov[k] & 0xffff0000u)) << 16);
          val = make_uint4(rv[0], rv[1], rv[2], rv[3]);
        }
        *dst = val;
      }
    }
  }
  if (a.pool) {
    // the tile (16 x 4 x MT, all even) holds whole 2x2x2 cells; wave w's rows
    // are at y0 + w, so a cell reads the buffers of waves 2Y and 2Y + 1
    __syncthreads();
    constexpr int C8 = COUT / 8;
    constexpr int CELLS = 8 * 2 * (MT / 2);     // x cells, y cells, z cells
    const int Dh = D >> 1;
    for (int u = threadIdx.x; u < CELLS * C8; u += 256) {
      const int c8 = u % C8, cell = u / C8;
      const int cx = cell % 8, cy = (cell / 8) % 2, cz = cell / 16;
      float sum[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int zz = 2 * cz + (q >> 2), yy = 2 * cy + ((q >> 1) & 1), xx = 2 * cx + (q & 1);
        const uint4 v = *reinterpret_cast<const uint4 *>(lds + (size_t)yy * MT * 16 * COUT +
                                                        (zz * 16 + xx) * COUT + 8 * c8);
        const uint32_t w4[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          sum[2 * k] += __uint_as_float(w4[k] << 16);
          sum[2 * k + 1] += __uint_as_float(w4[k] & 0xffff0000u);
        }
      }
      uint32_t o[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        o[k] = (uint32_t)f2bf(0.125f * sum[2 * k]) | ((uint32_t)f2bf(0.125f * sum[2 * k + 1]) << 16);
      const size_t vo = ((((size_t)n * Dh + (z0 >> 1) + cz) * Dh + (y0 >> 1) + cy) * Dh) +
                        (x0 >> 1) + cx;
      *reinterpret_cast<uint4 *>(a.pool + vo * a.ldp + a.pofs + 8 * c8) = make_uint4(o[0], o[1], o[2], o[3]);
    }
  }
  if constexpr (COUT == 16) {
    if (a.hout) {
      // one voxel per lane: its 16 channels (bfloat16) dotted with the head
#pragma unroll
      for (int v0 = 0; v0 < MT * 16; v0 += 64) {
        const int vv = v0 + lane;
        if ((MT * 16) % 64 == 0 || vv < MT * 16) {
          const int vx = vv % 16, r = vv / 16;
          const uint4 *hv = reinterpret_cast<const uint4 *>(ob + (r * 16 + vx) * 16);
          const size_t vo = (((size_t)n * D + z0 + r) * D + y0 + w) * D + x0 + vx;
          float sum = (PRE ? hin[v0 / 64] : a.hin[vo]) + a.hb[0];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const uint4 hq = hv[q];
            const uint32_t hw4[4] = { hq.x, hq.y, hq.z, hq.w };
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              sum = fmaf(a.hw[8 * q + 2 * k], __uint_as_float(hw4[k] << 16), sum);
              sum = fmaf(a.hw[8 * q + 2 * k + 1], __uint_as_float(hw4[k] & 0xffff0000u), sum);
            }
          }
          a.hout[vo] = sum;
        }
      }
    }
  }
}

// Epilogue of the wide-output forms (NT > 2), whose MFMAs take the weights
// as the A operand and the activations as the B operand: lane (m, kg) of
// wave w holds acc[r][t][i] = output voxel (x0 + m, y0 + w, z0 + r),
// channel 16 t + 4 kg + i, four consecutive channels (8 bytes in bfloat16),
// stored from registers without the LDS transpose, which costs these forms
// NT x MT x 4 two-byte LDS writes per lane.  The fused pooling reads the
// same [MT][16 voxels][COUT] LDS buffer as in conv_epilogue (8-byte writes).
template <int NT, int MT>
__device__ __forceinline__ void conv_epilogue_wide(const ConvLdsArgs &a, f32x4 (&acc)[MT][NT],
    uint16_t *lds, int n, int x0, int y0, int z0, int w, int m, int kg)
{
  const int D = a.D;
  constexpr int COUT = NT * 16;
  uint2 h[MT][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    float bias[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      bias[i] = a.bias ? a.bias[16 * t + 4 * kg + i] : 0.0f;
#pragma unroll
    for (int r = 0; r < MT; ++r) {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = acc[r][t][i] + bias[i];
        if (a.relu)
          v[i] = fmaxf(v[i], 0.0f);
      }
      h[r][t] = make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                           (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
    }
  }
  if (a.y) {
#pragma unroll
    for (int r = 0; r < MT; ++r) {
      const size_t vo = (((size_t)n * D + z0 + r) * D + y0 + w) * D + x0 + m;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        uint2 *dst = reinterpret_cast<uint2 *>(a.y + vo * a.ldy + a.yofs + 16 * t + 4 * kg);
        uint2 val = h[r][t];
        if (a.accumulate) {
          const uint2 old = *dst;
          const uint32_t nv[2] = { val.x, val.y };
          const uint32_t ov[2] = { old.x, old.y };
          uint32_t rv[2];
#pragma unroll
          for (int k = 0; k < 2; ++k)
            rv[k] = (uint32_t)f2bf(__uint_as_float(nv[k] << 16) + __uint_as_float(ov[k] << 16)) |
                    ((uint32_t)f2bf(__uint_as_float(nv[k] & 0xffff0000u) +
                                    __uint_as_float(ov[k] & 0xffff0000u)) << 16);
          val = make_uint2(rv[0], rv[1]);
        }
        *dst = val;
      }
    }
  }
  if (a.pool) {
    // every wave is done reading the tile's input from lds
    __syncthreads();
    uint16_t *ob = lds + (size_t)w * MT * 16 * COUT;
#pragma unroll
    for (int r = 0; r < MT; ++r)
#pragma unroll
      for (int t = 0; t < NT; ++t)
        *reinterpret_cast<uint2 *>(ob + (r * 16 + m) * COUT + 16 * t + 4 * kg) = h[r][t];
    __syncthreads();
    constexpr int C8 = COUT / 8;
    constexpr int CELLS = 8 * 2 * (MT / 2);
    const int Dh = D >> 1;
    for (int u = threadIdx.x; u < CELLS * C8; u += 256) {
      const int c8 = u % C8, cell = u / C8;
      const int cx = cell % 8, cy = (cell / 8) % 2, cz = cell / 16;
      float sum[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int zz = 2 * cz + (q >> 2), yy = 2 * cy + ((q >> 1) & 1), xx = 2 * cx + (q & 1);
        const uint4 v = *reinterpret_cast<const uint4 *>(lds + (size_t)yy * MT * 16 * COUT +
                                                        (zz * 16 + xx) * COUT + 8 * c8);
        const uint32_t w4[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          sum[2 * k] += __uint_as_float(w4[k] << 16);
          sum[2 * k + 1] += __uint_as_float(w4[k] & 0xffff0000u);
        }
      }
      uint32_t o[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        o[k] = (uint32_t)f2bf(0.125f * sum[2 * k]) | ((uint32_t)f2bf(0.125f * sum[2 * k + 1]) << 16);
      const size_t vo = ((((size_t)n * Dh + (z0 >> 1) + cz) * Dh + (y0 >> 1) + cy) * Dh) +
                        (x0 >> 1) + cx;
      *reinterpret_cast<uint4 *>(a.pool + vo * a.ldp + a.pofs + 8 * c8) = make_uint4(o[0], o[1], o[2], o[3]);
    }
  }
}

// Input channels are staged in passes: F = cin/32 passes of 32 channels
// (one tap per K chunk), then with TAIL a pass of the last 16 channels (tap
// pairs per K chunk), matching the weight packing.
template <int NT, int MT, bool TAIL, bool UPC = false>
__global__ __launch_bounds__(256) void conv3_lds_kernel(ConvLdsArgs a)
{
  extern __shared__ __align__(16) uint16_t lds[];
  const int D = a.D;
  const int txN = D / 16, tyN = D / 4, tzN = D / MT;
  int tile = blockIdx.x;
  if (a.tilesPerXcd > 0)
    tile = (blockIdx.x & 7) * a.tilesPerXcd + (blockIdx.x >> 3);
  const int tx = tile % txN;
  int rest = tile / txN;
  const int ty = rest % tyN;
  rest /= tyN;
  const int tz = rest % tzN, n = rest / tzN;
  const int x0 = tx * 16, y0 = ty * 4, z0 = tz * MT;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const uint16_t *wrow = a.w + (size_t)m * a.kp + 8 * kg;
  const size_t wtile = (size_t)16 * a.kp;
  const int F = a.cin / 32;
  // with the upsampled channels staged coarse (dec1a: four workgroups per
  // CU, two short staging phases) the epilogue's bias is loaded ahead of the
  // staging, so that its latency is not exposed after the last MFMA
  // (measured 2-7% faster there; the other forms load it in the epilogue and
  // keep their register counts)
  float biasv[NT] = {};
  if constexpr (UPC)
    load_bias<NT>(a, m, biasv);
  f32x4 acc[MT][NT];
#pragma unroll
  for (int r = 0; r < MT; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t)
      acc[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  constexpr int RS32 = kHX * kHY * 32, RS16 = kHX * kHY * 16; // z-row strides
  for (int pass = 0; pass < F; ++pass) {
    if (pass)
      __syncthreads();
    // UPC: every 32-channel pass holds upsampled channels, staged coarse;
    // the pass's first weight operands are loaded before the staging, whose
    // wait then covers them
    bf16x8 Bw[UPC ? 3 : 1][NT];
    if constexpr (UPC) {
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
        load_b(Bw[dz], wrow, wtile, (9 * dz) * F + pass);
      stage_coarse<MT>(a, lds, 32 * pass, n, x0, y0, z0);
    } else
      stage_halo<32, MT>(a, lds, 32 * pass, n, x0, y0, z0);
    __syncthreads();
    if constexpr (UPC) {
      static_assert(NT <= 2, "coarse staging is built for the plane-reuse layout");
      // as below, with fine halo voxel (h) read at coarse halo voxel
      // ((h - 1) >> 1) + 1 per axis; pairs of lanes read the same voxel
      constexpr int CRS = kCHX * kCHY * 32;  // coarse z-plane stride
      bf16x8 Bn[3][NT];
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int t = 0; t < NT; ++t)
          Bn[dz][t] = Bw[dz][t];
      for (int dyx = 0; dyx < 9; ++dyx) {
        bf16x8 B[3][NT];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            B[dz][t] = Bn[dz][t];
        if (dyx + 1 < 9)
#pragma unroll
          for (int dz = 0; dz < 3; ++dz)
            load_b(Bn[dz], wrow, wtile, (9 * dz + dyx + 1) * F + pass);
        const int dy = dyx / 3, dx = dyx % 3;
        const int cx = ((m + dx - 1) >> 1) + 1, cy = ((w + dy - 1) >> 1) + 1;
        const uint16_t *base = lds + (cy * kCHX + cx) * 32 + 8 * kg;
#pragma unroll
        for (int p = 0; p < MT + 2; ++p) {
          const int cz = ((p - 1) >> 1) + 1;
          const bf16x8 A = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + cz * CRS));
#pragma unroll
          for (int dz = 0; dz < 3; ++dz) {
            const int r = p - dz;
            if (r >= 0 && r < MT)
#pragma unroll
              for (int t = 0; t < NT; ++t)
                acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[dz][t], acc[r][t], 0, 0, 0);
          }
        }
      }
      continue;
    }
    if constexpr (NT <= 2) {
      // Input planes outer: the A operand of halo plane p at (dy, dx) serves
      // output rows p - dz for dz = 0, 1, 2 (taps 9 dz + 3 dy + dx), so each
      // is read from LDS once instead of up to three times.
      bf16x8 Bn[3][NT];
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
        load_b(Bn[dz], wrow, wtile, (9 * dz) * F + pass);
      for (int dyx = 0; dyx < 9; ++dyx) {
        bf16x8 B[3][NT];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            B[dz][t] = Bn[dz][t];
        if (dyx + 1 < 9)
#pragma unroll
          for (int dz = 0; dz < 3; ++dz)
            load_b(Bn[dz], wrow, wtile, (9 * dz + dyx + 1) * F + pass);
        const int dy = dyx / 3, dx = dyx % 3;
        const uint16_t *base = lds + ((w + dy) * kHX + m + dx) * 32 + 8 * kg;
#pragma unroll
        for (int p = 0; p < MT + 2; ++p) {
          const bf16x8 A = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + p * RS32));
#pragma unroll
          for (int dz = 0; dz < 3; ++dz) {
            const int r = p - dz;
            if (r >= 0 && r < MT)
#pragma unroll
              for (int t = 0; t < NT; ++t)
                acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[dz][t], acc[r][t], 0, 0, 0);
          }
        }
      }
    } else {
      bf16x8 Bn[NT];
      load_b(Bn, wrow, wtile, pass);
      for (int tap = 0; tap < 27; ++tap) {
        bf16x8 B[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
          B[t] = Bn[t];
        if (tap + 1 < 27)
          load_b(Bn, wrow, wtile, (tap + 1) * F + pass);
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        const uint16_t *base = lds + ((dz * kHY + w + dy) * kHX + m + dx) * 32 + 8 * kg;
#pragma unroll
        for (int r = 0; r < MT; ++r) {
          const bf16x8 A = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + r * RS32));
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t], A, acc[r][t], 0, 0, 0);
        }
      }
    }
  }
  if constexpr (TAIL) {
    if (F)
      __syncthreads();
    // UPC: the tail's first weight operands loaded before its staging, as above
    bf16x8 Bt[UPC ? 3 : 1][NT];
    if constexpr (UPC)
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
        load_b(Bt[dz], wrow, wtile, 27 * F + 5 * dz);
    stage_halo<16, MT>(a, lds, 32 * F, n, x0, y0, z0);
    __syncthreads();
    // chunk 27F + 5 dz + j: lanes kg < 2 hold tap (dz, 2j), kg >= 2 tap
    // (dz, 2j+1) (e = 9 is the zero pad); each lane reads 8 of the 16 channels
    const int half = kg >> 1, cofs = 8 * (kg & 1);
    if constexpr (NT <= 2) {
      tail16_mfma<NT, MT, UPC>(acc, lds, wrow, wtile, 27 * F, w, m, kg, Bt);
    } else {
      bf16x8 Bn[NT];
      load_b(Bn, wrow, wtile, 27 * F);
      for (int dz = 0; dz < 3; ++dz) {
        for (int j = 0; j < 5; ++j) {
          bf16x8 B[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t)
            B[t] = Bn[t];
          if (5 * dz + j + 1 < 15)
            load_b(Bn, wrow, wtile, 27 * F + 5 * dz + j + 1);
          const int e  = 2 * j + half;
          const int ee = e >= 9 ? 8 : e;     // the pad: zero weights, as above
          const uint16_t *base = lds + ((dz * kHY + w + ee / 3) * kHX + m + ee % 3) * 16 + cofs;
#pragma unroll
          for (int r = 0; r < MT; ++r) {
            const bf16x8 A =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(base + r * RS16));
#pragma unroll
            for (int t = 0; t < NT; ++t)
              acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t], A, acc[r][t], 0, 0, 0);
          }
        }
      }
    }
  }
  float hin[head_per_lane<MT>()] = {};
  if constexpr (UPC)
    load_hin<NT, MT>(a, n, x0, y0, z0, w, lane, hin);
  if constexpr (NT > 2)
    conv_epilogue_wide<NT, MT>(a, acc, lds, n, x0, y0, z0, w, m, kg);
  else
    conv_epilogue<NT, MT, UPC>(a, acc, lds, n, x0, y0, z0, w, lane, m, kg, biasv, hin);
}

// Persistent form of the 16-input-channel convolution (one 16-channel pass;
// a single full-resolution source, a.s0).  A grid sized to the chip walks the tiles
// (8 contiguous ranges, one per XCD); each tile's halo is staged by
// direct-to-LDS buffer loads (unit i of [MT+2][kHY][kHX][16] lands 16 bytes
// at lds + 8 i, 1 KB per wave instruction; units outside the volume read past
// the buffer range and land zeros) into one of two LDS buffers while the
// previous tile computes and runs its epilogue in the other.
template <int MT> constexpr int stage16_iter() { return (kHX * kHY * (MT + 2) * 2 + 255) / 256; }

// Byte offsets of this thread's halo units from the halo's first voxel
// (channel offset included), the same for every tile of a launch.
template <int MT>
__device__ __forceinline__ void stage16_rel(const ConvLdsArgs &a, uint32_t (&rel)[stage16_iter<MT>()])
{
  constexpr int TOTAL = kHX * kHY * (MT + 2) * 2;
  const uint32_t D = (uint32_t)a.D, ld = (uint32_t)a.s0.ld, c0 = (uint32_t)a.s0.ofs;
#pragma unroll
  for (int k = 0; k < stage16_iter<MT>(); ++k) {
    const int i = threadIdx.x + k * 256;
    const int q = i & 1, v = i >> 1;
    const uint32_t hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
    rel[k] = i < TOTAL ? (((hz * D + hy) * D + hx) * ld + c0 + 8 * q) * 2u : 0u;
  }
}

template <int MT>
__device__ __forceinline__ void stage16_dma(const ConvLdsArgs &a, uint16_t *buf, int n, int x0,
    int y0, int z0, const uint32_t (&rel)[stage16_iter<MT>()])
{
  constexpr int TOTAL = kHX * kHY * (MT + 2) * 2;  // 16-byte units
  constexpr int ITER  = stage16_iter<MT>();
  const int D         = a.D;
  const ConvSrc &sr   = a.s0;                      // channels [0, c0) = all 16
  const uint32_t ld   = (uint32_t)sr.ld;
  const uint32_t c0   = (uint32_t)sr.ofs;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)sr.p, (short)0, sr.bytes, kRsrcWord3);
  const int wbase = threadIdx.x & ~63;
  // a halo inside the volume (most tiles): the tile's base offset plus the
  // thread's precomputed unit offsets
  if (x0 > 0 && y0 > 0 && z0 > 0 && x0 + 16 < D && y0 + 4 < D && z0 + MT < D) {
    const uint32_t base =
        ((((uint32_t)n * D + z0 - 1) * D + y0 - 1) * D + x0 - 1) * ld * 2u;
#pragma unroll
    for (int k = 0; k < ITER; ++k) {
      const int i = threadIdx.x + k * 256;
      if (i < TOTAL)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
            (__attribute__((address_space(3))) void *)(buf + 8 * (wbase + k * 256)), 16,
            (int)(base + rel[k]), 0, 0, 0);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int i = threadIdx.x + k * 256;
    if (i < TOTAL) {
      const int q = i & 1, v = i >> 1;
      const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      const bool in = (unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D &&
                      (unsigned)gz < (unsigned)D;
      const uint32_t off =
          in ? ((((uint32_t)n * D + gz) * D + gy) * D + gx) * ld * 2u + (c0 + 8 * q) * 2u
             : kMissOffset;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
          (__attribute__((address_space(3))) void *)(buf + 8 * (wbase + k * 256)), 16, (int)off,
          0, 0, 0);
    }
  }
}

template <int NT, int MT>
constexpr int conv16_buf_elems()
{
  const int halo = (MT + 2) * kHY * kHX * 16;
  const int outb = 4 * MT * 16 * NT * 16;
  return halo > outb ? halo : outb;
}

template <int NT, int MT>
__global__ __launch_bounds__(256) void conv16_persist_kernel(ConvLdsArgs a)
{
  extern __shared__ __align__(16) uint16_t lds[];
  constexpr int BUF = conv16_buf_elems<NT, MT>();
  const int D = a.D;
  const int txN = D / 16, tyN = D / 4, tzN = D / MT;
  const int nTiles = a.n * txN * tyN * tzN;
  const int perXcd = (nTiles + 7) / 8;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int stride = (gridDim.x + 7 - xcd) / 8;   // workgroups of this XCD
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const uint16_t *wrow = a.w + (size_t)m * a.kp + 8 * kg;
  const size_t wtile = (size_t)16 * a.kp;
  auto coords = [&](int t, int &n, int &x0, int &y0, int &z0) {
    x0 = (t % txN) * 16;
    int rest = t / txN;
    y0 = (rest % tyN) * 4;
    rest /= tyN;
    z0 = (rest % tzN) * MT;
    n  = rest / tzN;
  };
  int k = local, t = xcd * perXcd + k;
  if (k >= perXcd || t >= nTiles)
    return;
  // one output tile (NT = 1): the weight operands of every tile in 60
  // registers, loaded before the first halo copy (two workgroups per CU,
  // which the halo buffers' LDS sets, leave room for them; NT = 2 would
  // halve that occupancy and loads them per tile)
  constexpr bool kRegW = NT == 1;
  bf16x8 W[kRegW ? 15 : 1][NT];
  if constexpr (kRegW) {
#pragma unroll
    for (int c = 0; c < 15; ++c)
      load_b(W[c], wrow, wtile, c);
  }
  float biasv[NT];
  load_bias<NT>(a, m, biasv);
  uint32_t rel[stage16_iter<MT>()];
  stage16_rel<MT>(a, rel);
  int n, x0, y0, z0;
  coords(t, n, x0, y0, z0);
  stage16_dma<MT>(a, lds, n, x0, y0, z0, rel);
  int buf = 0;
  for (;;) {
    // this tile's halo has landed, and every wave is done with the other
    // buffer (the previous tile's epilogue)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // the fused head's inputs of this tile, loaded before the prefetch
    float hin[head_per_lane<MT>()];
    load_hin<NT, MT>(a, n, x0, y0, z0, w, lane, hin);
    const int k2 = k + stride, t2 = xcd * perXcd + k2;
    const bool next = k2 < perXcd && t2 < nTiles;
    int n2 = 0, x2 = 0, y2 = 0, z2 = 0;
    if (next) {
      coords(t2, n2, x2, y2, z2);
      stage16_dma<MT>(a, lds + (buf ^ 1) * BUF, n2, x2, y2, z2, rel);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int r = 0; r < MT; ++r)
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
        acc[r][tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (kRegW)
      tail16_mfma_reg<NT, MT, 15>(acc, lds + buf * BUF, W, w, m, kg);
    else
      tail16_mfma<NT, MT>(acc, lds + buf * BUF, wrow, wtile, 0, w, m, kg, W);
    conv_epilogue<NT, MT, true>(a, acc, lds + buf * BUF, n, x0, y0, z0, w, lane, m, kg, biasv,
        hin);
    if (!next)
      break;
    k = k2;
    n = n2, x0 = x2, y0 = y2, z0 = z2;
    buf ^= 1;
  }
}

// First layer: one live input channel (the FP32 volume, rounded to bfloat16),
// K = 27 taps (padded to 32), 16 output channels: one MFMA per 16 voxels.
// Block tile 16 x 4 x 8 voxels as in conv3_lds_kernel; the input halo
// (18 x 6 x 10 voxels) is staged in LDS as bfloat16 and each lane gathers its
// 8 taps from it.  w0 = bfloat16 [16][32] (taps 27..31 zero).
__global__ __launch_bounds__(256) void conv_first_kernel(const float *__restrict__ in,
    const uint16_t *__restrict__ w0, const float *__restrict__ b0, uint16_t *__restrict__ y,
    int D, int tilesPerXcd)
{
  constexpr int MT = 8;
  __shared__ uint16_t halo[kHX * kHY * (MT + 2)];
  __shared__ __align__(16) uint16_t obuf[4 * MT * 16 * 16];
  const int txN = D / 16, tyN = D / 4, tzN = D / MT;
  int tile = blockIdx.x;
  if (tilesPerXcd > 0)
    tile = (blockIdx.x & 7) * tilesPerXcd + (blockIdx.x >> 3);
  const int tx = tile % txN;
  int rest = tile / txN;
  const int ty = rest % tyN;
  rest /= tyN;
  const int tz = rest % tzN, n = rest / tzN;
  const int x0 = tx * 16, y0 = ty * 4, z0 = tz * MT;
  // all of a thread's halo loads are issued before the first is used
  constexpr int HV = kHX * kHY * (MT + 2), HITER = (HV + 255) / 256;
  float hv[HITER];
#pragma unroll
  for (int k = 0; k < HITER; ++k) {
    const int i = threadIdx.x + 256 * k;
    hv[k] = 0.0f;
    if (i < HV) {
      const int hx = i % kHX, hy = (i / kHX) % kHY, hz = i / (kHX * kHY);
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      if ((unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D && (unsigned)gz < (unsigned)D)
        hv[k] = in[(((size_t)n * D + gz) * D + gy) * D + gx];
    }
  }
#pragma unroll
  for (int k = 0; k < HITER; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < HV)
      halo[i] = f2bf(hv[k]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const bf16x8 B = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(w0 + m * 32 + 8 * kg));
  int hoff[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tap = 8 * kg + j;
    const int tp = tap < 27 ? tap : 0;
    hoff[j] = ((tp / 9) * kHY + w + (tp / 3) % 3) * kHX + m + tp % 3;
  }
  uint16_t *ob = obuf + w * MT * 16 * 16;
  const float bias = b0[m];
#pragma unroll
  for (int r = 0; r < MT; ++r) {
    uint16_t av[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      av[j] = halo[hoff[j] + r * kHX * kHY];   // pad taps read tap 0 (zero weights)
    const uint4 a4 = make_uint4((uint32_t)av[0] | ((uint32_t)av[1] << 16),
        (uint32_t)av[2] | ((uint32_t)av[3] << 16), (uint32_t)av[4] | ((uint32_t)av[5] << 16),
        (uint32_t)av[6] | ((uint32_t)av[7] << 16));
    const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a4), B,
        f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      ob[(r * 16 + 4 * kg + i) * 16 + m] = f2bf(fmaxf(acc[i] + bias, 0.0f));
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u0 = 0; u0 < MT * 16 * 2; u0 += 64) {
    const int u = u0 + lane;
    const int c8 = u & 1, vx = (u >> 1) % 16, r = u >> 5;
    const size_t vo = (((size_t)n * D + z0 + r) * D + y0 + w) * D + x0 + vx;
    *reinterpret_cast<uint4 *>(y + vo * 16 + 8 * c8) =
        *reinterpret_cast<const uint4 *>(ob + (r * 16 + vx) * 16 + 8 * c8);
  }
}

// The network's first two convolutions at inference (enc1a: 1 -> 16
// channels, enc1b: 16 -> 16 with the 2x2x2 pool fused) in one kernel per
// 16 x 4 x 8 tile.  enc1a is evaluated on enc1b's whole halo (18 x 6 x 10
// voxels, 2.1x the tile) from the input's halo of that halo (20 x 8 x 12
// voxels, staged in LDS as bfloat16) and written into the halo buffer enc1b
// reads (zeros outside the volume: enc1b's padding), so enc1a's 32 bytes
// per voxel never go to HBM and come back 2.1 times as enc1b's halo.  Each
// MFMA has the operands of conv_first_kernel (one halo voxel per A row, its
// 27 taps then 5 copies of tap 0 against zero weights) and of the
// 16-channel path, and the roundings are theirs: the outputs are the
// two-kernel form's bit for bit.
__global__ __launch_bounds__(256) void conv_first16_kernel(const float *__restrict__ in,
    const uint16_t *__restrict__ w0, const float *__restrict__ b0, ConvLdsArgs a)
{
  constexpr int MT = 8;
  constexpr int IX = kHX + 2, IY = kHY + 2, IZ = MT + 4;  // input halo of the halo
  constexpr int IV = IX * IY * IZ;
  constexpr int HV = kHX * kHY * (MT + 2);                // enc1b's halo voxels
  constexpr int GROUPS = (HV + 15) / 16;                  // MFMAs of enc1a
  static_assert(GROUPS % 4 == 0, "enc1a's groups split evenly over the four waves");
  __shared__ uint16_t inh[IV];
  // enc1b's halo [MT+2][kHY][kHX][16] (and the last enc1a group's pad
  // voxels, never read), then the epilogue's transpose buffer
  __shared__ __align__(16) uint16_t halo[GROUPS * 16 * 16];
  const int D = a.D;
  const int txN = D / 16, tyN = D / 4, tzN = D / MT;
  int tile = blockIdx.x;
  if (a.tilesPerXcd > 0)
    tile = (blockIdx.x & 7) * a.tilesPerXcd + (blockIdx.x >> 3);
  const int tx = tile % txN;
  int rest = tile / txN;
  const int ty = rest % tyN;
  rest /= tyN;
  const int tz = rest % tzN, n = rest / tzN;
  const int x0 = tx * 16, y0 = ty * 4, z0 = tz * MT;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  // enc1b's bias and first weight operands, loaded ahead of the input so
  // that their latency is covered by the input's
  const uint16_t *wrow = a.w + (size_t)m * a.kp + 8 * kg;
  float biasv[1];
  load_bias<1>(a, m, biasv);
  bf16x8 Bt[3][1];
#pragma unroll
  for (int dz = 0; dz < 3; ++dz)
    load_b(Bt[dz], wrow, (size_t)16 * a.kp, 5 * dz);
  // the input around the halo, all of a thread's loads issued before the
  // first is used; zeros outside the volume
  constexpr int ITER = (IV + 255) / 256;
  float iv[ITER];
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int i = threadIdx.x + 256 * k;
    iv[k] = 0.0f;
    if (i < IV) {
      const int gx = x0 - 2 + i % IX, gy = y0 - 2 + (i / IX) % IY, gz = z0 - 2 + i / (IX * IY);
      if ((unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D && (unsigned)gz < (unsigned)D)
        iv[k] = in[(((size_t)n * D + gz) * D + gy) * D + gx];
    }
  }
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < IV)
      inh[i] = f2bf(iv[k]);
  }
  __syncthreads();
  {
    // enc1a: halo voxels 16 g .. 16 g + 15 per MFMA (lane m's A row is voxel
    // 16 g + m, its taps 8 kg .. 8 kg + 7; the pad voxels of the last group
    // repeat the last voxel, and the pad taps 27..31 read tap 0, whose
    // products meet zero weights); wave w takes groups w, w + 4, ..., all
    // unrolled so that the gathers of later groups overlap earlier MFMAs
    const bf16x8 B =
        __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(w0 + m * 32 + 8 * kg));
    int toff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int tap = 8 * kg + j;
      const int tp  = tap < 27 ? tap : 0;
      toff[j] = ((tp / 9) * IY + (tp / 3) % 3) * IX + tp % 3;
    }
    const float bias = b0[m];
#pragma unroll
    for (int gi = 0; gi < GROUPS / 4; ++gi) {
      const int g    = w + 4 * gi;
      const int v    = min(16 * g + m, HV - 1);
      const int base = ((v / (kHX * kHY)) * IY + (v / kHX) % kHY) * IX + v % kHX;
      uint16_t av[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        av[j] = inh[base + toff[j]];
      const uint4 a4 = make_uint4((uint32_t)av[0] | ((uint32_t)av[1] << 16),
          (uint32_t)av[2] | ((uint32_t)av[3] << 16), (uint32_t)av[4] | ((uint32_t)av[5] << 16),
          (uint32_t)av[6] | ((uint32_t)av[7] << 16));
      const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a4),
          B, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
      uint16_t *hp = halo + (16 * g + 4 * kg) * 16 + m;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        hp[16 * i] = f2bf(fmaxf(acc[i] + bias, 0.0f));
    }
  }
  // halo voxels outside the volume hold enc1b's zero padding, not enc1a's
  // output there (tiles at the volume's faces only)
  if (x0 == 0 || y0 == 0 || z0 == 0 || x0 + 16 == D || y0 + 4 == D || z0 + MT == D) {
    __syncthreads();
    for (int i = threadIdx.x; i < HV * 2; i += 256) {
      const int v  = i >> 1;
      const int gx = x0 - 1 + v % kHX, gy = y0 - 1 + (v / kHX) % kHY,
                gz = z0 - 1 + v / (kHX * kHY);
      if (!((unsigned)gx < (unsigned)D && (unsigned)gy < (unsigned)D &&
              (unsigned)gz < (unsigned)D))
        *reinterpret_cast<uint4 *>(halo + v * 16 + 8 * (i & 1)) = make_uint4(0, 0, 0, 0);
    }
  }
  __syncthreads();
  // enc1b: the 16-channel path on the computed halo, then bias, ReLU, the
  // stores and the fused pool
  f32x4 acc[MT][1];
#pragma unroll
  for (int r = 0; r < MT; ++r)
    acc[r][0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  tail16_mfma<1, MT, true>(acc, halo, wrow, (size_t)16 * a.kp, 0, w, m, kg, Bt);
  float hin[head_per_lane<MT>()];
  load_hin<1, MT>(a, n, x0, y0, z0, w, lane, hin);
  conv_epilogue<1, MT, true>(a, acc, halo, n, x0, y0, z0, w, lane, m, kg, biasv, hin);
}

// Weight gradient dW[co][tap][ci] += sum_v g[v][co] * x[v + off(tap)][ci].
// Grid: (blocks over x-rows) x 9 (dz, dy).  Per 32-voxel x segment, LDS
// holds g transposed (gT[co][v]) and the input row at the three x shifts,
// transposed (xT[dx][ci][v]); every MFMA operand is one aligned 16-byte LDS
// read (M = 16 output channels, N = 16 input channels, K = 32 voxels).
// The 3*COT*CIT tiles are spread over the 4 waves; partial sums are added to
// dW atomically.
struct WgradArgs {
  ConvSrc s0, s1; int c0;   // input channels [0, c0) from s0 (maybe half resolution), rest s1
  const uint16_t *g; int ldg, gofs;
  float *dw;           // [cout][27][cin]
  int D, nrows;        // rows = n*D*D x-rows of D voxels
  int rowsPerBlock;
};

constexpr int kSeg  = 32;
constexpr int kSegP = 40; // padded LDS row (80 bytes)

template <int COT, int CIT>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a)
{
  constexpr int CIN = 16 * CIT, COUT = 16 * COT;
  constexpr int NTILES = 3 * COT * CIT, PER = (NTILES + 3) / 4;
  __shared__ __align__(16) uint16_t gT[COUT * kSegP];
  __shared__ __align__(16) uint16_t xT[3 * CIN * kSegP];
  const int dz = blockIdx.y / 3, dy = blockIdx.y % 3;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int D = a.D;
  f32x4 acc[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s)
    acc[s] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int r0 = blockIdx.x * a.rowsPerBlock;
  const int r1 = min(r0 + a.rowsPerBlock, a.nrows);
  for (int row = r0; row < r1; ++row) {
    const int yy = row % D, z = (row / D) % D, n = row / (D * D);
    const int sz = z + dz - 1, sy = yy + dy - 1;
    const bool rowOk = (unsigned)sz < (unsigned)D && (unsigned)sy < (unsigned)D;
    const size_t gBase = (((size_t)n * D + z) * D + yy) * D;
    const size_t xBase = (((size_t)n * D + sz) * D + sy) * D;
    for (int x0 = 0; x0 < D; x0 += kSeg) {
      const int seg = min(kSeg, D - x0); // voxels of this segment; the rest of K is zero
      __syncthreads();
      for (int i = threadIdx.x; i < COUT * kSeg; i += 256) {
        const int co = i % COUT, xv = i / COUT;
        gT[co * kSegP + xv] = xv < seg ? a.g[(gBase + x0 + xv) * a.ldg + a.gofs + co] : (uint16_t)0;
      }
      for (int i = threadIdx.x; i < 3 * CIN * kSeg; i += 256) {
        const int ci = i % CIN, rem = i / CIN;
        const int xv = rem % kSeg, dx = rem / kSeg;
        const int sx = x0 + xv + dx - 1;
        uint16_t val = 0;
        if (rowOk && xv < seg && (unsigned)sx < (unsigned)D) {
          const bool first = ci < a.c0;
          const ConvSrc &sr = first ? a.s0 : a.s1;
          const int cc = first ? ci : ci - a.c0;
          if (sr.half) {
            const int Dh = D >> 1;
            val = sr.p[((((size_t)n * Dh + (sz >> 1)) * Dh + (sy >> 1)) * Dh + (sx >> 1)) * sr.ld +
                       sr.ofs + cc];
          } else {
            val = sr.p[(xBase + sx) * sr.ld + sr.ofs + cc];
          }
        }
        xT[(dx * CIN + ci) * kSegP + xv] = val;
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < PER; ++s) {
        const int t = wv + 4 * s;
        if (t < NTILES) {
          const int dx = t / (COT * CIT), rem = t % (COT * CIT);
          const int cot = rem / CIT, cit = rem % CIT;
          const bf16x8 A = __builtin_bit_cast(bf16x8,
              *reinterpret_cast<const uint4 *>(gT + (cot * 16 + m) * kSegP + 8 * kg));
          const bf16x8 B = __builtin_bit_cast(bf16x8,
              *reinterpret_cast<const uint4 *>(xT + (dx * CIN + cit * 16 + m) * kSegP + 8 * kg));
          acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[s], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = wv + 4 * s;
    if (t < NTILES) {
      const int dx = t / (COT * CIT), rem = t % (COT * CIT);
      const int cot = rem / CIT, cit = rem % CIT;
      const int tap = blockIdx.y * 3 + dx;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = cot * 16 + 4 * kg + i, ci = cit * 16 + m;
        atomicAdd(a.dw + ((size_t)co * 27 + tap) * CIN + ci, acc[s][i]);
      }
    }
  }
}

// bias gradient: db[c] += sum_v g[v][c] (C divides 256)
__global__ __launch_bounds__(256) void bgrad_kernel(const uint16_t *g, int ldg, int gofs, int C,
    int nvox, float *db)
{
  __shared__ float red[256];
  const int c = threadIdx.x % C;
  const int lanesPerC = blockDim.x / C;
  const int sub = threadIdx.x / C;
  float s = 0.0f;
  for (int v = blockIdx.x * lanesPerC + sub; v < nvox; v += gridDim.x * lanesPerC)
    s += bf2f(g[(size_t)v * ldg + gofs + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.0f;
    for (int k = 0; k < lanesPerC; ++k)
      t += red[k * C + threadIdx.x];
    atomicAdd(db + threadIdx.x, t);
  }
}

// g *= (y > 0) over a channel slice
__global__ void relu_bwd_kernel(uint16_t *g, int ldg, int gofs, const uint16_t *y, int ldy,
    int yofs, int C, int nvox)
{
  const size_t n = (size_t)nvox * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t v = i / C, c = i % C;
    if (!(bf2f(y[v * ldy + yofs + c]) > 0.0f))
      g[v * ldg + gofs + c] = 0;
  }
}

// pooling backward: gx += gy(parent) / 8
__global__ void pool_bwd_kernel(const uint16_t *gy, int ldy, int yofs, uint16_t *gx, int ldx,
    int xofs, int C, int Di, int n)
{
  const size_t total = (size_t)n * Di * Di * Di * C;
  const int Do = Di / 2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    const size_t v = i / C;
    const int xi = v % Di, yi = (v / Di) % Di, zi = (v / (Di * Di)) % Di, b = v / ((size_t)Di * Di * Di);
    const size_t vo = (((size_t)b * Do + zi / 2) * Do + yi / 2) * Do + xi / 2;
    uint16_t *dst = gx + v * ldx + xofs + c;
    *dst = f2bf(bf2f(*dst) + 0.125f * bf2f(gy[vo * ldy + yofs + c]));
  }
}

// upsampling backward: gx = sum of the 8 children
__global__ void up_bwd_kernel(const uint16_t *gy, int ldy, int yofs, uint16_t *gx, int ldx, int xofs,
    int C, int Di, int n)
{
  const size_t total = (size_t)n * Di * Di * Di * C;
  const int Do = 2 * Di;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    const size_t v = i / C;
    const int xi = v % Di, yi = (v / Di) % Di, zi = (v / (Di * Di)) % Di, b = v / ((size_t)Di * Di * Di);
    float s = 0.0f;
    for (int q = 0; q < 8; ++q) {
      const size_t vo = (((size_t)b * Do + 2 * zi + (q >> 2)) * Do + 2 * yi + ((q >> 1) & 1)) * Do + 2 * xi + (q & 1);
      s += bf2f(gy[vo * ldy + yofs + c]);
    }
    gx[v * ldx + xofs + c] = f2bf(s);
  }
}

// input packing: FP32 volume -> BF16 [v][16], channel 0 = value
__global__ void pack_input_kernel(const float *in, uint16_t *x, size_t nvox)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nvox * 16; i += (size_t)gridDim.x * blockDim.x)
    x[i] = (i % 16 == 0) ? f2bf(in[i / 16]) : (uint16_t)0;
}

// 1x1x1 head with residual: out[v] = in[v] + b + sum_c w[c] h[v][c]
__global__ void head_kernel(const uint16_t *h, const float *in, const float *w, const float *b,
    float *out, size_t nvox)
{
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    float s = in[v] + b[0];
    const uint16_t *hv = h + v * 16;
    for (int c = 0; c < 16; ++c)
      s = fmaf(w[c], bf2f(hv[c]), s);
    out[v] = s;
  }
}

// Block-wide sum of NV per-thread values (256 threads); thread 0 receives the
// totals in v[].
template <int NV> __device__ __forceinline__ void block_sum(float (&v)[NV])
{
  __shared__ float red[4][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int o = 32; o > 0; o >>= 1)
      v[k] += __shfl_xor(v[k], o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      red[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      v[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// MSE: loss += sum (out-t)^2 / N;  gy = 2 (out - t) / N
__global__ __launch_bounds__(256) void mse_kernel(const float *out, const float *t, float *gy,
    size_t nvox, float *loss)
{
  float s[1] = { 0.0f };
  const float inv = 1.0f / (float)nvox;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    const float d = out[v] - t[v];
    s[0] += d * d;
    gy[v] = 2.0f * d * inv;
  }
  block_sum(s);
  if (threadIdx.x == 0)
    atomicAdd(loss, s[0] * inv);
}

// head backward: gh[v][c] = gy[v] w[c]; gw[c] += sum gy h; gb += sum gy
__global__ __launch_bounds__(256) void head_bwd_kernel(const float *gy, const uint16_t *h,
    const float *w, uint16_t *gh, float *gw, float *gb, size_t nvox)
{
  float acc[17];
  float wl[16];
#pragma unroll
  for (int c = 0; c < 17; ++c)
    acc[c] = 0.0f;
#pragma unroll
  for (int c = 0; c < 16; ++c)
    wl[c] = w[c];
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    const float g = gy[v];
    acc[16] += g;
    const uint4 *hv = reinterpret_cast<const uint4 *>(h + v * 16);
    uint4 *gv = reinterpret_cast<uint4 *>(gh + v * 16);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const uint4 hq = hv[q];
      const uint32_t hw[4] = { hq.x, hq.y, hq.z, hq.w };
      uint32_t ow[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = 8 * q + 2 * k;
        acc[c] = fmaf(g, __uint_as_float(hw[k] << 16), acc[c]);
        acc[c + 1] = fmaf(g, __uint_as_float(hw[k] & 0xffff0000u), acc[c + 1]);
        ow[k] = (uint32_t)f2bf(g * wl[c]) | ((uint32_t)f2bf(g * wl[c + 1]) << 16);
      }
      gv[q] = make_uint4(ow[0], ow[1], ow[2], ow[3]);
    }
  }
  block_sum(acc);
  if (threadIdx.x == 0) {
    atomicAdd(gb, acc[16]);
#pragma unroll
    for (int c = 0; c < 16; ++c)
      atomicAdd(gw + c, acc[c]);
  }
}

// Adam on a flat FP32 parameter array (bias-corrected, PyTorch's form)
__global__ void adam_kernel(float *p, const float *g, float *m, float *v, size_t n, float lr,
    float b1, float b2, float eps, float c1, float c2)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// master W[co][tap][ci] (FP32) -> forward packing wf[co][kpos(cin, tap, ci)]
// and backward-data packing wb[ci][kpos(cout, 26 - tap, co)] (BF16)
__global__ void pack_weights_kernel(const float *W, int cin, int cout, uint16_t *wf, int kpf,
    uint16_t *wb, int kpb)
{
  const int total = cout * 27 * cin;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int ci = i % cin, tap = (i / cin) % 27, co = i / (27 * cin);
    const uint16_t h = f2bf(W[i]);
    wf[(size_t)co * kpf + kpos(cin, tap, ci)] = h;
    wb[(size_t)ci * kpb + kpos(cout, 26 - tap, co)] = h;
  }
}

// first-layer weights for conv_first_kernel: w0[co][tap] = bf16(W[co][tap][0]),
// rows of 32 (taps 27..31 stay zero)
__global__ void pack_first_kernel(const float *W, int cin, uint16_t *w0)
{
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i < 16 * 27)
    w0[(i / 27) * 32 + i % 27] = f2bf(W[(size_t)i * cin]);
}

} // namespace

// ---------------------------------------------------------------------------
// host side

namespace {

int grid_for(size_t n, int block = 256, int cap = 16384)
{
  return (int)std::max<size_t>(1, std::min<size_t>((n + block - 1) / block, cap));
}

struct Conv {
  int cin, cout, kpf, kpb;
  float *W = nullptr, *b = nullptr;     // master FP32
  float *gW = nullptr, *gb = nullptr;
  float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
  uint16_t *wf = nullptr, *wb = nullptr;
};

struct Act {
  uint16_t *p = nullptr;
  int ld = 0, ofs = 0;
};

} // namespace

struct LhUnet {
  Conv c[10];
  float *hw = nullptr, *hb = nullptr, *ghw = nullptr, *ghb = nullptr;
  uint16_t *w0 = nullptr;  // first layer, bfloat16 [16][32]
  float *mhw = nullptr, *vhw = nullptr, *mhb = nullptr, *vhb = nullptr;
  int step = 0;
  int fuseFirst = 1;       // inference: enc1a + enc1b in conv_first16_kernel
  hipStream_t stream = nullptr;
  // workspace for one (n, D)
  int wsN = 0, wsD = 0;
  std::vector<void *> ws;
  uint16_t *x16, *a1, *e1, *p1, *a2, *e2, *p2, *a3, *b3, *d2a, *d2b, *d1a, *d1b;
  uint16_t *g_a1, *g_cat1, *g_p1, *g_a2, *g_cat2, *g_p2, *g_a3, *g_b3, *g_d2a, *g_d2b, *g_d1a,
      *g_d1b;
  float *out, *gy, *loss;
};

namespace {

// layer table: (cin, cout); order enc1a enc1b enc2a enc2b bot_a bot_b dec2a dec2b dec1a dec1b
const int kLayers[10][2] = { { 16, 16 }, { 16, 16 }, { 16, 32 }, { 32, 32 }, { 32, 64 },
  { 64, 64 }, { 96, 32 }, { 32, 32 }, { 48, 16 }, { 16, 16 } };

#define UN_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return lh_fail(e_ == hipErrorOutOfMemory ? LH_E_NOMEM : LH_E_HIP, "%s:%d %s: %s",    \
          __FILE__, __LINE__, #call, hipGetErrorString(e_));                             \
  } while (0)

template <typename T> int dalloc(T **p, size_t n)
{
  UN_HIP(hipMalloc((void **)p, n * sizeof(T)));
  UN_HIP(hipMemset(*p, 0, n * sizeof(T)));
  return LH_OK;
}

// Convolution input: channels [0, c0) from a0 (at half resolution when
// half0), channels [c0, cin) from a1.
struct ConvIn {
  Act a0;
  int half0 = 0;
  int c0 = 0;  // 0: all channels from a0
  Act a1;
};

ConvIn plain_in(Act a)
{
  ConvIn c;
  c.a0 = a;
  c.a1 = a;
  return c;
}

template <int NT, int MT, bool TAIL, bool UPC = false> int launch_lds(hipStream_t s, ConvLdsArgs &a)
{
  constexpr size_t ldsMax = conv_lds_bytes<NT, MT>();
  static bool attr = false;
  if (!attr) {
    UN_HIP(hipFuncSetAttribute((const void *)conv3_lds_kernel<NT, MT, TAIL, UPC>,
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsMax));
    attr = true;
  }
  // a 16-channel input stages a 16-channel halo only; with every 32-channel
  // pass staged coarse, the coarse halo or the 16-channel tail sizes it
  const size_t halo16 = (size_t)(MT + 2) * kHY * kHX * 16 * 2;
  const size_t outb = (size_t)4 * MT * 16 * NT * 16 * 2;
  const size_t coarse = (size_t)(MT / 2 + 2) * kCHY * kCHX * 32 * 2;
  size_t lds = (a.cin >= 32) ? ldsMax : std::max(halo16, outb);
  if (UPC)
    lds = std::max(std::max(coarse, TAIL ? halo16 : 0), outb);
  if (a.D % MT)
    return lh_fail(LH_E_ARG, "conv: D=%d is not a multiple of the z tile %d", a.D, MT);
  const size_t tiles = (size_t)a.n * (a.D / 16) * (a.D / 4) * (a.D / MT);
  a.tilesPerXcd = (tiles % 8 == 0) ? (int)(tiles / 8) : 0;
  conv3_lds_kernel<NT, MT, TAIL, UPC><<<(unsigned)tiles, 256, lds, s>>>(a);
  return LH_OK;
}

// z tile of the 16 -> 16 channel convolutions (full resolution, halo-bound)
#ifndef LH_MT16
#define LH_MT16 8
#endif

// Persistent launch of the 16-input-channel convolutions: two workgroups per
// CU (two halo buffers each), every CU of the device.
#ifndef LH_CONV16_PERSIST
#define LH_CONV16_PERSIST 1
#endif
template <int NT, int MT> int launch_conv16(hipStream_t s, ConvLdsArgs &a)
{
  constexpr size_t lds = (size_t)2 * conv16_buf_elems<NT, MT>() * sizeof(uint16_t);
  static int grid = 0;
  if (grid == 0) {
    UN_HIP(hipFuncSetAttribute((const void *)conv16_persist_kernel<NT, MT>,
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int dev = 0, cus = 0, perCU = 0;
    UN_HIP(hipGetDevice(&dev));
    UN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    UN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, conv16_persist_kernel<NT, MT>,
        256, lds));
    grid = cus * std::max(perCU, 1);
  }
  if (a.D % MT)
    return lh_fail(LH_E_ARG, "conv: D=%d is not a multiple of the z tile %d", a.D, MT);
  const long tiles = (long)a.n * (a.D / 16) * (a.D / 4) * (a.D / MT);
  conv16_persist_kernel<NT, MT><<<(unsigned)std::min<long>(grid, tiles), 256, lds, s>>>(a);
  return LH_OK;
}

template <bool TAIL> int launch_cout(hipStream_t s, ConvLdsArgs &a, int cout, bool upc)
{
  if (LH_CONV16_PERSIST && TAIL && a.cin == 16 && a.c0 == 16 && !a.s0.half && !upc) {
    if (cout == 16)
      return launch_conv16<1, LH_MT16>(s, a);
    if (cout == 32)
      return launch_conv16<2, 8>(s, a);
  }
  if (upc && cout == 16)
    return launch_lds<1, 8, TAIL, true>(s, a);
  if (upc && cout == 32)
    return launch_lds<2, 8, TAIL, true>(s, a);
  switch (cout) {
  case 16: return launch_lds<1, 8, TAIL>(s, a);
  case 32: return launch_lds<2, 8, TAIL>(s, a);
  case 48: return launch_lds<3, 4, TAIL>(s, a);
  case 64: return launch_lds<4, 4, TAIL>(s, a);
  case 96: return launch_lds<6, 4, TAIL>(s, a);
  default: return lh_fail(LH_E_ARG, "conv: unsupported output channels %d", cout);
  }
}

int make_src(ConvSrc &sr, Act a, int half, int D, int n)
{
  const int Ds = half ? D / 2 : D;
  const uint64_t bytes = (uint64_t)n * Ds * Ds * Ds * a.ld * 2;
  if (bytes >= kMissOffset)
    return lh_fail(LH_E_ARG, "conv: input of %llu bytes exceeds the 2 GiB buffer range",
        (unsigned long long)bytes);
  if (a.ld % 8 || a.ofs % 8)
    return lh_fail(LH_E_ARG, "conv: channel stride/offset must be multiples of 8");
  sr.p = a.p;
  sr.ld = a.ld;
  sr.ofs = a.ofs;
  sr.half = half;
  sr.bytes = (int)bytes;
  return LH_OK;
}

struct HeadFuse {
  const float *in = nullptr, *w = nullptr, *b = nullptr;
  float *out = nullptr;
};

struct PoolFuse {
  uint16_t *p = nullptr;
  int ld = 0, ofs = 0;
};

int conv_fwd(hipStream_t s, const Conv &L, ConvIn in, Act out, int D, int n, bool relu,
    bool bias, bool accumulate, bool backward, HeadFuse head = HeadFuse(),
    PoolFuse pool = PoolFuse())
{
  ConvLdsArgs a;
  a.pool = pool.p;
  a.ldp = pool.ld;
  a.pofs = pool.ofs;
  if (pool.p && (pool.ld % 8 || pool.ofs % 8))
    return lh_fail(LH_E_ARG, "conv: pooled output stride/offset must be multiples of 8");
  a.hin = head.in;
  a.hw = head.w;
  a.hb = head.b;
  a.hout = head.out;
  if (head.out && (backward ? L.cin : L.cout) != 16)
    return lh_fail(LH_E_ARG, "conv: the fused head needs 16 output channels");
  int cout;
  if (!backward) {
    a.w = L.wf; a.cin = L.cin; cout = L.cout; a.kp = L.kpf; a.bias = bias ? L.b : nullptr;
  } else {
    a.w = L.wb; a.cin = L.cout; cout = L.cin; a.kp = L.kpb; a.bias = nullptr;
  }
  if (D % 16 || D < 16)
    return lh_fail(LH_E_ARG, "conv: D must be a multiple of 16 (got %d)", D);
  if (a.cin % 16)
    return lh_fail(LH_E_ARG, "conv: input channels must be a multiple of 16");
  a.c0 = in.c0 ? in.c0 : a.cin;
  if (a.c0 % 8 || a.c0 > a.cin || (in.half0 && D % 2))
    return lh_fail(LH_E_ARG, "conv: bad channel split %d", a.c0);
  int rc = make_src(a.s0, in.a0, in.half0, D, n);
  if (rc == LH_OK)
    rc = make_src(a.s1, in.c0 ? in.a1 : in.a0, 0, D, n);
  if (rc != LH_OK)
    return rc;
  a.y = out.p;
  a.ldy = out.ld;
  a.yofs = out.ofs;
  a.D = D;
  a.n = n;
  a.relu = relu;
  a.accumulate = accumulate;
  if (a.cin > 96 || (out.p && ((out.ld % 8) || (out.ofs % 8))))
    return lh_fail(LH_E_ARG, "conv: unsupported channels (in %d, out stride %d offset %d)", a.cin,
        out.ld, out.ofs);
  // Upsampled channels (a half-resolution a.s0 filling every 32-channel
  // pass) are staged once at coarse resolution; plane-reuse layouts only.
  const bool upc = in.half0 && a.c0 == 32 * (a.cin / 32) && (cout == 16 || cout == 32) && LH_UPC;
  rc = (a.cin % 32) ? launch_cout<true>(s, a, cout, upc) : launch_cout<false>(s, a, cout, upc);
  if (rc != LH_OK)
    return rc;
  UN_HIP(hipGetLastError());
  return LH_OK;
}

int conv_fwd(hipStream_t s, const Conv &L, Act in, Act out, int D, int n, bool relu, bool bias,
    bool accumulate, bool backward)
{
  return conv_fwd(s, L, plain_in(in), out, D, n, relu, bias, accumulate, backward);
}

int conv_wgrad(hipStream_t s, const Conv &L, ConvIn x, Act g, int D, int n)
{
  WgradArgs a;
  a.c0 = x.c0 ? x.c0 : L.cin;
  int rc = make_src(a.s0, x.a0, x.half0, D, n);
  if (rc == LH_OK)
    rc = make_src(a.s1, x.c0 ? x.a1 : x.a0, 0, D, n);
  if (rc != LH_OK)
    return rc;
  a.g = g.p; a.ldg = g.ld; a.gofs = g.ofs;
  a.dw = L.gW;
  a.D = D;
  a.nrows = n * D * D;
  if (D % 16 || D < 16)
    return lh_fail(LH_E_ARG, "wgrad: D must be a multiple of 16 (got %d)", D);
  const int blocksX = std::max(1, std::min(a.nrows, 1024));
  a.rowsPerBlock = (a.nrows + blocksX - 1) / blocksX;
  const dim3 grid((a.nrows + a.rowsPerBlock - 1) / a.rowsPerBlock, 9);
  const int key = L.cout / 16 * 16 + L.cin / 16;
  switch (key) {
  case 1 * 16 + 1: wgrad_kernel<1, 1><<<grid, 256, 0, s>>>(a); break;
  case 2 * 16 + 1: wgrad_kernel<2, 1><<<grid, 256, 0, s>>>(a); break;
  case 2 * 16 + 2: wgrad_kernel<2, 2><<<grid, 256, 0, s>>>(a); break;
  case 4 * 16 + 2: wgrad_kernel<4, 2><<<grid, 256, 0, s>>>(a); break;
  case 4 * 16 + 4: wgrad_kernel<4, 4><<<grid, 256, 0, s>>>(a); break;
  case 2 * 16 + 6: wgrad_kernel<2, 6><<<grid, 256, 0, s>>>(a); break;
  case 1 * 16 + 3: wgrad_kernel<1, 3><<<grid, 256, 0, s>>>(a); break;
  default:
    return lh_fail(LH_E_ARG, "wgrad: unsupported channels %d -> %d", L.cin, L.cout);
  }
  UN_HIP(hipGetLastError());
  bgrad_kernel<<<256, 256, 0, s>>>(g.p, g.ld, g.ofs, L.cout, n * D * D * D, L.gb);
  UN_HIP(hipGetLastError());
  return LH_OK;
}

int conv_wgrad(hipStream_t s, const Conv &L, Act x, Act g, int D, int n)
{
  return conv_wgrad(s, L, plain_in(x), g, D, n);
}

int pack_all(LhUnet *u)
{
  for (auto &L : u->c) {
    pack_weights_kernel<<<grid_for((size_t)L.cout * 27 * L.cin), 256, 0, u->stream>>>(L.W, L.cin,
        L.cout, L.wf, L.kpf, L.wb, L.kpb);
    UN_HIP(hipGetLastError());
  }
  pack_first_kernel<<<2, 256, 0, u->stream>>>(u->c[0].W, u->c[0].cin, u->w0);
  UN_HIP(hipGetLastError());
  return LH_OK;
}

int ensure_ws(LhUnet *u, int n, int D)
{
  if (u->wsN == n && u->wsD == D)
    return LH_OK;
  for (void *p : u->ws)
    (void)hipFree(p);
  u->ws.clear();
  u->wsN = u->wsD = 0;
  const size_t v1 = (size_t)n * D * D * D, v2 = v1 / 8, v3 = v2 / 8;
  struct { uint16_t **p; size_t n; } bufs[] = {
    { &u->x16, v1 * 16 }, { &u->a1, v1 * 16 }, { &u->e1, v1 * 16 }, { &u->p1, v2 * 16 },
    { &u->a2, v2 * 32 }, { &u->e2, v2 * 32 }, { &u->p2, v3 * 32 }, { &u->a3, v3 * 64 },
    { &u->b3, v3 * 64 }, { &u->d2a, v2 * 32 }, { &u->d2b, v2 * 32 }, { &u->d1a, v1 * 16 },
    { &u->d1b, v1 * 16 }, { &u->g_a1, v1 * 16 }, { &u->g_cat1, v1 * 48 }, { &u->g_p1, v2 * 16 },
    { &u->g_a2, v2 * 32 }, { &u->g_cat2, v2 * 96 }, { &u->g_p2, v3 * 32 }, { &u->g_a3, v3 * 64 },
    { &u->g_b3, v3 * 64 }, { &u->g_d2a, v2 * 32 }, { &u->g_d2b, v2 * 32 }, { &u->g_d1a, v1 * 16 },
    { &u->g_d1b, v1 * 16 } };
  for (auto &b : bufs) {
    int rc = dalloc(b.p, b.n);
    if (rc != LH_OK)
      return rc;
    u->ws.push_back(*b.p);
  }
  int rc = dalloc(&u->out, v1);
  if (rc == LH_OK) { u->ws.push_back(u->out); rc = dalloc(&u->gy, v1); }
  if (rc == LH_OK) { u->ws.push_back(u->gy); rc = dalloc(&u->loss, 1); }
  if (rc == LH_OK) u->ws.push_back(u->loss);
  if (rc != LH_OK)
    return rc;
  u->wsN = n;
  u->wsD = D;
  return LH_OK;
}

// Input of the first convolution of decoder level `lev` (2: up(b3) ++ e2,
// 1: up(d2b) ++ e1), read in place.
ConvIn dec_in(LhUnet *u, int lev)
{
  ConvIn c;
  c.half0 = 1;
  if (lev == 2) {
    c.a0.p = u->b3; c.a0.ld = 64; c.a0.ofs = 0; c.c0 = 64;
    c.a1.p = u->e2; c.a1.ld = 32; c.a1.ofs = 0;
  } else {
    c.a0.p = u->d2b; c.a0.ld = 32; c.a0.ofs = 0; c.c0 = 32;
    c.a1.p = u->e1; c.a1.ld = 16; c.a1.ofs = 0;
  }
  return c;
}

// forward through the whole net; activations stay in the workspace.  The
// decoder reads its upsampled inputs and the skip tensors in place (no
// concatenation is materialised).  Inference fuses the head into the last
// convolution; `training` keeps every activation the backward pass reads
// (and the padded first-layer input for its weight gradient).
int forward(LhUnet *u, const float *in, int n, int D, bool training, float *outp = nullptr)
{
  if (D % 64)
    return lh_fail(LH_E_ARG, "U-Net input edge must be a multiple of 64 (got %d)", D);
  int rc = ensure_ws(u, n, D);
  if (rc != LH_OK)
    return rc;
  hipStream_t s = u->stream;
  const int D1 = D, D2 = D / 2, D3 = D / 4;
  const int n1 = n * D1 * D1 * D1;
  auto A = [](uint16_t *p, int ld, int ofs = 0) { Act a; a.p = p; a.ld = ld; a.ofs = ofs; return a; };
  if (training)
    pack_input_kernel<<<grid_for((size_t)n1 * 16), 256, 0, s>>>(in, u->x16, (size_t)n1);
#define FW(L, i, o, Dd) if ((rc = conv_fwd(s, u->c[L], i, o, Dd, n, true, true, false, false)) != LH_OK) return rc
  // the encoders' convolutions also write the pooled tensors
  PoolFuse pf1, pf2;
  pf1.p = u->p1;
  pf1.ld = 16;
  pf2.p = u->p2;
  pf2.ld = 32;
  const size_t tiles1 = (size_t)n * (D1 / 16) * (D1 / 4) * (D1 / 8);
  const int perXcd1   = tiles1 % 8 == 0 ? (int)(tiles1 / 8) : 0;
  if (!training && u->fuseFirst) {
    // enc1a + enc1b (+ pool) in one kernel; enc1a's activations are not kept
    ConvLdsArgs a;
    std::memset(&a, 0, sizeof a);
    a.w = u->c[1].wf;
    a.kp = u->c[1].kpf;
    a.bias = u->c[1].b;
    a.cin = 16;
    a.c0 = 16;
    a.y = u->e1;
    a.ldy = 16;
    a.D = D1;
    a.n = n;
    a.relu = 1;
    a.pool = pf1.p;
    a.ldp = pf1.ld;
    a.tilesPerXcd = perXcd1;
    conv_first16_kernel<<<(unsigned)tiles1, 256, 0, s>>>(in, u->w0, u->c[0].b, a);
  } else {
    conv_first_kernel<<<(unsigned)tiles1, 256, 0, s>>>(in, u->w0, u->c[0].b, u->a1, D1, perXcd1);
    if ((rc = conv_fwd(s, u->c[1], plain_in(A(u->a1, 16)), A(u->e1, 16), D1, n, true, true,
             false, false, HeadFuse(), pf1)) != LH_OK)
      return rc;
  }
  FW(2, A(u->p1, 16), A(u->a2, 32), D2);
  if ((rc = conv_fwd(s, u->c[3], plain_in(A(u->a2, 32)), A(u->e2, 32), D2, n, true, true, false,
           false, HeadFuse(), pf2)) != LH_OK)
    return rc;
  FW(4, A(u->p2, 32), A(u->a3, 64), D3);
  FW(5, A(u->a3, 64), A(u->b3, 64), D3);
  FW(6, dec_in(u, 2), A(u->d2a, 32), D2);
  FW(7, A(u->d2a, 32), A(u->d2b, 32), D2);
  FW(8, dec_in(u, 1), A(u->d1a, 16), D1);
#undef FW
  if (training) {
    rc = conv_fwd(s, u->c[9], A(u->d1a, 16), A(u->d1b, 16), D1, n, true, true, false, false);
    if (rc != LH_OK)
      return rc;
    head_kernel<<<grid_for((size_t)n1), 256, 0, s>>>(u->d1b, in, u->hw, u->hb, u->out, (size_t)n1);
  } else {
    HeadFuse h;
    h.in = in;
    h.w = u->hw;
    h.b = u->hb;
    h.out = outp ? outp : u->out;
    Act none;
    none.p = nullptr;
    none.ld = 16;
    none.ofs = 0;
    rc = conv_fwd(s, u->c[9], plain_in(A(u->d1a, 16)), none, D1, n, true, true, false, false, h);
    if (rc != LH_OK)
      return rc;
  }
  UN_HIP(hipGetLastError());
  return LH_OK;
}

int zero_grads(LhUnet *u)
{
  for (auto &L : u->c) {
    UN_HIP(hipMemsetAsync(L.gW, 0, (size_t)L.cout * 27 * L.cin * sizeof(float), u->stream));
    UN_HIP(hipMemsetAsync(L.gb, 0, L.cout * sizeof(float), u->stream));
  }
  UN_HIP(hipMemsetAsync(u->ghw, 0, 16 * sizeof(float), u->stream));
  UN_HIP(hipMemsetAsync(u->ghb, 0, sizeof(float), u->stream));
  return LH_OK;
}

// backward from gy (already in u->gy) into the parameter gradients
int backward(LhUnet *u, const float *in, int n, int D)
{
  hipStream_t s = u->stream;
  const int D1 = D, D2 = D / 2, D3 = D / 4;
  const int n1 = n * D1 * D1 * D1, n2 = n * D2 * D2 * D2, n3 = n * D3 * D3 * D3;
  const size_t v1 = n1, v2 = n2;
  int rc = zero_grads(u);
  if (rc != LH_OK)
    return rc;
  auto A = [](uint16_t *p, int ld, int ofs = 0) { Act a; a.p = p; a.ld = ld; a.ofs = ofs; return a; };
  // g (channel slice of a gradient buffer) *= (y > 0)
  auto relub = [&](uint16_t *g, int ldg, int gofs, const uint16_t *y, int ldy, int C, int nv) {
    relu_bwd_kernel<<<grid_for((size_t)nv * C), 256, 0, s>>>(g, ldg, gofs, y, ldy, 0, C, nv);
  };
#define WG(L, x, g, Dd) if ((rc = conv_wgrad(s, u->c[L], x, g, Dd, n)) != LH_OK) return rc
#define BD(L, g, o, Dd, nv, acc) if ((rc = conv_fwd(s, u->c[L], g, o, Dd, n, false, false, acc, true)) != LH_OK) return rc
  head_bwd_kernel<<<grid_for(v1, 256, 1024), 256, 0, s>>>(u->gy, u->d1b, u->hw, u->g_d1b, u->ghw,
      u->ghb, v1);
  relub(u->g_d1b, 16, 0, u->d1b, 16, 16, n1);
  WG(9, A(u->d1a, 16), A(u->g_d1b, 16), D1);
  BD(9, A(u->g_d1b, 16), A(u->g_d1a, 16), D1, n1, false);
  relub(u->g_d1a, 16, 0, u->d1a, 16, 16, n1);
  WG(8, dec_in(u, 1), A(u->g_d1a, 16), D1);
  BD(8, A(u->g_d1a, 16), A(u->g_cat1, 48), D1, n1, false);
  up_bwd_kernel<<<grid_for(v2 * 32), 256, 0, s>>>(u->g_cat1, 48, 0, u->g_d2b, 32, 0, 32, D2, n);
  relub(u->g_d2b, 32, 0, u->d2b, 32, 32, n2);
  WG(7, A(u->d2a, 32), A(u->g_d2b, 32), D2);
  BD(7, A(u->g_d2b, 32), A(u->g_d2a, 32), D2, n2, false);
  relub(u->g_d2a, 32, 0, u->d2a, 32, 32, n2);
  WG(6, dec_in(u, 2), A(u->g_d2a, 32), D2);
  BD(6, A(u->g_d2a, 32), A(u->g_cat2, 96), D2, n2, false);
  up_bwd_kernel<<<grid_for((size_t)n3 * 64), 256, 0, s>>>(u->g_cat2, 96, 0, u->g_b3, 64, 0, 64, D3, n);
  relub(u->g_b3, 64, 0, u->b3, 64, 64, n3);
  WG(5, A(u->a3, 64), A(u->g_b3, 64), D3);
  BD(5, A(u->g_b3, 64), A(u->g_a3, 64), D3, n3, false);
  relub(u->g_a3, 64, 0, u->a3, 64, 64, n3);
  WG(4, A(u->p2, 32), A(u->g_a3, 64), D3);
  BD(4, A(u->g_a3, 64), A(u->g_p2, 32), D3, n3, false);
  pool_bwd_kernel<<<grid_for(v2 * 32), 256, 0, s>>>(u->g_p2, 32, 0, u->g_cat2, 96, 64, 32, D2, n);
  relub(u->g_cat2, 96, 64, u->e2, 32, 32, n2);
  WG(3, A(u->a2, 32), A(u->g_cat2, 96, 64), D2);
  BD(3, A(u->g_cat2, 96, 64), A(u->g_a2, 32), D2, n2, false);
  relub(u->g_a2, 32, 0, u->a2, 32, 32, n2);
  WG(2, A(u->p1, 16), A(u->g_a2, 32), D2);
  BD(2, A(u->g_a2, 32), A(u->g_p1, 16), D2, n2, false);
  pool_bwd_kernel<<<grid_for(v1 * 16), 256, 0, s>>>(u->g_p1, 16, 0, u->g_cat1, 48, 32, 16, D1, n);
  relub(u->g_cat1, 48, 32, u->e1, 16, 16, n1);
  WG(1, A(u->a1, 16), A(u->g_cat1, 48, 32), D1);
  BD(1, A(u->g_cat1, 48, 32), A(u->g_a1, 16), D1, n1, false);
  relub(u->g_a1, 16, 0, u->a1, 16, 16, n1);
  WG(0, A(u->x16, 16), A(u->g_a1, 16), D1);
#undef WG
#undef BD
  (void)in;
  UN_HIP(hipGetLastError());
  return LH_OK;
}

int adam_all(LhUnet *u, float lr)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  u->step += 1;
  const float c1 = 1.0f - powf(b1, (float)u->step), c2 = 1.0f - powf(b2, (float)u->step);
  hipStream_t s = u->stream;
  for (auto &L : u->c) {
    const size_t nw = (size_t)L.cout * 27 * L.cin;
    adam_kernel<<<grid_for(nw), 256, 0, s>>>(L.W, L.gW, L.mW, L.vW, nw, lr, b1, b2, eps, c1, c2);
    adam_kernel<<<1, 256, 0, s>>>(L.b, L.gb, L.mb, L.vb, L.cout, lr, b1, b2, eps, c1, c2);
  }
  adam_kernel<<<1, 256, 0, s>>>(u->hw, u->ghw, u->mhw, u->vhw, 16, lr, b1, b2, eps, c1, c2);
  adam_kernel<<<1, 256, 0, s>>>(u->hb, u->ghb, u->mhb, u->vhb, 1, lr, b1, b2, eps, c1, c2);
  UN_HIP(hipGetLastError());
  return pack_all(u);
}

} // namespace

namespace {
// The layer of a test convolution: FP32 weights uploaded and packed; every
// device buffer is released on destruction, also after a failed step.
struct TestConv {
  Conv L;
  ~TestConv()
  {
    for (void *p : { (void *)L.W, (void *)L.b, (void *)L.wf, (void *)L.wb, (void *)L.gW,
             (void *)L.gb })
      if (p)
        (void)hipFree(p);
  }
  int init(int cin, int cout, const float *W, const float *b)
  {
    L.cin = cin;
    L.cout = cout;
    L.kpf = 32 * kchunks(cin);
    L.kpb = 32 * kchunks(cout);
    const size_t nw = (size_t)cout * 27 * cin;
    int rc;
    if ((rc = dalloc(&L.W, nw)) || (rc = dalloc(&L.b, cout)) ||
        (rc = dalloc(&L.wf, (size_t)cout * L.kpf)) || (rc = dalloc(&L.wb, (size_t)cin * L.kpb)))
      return rc;
    UN_HIP(hipMemcpy(L.W, W, nw * sizeof(float), hipMemcpyHostToDevice));
    if (b)
      UN_HIP(hipMemcpy(L.b, b, cout * sizeof(float), hipMemcpyHostToDevice));
    pack_weights_kernel<<<grid_for(nw), 256>>>(L.W, cin, cout, L.wf, L.kpf, L.wb, L.kpb);
    UN_HIP(hipGetLastError());
    return LH_OK;
  }
};
} // namespace

extern "C" {

int lhUnetCreate(LhUnet **out, uint64_t seed)
{
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return lh_fail(LH_E_NODEV, "no HIP device visible");
  LhUnet *u = new LhUnet;
  *out = u;
  UN_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  std::mt19937_64 rng(seed);
  for (int i = 0; i < 10; ++i) {
    Conv &L = u->c[i];
    L.cin = kLayers[i][0];
    L.cout = kLayers[i][1];
    L.kpf = 32 * kchunks(L.cin);
    L.kpb = 32 * kchunks(L.cout);
    const size_t nw = (size_t)L.cout * 27 * L.cin;
    int rc;
    if ((rc = dalloc(&L.W, nw)) || (rc = dalloc(&L.b, L.cout)) || (rc = dalloc(&L.gW, nw)) ||
        (rc = dalloc(&L.gb, L.cout)) || (rc = dalloc(&L.mW, nw)) || (rc = dalloc(&L.vW, nw)) ||
        (rc = dalloc(&L.mb, L.cout)) || (rc = dalloc(&L.vb, L.cout)) ||
        (rc = dalloc(&L.wf, (size_t)L.cout * L.kpf)) || (rc = dalloc(&L.wb, (size_t)L.cin * L.kpb)))
      return rc;
    // He-uniform initialisation; the first layer sees one live input channel
    const int fanIn = 27 * (i == 0 ? 1 : L.cin);
    const float bound = sqrtf(6.0f / (float)fanIn);
    std::uniform_real_distribution<float> dist(-bound, bound);
    std::vector<float> h(nw);
    for (size_t k = 0; k < nw; ++k)
      h[k] = (i == 0 && (k % L.cin) != 0) ? 0.0f : dist(rng);
    UN_HIP(hipMemcpy(L.W, h.data(), nw * sizeof(float), hipMemcpyHostToDevice));
  }
  int rc;
  if ((rc = dalloc(&u->w0, 16 * 32)) || (rc = dalloc(&u->hw, 16)) || (rc = dalloc(&u->hb, 1)) ||
      (rc = dalloc(&u->ghw, 16)) ||
      (rc = dalloc(&u->ghb, 1)) || (rc = dalloc(&u->mhw, 16)) || (rc = dalloc(&u->vhw, 16)) ||
      (rc = dalloc(&u->mhb, 1)) || (rc = dalloc(&u->vhb, 1)))
    return rc;
  // the head starts near zero so the untrained net is close to the identity
  std::vector<float> hw(16);
  std::uniform_real_distribution<float> hd(-1e-3f, 1e-3f);
  for (auto &w : hw)
    w = hd(rng);
  UN_HIP(hipMemcpy(u->hw, hw.data(), 16 * sizeof(float), hipMemcpyHostToDevice));
  if ((rc = pack_all(u)) != LH_OK)
    return rc;
  UN_HIP(hipStreamSynchronize(u->stream));
  return LH_OK;
}

int lhUnetDestroy(LhUnet *u)
{
  if (!u)
    return LH_OK;
  if (u->stream)
    (void)hipStreamSynchronize(u->stream);
  for (auto &L : u->c)
    for (void *p : { (void *)L.W, (void *)L.b, (void *)L.gW, (void *)L.gb, (void *)L.mW, (void *)L.vW,
             (void *)L.mb, (void *)L.vb, (void *)L.wf, (void *)L.wb })
      (void)hipFree(p);
  for (void *p : { (void *)u->w0, (void *)u->hw, (void *)u->hb, (void *)u->ghw, (void *)u->ghb, (void *)u->mhw,
           (void *)u->vhw, (void *)u->mhb, (void *)u->vhb })
    (void)hipFree(p);
  for (void *p : u->ws)
    (void)hipFree(p);
  if (u->stream)
    (void)hipStreamDestroy(u->stream);
  delete u;
  return LH_OK;
}

uint64_t lhUnetParamCount(void)
{
  uint64_t s = 17;
  for (auto &l : kLayers)
    s += (uint64_t)l[1] * 27 * l[0] + l[1];
  return s;
}

// Parameters in a flat FP32 host array, layer by layer: W[co][tap][ci] then b,
// then the head weights (16) and bias (1).
int lhUnetGetParams(LhUnet *u, float *host)
{
  if (!u || !host)
    return lh_fail(LH_E_ARG, "lhUnetGetParams: null argument");
  UN_HIP(hipStreamSynchronize(u->stream));
  size_t o = 0;
  for (auto &L : u->c) {
    const size_t nw = (size_t)L.cout * 27 * L.cin;
    UN_HIP(hipMemcpy(host + o, L.W, nw * sizeof(float), hipMemcpyDeviceToHost));
    o += nw;
    UN_HIP(hipMemcpy(host + o, L.b, L.cout * sizeof(float), hipMemcpyDeviceToHost));
    o += L.cout;
  }
  UN_HIP(hipMemcpy(host + o, u->hw, 16 * sizeof(float), hipMemcpyDeviceToHost));
  UN_HIP(hipMemcpy(host + o + 16, u->hb, sizeof(float), hipMemcpyDeviceToHost));
  return LH_OK;
}

int lhUnetSetParams(LhUnet *u, const float *host)
{
  if (!u || !host)
    return lh_fail(LH_E_ARG, "lhUnetSetParams: null argument");
  UN_HIP(hipStreamSynchronize(u->stream));
  size_t o = 0;
  for (auto &L : u->c) {
    const size_t nw = (size_t)L.cout * 27 * L.cin;
    UN_HIP(hipMemcpy(L.W, host + o, nw * sizeof(float), hipMemcpyHostToDevice));
    o += nw;
    UN_HIP(hipMemcpy(L.b, host + o, L.cout * sizeof(float), hipMemcpyHostToDevice));
    o += L.cout;
  }
  UN_HIP(hipMemcpy(u->hw, host + o, 16 * sizeof(float), hipMemcpyHostToDevice));
  UN_HIP(hipMemcpy(u->hb, host + o + 16, sizeof(float), hipMemcpyHostToDevice));
  int rc = pack_all(u);
  if (rc != LH_OK)
    return rc;
  UN_HIP(hipStreamSynchronize(u->stream));
  return LH_OK;
}

// Gradients of the last backward pass, same flat layout as the parameters.
int lhUnetGetGrads(LhUnet *u, float *host)
{
  if (!u || !host)
    return lh_fail(LH_E_ARG, "lhUnetGetGrads: null argument");
  UN_HIP(hipStreamSynchronize(u->stream));
  size_t o = 0;
  for (auto &L : u->c) {
    const size_t nw = (size_t)L.cout * 27 * L.cin;
    UN_HIP(hipMemcpy(host + o, L.gW, nw * sizeof(float), hipMemcpyDeviceToHost));
    o += nw;
    UN_HIP(hipMemcpy(host + o, L.gb, L.cout * sizeof(float), hipMemcpyDeviceToHost));
    o += L.cout;
  }
  UN_HIP(hipMemcpy(host + o, u->ghw, 16 * sizeof(float), hipMemcpyDeviceToHost));
  UN_HIP(hipMemcpy(host + o + 16, u->ghb, sizeof(float), hipMemcpyDeviceToHost));
  return LH_OK;
}

// Inference: dOut (device, n*D^3 floats) = net(dIn); D a multiple of 64.
int lhUnetSetFusion(LhUnet *u, int firstBlock)
{
  if (!u)
    return lh_fail(LH_E_ARG, "lhUnetSetFusion: null network");
  u->fuseFirst = firstBlock != 0;
  return LH_OK;
}

int lhUnetForward(LhUnet *u, const float *dIn, float *dOut, int n, int D)
{
  if (!u || !dIn || !dOut || n < 1)
    return lh_fail(LH_E_ARG, "lhUnetForward: bad arguments");
  // the fused head writes the caller's buffer (each voxel reads its input
  // before writing its output, so dOut may be dIn)
  int rc = forward(u, dIn, n, D, false, dOut);
  if (rc != LH_OK)
    return rc;
  UN_HIP(hipStreamSynchronize(u->stream));
  return LH_OK;
}

// Loss and gradients for inputs dIn, targets dTarget (device, n*D^3 floats);
// with lr > 0 also one Adam step.  *loss = mean squared error.
int lhUnetTrainStep(LhUnet *u, const float *dIn, const float *dTarget, int n, int D, float lr,
    float *loss)
{
  if (!u || !dIn || !dTarget || n < 1)
    return lh_fail(LH_E_ARG, "lhUnetTrainStep: bad arguments");
  int rc = forward(u, dIn, n, D, true);
  if (rc != LH_OK)
    return rc;
  const size_t nv = (size_t)n * D * D * D;
  UN_HIP(hipMemsetAsync(u->loss, 0, sizeof(float), u->stream));
  mse_kernel<<<grid_for(nv, 256, 1024), 256, 0, u->stream>>>(u->out, dTarget, u->gy, nv, u->loss);
  UN_HIP(hipGetLastError());
  if ((rc = backward(u, dIn, n, D)) != LH_OK)
    return rc;
  if (lr > 0.0f && (rc = adam_all(u, lr)) != LH_OK)
    return rc;
  float l = 0.0f;
  UN_HIP(hipMemcpyAsync(&l, u->loss, sizeof(float), hipMemcpyDeviceToHost, u->stream));
  UN_HIP(hipStreamSynchronize(u->stream));
  if (loss)
    *loss = l;
  return LH_OK;
}

// Single convolution entry for tests: x [n*D^3][cin] bf16 (device),
// W [cout][27][cin] FP32 host, bias FP32 host or null, y [n*D^3][cout] bf16.
// backward = 1 runs the backward-data pass (x holds cout channels, y cin).
int lhConv3dTest(const uint16_t *dx, int cin, int cout, const float *W, const float *b,
    uint16_t *dy, int n, int D, int relu, int backward)
{
  TestConv T;
  int rc = T.init(cin, cout, W, b);
  if (rc != LH_OK)
    return rc;
  Act in, out;
  in.p = (uint16_t *)dx; in.ld = backward ? cout : cin; in.ofs = 0;
  out.p = dy; out.ld = backward ? cin : cout; out.ofs = 0;
  rc = conv_fwd(0, T.L, in, out, D, n, relu && !backward, b != nullptr && !backward,
      false, backward != 0);
  UN_HIP(hipDeviceSynchronize());
  return rc;
}

// Decoder-form convolution entry for tests: input channels [0, c0) read from
// dx0 ([n*Ds^3][c0] bf16, Ds = D/2 with half0 = 1: nearest upsampling read in
// place), channels [c0, cin) from dx1 ([n*D^3][cin-c0] bf16); y [n*D^3][cout].
int lhConv3dSplitTest(const uint16_t *dx0, int c0, int half0, const uint16_t *dx1, int cin,
    int cout, const float *W, const float *b, uint16_t *dy, int n, int D, int relu)
{
  if (c0 <= 0 || c0 >= cin)
    return lh_fail(LH_E_ARG, "conv split test: need 0 < c0 < cin (got %d, %d)", c0, cin);
  TestConv T;
  int rc = T.init(cin, cout, W, b);
  if (rc != LH_OK)
    return rc;
  ConvIn in;
  in.a0.p = (uint16_t *)dx0; in.a0.ld = c0; in.a0.ofs = 0;
  in.half0 = half0 ? 1 : 0;
  in.c0 = c0;
  in.a1.p = (uint16_t *)dx1; in.a1.ld = cin - c0; in.a1.ofs = 0;
  Act out;
  out.p = dy; out.ld = cout; out.ofs = 0;
  rc = conv_fwd(0, T.L, in, out, D, n, relu != 0, b != nullptr, false, false);
  UN_HIP(hipDeviceSynchronize());
  return rc;
}

// Weight-gradient entry for tests: dW [cout][27][cin] FP32 host output.
int lhConv3dWgradTest(const uint16_t *dx, int cin, const uint16_t *dg, int cout, float *dW,
    int n, int D)
{
  TestConv T;
  Conv &L = T.L;
  L.cin = cin;
  L.cout = cout;
  const size_t nw = (size_t)cout * 27 * cin;
  int rc;
  if ((rc = dalloc(&L.gW, nw)) || (rc = dalloc(&L.gb, cout)))
    return rc;
  Act x, g;
  x.p = (uint16_t *)dx; x.ld = cin; x.ofs = 0;
  g.p = (uint16_t *)dg; g.ld = cout; g.ofs = 0;
  if ((rc = conv_wgrad(0, L, x, g, D, n)) != LH_OK)
    return rc;
  UN_HIP(hipDeviceSynchronize());
  UN_HIP(hipMemcpy(dW, L.gW, nw * sizeof(float), hipMemcpyDeviceToHost));
  return LH_OK;
}

// Device memory helpers for callers without another allocator.
int lhDeviceAlloc(void **p, uint64_t bytes)
{
  UN_HIP(hipMalloc(p, bytes));
  return LH_OK;
}
int lhDeviceFree(void *p)
{
  UN_HIP(hipFree(p));
  return LH_OK;
}
int lhCopyToDevice(void *dst, const void *src, uint64_t bytes)
{
  UN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return LH_OK;
}
int lhCopyToHost(void *dst, const void *src, uint64_t bytes)
{
  UN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return LH_OK;
}

} // extern "C"
