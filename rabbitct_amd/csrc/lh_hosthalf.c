/* lh_hosthalf.c -- projections converted to half precision on the host, for
 * uploads of half the bytes (lhReconstructHostHalf).  No counterpart in the
 * reference: its projections cross into the module as float32 images
 * (rabbitCt.h:20-27).
 *
 * The conversion is the one the device staging applies to FP32 uploads
 * (stage_element in lh_device.hip): FP16 scales each view by 4^-m, m chosen
 * so that max|p| 4^-m lies in (4096, 16384] (forward differences of up to
 * twice that stay finite), then rounds to nearest even; BF16 is unscaled
 * (m = 0) and rounds to nearest even.  Both conversions are written out bit by
 * bit here so the host and the device round alike; NaN inputs become quiet
 * NaNs.
 */
#include <stdint.h>
#include <string.h>

#include "lolahip.h"

static uint32_t bits_of(float f)
{
  uint32_t x;
  memcpy(&x, &f, sizeof x);
  return x;
}

/* float -> IEEE binary16, round to nearest even (overflow to infinity) */
static uint16_t to_half(float f)
{
  const uint32_t x    = bits_of(f);
  const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  const uint32_t ax   = x & 0x7fffffffu;
  if (ax > 0x7f800000u)                     /* NaN */
    return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
  if (ax >= 0x477ff000u)                    /* >= 65520: rounds to infinity */
    return (uint16_t)(sign | 0x7c00u);
  if (ax >= 0x38800000u) {                  /* normal: rebias, round on bit 13 */
    const uint32_t r = ax - ((127u - 15u) << 23);
    return (uint16_t)(sign | ((r + 0xfffu + ((r >> 13) & 1u)) >> 13));
  }
  if (ax < 0x33000000u)                     /* below 2^-25: zero */
    return sign;
  /* subnormal: units of 2^-24 */
  const uint32_t e     = ax >> 23;
  const uint32_t mant  = (ax & 0x7fffffu) | 0x800000u;
  const uint32_t shift = 126u - e;          /* 14 .. 24 */
  uint32_t q           = mant >> shift;
  const uint32_t rem   = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (q & 1u)))
    ++q;
  return (uint16_t)(sign | q);
}

/* float -> bfloat16, round to nearest even */
static uint16_t to_bfloat(float f)
{
  uint32_t b = bits_of(f);
  if ((b & 0x7fffffffu) > 0x7f800000u)
    return (uint16_t)((b >> 16) | 0x40u);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

/* the storage exponent of a view whose largest |pixel| is mx (FP16 only) */
static int storage_exponent(int precision, float mx)
{
  if (precision != LH_PREC_F16 || !(mx > 0.0f))
    return 0;
  int m   = 0;
  float s = 1.0f;
  while (mx / s > 16384.0f && m < 60) {
    s *= 4.0f;
    ++m;
  }
  while (mx / s <= 4096.0f && m > -60) {
    s *= 0.25f;
    --m;
  }
  return m;
}

int lhHostHalfViews(int precision, const float *images, uint32_t count, uint32_t W, uint32_t H,
    uint16_t *halves, int32_t *exponents)
{
  if ((precision != LH_PREC_F16 && precision != LH_PREC_BF16) || !images || !halves ||
      !exponents || W == 0 || H == 0)
    return LH_E_ARG;
  const size_t px = (size_t)W * H;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t v = 0; v < (int64_t)count; ++v) {
    const float *img = images + (size_t)v * px;
    uint16_t *out    = halves + (size_t)v * px;
    float mx         = 0.0f;
    for (size_t i = 0; i < px; ++i) {
      const float a = img[i] < 0.0f ? -img[i] : img[i];
      mx            = a > mx ? a : mx;
    }
    const int m = storage_exponent(precision, mx);
    exponents[v] = m;
    /* 4^-m as a float (exact; |m| <= 60 keeps it normal) */
    float is = 1.0f;
    for (int k = 0; k < (m > 0 ? m : -m); ++k)
      is *= m > 0 ? 0.25f : 4.0f;
    if (precision == LH_PREC_F16)
      for (size_t i = 0; i < px; ++i)
        out[i] = to_half(img[i] * is);
    else
      for (size_t i = 0; i < px; ++i)
        out[i] = to_bfloat(img[i]);
  }
  return LH_OK;
}
