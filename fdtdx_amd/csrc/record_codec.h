// Narrow float formats of the interface record (interfaces/modules.py:98-161 DtypeConversion):
// float32 -> format by round to nearest, ties to even, on the format's lattice extended past its
// largest finite value, then overflow to inf (IEEE-like formats) or NaN (formats without inf);
// the FNUZ formats have no negative zero and one NaN code (0x80). Integer arithmetic only, so the
// result does not depend on the float unit's rounding mode. The functions compile for the host
// too, so the codec can be checked exhaustively on the CPU.
#pragma once

#include <cmath>
#include <cstdint>

#include "engine_common.h"

namespace fdtdx {

#define FDTDX_HD __host__ __device__ __forceinline__

FDTDX_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
FDTDX_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// KIND: 0 IEEE-like (inf, NaN), 1 "fn" (no inf, S.1..1.1..1 is NaN), 2 "fnuz" (no inf, 0x80 NaN)
template <int EB, int MB, int BIAS, int KIND>
FDTDX_HD uint32_t enc_small(float x) {
  constexpr int NB = 1 + EB + MB;
  constexpr uint32_t SB = 1u << (NB - 1);
  constexpr uint32_t INF = ((1u << EB) - 1u) << MB;
  constexpr int EMIN = 1 - BIAS;
  constexpr uint32_t TOP = KIND == 0 ? (1u << EB) - 2u : (1u << EB) - 1u;
  constexpr uint32_t FRAC = KIND == 1 ? (1u << MB) - 2u : (1u << MB) - 1u;
  const uint32_t bits = f2u(x);
  const uint32_t sbit = (bits >> 31) ? SB : 0u;
  const uint32_t a = bits & 0x7fffffffu;
  const uint32_t nan_code = KIND == 0 ? (sbit | INF | (1u << (MB - 1))) : (KIND == 2 ? SB : (sbit | (SB - 1u)));
  if (a > 0x7f800000u) return nan_code;
  if (a == 0x7f800000u) return KIND == 0 ? (sbit | INF) : nan_code;
  const int e32 = (int)(a >> 23);
  const uint32_t mant = e32 == 0 ? (a & 0x7fffffu) : ((a & 0x7fffffu) | 0x800000u);
  const int exp2 = e32 == 0 ? -126 : e32 - 127;
  const int sh = 23 - MB + (EMIN > exp2 ? EMIN - exp2 : 0);  // >= 13
  uint32_t q = 0u;
  if (sh <= 24) {
    q = mant >> sh;
    const uint32_t rem = mant & ((1u << sh) - 1u);
    const uint32_t half = 1u << (sh - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
  }
  int qexp = exp2 > EMIN ? exp2 : EMIN;
  if (q >= (1u << (MB + 1))) {
    q >>= 1;
    ++qexp;
  }
  uint32_t ex = 0u, fr = q;
  if (q >= (1u << MB)) {
    ex = (uint32_t)(qexp - EMIN + 1);
    fr = q - (1u << MB);
  }
  const bool over = ex > TOP || (ex == TOP && fr > FRAC);
  const uint32_t mag = (ex << MB) | fr;
  if (KIND == 0) return over ? (sbit | INF) : (sbit | mag);
  if (KIND == 2) return over ? SB : (mag == 0u ? 0u : (sbit | mag));
  return over ? (sbit | (SB - 1u)) : (sbit | mag);
}

template <int EB, int MB, int BIAS, int KIND>
FDTDX_HD float dec_small(uint32_t c) {
  constexpr int NB = 1 + EB + MB;
  constexpr uint32_t SB = 1u << (NB - 1);
  constexpr uint32_t EMAX = (1u << EB) - 1u;
  constexpr uint32_t FMAX = (1u << MB) - 1u;
  const bool neg = (c & SB) != 0u;
  const uint32_t ex = (c >> MB) & EMAX;
  const uint32_t fr = c & FMAX;
  if (KIND == 0 && ex == EMAX) return fr ? u2f(0x7fc00000u) : (neg ? -u2f(0x7f800000u) : u2f(0x7f800000u));
  if (KIND == 2 && c == SB) return u2f(0x7fc00000u);
  if (KIND == 1 && ex == EMAX && fr == FMAX) return u2f(0x7fc00000u);
  float mag;
  if (ex == 0u) {
    mag = ldexpf((float)fr, 1 - BIAS - MB);  // exact: a small integer times a power of two
  } else {
    mag = u2f(((uint32_t)((int)ex - BIAS + 127) << 23) | (fr << (23 - MB)));
  }
  return neg ? -mag : mag;
}

FDTDX_HD int rec_width(int fmt) {
  return fmt == FDTDX_REC_F32 ? 4 : (fmt <= FDTDX_REC_F16 ? 2 : 1);
}

FDTDX_HD void rec_store(char* base, int fmt, long long idx, float v) {
  switch (fmt) {
    case FDTDX_REC_BF16: ((uint16_t*)base)[idx] = (uint16_t)enc_small<8, 7, 127, 0>(v); break;
    case FDTDX_REC_F16: ((uint16_t*)base)[idx] = (uint16_t)enc_small<5, 10, 15, 0>(v); break;
    case FDTDX_REC_E4M3FN: ((uint8_t*)base)[idx] = (uint8_t)enc_small<4, 3, 7, 1>(v); break;
    case FDTDX_REC_E4M3FNUZ: ((uint8_t*)base)[idx] = (uint8_t)enc_small<4, 3, 8, 2>(v); break;
    case FDTDX_REC_E5M2: ((uint8_t*)base)[idx] = (uint8_t)enc_small<5, 2, 15, 0>(v); break;
    case FDTDX_REC_E5M2FNUZ: ((uint8_t*)base)[idx] = (uint8_t)enc_small<5, 2, 16, 2>(v); break;
    default: ((float*)base)[idx] = v; break;
  }
}

FDTDX_HD float rec_load(const char* base, int fmt, long long idx) {
  switch (fmt) {
    case FDTDX_REC_BF16: return u2f((uint32_t)((const uint16_t*)base)[idx] << 16);
    case FDTDX_REC_F16: return dec_small<5, 10, 15, 0>(((const uint16_t*)base)[idx]);
    case FDTDX_REC_E4M3FN: return dec_small<4, 3, 7, 1>(((const uint8_t*)base)[idx]);
    case FDTDX_REC_E4M3FNUZ: return dec_small<4, 3, 8, 2>(((const uint8_t*)base)[idx]);
    case FDTDX_REC_E5M2: return dec_small<5, 2, 15, 0>(((const uint8_t*)base)[idx]);
    case FDTDX_REC_E5M2FNUZ: return dec_small<5, 2, 16, 2>(((const uint8_t*)base)[idx]);
    default: return ((const float*)base)[idx];
  }
}

// float32 -> format -> float32
FDTDX_HD float rec_round(float v, int fmt) {
  switch (fmt) {
    case FDTDX_REC_BF16: return u2f(enc_small<8, 7, 127, 0>(v) << 16);
    case FDTDX_REC_F16: return dec_small<5, 10, 15, 0>(enc_small<5, 10, 15, 0>(v));
    case FDTDX_REC_E4M3FN: return dec_small<4, 3, 7, 1>(enc_small<4, 3, 7, 1>(v));
    case FDTDX_REC_E4M3FNUZ: return dec_small<4, 3, 8, 2>(enc_small<4, 3, 8, 2>(v));
    case FDTDX_REC_E5M2: return dec_small<5, 2, 15, 0>(enc_small<5, 2, 15, 0>(v));
    case FDTDX_REC_E5M2FNUZ: return dec_small<5, 2, 16, 2>(enc_small<5, 2, 16, 2>(v));
    default: return v;
  }
}

// LinearReconstructEveryK.decompress (time_filter.py:233-257): prev + f * (next - prev), in
// float32 (fmt = FDTDX_REC_F32) or with every operation rounded to the narrow format.
FDTDX_HD float rec_interp(float prev, float next, float f, int fmt) {
  if (fmt == FDTDX_REC_F32) return prev + f * (next - prev);
  const float ff = rec_round(f, fmt);
  const float d = rec_round(next - prev, fmt);
  return rec_round(prev + rec_round(ff * d, fmt), fmt);
}

#undef FDTDX_HD

}  // namespace fdtdx
