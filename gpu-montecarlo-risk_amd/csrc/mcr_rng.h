// mcr_rng.h -- the counter-based normal stream every kernel draws from.
//
// Replaces the per-thread cuRAND XORWOW state of the reference GPU model
// (src/05-GPU/monte_carlos_cuda.cu:30-31 curand_init(seed, idx, 0) and :45-47
// curand_normal_double) and the sequential GSL MT19937 stream of the CPU models
// (src/02-C-serial/monte_carlo_serial.c:46,54,67-69).  Nothing of cuRAND is
// reproduced: the generator is Philox4x32-10 as published by Salmon, Moraes,
// Dror and Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3" (SC'11), and the
// normal transform is Box-Muller with this file's own polynomial kernels.
//
// Stream definition (the contract a host mirror must follow bit for bit):
//   key     = (seed & 0xffffffff, seed >> 32)
//   counter = (block, 0, path & 0xffffffff, path >> 32)
//   (x0,x1,x2,x3) = Philox4x32-10(counter, key)
//   z[4*block+0], z[4*block+1] = BoxMuller(x0, x1)
//   z[4*block+2], z[4*block+3] = BoxMuller(x2, x3)
//   BoxMuller(a, b):  u = ((a >> 8) + 1) * 2^-24  in [2^-24, 1]
//                     r = sqrt(-2 ln u),   theta = 2*pi * b / 2^32
//                     (r cos theta, r sin theta)
// Every step uses only integer ops and IEEE fp32 add/mul/fma (no libm, no
// hardware transcendental), so a host compiled with fp contraction off
// reproduces each normal exactly.  Must be compiled with -ffp-contract=off.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MCR_HD __host__ __device__ __forceinline__
#else
#define MCR_HD static inline
#endif

namespace mcr {

// Philox4x32 round constants (Salmon et al. SC'11, Table 2 / Random123).
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;  // golden ratio
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;  // sqrt(3) - 1

struct Ctr4 {
  uint32_t x, y, z, w;
};

// a ^ b ^ c in one v_bitop3_b32 on gfx950 (truth table 0x96); plain xors on the host
MCR_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

MCR_HD Ctr4 philox4x32_10(Ctr4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
    Ctr4 n;
    n.x = xor3((uint32_t)(p1 >> 32), c.y, k0);
    n.y = (uint32_t)p1;
    n.z = xor3((uint32_t)(p0 >> 32), c.w, k1);
    n.w = (uint32_t)p0;
    c = n;
  }
  return c;
}

MCR_HD float as_f32(uint32_t b) { return __builtin_bit_cast(float, b); }
MCR_HD uint32_t as_u32(float f) { return __builtin_bit_cast(uint32_t, f); }

// Near-minimax coefficients (Lawson-reweighted least squares on Chebyshev
// nodes, rounded to fp32; the fit script is derive_normal_poly.py beside this
// file).  ln(1+f) = f - f^2/2 + f^3 P(f) on f in [sqrt(1/2)-1, sqrt(2)-1],
// max relative error 3.5e-8.
constexpr float kLogP0 = 0x1.5556d8p-2f;
constexpr float kLogP1 = -0x1.000382p-2f;
constexpr float kLogP2 = 0x1.98d7f0p-3f;
constexpr float kLogP3 = -0x1.53824cp-3f;
constexpr float kLogP4 = 0x1.31747cp-3f;
constexpr float kLogP5 = -0x1.2432a4p-3f;
constexpr float kLogP6 = 0x1.645d88p-4f;
constexpr float kLn2Hi = 0x1.62e400p-1f;  // 16 significant bits: e*kLn2Hi exact
constexpr float kLn2Lo = 0x1.7f7d1cp-20f;
// sin(2 pi a) = a S(a^2), cos(2 pi a) = 1 + a^2 C(a^2) on a in [-1/8, 1/8]
// (turns); max relative error 3.6e-8 (sin), absolute 6.7e-9 (cos).
constexpr float kSin0 = 0x1.921fb6p+2f;
constexpr float kSin1 = -0x1.4abbbap+5f;
constexpr float kSin2 = 0x1.465ec8p+6f;
constexpr float kSin3 = -0x1.2d9bf2p+6f;
constexpr float kCos1 = -0x1.3bd3ccp+4f;
constexpr float kCos2 = 0x1.03c1dep+6f;
constexpr float kCos3 = -0x1.55c664p+6f;
constexpr float kCos4 = 0x1.d9f7bep+5f;

// -ln(u) for u = ((a >> 8) + 1) * 2^-24.  Range reduction u = 2^e m with
// m in [sqrt(1/2), sqrt(2)) done on the fp32 bit pattern.
MCR_HD float neg_log_u24(uint32_t a) {
  const float u = __builtin_fmaf((float)(a >> 8), 0x1p-24f, 0x1p-24f);  // exact
  const int32_t ix = (int32_t)as_u32(u) - 0x3f3504f3;
  const int32_t e = ix >> 23;
  const float m = as_f32((uint32_t)(ix & 0x007fffff) + 0x3f3504f3u);
  const float f = m - 1.0f;  // exact (Sterbenz)
  float p = kLogP6;
  p = __builtin_fmaf(p, f, kLogP5);
  p = __builtin_fmaf(p, f, kLogP4);
  p = __builtin_fmaf(p, f, kLogP3);
  p = __builtin_fmaf(p, f, kLogP2);
  p = __builtin_fmaf(p, f, kLogP1);
  p = __builtin_fmaf(p, f, kLogP0);
  const float q = __builtin_fmaf(p, f, -0.5f);
  const float t = f * f;
  const float l1p = __builtin_fmaf(t, q, f);
  const float ef = (float)e;
  const float lo = __builtin_fmaf(ef, kLn2Lo, l1p);
  return -__builtin_fmaf(ef, kLn2Hi, lo);
}

// sqrt(2 h) for h in [0, 17]: bit-pattern seed for 1/sqrt(2h) (Lomont's
// constant 0x5f375a86 with the factor 2 folded into the exponent; 3.4% worst
// error), three Newton steps g <- g (3/2 - h g^2) (error 1.7e-3, 4.6e-6, then
// fp32 rounding), then 2h * g.  The iteration runs on hn = h + 2^-60, which
// equals h for every nonzero h here (h >= 5.9e-8) and keeps g finite at h = 0,
// where the result is 0.
MCR_HD float sqrt_twice(float h) {
  const float hn = h + 0x1p-60f;
  float g = as_f32(0x5ef75a86u - (as_u32(hn) >> 1));
#pragma unroll
  for (int it = 0; it < 3; ++it) {
    float t = g * g;
    t = __builtin_fmaf(-hn, t, 1.5f);
    g = g * t;
  }
  return (h + h) * g;
}

// cos and sin of 2*pi*b/2^32.  The nearest quarter turn q comes from the top
// bits of b, the signed remainder (30 bits) is the reduced argument in turns.
MCR_HD void sincos_turn(uint32_t b, float* c_out, float* s_out) {
  const int32_t rem = (int32_t)(b << 2) >> 2;
  const uint32_t q = (b - (uint32_t)rem) >> 30;
  const float a = (float)rem * 0x1p-32f;
  const float t = a * a;
  float s = __builtin_fmaf(kSin3, t, kSin2);
  s = __builtin_fmaf(s, t, kSin1);
  s = __builtin_fmaf(s, t, kSin0);
  s = s * a;
  float c = __builtin_fmaf(kCos4, t, kCos3);
  c = __builtin_fmaf(c, t, kCos2);
  c = __builtin_fmaf(c, t, kCos1);
  c = __builtin_fmaf(c, t, 1.0f);
  const float sw_s = (q & 1u) ? c : s;
  const float sw_c = (q & 1u) ? s : c;
  *s_out = as_f32(as_u32(sw_s) ^ ((q & 2u) << 30));
  *c_out = as_f32(as_u32(sw_c) ^ (((q + 1u) & 2u) << 30));
}

MCR_HD void box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
  const float r = sqrt_twice(neg_log_u24(a));
  float c, s;
  sincos_turn(b, &c, &s);
  *z0 = r * c;
  *z1 = r * s;
}

struct Key2 {
  uint32_t k0, k1;
};
MCR_HD Key2 key_of(uint64_t seed) { return Key2{(uint32_t)seed, (uint32_t)(seed >> 32)}; }

// Normals z[4*blk .. 4*blk+3] of one path.
MCR_HD void normals4(Key2 key, uint64_t path, uint32_t blk, float z[4]) {
  const Ctr4 x = philox4x32_10(Ctr4{blk, 0u, (uint32_t)path, (uint32_t)(path >> 32)}, key.k0, key.k1);
  box_muller(x.x, x.y, &z[0], &z[1]);
  box_muller(x.z, x.w, &z[2], &z[3]);
}

// normals4's z[0], z[1] alone (the same bits): for a path's last block when
// n leaves its second pair unused.
MCR_HD void normals2(Key2 key, uint64_t path, uint32_t blk, float z[2]) {
  const Ctr4 x = philox4x32_10(Ctr4{blk, 0u, (uint32_t)path, (uint32_t)(path >> 32)}, key.k0, key.k1);
  box_muller(x.x, x.y, &z[0], &z[1]);
}

}  // namespace mcr
