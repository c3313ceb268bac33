// common.h -- device helpers shared by the node kernels of the
// raw-development path (gfx950, wave64).
//
// Texture semantics of the reference (qvk/qvk.c:596-611, LINEAR +
// MIRRORED_REPEAT; graph-run-nodes-allocate.h:39-96, R16_UNORM / R16F /
// RGBA16F / RGBA32F) are expressed with plain loads:
//   texelFetch     -> fetch_*  (out of bounds reads 0)
//   texture() at a texel centre -> mirrored index + load
//   texture() elsewhere         -> two/four mirrored taps + lerp
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define DEV static __device__ __forceinline__

// device view of one connector (pitch in texels, plane stride in texels)
struct dimg
{
  void    *p;
  int      wd, ht;
  int      pitch;
  int      chan;
  long     pstride;
  int      fmt;       // 1 ui16 (UNORM), 2 f16, 3 f32, 4 FP64 accumulators (the VKDT_FMT_* codes)
};
#define DIMG_UI16 1
#define DIMG_F16  2
#define DIMG_F32  3
#define DIMG_ATOM 4

DEV float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
DEV float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
DEV float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
DEV float smoothstepf(float e0, float e1, float x)
{
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// MIRRORED_REPEAT index (period 2n, edge texels repeated once)
DEV int mirror(int i, int n)
{
  if((unsigned)i < (unsigned)n) return i;
  int t = i % (2 * n);
  if(t < 0) t += 2 * n;
  return t < n ? t : 2 * n - 1 - t;
}

// hardware transcendentals (v_exp_f32, v_log_f32, v_rcp_f32, v_sqrt_f32,
// v_rsq_f32: about 1 ulp).  The reference's GLSL exp/pow/division carry no
// stronger guarantee (SPIR-V precision of exp2/log2 is 3 ulp); every use
// is covered by the per-node parity tests against the float oracle.
DEV float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
DEV float flog2(float x) { return __builtin_amdgcn_logf(x); }
DEV float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DEV float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
// x^y for x >= 0 (0^y = 0 for y > 0)
DEV float fpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }

DEV float h2f(__half h) { return __half2float(h); }
DEV __half f2h(float f) { return __float2half_rn(f); }

// ---- loads -------------------------------------------------------------

DEV float ld_u16(const dimg &b, int x, int y)
{ // R16_UNORM: the float32 value v/65535, correctly rounded.  The product
  // with the rounded reciprocal and one FMA correction step equal the IEEE
  // quotient for each of the 65536 codes (an exhaustive check), in three
  // instructions instead of the division's ten
  const float v = (float)((const uint16_t *)b.p)[(long)y * b.pitch + x];
  const float r = 1.0f / 65535.0f;
  const float q = v * r;
  return __builtin_fmaf(__builtin_fmaf(-q, 65535.0f, v), r, q);
}

DEV void st_u16(const dimg &b, int x, int y, float v)
{ // float -> R16_UNORM store: round(clamp(v,0,1) * 65535), ties to even
  const float c = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;   // NaN -> 0
  ((uint16_t *)b.p)[(long)y * b.pitch + x] = (uint16_t)__builtin_rintf(c * 65535.0f);
}

DEV float ld_h1(const dimg &b, int x, int y, int pl = 0)
{
  return h2f(((const __half *)b.p)[pl * b.pstride + (long)y * b.pitch + x]);
}

DEV float4 ld_h4(const dimg &b, int x, int y)
{ // one 8-byte load of an rgba f16 texel
  const uint2 u = *(const uint2 *)((const __half *)b.p + 4 * ((long)y * b.pitch + x));
  __half2 a = *(const __half2 *)&u.x, c = *(const __half2 *)&u.y;
  return make_float4(__low2float(a), __high2float(a), __low2float(c), __high2float(c));
}

DEV void st_h4(const dimg &b, int x, int y, float4 v)
{
  __half2 a = __floats2half2_rn(v.x, v.y), c = __floats2half2_rn(v.z, v.w);
  uint2 u;
  u.x = *(const uint32_t *)&a;
  u.y = *(const uint32_t *)&c;
  *(uint2 *)((__half *)b.p + 4 * ((long)y * b.pitch + x)) = u;
}

DEV void st_h1(const dimg &b, int x, int y, float v, int pl = 0)
{
  ((__half *)b.p)[pl * b.pstride + (long)y * b.pitch + x] = f2h(v);
}

DEV void st_f4(const dimg &b, int x, int y, float4 v)
{
  *(float4 *)((float *)b.p + 4 * ((long)y * b.pitch + x)) = v;
}

// texelFetch with zero outside
DEV float4 fetch_h4(const dimg &b, int x, int y)
{
  if(x < 0 || y < 0 || x >= b.wd || y >= b.ht) return make_float4(0.f, 0.f, 0.f, 0.f);
  return ld_h4(b, x, y);
}
DEV float fetch_h1(const dimg &b, int x, int y, int pl = 0)
{
  if(x < 0 || y < 0 || x >= b.wd || y >= b.ht) return 0.0f;
  return ld_h1(b, x, y, pl);
}
DEV float fetch_u16(const dimg &b, int x, int y)
{
  if(x < 0 || y < 0 || x >= b.wd || y >= b.ht) return 0.0f;
  return ld_u16(b, x, y);
}

// mirror() without branches for coordinates within one extent of the image
// (-n <= i < 2n, where it equals mirror()); beyond that it clamps into the
// image, so that a load stays in bounds wherever its value is discarded.
// Branch-free addressing lets a kernel keep several loads in flight.
DEV int mirror_near(int i, int n)
{
  int r = i < 0 ? -i - 1 : i;
  r = r >= n ? 2 * n - 1 - r : r;
  return min(max(r, 0), n - 1);
}

// texture() at a texel centre
DEV float4 centre_h4(const dimg &b, int x, int y) { return ld_h4(b, mirror(x, b.wd), mirror(y, b.ht)); }
DEV float  centre_h1(const dimg &b, int x, int y, int pl = 0) { return ld_h1(b, mirror(x, b.wd), mirror(y, b.ht), pl); }

// a single-channel mosaic stored as f16 or ui16 (a hilite output is UNORM16)
DEV float ld_m(const dimg &b, int x, int y) { return b.fmt == DIMG_UI16 ? ld_u16(b, x, y) : ld_h1(b, x, y); }
DEV float fetch_m(const dimg &b, int x, int y)
{
  if(x < 0 || y < 0 || x >= b.wd || y >= b.ht) return 0.0f;
  return ld_m(b, x, y);
}
DEV float centre_m(const dimg &b, int x, int y) { return ld_m(b, mirror(x, b.wd), mirror(y, b.ht)); }
DEV float  centre_u16(const dimg &b, int x, int y) { return ld_u16(b, mirror(x, b.wd), mirror(y, b.ht)); }

// the same with the format fixed at compile time (U16: R16_UNORM, else f16),
// for the hot kernels that are instantiated per mosaic format
template<int U16> DEV float ld_mt(const dimg &b, int x, int y) { return U16 ? ld_u16(b, x, y) : ld_h1(b, x, y); }
template<int U16> DEV float fetch_mt(const dimg &b, int x, int y)
{
  if(x < 0 || y < 0 || x >= b.wd || y >= b.ht) return 0.0f;
  return ld_mt<U16>(b, x, y);
}
template<int U16> DEV float centre_mt(const dimg &b, int x, int y)
{
  return ld_mt<U16>(b, mirror(x, b.wd), mirror(y, b.ht));
}

// texture() at texture-space (s,t): taps floor(s-.5), +1, weights frac
DEV float4 bilin_h4(const dimg &b, float s, float t)
{
  const float fs = s - 0.5f, ft = t - 0.5f;
  const float fls = floorf(fs), flt = floorf(ft);
  const int i0 = (int)fls, j0 = (int)flt;
  const float a = fs - fls, c = ft - flt;
  const int x0 = mirror(i0, b.wd), x1 = mirror(i0 + 1, b.wd);
  const int y0 = mirror(j0, b.ht), y1 = mirror(j0 + 1, b.ht);
  const float4 p00 = ld_h4(b, x0, y0), p10 = ld_h4(b, x1, y0);
  const float4 p01 = ld_h4(b, x0, y1), p11 = ld_h4(b, x1, y1);
  float4 r;
  r.x = (p00.x * (1.f - a) + p10.x * a) * (1.f - c) + (p01.x * (1.f - a) + p11.x * a) * c;
  r.y = (p00.y * (1.f - a) + p10.y * a) * (1.f - c) + (p01.y * (1.f - a) + p11.y * a) * c;
  r.z = (p00.z * (1.f - a) + p10.z * a) * (1.f - c) + (p01.z * (1.f - a) + p11.z * a) * c;
  r.w = (p00.w * (1.f - a) + p10.w * a) * (1.f - c) + (p01.w * (1.f - a) + p11.w * a) * c;
  return r;
}

DEV float4 ld_f4(const dimg &b, int x, int y)
{ // one 16-byte load of an rgba f32 texel
  return *(const float4 *)((const float *)b.p + 4 * ((long)y * b.pitch + x));
}

// texture() of an rgba f32 image at texture-space (s,t), as bilin_h4
DEV float4 bilin_f4(const dimg &b, float s, float t)
{
  const float fs = s - 0.5f, ft = t - 0.5f;
  const float fls = floorf(fs), flt = floorf(ft);
  const int i0 = (int)fls, j0 = (int)flt;
  const float a = fs - fls, c = ft - flt;
  const int x0 = mirror(i0, b.wd), x1 = mirror(i0 + 1, b.wd);
  const int y0 = mirror(j0, b.ht), y1 = mirror(j0 + 1, b.ht);
  const float4 p00 = ld_f4(b, x0, y0), p10 = ld_f4(b, x1, y0);
  const float4 p01 = ld_f4(b, x0, y1), p11 = ld_f4(b, x1, y1);
  float4 r;
  r.x = (p00.x * (1.f - a) + p10.x * a) * (1.f - c) + (p01.x * (1.f - a) + p11.x * a) * c;
  r.y = (p00.y * (1.f - a) + p10.y * a) * (1.f - c) + (p01.y * (1.f - a) + p11.y * a) * c;
  r.z = (p00.z * (1.f - a) + p10.z * a) * (1.f - c) + (p01.z * (1.f - a) + p11.z * a) * c;
  r.w = (p00.w * (1.f - a) + p10.w * a) * (1.f - c) + (p01.w * (1.f - a) + p11.w * a) * c;
  return r;
}

// bilinear with an integer base texel (x0,y0) and fixed weights (a,c):
// the exact form of texture() at texel-space (x0 + .5 + a, y0 + .5 + c)
DEV float4 bilin_fixed_h4(const dimg &b, int x0, int y0, float a, float c)
{
  const int xa = mirror(x0, b.wd), xb = mirror(x0 + 1, b.wd);
  const int ya = mirror(y0, b.ht), yb = mirror(y0 + 1, b.ht);
  const float4 p00 = ld_h4(b, xa, ya), p10 = ld_h4(b, xb, ya);
  const float4 p01 = ld_h4(b, xa, yb), p11 = ld_h4(b, xb, yb);
  float4 r;
  r.x = (p00.x * (1.f - a) + p10.x * a) * (1.f - c) + (p01.x * (1.f - a) + p11.x * a) * c;
  r.y = (p00.y * (1.f - a) + p10.y * a) * (1.f - c) + (p01.y * (1.f - a) + p11.y * a) * c;
  r.z = (p00.z * (1.f - a) + p10.z * a) * (1.f - c) + (p01.z * (1.f - a) + p11.z * a) * c;
  r.w = (p00.w * (1.f - a) + p10.w * a) * (1.f - c) + (p01.w * (1.f - a) + p11.w * a) * c;
  return r;
}

DEV float bilin_h1(const dimg &b, float s, float t, int pl = 0)
{
  const float fs = s - 0.5f, ft = t - 0.5f;
  const float fls = floorf(fs), flt = floorf(ft);
  const int i0 = (int)fls, j0 = (int)flt;
  const float a = fs - fls, c = ft - flt;
  const int x0 = mirror(i0, b.wd), x1 = mirror(i0 + 1, b.wd);
  const int y0 = mirror(j0, b.ht), y1 = mirror(j0 + 1, b.ht);
  const float p00 = ld_h1(b, x0, y0, pl), p10 = ld_h1(b, x1, y0, pl);
  const float p01 = ld_h1(b, x0, y1, pl), p11 = ld_h1(b, x1, y1, pl);
  return (p00 * (1.f - a) + p10 * a) * (1.f - c) + (p01 * (1.f - a) + p11 * a) * c;
}

// ---- colour helpers (shared.glsl) ---------------------------------------

DEV float lum2020(float r, float g, float b)
{ // shared.glsl:150-154
  return 2.62700212e-01f * r + 6.77998072e-01f * g + 5.93017165e-02f * b;
}

// shared.glsl:244-293; M = {M[0][0], M[0][1], M[1][0], M[1][1]}
DEV void evd2x2(float eval[2], float evec0[2], float evec1[2], const float M[4])
{
  const float a = M[0], b = M[2], c = M[3];
  const float pHalf = -0.5f * (a + c);
  const float q = a * c - b * b;
  const float droot = sqrtf(pHalf * pHalf - q);
  eval[0] = -pHalf + droot;
  eval[1] = -pHalf - droot;
  const float a0 = a - eval[0], b0 = b, c0 = c - eval[0];
  const float sl0 = a0 * a0 + b0 * b0;
  const float sl1 = b0 * b0 + c0 * c0;
  float sl;
  if(sl0 > sl1) { evec1[0] = a0; evec1[1] = b0; sl = sl0; }
  else          { evec1[0] = b0; evec1[1] = c0; sl = sl1; }
  evec1[0] = (sl == 0.0f) ? 1.0f : evec1[0];
  sl = (sl == 0.0f) ? 1.0f : sl;
  const float il = 1.0f / sqrtf(sl);
  evec1[0] *= il;
  evec1[1] *= il;
  evec0[0] = evec1[1];
  evec0[1] = -evec1[0];
}

DEV float det2(const float M[4]) { return M[0] * M[3] - M[2] * M[1]; }

// canonical x-trans colour at (x,y) >= 0: 0 red, 1 green, 2 blue
// (demosaic/splat.comp:52-68)
DEV int xtrans_colour(int x, int y)
{
  const int blue_top = ((x / 3 + y / 3) & 1) > 0;
  const int px = x - (x / 3) * 3, py = y - (y / 3) * 3;
  if(((px + py) & 1) == 0) return 1;
  if(blue_top ^ (py == 1)) return 2;
  return 0;
}

// denoise polyphase write position (denoise/down.comp:102-104)
DEV int sorted1(int x, int n) { return x / 2 + (x & 1) * ((n + 1) / 2); }

// grid helpers: 64x4 work groups (one wave per row segment)
#define BX 64
#define BY 4
// XCD-aware tile order.  Work groups are handed to the 8 XCDs round robin
// by linear id, and each XCD has its own L2: consecutive ids therefore land
// on different L2s, and the halo rows a stencil tile shares with its
// neighbours are fetched once per XCD.  xcd_block() renumbers so that XCD
// k runs one contiguous run of tiles (raster order, k-th eighth of the
// grid): vertical neighbours share an L2.  A bijection on the grid.
#define NUM_XCD 8
DEV int2 xcd_block()
{
  const int b = blockIdx.x + blockIdx.y * gridDim.x;
  const int n = gridDim.x * gridDim.y;
  const int q = n / NUM_XCD, r = n - q * NUM_XCD;
  const int k = b % NUM_XCD, i = b / NUM_XCD;
  const int t = k * q + min(k, r) + i;
  return make_int2(t % gridDim.x, t / gridDim.x);
}
DEV int tid_x() { return xcd_block().x * BX + threadIdx.x; }
DEV int tid_y() { return xcd_block().y * BY + threadIdx.y; }

// Bayer-uniform waves: the 64x4 texels of a tile at (X0,Y0) handed out so
// that each wave (threadIdx.y) holds one CFA colour, 32 texels of each of
// two rows of equal parity.  Per-tap colour tests are then uniform across
// the wave and the taps a colour skips are branched over.  A permutation
// of the tile's texels.
DEV int2 bayer_wave_xy(int X0, int Y0)
{
  const int l = threadIdx.x, w = threadIdx.y;
  return make_int2(X0 + 2 * (l & 31) + (w & 1), Y0 + (w >> 1) + 2 * (l >> 5));
}
