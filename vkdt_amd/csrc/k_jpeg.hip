// k_jpeg.hip -- baseline JPEG encoding on the GPU for the o-jpg sink
// (reference o-jpg/main.c:101-160, which hands rgb scanlines to libjpeg with
// optimised Huffman coding).  The steps are libjpeg's published ones (ITU-T
// T.81 baseline with IJG's choices), so the file's tables and entropy-coded
// data are those libjpeg writes for the same image and settings:
//   k_jpg_coef    one lane per 8x8 block in MCU order: fixed-point YCbCr,
//                 2:1 chroma boxes with the alternating bias (or libjpeg's
//                 smoothing filter), samples past the edges replicated,
//                 the islow, ifast or float DCT and its quantisation, all
//                 in registers; stored zigzag ordered, so the passes below
//                 load a block in 8 vectors
//   k_jpg_dummy   blocks of the last MCUs past a component's own extent:
//                 zero, with the DC of the block to the left / above
//   k_jpg_stats   symbol counts of the four tables (LDS, then global)
//   (host)        optimal code lengths per table (T.81 K.2, K.3)
//   k_jpg_len     bits per block; k_scan_* their offsets
//   k_jpg_emit    each block writes its codes at its bit offset (atomic OR
//                 into zeroed words, most significant bit first)
//   k_stuff_*     a zero byte after every 0xFF: counts per chunk, offsets,
//                 scatter
#include "launch.h"

#include <algorithm>
#include <stdlib.h>
#include <string.h>
#include <vector>

namespace {

static const uint8_t jz_host[64] = {
  0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21,
  28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
  47, 55, 62, 63 };

struct jpg_geom
{
  const uint8_t *rgba;       // device, 4 bytes a pixel
  int W, H, pitch;           // pitch in pixels
  int hs, vs, bpm;           // luma sampling, blocks per MCU
  int mx, my, nb;            // MCUs across, down; blocks
  int wib[3], hib[3];        // each component's own extent in blocks
  int smooth;                // libjpeg's smoothing_factor, 0..100
};

// block b -> component, block column and row
DEV void jpg_block(const jpg_geom &g, int b, int &c, int &bx, int &by, int &i)
{
  const int m = b / g.bpm, r = b - m * g.bpm, n0 = g.hs * g.vs;
  const int ux = m % g.mx, uy = m / g.mx;
  if(r < n0) { c = 0; i = r % g.hs; bx = ux * g.hs + i; by = uy * g.vs + r / g.hs; }
  else { c = r - n0 + 1; i = 0; bx = ux; by = uy; }
}

DEV void jpg_ycc(const jpg_geom &g, int x, int y, int &Y, int &cb, int &cr)
{
  x = x < 0 ? 0 : (x < g.W ? x : g.W - 1);
  y = y < 0 ? 0 : (y < g.H ? y : g.H - 1);
  const uchar4 p = *(const uchar4 *)(g.rgba + 4 * ((long)y * g.pitch + x));
  const int r = p.x, gg = p.y, b = p.z;
  // 16-bit fixed point: round(0.299 * 65536) etc.
  Y  = (19595 * r + 38470 * gg + 7471 * b + 32768) >> 16;
  cb = (-11059 * r - 21709 * gg + 32768 * b + (128 << 16) + 32767) >> 16;
  cr = (32768 * r - 27439 * gg - 5329 * b + (128 << 16) + 32767) >> 16;
}

DEV int jpg_comp(const jpg_geom &g, int c, int x, int y)
{ // full-resolution sample of component c, replicated past the edges
  int v[3];
  jpg_ycc(g, x, y, v[0], v[1], v[2]);
  return v[c];
}

// libjpeg's input smoothing (smoothing_factor sf): a full-size component
// mixes a sample with its 8 neighbours, weights 1 - 8 sf/1024 and sf/1024
// in 16-bit fixed point; 2x2 chroma forms an output from its 4 members
// ((1 - 5 sf/1024)/4 each), their 8 edge neighbours (sf/512) and 4 corner
// neighbours (sf/1024).  2x1 chroma is not smoothed.
DEV int jpg_sample_smooth(const jpg_geom &g, int c, int x, int y)
{
  const int sf = g.smooth;
  if(c == 0 || g.hs * g.vs == 1)
  {
    int n = 0;
#pragma unroll
    for(int j = -1; j <= 1; j++)
#pragma unroll
      for(int i = -1; i <= 1; i++)
        if(i || j) n += jpg_comp(g, c, x + i, y + j);
    return (jpg_comp(g, c, x, y) * (65536 - 512 * sf) + n * (64 * sf) + 32768) >> 16;
  }
  const int X = 2 * x, Y = 2 * y;
  int m = 0, e = 0, k = 0;
#pragma unroll
  for(int j = -1; j <= 2; j++)
#pragma unroll
    for(int i = -1; i <= 2; i++)
    {
      const int v = jpg_comp(g, c, X + i, Y + j);
      const bool ix = i == 0 || i == 1, iy = j == 0 || j == 1;
      if(ix && iy) m += v;
      else if(ix || iy) e += v;
      else k += v;
    }
  return (m * (16384 - 80 * sf) + (2 * e + k) * (16 * sf) + 32768) >> 16;
}

DEV int jpg_sample(const jpg_geom &g, int c, int x, int y)
{
  if(g.smooth && !(c > 0 && g.hs == 2 && g.vs == 1)) return jpg_sample_smooth(g, c, x, y);
  int Y, cb, cr;
  if(c == 0) { jpg_ycc(g, x, y, Y, cb, cr); return Y; }
  int s = 0;
  for(int j = 0; j < g.vs; j++)
    for(int i = 0; i < g.hs; i++)
    {
      jpg_ycc(g, g.hs * x + i, g.vs * y + j, Y, cb, cr);
      s += c == 1 ? cb : cr;
    }
  if(g.hs * g.vs == 1) return s;
  if(g.vs == 2) return (s + ((x & 1) ? 2 : 1)) >> 2;   // 2x2: bias 1, 2, 1, 2, ...
  return (s + (x & 1)) >> 1;                          // 2x1: bias 0, 1, 0, 1, ...
}

#define JC(x) ((int32_t)((x) * 8192.0 + 0.5))
#define JDESC(x, n) (((x) + (1 << ((n) - 1))) >> (n))

// islow forward DCT of 8 values a stride apart (pass 0 rows, 1 columns)
template<int PASS, int S>
DEV void jpg_fdct8(int32_t *p)
{
  const int32_t t0 = p[0] + p[7 * S], t7 = p[0] - p[7 * S], t1 = p[S] + p[6 * S], t6 = p[S] - p[6 * S];
  const int32_t t2 = p[2 * S] + p[5 * S], t5 = p[2 * S] - p[5 * S], t3 = p[3 * S] + p[4 * S], t4 = p[3 * S] - p[4 * S];
  const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
  constexpr int sh = PASS == 0 ? 13 - 2 : 13 + 2;
  if(PASS == 0) { p[0] = (t10 + t11) << 2; p[4 * S] = (t10 - t11) << 2; }
  else { p[0] = JDESC(t10 + t11, 2); p[4 * S] = JDESC(t10 - t11, 2); }
  const int32_t z1 = (t12 + t13) * JC(0.541196100);
  p[2 * S] = JDESC(z1 + t13 * JC(0.765366865), sh);
  p[6 * S] = JDESC(z1 + t12 * (-JC(1.847759065)), sh);
  int32_t a1 = t4 + t7, a2 = t5 + t6, a3 = t4 + t6, a4 = t5 + t7;
  const int32_t z5 = (a3 + a4) * JC(1.175875602);
  const int32_t b4 = t4 * JC(0.298631336), b5 = t5 * JC(2.053119869), b6 = t6 * JC(3.072711026), b7 = t7 * JC(1.501321110);
  a1 *= -JC(0.899976223); a2 *= -JC(2.562915447); a3 *= -JC(1.961570560); a4 *= -JC(0.390180644);
  a3 += z5; a4 += z5;
  p[7 * S] = JDESC(b4 + a1 + a3, sh);
  p[5 * S] = JDESC(b5 + a2 + a4, sh);
  p[3 * S] = JDESC(b6 + a2 + a3, sh);
  p[S] = JDESC(b7 + a1 + a4, sh);
}

// the scaled DCT of Arai, Agui and Nakajima, libjpeg's "ifast" (integer,
// 8 fractional bits, products truncated) and "float": outputs carry the
// factors 8 s_u s_v (s_0 = 1, s_k = sqrt(2) cos(k pi/16)), which the
// divisors hold
#define AAN8(v, c) ((int32_t)(((v) * (c)) >> 8))
template<int S>
DEV void jpg_fdct8_ifast(int32_t *p)
{
  const int32_t a0 = p[0] + p[7 * S], a7 = p[0] - p[7 * S], a1 = p[S] + p[6 * S], a6 = p[S] - p[6 * S];
  const int32_t a2 = p[2 * S] + p[5 * S], a5 = p[2 * S] - p[5 * S], a3 = p[3 * S] + p[4 * S], a4 = p[3 * S] - p[4 * S];
  const int32_t e0 = a0 + a3, e3 = a0 - a3, e1 = a1 + a2, e2 = a1 - a2;
  p[0] = e0 + e1;
  p[4 * S] = e0 - e1;
  const int32_t r = AAN8(e2 + e3, 181);
  p[2 * S] = e3 + r;
  p[6 * S] = e3 - r;
  const int32_t o0 = a4 + a5, o1 = a5 + a6, o2 = a6 + a7;
  const int32_t rot = AAN8(o0 - o2, 98);
  const int32_t q2 = AAN8(o0, 139) + rot, q4 = AAN8(o2, 334) + rot, q3 = AAN8(o1, 181);
  const int32_t h1 = a7 + q3, h3 = a7 - q3;
  p[5 * S] = h3 + q2;
  p[3 * S] = h3 - q2;
  p[S] = h1 + q4;
  p[7 * S] = h1 - q4;
}

template<int S>
DEV void jpg_fdct8_float(float *p)
{
  const float a0 = p[0] + p[7 * S], a7 = p[0] - p[7 * S], a1 = p[S] + p[6 * S], a6 = p[S] - p[6 * S];
  const float a2 = p[2 * S] + p[5 * S], a5 = p[2 * S] - p[5 * S], a3 = p[3 * S] + p[4 * S], a4 = p[3 * S] - p[4 * S];
  const float e0 = a0 + a3, e3 = a0 - a3, e1 = a1 + a2, e2 = a1 - a2;
  p[0] = e0 + e1;
  p[4 * S] = e0 - e1;
  const float r = (e2 + e3) * 0.707106781f;
  p[2 * S] = e3 + r;
  p[6 * S] = e3 - r;
  const float o0 = a4 + a5, o1 = a5 + a6, o2 = a6 + a7;
  const float rot = (o0 - o2) * 0.382683433f;
  const float q2 = 0.541196100f * o0 + rot, q4 = 1.306562965f * o2 + rot, q3 = o1 * 0.707106781f;
  const float h1 = a7 + q3, h3 = a7 - q3;
  p[5 * S] = h3 + q2;
  p[3 * S] = h3 - q2;
  p[S] = h1 + q4;
  p[7 * S] = h1 - q4;
}

DEV int jpg_quant(int32_t v, int q8)
{ // v / q8 rounded half away from zero
  if(v < 0) { v = -v + (q8 >> 1); return -(v >= q8 ? v / q8 : 0); }
  v += q8 >> 1;
  return v >= q8 ? v / q8 : 0;
}

// coefficients in natural order; q8: the divisors, luma then chroma: 8 x
// quantiser (DCT 0, islow), 8 x quantiser x the AAN factors in integer
// (1, ifast), or the float reciprocal of that (2, float; its bits)
template<int DCT>
__global__ void __launch_bounds__(64) k_jpg_coef(jpg_geom g, const int32_t *q8, int16_t *coef, uint8_t *dummy)
{
  const int b = blockIdx.x * 64 + threadIdx.x;
  if(b >= g.nb) return;
  int c, bx, by, i;
  jpg_block(g, b, c, bx, by, i);
  int16_t *o = coef + (long)b * 64;
  const bool dm = bx >= g.wib[c] || by >= g.hib[c];
  dummy[b] = dm ? (by >= g.hib[c] ? 2 : 1) : 0;
  if(dm)
  {
    for(int k = 0; k < 64; k += 4) *(uint2 *)(o + k) = make_uint2(0, 0);
    return;
  }
  int32_t d[64];
#pragma unroll
  for(int k = 0; k < 64; k++) d[k] = jpg_sample(g, c, bx * 8 + (k & 7), by * 8 + (k >> 3)) - 128;
  const int32_t *q = q8 + (c ? 64 : 0);
  if(DCT == 2)
  { // float: multiply by the reciprocal, round to nearest even
    float f[64];
#pragma unroll
    for(int k = 0; k < 64; k++) f[k] = (float)d[k];
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8_float<1>(f + 8 * k);
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8_float<8>(f + k);
#pragma unroll
    for(int k = 0; k < 64; k++) d[k] = (int32_t)rintf(f[k] * __int_as_float(q[k]));
  }
  else if(DCT == 1)
  {
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8_ifast<1>(d + 8 * k);
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8_ifast<8>(d + k);
  }
  else
  {
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8<0, 1>(d + 8 * k);
#pragma unroll
    for(int k = 0; k < 8; k++) jpg_fdct8<1, 8>(d + k);
  }
  // stored in zigzag order, the order the entropy coder walks them
  constexpr uint8_t zz[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21,
    28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
    47, 55, 62, 63 };
#pragma unroll
  for(int k = 0; k < 64; k += 2)
  {
    const int a = DCT == 2 ? d[zz[k]] : jpg_quant(d[zz[k]], q[zz[k]]);
    const int e = DCT == 2 ? d[zz[k + 1]] : jpg_quant(d[zz[k + 1]], q[zz[k + 1]]);
    *(uint32_t *)(o + k) = (uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)e << 16);
  }
}

// a block's 64 coefficients (zigzag order) in registers: eight 16-byte loads
struct jpg_blk
{
  uint4 w[8];
  __device__ __forceinline__ int at(int k) const
  {
    const uint4 &v = w[k >> 3];
    const int j = (k >> 1) & 3;
    const uint32_t u = j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
    return (int)(int16_t)(uint16_t)(u >> (16 * (k & 1)));
  }
};

DEV void jpg_load(jpg_blk &z, const int16_t *coef, int b)
{
  const uint4 *p = (const uint4 *)(coef + (long)b * 64);
#pragma unroll
  for(int i = 0; i < 8; i++) z.w[i] = p[i];
}

// the dummy blocks' DC: pass 1 the right edge (block to the left), pass 2 the
// bottom rows (the last block of the row above in the MCU, itself patched)
__global__ void __launch_bounds__(256) k_jpg_dummy(jpg_geom g, int16_t *coef, const uint8_t *dummy, int pass)
{
  const int b = blockIdx.x * 256 + threadIdx.x;
  if(b >= g.nb || dummy[b] != pass) return;
  int c, bx, by, i;
  jpg_block(g, b, c, bx, by, i);
  coef[(long)b * 64] = pass == 1 ? coef[(long)(b - 1) * 64] : coef[(long)(b - i - 1) * 64];
}

// the previous block of the same component in scan order, -1 for the first
DEV int jpg_prev(const jpg_geom &g, int b)
{
  const int m = b / g.bpm, r = b - m * g.bpm, n0 = g.hs * g.vs;
  if(r > 0 && r < n0) return b - 1;
  if(m == 0) return -1;
  return r == 0 ? b - g.bpm + n0 - 1 : b - g.bpm;
}

DEV int jpg_nbits(int v)
{
  v = v < 0 ? -v : v;
  return v ? 32 - __clz(v) : 0;
}

__global__ void __launch_bounds__(256) k_jpg_stats(jpg_geom g, const int16_t *coef, uint32_t *freq)
{
  __shared__ uint32_t h[4 * 257];
  for(int k = threadIdx.x; k < 4 * 257; k += 256) h[k] = 0;
  __syncthreads();
  const int b = blockIdx.x * 256 + threadIdx.x;
  if(b < g.nb)
  {
    int c, bx, by, i;
    jpg_block(g, b, c, bx, by, i);
    const int t = c ? 1 : 0;
    jpg_blk z;
    jpg_load(z, coef, b);
    const int p = jpg_prev(g, b);
    const int last = p < 0 ? 0 : coef[(long)p * 64];
    atomicAdd(h + (2 * t) * 257 + jpg_nbits(z.at(0) - last), 1u);
    uint32_t *ac = h + (2 * t + 1) * 257;
    int run = 0;
#pragma unroll
    for(int k = 1; k < 64; k++)
    {
      const int v = z.at(k);
      if(!v) { run++; continue; }
      while(run > 15) { atomicAdd(ac + 0xf0, 1u); run -= 16; }
      atomicAdd(ac + ((run << 4) | jpg_nbits(v)), 1u);
      run = 0;
    }
    if(run) atomicAdd(ac, 1u);
  }
  __syncthreads();
  for(int k = threadIdx.x; k < 4 * 257; k += 256)
    if(h[k]) atomicAdd(freq + k, h[k]);
}

struct jpg_codes { uint16_t code[4][256]; uint8_t len[4][256]; };

// bits per block (mode 0) or the block's codes at its offset (mode 1)
template<int EMIT>
__global__ void __launch_bounds__(256) k_jpg_code(jpg_geom g, const int16_t *coef, const jpg_codes *hc,
                                                  uint64_t *bits, uint32_t *words)
{
  const int b = blockIdx.x * 256 + threadIdx.x;
  if(b >= g.nb) return;
  int c, bx, by, i;
  jpg_block(g, b, c, bx, by, i);
  const int t = c ? 1 : 0;
  const uint16_t *dcc = hc->code[2 * t], *acc_c = hc->code[2 * t + 1];
  const uint8_t *dcl = hc->len[2 * t], *acl = hc->len[2 * t + 1];
  jpg_blk z;
  jpg_load(z, coef, b);
  const int p = jpg_prev(g, b);
  const int last = p < 0 ? 0 : coef[(long)p * 64];
  // writer: acc holds the bits from word w on, `used` of them taken
  uint64_t total = 0, acc = 0;
  long w = 0;
  int used = 0;
  if(EMIT) { const uint64_t off = bits[b]; w = (long)(off >> 5); used = (int)(off & 31); }
  auto put = [&](uint32_t v, int n) {
    if(!EMIT) { total += n; return; }
    if(!n) return;
    acc |= (uint64_t)v << (64 - used - n);
    used += n;
    if(used >= 32)
    {
      atomicOr(words + w, (uint32_t)(acc >> 32));
      acc <<= 32;
      used -= 32;
      w++;
    }
  };
  const int diff = z.at(0) - last, n = jpg_nbits(diff);
  put(dcc[n], dcl[n]);
  put((uint32_t)(diff < 0 ? diff - 1 : diff) & ((1u << n) - 1), n);
  int run = 0;
#pragma unroll
  for(int k = 1; k < 64; k++)
  {
    const int v = z.at(k);
    if(!v) { run++; continue; }
    while(run > 15) { put(acc_c[0xf0], acl[0xf0]); run -= 16; }
    const int s = jpg_nbits(v), sym = (run << 4) | s;
    put(acc_c[sym], acl[sym]);
    put((uint32_t)(v < 0 ? v - 1 : v) & ((1u << s) - 1), s);
    run = 0;
  }
  if(run) put(acc_c[0], acl[0]);
  if(EMIT) { if(used) atomicOr(words + w, (uint32_t)(acc >> 32)); }
  else bits[b] = total;
}

// exclusive scan of n u64 in place: 1024 a work group, then the group sums
__global__ void __launch_bounds__(256) k_scan_local(uint64_t *v, long n, uint64_t *sums)
{
  __shared__ uint64_t s[256];
  const long base = (long)blockIdx.x * 1024 + 4 * threadIdx.x;
  uint64_t x[4], t = 0;
  for(int j = 0; j < 4; j++) { x[j] = base + j < n ? v[base + j] : 0; t += x[j]; }
  s[threadIdx.x] = t;
  __syncthreads();
  for(int o = 1; o < 256; o <<= 1)
  {
    const uint64_t a = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += a;
    __syncthreads();
  }
  uint64_t run = s[threadIdx.x] - t;
  for(int j = 0; j < 4; j++) { if(base + j < n) v[base + j] = run; run += x[j]; }
  if(threadIdx.x == 255) sums[blockIdx.x] = s[255];
}

__global__ void __launch_bounds__(256) k_scan_sums(uint64_t *sums, int n, uint64_t *total)
{ // one work group: exclusive scan of the group sums, and their total
  __shared__ uint64_t s[256];
  uint64_t carry = 0;
  for(int b0 = 0; b0 < n; b0 += 256)
  {
    const int k = b0 + threadIdx.x;
    const uint64_t x = k < n ? sums[k] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for(int o = 1; o < 256; o <<= 1)
    {
      const uint64_t a = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += a;
      __syncthreads();
    }
    if(k < n) sums[k] = carry + s[threadIdx.x] - x;
    carry += s[255];
    __syncthreads();
  }
  if(threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) k_scan_add(uint64_t *v, long n, const uint64_t *sums)
{
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if(k < n) v[k] += sums[k / 1024];
}

DEV uint32_t jpg_byte(const uint32_t *words, long i) { return (words[i >> 2] >> (24 - 8 * (i & 3))) & 0xffu; }

// 0xFF bytes per chunk of 1024
__global__ void __launch_bounds__(256) k_stuff_count(const uint32_t *words, long nbytes, uint64_t *cnt)
{
  __shared__ uint32_t s;
  if(threadIdx.x == 0) s = 0;
  __syncthreads();
  uint32_t c = 0;
  for(int j = 0; j < 4; j++)
  {
    const long i = (long)blockIdx.x * 1024 + 4 * threadIdx.x + j;
    if(i < nbytes && jpg_byte(words, i) == 0xffu) c++;
  }
  if(c) atomicAdd(&s, c);
  __syncthreads();
  if(threadIdx.x == 0) cnt[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_stuff_scatter(const uint32_t *words, long nbytes, const uint64_t *off,
                                                       uint8_t *out)
{
  __shared__ uint32_t s[256];
  uint32_t by[4], c = 0;
  const long i0 = (long)blockIdx.x * 1024 + 4 * threadIdx.x;
  for(int j = 0; j < 4; j++) { by[j] = i0 + j < nbytes ? jpg_byte(words, i0 + j) : 0u; c += (i0 + j < nbytes && by[j] == 0xffu); }
  s[threadIdx.x] = c;
  __syncthreads();
  for(int o = 1; o < 256; o <<= 1)
  {
    const uint32_t a = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += a;
    __syncthreads();
  }
  uint64_t at = (uint64_t)i0 + off[blockIdx.x] + s[threadIdx.x] - c;
  for(int j = 0; j < 4; j++)
  {
    if(i0 + j >= nbytes) break;
    out[at++] = (uint8_t)by[j];
    if(by[j] == 0xffu) out[at++] = 0;
  }
}

__global__ void k_jpg_zero(uint32_t *p, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0;
}

// ---- host ------------------------------------------------------------------

// optimal code lengths (T.81 K.2: the two smallest frequencies merged, ties to
// the larger symbol, a reserved symbol 256; K.3: lengths limited to 16)
void jpg_huff(const uint32_t *freq_in, uint8_t bits[17], uint8_t val[256], int *nval)
{
  long freq[257];
  int codesize[257], others[257];
  for(int i = 0; i < 256; i++) freq[i] = freq_in[i];
  freq[256] = 1;
  for(int i = 0; i < 257; i++) { codesize[i] = 0; others[i] = -1; }
  for(;;)
  {
    int c1 = -1, c2 = -1;
    long v = 1000000000L;
    for(int i = 0; i <= 256; i++) if(freq[i] && freq[i] <= v) { v = freq[i]; c1 = i; }
    v = 1000000000L;
    for(int i = 0; i <= 256; i++) if(freq[i] && freq[i] <= v && i != c1) { v = freq[i]; c2 = i; }
    if(c2 < 0) break;
    freq[c1] += freq[c2];
    freq[c2] = 0;
    codesize[c1]++;
    while(others[c1] >= 0) { c1 = others[c1]; codesize[c1]++; }
    others[c1] = c2;
    codesize[c2]++;
    while(others[c2] >= 0) { c2 = others[c2]; codesize[c2]++; }
  }
  int cnt[33] = { 0 };
  for(int i = 0; i <= 256; i++) if(codesize[i]) cnt[codesize[i] > 32 ? 32 : codesize[i]]++;
  for(int i = 32; i > 16; i--)
    while(cnt[i] > 0)
    {
      int j = i - 2;
      while(cnt[j] == 0) j--;
      cnt[i] -= 2;
      cnt[i - 1]++;
      cnt[j + 1] += 2;
      cnt[j]--;
    }
  int i = 16;
  while(cnt[i] == 0) i--;
  cnt[i]--;
  bits[0] = 0;
  for(int k = 1; k <= 16; k++) bits[k] = (uint8_t)cnt[k];
  int p = 0;
  for(int len = 1; len <= 32; len++)
    for(int s = 0; s < 256; s++) if(codesize[s] == len) val[p++] = (uint8_t)s;
  *nval = p;
}

struct out_t
{
  std::vector<uint8_t> d;
  void b(int v) { d.push_back((uint8_t)v); }
  void w(int v) { b(v >> 8); b(v & 255); }
};

} // namespace

extern "C" void vkdt_jpg_quant_tables(int quality, uint8_t luma[64], uint8_t chroma[64])
{
  static const int sl[64] = {
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
    14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99 };
  static const int sc[64] = {
    17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
    47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99 };
  quality = quality < 1 ? 1 : (quality > 100 ? 100 : quality);
  const long scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;   // IJG quality scaling
  for(int i = 0; i < 64; i++)
  {
    const long l = (sl[i] * scale + 50) / 100, c = (sc[i] * scale + 50) / 100;
    luma[i] = (uint8_t)(l < 1 ? 1 : (l > 255 ? 255 : l));
    chroma[i] = (uint8_t)(c < 1 ? 1 : (c > 255 ? 255 : c));
  }
}

extern "C" int vkdt_hip_jpg_encode(const void *rgba, uint32_t wd, uint32_t ht, uint32_t pitch, int quality, int hs,
                                   int vs, uint8_t **out, size_t *out_len, void *stream)
{
  return vkdt_hip_jpg_encode_ex(rgba, wd, ht, pitch, quality, hs, vs, VKDT_JPG_DCT_ISLOW, 0, nullptr, 0, out, out_len,
                                stream);
}

extern "C" int vkdt_hip_jpg_encode_ex(const void *rgba, uint32_t wd, uint32_t ht, uint32_t pitch, int quality, int hs,
                                      int vs, int dct, int smoothing, const uint8_t *icc, size_t icc_len, uint8_t **out,
                                      size_t *out_len, void *stream)
{
  *out = nullptr;
  *out_len = 0;
  if(!rgba || wd == 0 || ht == 0 || wd > 65535 || ht > 65535 || pitch < wd)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: extent %ux%u (pitch %u)", wd, ht, pitch);
  if(!((hs == 1 && vs == 1) || (hs == 2 && vs == 1) || (hs == 2 && vs == 2)))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: sampling %dx%d (4:4:4, 4:2:2 or 4:2:0)", hs, vs);
  if(dct < VKDT_JPG_DCT_ISLOW || dct > VKDT_JPG_DCT_FLOAT)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: dct method %d (0 islow, 1 ifast, 2 float)", dct);
  if(smoothing < 0 || smoothing > 100)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: smoothing factor %d (0..100)", smoothing);
  // ICC.1 Annex B.4: at most 255 APP2 markers of 65519 profile bytes
  const size_t icc_chunk = 65533 - 14;
  if(icc_len && (!icc || (icc_len + icc_chunk - 1) / icc_chunk > 255))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: icc profile of %zu bytes", icc_len);
  jpg_geom g;
  g.rgba = (const uint8_t *)rgba;
  g.W = (int)wd; g.H = (int)ht; g.pitch = (int)pitch;
  g.hs = hs; g.vs = vs; g.bpm = hs * vs + 2;
  g.mx = (g.W + 8 * hs - 1) / (8 * hs);
  g.my = (g.H + 8 * vs - 1) / (8 * vs);
  const long nbl = (long)g.mx * g.my * g.bpm;
  if(nbl >= (1l << 31) / 64) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "jpg: %ld blocks", nbl);
  g.nb = (int)nbl;
  g.wib[0] = (g.W + 7) / 8;
  g.hib[0] = (g.H + 7) / 8;
  g.wib[1] = g.wib[2] = ((g.W + hs - 1) / hs + 7) / 8;
  g.hib[1] = g.hib[2] = ((g.H + vs - 1) / vs + 7) / 8;
  g.smooth = smoothing;
  uint8_t ql[64], qc[64];
  vkdt_jpg_quant_tables(quality, ql, qc);
  // divisors by DCT method; the AAN factors s_0 = 1, s_k = sqrt(2) cos(k pi/16)
  static const double aan[8] = { 1.0, 1.387039845, 1.306562965, 1.175875602, 1.0, 0.785694958, 0.541196100,
                                 0.275899379 };
  int32_t q8[128];
  for(int t = 0; t < 2; t++)
    for(int k = 0; k < 64; k++)
    {
      const int q = t ? qc[k] : ql[k];
      const double f = aan[k / 8] * aan[k % 8];
      if(dct == VKDT_JPG_DCT_FLOAT)
      {
        const float r = (float)(1.0 / ((double)q * aan[k / 8] * aan[k % 8] * 8.0));
        memcpy(q8 + 64 * t + k, &r, 4);
      }
      else if(dct == VKDT_JPG_DCT_IFAST)
        q8[64 * t + k] = (int32_t)(((int64_t)q * (int32_t)(16384.0 * f + 0.5) + 1024) >> 11);
      else
        q8[64 * t + k] = q << 3;
    }

  // device: q8 | codes | freq | coef | dummy flags | bits | sums | total
  const long ngroups = (nbl + 1023) / 1024;
  const size_t o_q = 0, o_codes = 512, o_freq = o_codes + ((sizeof(jpg_codes) + 255) & ~(size_t)255);
  const size_t o_coef = o_freq + 4 * 257 * 4 + 256 - (4 * 257 * 4) % 256;
  const size_t o_dummy = o_coef + (size_t)nbl * 128;
  const size_t o_bits = (o_dummy + nbl + 255) & ~(size_t)255;
  const size_t o_sums = o_bits + (size_t)nbl * 8;
  const size_t o_total = o_sums + (size_t)(ngroups + 1) * 8;
  const size_t total = o_total + 256;
  uint8_t *dev = nullptr;
  if(int e = hip_fail(hipMalloc((void **)&dev, total), "jpg scratch")) return e;
  hipStream_t s = (hipStream_t)stream;
  int16_t *coef = (int16_t *)(dev + o_coef);
  uint32_t *freq = (uint32_t *)(dev + o_freq);
  uint64_t *bits = (uint64_t *)(dev + o_bits), *sums = (uint64_t *)(dev + o_sums), *tot = (uint64_t *)(dev + o_total);
  uint8_t *words8 = nullptr, *stuffed = nullptr;
  int err = hip_fail(hipMemcpyAsync(dev + o_q, q8, sizeof(q8), hipMemcpyHostToDevice, s), "jpg tables");
  uint32_t fr[4 * 257];
  if(!err)
  {
    hipLaunchKernelGGL(k_jpg_zero, dim3((4 * 257 + 255) / 256), dim3(256), 0, s, freq, (long)(4 * 257));
    const auto coef_kernel = dct == VKDT_JPG_DCT_FLOAT ? k_jpg_coef<2> : (dct == VKDT_JPG_DCT_IFAST ? k_jpg_coef<1> : k_jpg_coef<0>);
    hipLaunchKernelGGL(coef_kernel, dim3((g.nb + 63) / 64), dim3(64), 0, s, g, (const int32_t *)(dev + o_q), coef,
                       dev + o_dummy);
    hipLaunchKernelGGL(k_jpg_dummy, dim3((g.nb + 255) / 256), dim3(256), 0, s, g, coef, (const uint8_t *)(dev + o_dummy), 1);
    hipLaunchKernelGGL(k_jpg_dummy, dim3((g.nb + 255) / 256), dim3(256), 0, s, g, coef, (const uint8_t *)(dev + o_dummy), 2);
    hipLaunchKernelGGL(k_jpg_stats, dim3((g.nb + 255) / 256), dim3(256), 0, s, g, (const int16_t *)coef, freq);
    err = launch_post("jpg coefficients");
    if(!err) err = hip_fail(hipMemcpyAsync(fr, freq, sizeof(fr), hipMemcpyDeviceToHost, s), "jpg counts");
    if(!err) err = hip_fail(hipStreamSynchronize(s), "jpg sync");
  }
  uint8_t hb[4][17], hv[4][256];
  int nv[4] = { 0, 0, 0, 0 };
  jpg_codes hc;
  memset(&hc, 0, sizeof(hc));
  if(!err)
  {
    for(int t = 0; t < 4; t++)
    {
      jpg_huff(fr + 257 * t, hb[t], hv[t], nv + t);
      int code = 0, k = 0;
      for(int len = 1; len <= 16; len++)
      {
        for(int j = 0; j < hb[t][len]; j++, k++) { hc.code[t][hv[t][k]] = (uint16_t)code++; hc.len[t][hv[t][k]] = (uint8_t)len; }
        code <<= 1;
      }
    }
    err = hip_fail(hipMemcpyAsync(dev + o_codes, &hc, sizeof(hc), hipMemcpyHostToDevice, s), "jpg codes");
  }
  uint64_t nbits = 0;
  if(!err)
  {
    const jpg_codes *dhc = (const jpg_codes *)(dev + o_codes);
    hipLaunchKernelGGL(k_jpg_code<0>, dim3((g.nb + 255) / 256), dim3(256), 0, s, g, (const int16_t *)coef, dhc, bits,
                       (uint32_t *)nullptr);
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)ngroups), dim3(256), 0, s, bits, nbl, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, (int)ngroups, tot);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((nbl + 255) / 256)), dim3(256), 0, s, bits, nbl, (const uint64_t *)sums);
    err = launch_post("jpg lengths");
    if(!err) err = hip_fail(hipMemcpyAsync(&nbits, tot, 8, hipMemcpyDeviceToHost, s), "jpg total");
    if(!err) err = hip_fail(hipStreamSynchronize(s), "jpg sync");
  }
  const long nbytes = (long)((nbits + 7) / 8);
  const long nwords = (long)((nbits + 31) / 32) + 1, nchunks = (nbytes + 1023) / 1024;
  uint64_t nff = 0;
  if(!err)
  { // words | chunk counts and their sums | stuffed bytes (at most twice)
    const size_t wbytes = (size_t)nwords * 4, cbytes = (size_t)(nchunks + 1) * 8, gsum = (size_t)((nchunks + 1023) / 1024 + 1) * 8;
    err = hip_fail(hipMalloc((void **)&words8, wbytes + 2 * cbytes + gsum + 256 + 2 * (size_t)nbytes + 16), "jpg stream");
    if(!err)
    {
      uint32_t *words = (uint32_t *)words8;
      uint64_t *cnt = (uint64_t *)(words8 + wbytes), *csum = cnt + nchunks + 1, *ctot = csum + (nchunks + 1024) / 1024 + 1;
      stuffed = words8 + wbytes + 2 * cbytes + gsum + 256;
      hipLaunchKernelGGL(k_jpg_zero, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, words, nwords);
      hipLaunchKernelGGL(k_jpg_code<1>, dim3((g.nb + 255) / 256), dim3(256), 0, s, g, (const int16_t *)coef,
                         (const jpg_codes *)(dev + o_codes), bits, words);
      err = launch_post("jpg emit");
      if(!err && (nbits & 7))
      { // the last byte's free bits are ones
        const uint32_t pad = (1u << (8 - (nbits & 7))) - 1u;
        uint32_t wlast = 0;
        const long wi = (long)((nbits - 1) >> 5);
        err = hip_fail(hipMemcpyAsync(&wlast, words + wi, 4, hipMemcpyDeviceToHost, s), "jpg pad");
        if(!err) err = hip_fail(hipStreamSynchronize(s), "jpg sync");
        const int byte_in_word = (int)(((nbits - 1) >> 3) & 3);
        wlast |= pad << (24 - 8 * byte_in_word);
        if(!err) err = hip_fail(hipMemcpyAsync(words + wi, &wlast, 4, hipMemcpyHostToDevice, s), "jpg pad");
      }
      if(!err && nbytes > 0)
      {
        hipLaunchKernelGGL(k_stuff_count, dim3((unsigned)nchunks), dim3(256), 0, s, (const uint32_t *)words, nbytes, cnt);
        const long ng2 = (nchunks + 1023) / 1024;
        hipLaunchKernelGGL(k_scan_local, dim3((unsigned)ng2), dim3(256), 0, s, cnt, nchunks, csum);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, csum, (int)ng2, ctot);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((nchunks + 255) / 256)), dim3(256), 0, s, cnt, nchunks,
                           (const uint64_t *)csum);
        hipLaunchKernelGGL(k_stuff_scatter, dim3((unsigned)nchunks), dim3(256), 0, s, (const uint32_t *)words, nbytes,
                           (const uint64_t *)cnt, stuffed);
        err = launch_post("jpg stuffing");
        if(!err) err = hip_fail(hipMemcpyAsync(&nff, ctot, 8, hipMemcpyDeviceToHost, s), "jpg stuffed");
        if(!err) err = hip_fail(hipStreamSynchronize(s), "jpg sync");
      }
    }
  }
  out_t o;
  if(!err)
  {
    o.w(0xffd8);
    o.w(0xffe0); o.w(16);                                     // JFIF, 300 dpi (o-jpg/main.c:157-159)
    o.b('J'); o.b('F'); o.b('I'); o.b('F'); o.b(0);
    o.b(1); o.b(1); o.b(1); o.w(300); o.w(300); o.b(0); o.b(0);
    // the ICC profile in APP2 markers: "ICC_PROFILE\0", sequence number
    // from 1, marker count, then up to 65519 profile bytes (ICC.1 B.4)
    const int nicc = (int)((icc_len + icc_chunk - 1) / icc_chunk);
    for(int m = 0; m < nicc; m++)
    {
      const size_t n = std::min(icc_chunk, icc_len - m * icc_chunk);
      o.w(0xffe2); o.w((int)(2 + 14 + n));
      for(const char *c = "ICC_PROFILE"; *c; c++) o.b(*c);
      o.b(0); o.b(m + 1); o.b(nicc);
      o.d.insert(o.d.end(), icc + m * icc_chunk, icc + m * icc_chunk + n);
    }
    for(int t = 0; t < 2; t++)
    {
      o.w(0xffdb); o.w(67); o.b(t);
      for(int k = 0; k < 64; k++) o.b((t ? qc : ql)[jz_host[k]]);
    }
    o.w(0xffc0); o.w(17); o.b(8); o.w(g.H); o.w(g.W); o.b(3);
    o.b(1); o.b((hs << 4) | vs); o.b(0);
    o.b(2); o.b(0x11); o.b(1);
    o.b(3); o.b(0x11); o.b(1);
    for(int t = 0; t < 4; t++)
    {
      o.w(0xffc4); o.w(2 + 1 + 16 + nv[t]); o.b(((t & 1) << 4) | (t >> 1));
      for(int k = 1; k <= 16; k++) o.b(hb[t][k]);
      for(int k = 0; k < nv[t]; k++) o.b(hv[t][k]);
    }
    o.w(0xffda); o.w(12); o.b(3);
    o.b(1); o.b(0x00); o.b(2); o.b(0x11); o.b(3); o.b(0x11);
    o.b(0); o.b(63); o.b(0);
    const size_t head = o.d.size(), body = (size_t)nbytes + (size_t)nff;
    o.d.resize(head + body);
    if(body) err = hip_fail(hipMemcpyAsync(o.d.data() + head, stuffed, body, hipMemcpyDeviceToHost, s), "jpg download");
    if(!err) err = hip_fail(hipStreamSynchronize(s), "jpg sync");
    o.w(0xffd9);
  }
  (void)hipFree(dev);
  if(words8) (void)hipFree(words8);
  if(err) return err;
  *out = (uint8_t *)malloc(o.d.size());
  if(!*out) return vkdt_set_error(VKDT_HIP_OOM, "jpg: out of memory");
  memcpy(*out, o.d.data(), o.d.size());
  *out_len = o.d.size();
  return VKDT_HIP_OK;
}

extern "C" void vkdt_jpg_free(uint8_t *p) { free(p); }
