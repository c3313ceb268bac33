/*
 * jpeg.c -- the oracle's baseline JPEG encoder (test infrastructure, not
 * the product): what the reference's o-jpg gets from libjpeg
 * (o-jpg/main.c:101-160 sets quality, sampling, optimised Huffman codes and
 * 300 dpi, then writes scanlines of rgb).  libjpeg is a dependency outside
 * the reference tree (IJG libjpeg / libjpeg-turbo, the "libjpeg8" API); this
 * restates its published algorithm (ITU-T T.81 baseline with IJG's choices):
 *   colour      JFIF YCbCr in 16-bit fixed point, rounding as IJG's rgb_ycc
 *   sampling    4:4:4, 4:2:2 or 4:2:0; 2:1 boxes averaged with a bias that
 *               alternates per output column (1,2 for 2x2, 0,1 for 2x1);
 *               samples replicated past the edges; blocks of the last MCUs
 *               beyond a component's own block extent are dummies (zero, the
 *               DC of the block to the left or of the row above's last one)
 *   smoothing   libjpeg's smoothing_factor filter (or_jpg_encode_ex): full-size
 *               components 3x3, 2x2-subsampled chroma 4x4 (members, edge
 *               and corner neighbours), 16-bit fixed point
 *   dct         the "islow" integer DCT: Loeffler-Ligtenberg-Moschytz with
 *               13-bit constants, 2 extra bits after the row pass; or
 *               (or_jpg_encode_ex) the "ifast" and "float" scaled DCTs of
 *               Arai, Agui and Nakajima
 *   quant       IJG quality scaling of the Annex K tables, coefficients
 *               divided by 8 q with rounding half away from zero (islow),
 *               by 8 q times the AAN scale factors (ifast), or multiplied by
 *               that divisor's float reciprocal and rounded to nearest even
 *               (float)
 *   huffman     optimal tables per image: Annex K.2 code lengths (the
 *               smallest frequencies merged first, ties to the larger
 *               symbol, a reserved symbol 256), K.3 length limit 16, values
 *               by length then symbol
 * It returns the whole file (SOI, JFIF APP0 at 300 dpi, DQT, SOF0, DHT, SOS,
 * entropy-coded data with 0xFF stuffing, EOI).
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int zz[64] = {   // zigzag position -> natural index
  0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21,
  28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
  47, 55, 62, 63 };

// T.81 Annex K.1 tables, natural order
static const int std_luma[64] = {
  16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
  14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
  49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99 };
static const int std_chroma[64] = {
  17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
  47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
  99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99 };

void or_jpg_qtables(int quality, int luma[64], int chroma[64])
{
  if(quality <= 0) quality = 1;
  if(quality > 100) quality = 100;
  const long scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  for(int i = 0; i < 64; i++)
  {
    long l = (std_luma[i] * scale + 50) / 100, c = (std_chroma[i] * scale + 50) / 100;
    luma[i] = l < 1 ? 1 : (l > 255 ? 255 : (int)l);
    chroma[i] = c < 1 ? 1 : (c > 255 ? 255 : (int)c);
  }
}

// fixed-point rgb -> ycc (16 fractional bits)
#define YC_FIX(x) ((int32_t)((x) * 65536.0 + 0.5))
void or_jpg_ycc(int r, int g, int b, int *y, int *cb, int *cr)
{
  const int32_t half = 1 << 15, off = 128 << 16;
  *y = (int)((YC_FIX(0.29900) * r + YC_FIX(0.58700) * g + YC_FIX(0.11400) * b + half) >> 16);
  *cb = (int)((-YC_FIX(0.16874) * r - YC_FIX(0.33126) * g + YC_FIX(0.50000) * b + off + half - 1) >> 16);
  *cr = (int)((YC_FIX(0.50000) * r - YC_FIX(0.41869) * g - YC_FIX(0.08131) * b + off + half - 1) >> 16);
}

#define CB 13
#define P1 2
#define FX(x) ((int32_t)((x) * (1 << CB) + 0.5))
#define DESC(x, n) (((x) + (1 << ((n) - 1))) >> (n))

// islow forward DCT of one block of level-shifted samples, in place (8x scaled)
void or_jpg_fdct(int32_t *d)
{
  for(int pass = 0; pass < 2; pass++)
    for(int k = 0; k < 8; k++)
    {
      int32_t *p = pass == 0 ? d + 8 * k : d + k;
      const int s = pass == 0 ? 1 : 8;
      const int32_t t0 = p[0] + p[7 * s], t7 = p[0] - p[7 * s], t1 = p[s] + p[6 * s], t6 = p[s] - p[6 * s];
      const int32_t t2 = p[2 * s] + p[5 * s], t5 = p[2 * s] - p[5 * s], t3 = p[3 * s] + p[4 * s], t4 = p[3 * s] - p[4 * s];
      const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
      const int sh = pass == 0 ? CB - P1 : CB + P1;
      if(pass == 0) { p[0] = (t10 + t11) << P1; p[4 * s] = (t10 - t11) << P1; }
      else { p[0] = DESC(t10 + t11, P1); p[4 * s] = DESC(t10 - t11, P1); }
      const int32_t z1 = (t12 + t13) * FX(0.541196100);
      p[2 * s] = DESC(z1 + t13 * FX(0.765366865), sh);
      p[6 * s] = DESC(z1 + t12 * (-FX(1.847759065)), sh);
      int32_t a1 = t4 + t7, a2 = t5 + t6, a3 = t4 + t6, a4 = t5 + t7;
      const int32_t z5 = (a3 + a4) * FX(1.175875602);
      const int32_t b4 = t4 * FX(0.298631336), b5 = t5 * FX(2.053119869), b6 = t6 * FX(3.072711026), b7 = t7 * FX(1.501321110);
      a1 *= -FX(0.899976223); a2 *= -FX(2.562915447); a3 *= -FX(1.961570560); a4 *= -FX(0.390180644);
      a3 += z5; a4 += z5;
      p[7 * s] = DESC(b4 + a1 + a3, sh);
      p[5 * s] = DESC(b5 + a2 + a4, sh);
      p[3 * s] = DESC(b6 + a2 + a3, sh);
      p[s] = DESC(b7 + a1 + a4, sh);
    }
}

// The other two DCTs libjpeg offers (dct_method JDCT_IFAST and JDCT_FLOAT):
// the scaled DCT of Arai, Agui and Nakajima (one flow graph, 5 multiplies a
// row pass), whose outputs carry the per-frequency factors
// s_u s_v 8, s_0 = 1, s_k = sqrt(2) cos(k pi/16), which the quantiser
// divides out.  ifast: 8 fractional bits, products shifted right without
// rounding, no extra precision between the passes.  float: single precision
// in the same order.
static const double aan_s[8] = { 1.0, 1.387039845, 1.306562965, 1.175875602, 1.0, 0.785694958, 0.541196100,
                                 0.275899379 };

#define AAN8(v, c) ((int32_t)(((v) * (c)) >> 8))
void or_jpg_fdct_ifast(int32_t *d)
{
  for(int pass = 0; pass < 2; pass++)
    for(int k = 0; k < 8; k++)
    {
      int32_t *p = pass == 0 ? d + 8 * k : d + k;
      const int s = pass == 0 ? 1 : 8;
      const int32_t a0 = p[0] + p[7 * s], a7 = p[0] - p[7 * s], a1 = p[s] + p[6 * s], a6 = p[s] - p[6 * s];
      const int32_t a2 = p[2 * s] + p[5 * s], a5 = p[2 * s] - p[5 * s], a3 = p[3 * s] + p[4 * s], a4 = p[3 * s] - p[4 * s];
      // even half
      const int32_t e0 = a0 + a3, e3 = a0 - a3, e1 = a1 + a2, e2 = a1 - a2;
      p[0] = e0 + e1;
      p[4 * s] = e0 - e1;
      const int32_t r = AAN8(e2 + e3, 181);                     // cos(pi/4)
      p[2 * s] = e3 + r;
      p[6 * s] = e3 - r;
      // odd half
      const int32_t o0 = a4 + a5, o1 = a5 + a6, o2 = a6 + a7;
      const int32_t rot = AAN8(o0 - o2, 98);                    // cos(3pi/8)
      const int32_t q2 = AAN8(o0, 139) + rot;                   // cos(pi/8) - cos(3pi/8)
      const int32_t q4 = AAN8(o2, 334) + rot;                   // cos(pi/8) + cos(3pi/8)
      const int32_t q3 = AAN8(o1, 181);
      const int32_t h1 = a7 + q3, h3 = a7 - q3;
      p[5 * s] = h3 + q2;
      p[3 * s] = h3 - q2;
      p[s] = h1 + q4;
      p[7 * s] = h1 - q4;
    }
}

void or_jpg_fdct_float(float *d)
{
  for(int pass = 0; pass < 2; pass++)
    for(int k = 0; k < 8; k++)
    {
      float *p = pass == 0 ? d + 8 * k : d + k;
      const int s = pass == 0 ? 1 : 8;
      const float a0 = p[0] + p[7 * s], a7 = p[0] - p[7 * s], a1 = p[s] + p[6 * s], a6 = p[s] - p[6 * s];
      const float a2 = p[2 * s] + p[5 * s], a5 = p[2 * s] - p[5 * s], a3 = p[3 * s] + p[4 * s], a4 = p[3 * s] - p[4 * s];
      const float e0 = a0 + a3, e3 = a0 - a3, e1 = a1 + a2, e2 = a1 - a2;
      p[0] = e0 + e1;
      p[4 * s] = e0 - e1;
      const float r = (e2 + e3) * 0.707106781f;
      p[2 * s] = e3 + r;
      p[6 * s] = e3 - r;
      const float o0 = a4 + a5, o1 = a5 + a6, o2 = a6 + a7;
      const float rot = (o0 - o2) * 0.382683433f;
      const float q2 = 0.541196100f * o0 + rot;
      const float q4 = 1.306562965f * o2 + rot;
      const float q3 = o1 * 0.707106781f;
      const float h1 = a7 + q3, h3 = a7 - q3;
      p[5 * s] = h3 + q2;
      p[3 * s] = h3 - q2;
      p[s] = h1 + q4;
      p[7 * s] = h1 - q4;
    }
}

// divisors of the DCT methods for quantiser q at natural index i:
// islow 8 q; ifast 8 q s_u s_v in integer (14-bit factors, rounded to
// 3 fractional bits); float the reciprocal of 8 q s_u s_v
int or_jpg_divisor_ifast(int q, int i)
{
  const int32_t f = (int32_t)(16384.0 * aan_s[i / 8] * aan_s[i % 8] + 0.5);
  return (int)(((int64_t)q * f + (1 << 10)) >> 11);
}
float or_jpg_divisor_float(int q, int i)
{
  return (float)(1.0 / ((double)q * aan_s[i / 8] * aan_s[i % 8] * 8.0));
}

// a float coefficient times its reciprocal divisor, rounded to nearest with
// ties to even (the SIMD quantiser's conversion)
int or_jpg_quant_float(float v, float rdiv)
{
  return (int)nearbyintf(v * rdiv);
}

int or_jpg_quant(int32_t v, int q8)
{ // rounding half away from zero of v / q8
  if(v < 0) { v = -v + (q8 >> 1); return -(v >= q8 ? v / q8 : 0); }
  v += q8 >> 1;
  return v >= q8 ? v / q8 : 0;
}

// optimal code lengths (K.2 with IJG tie-breaking, K.3 limit): bits[1..16], val
void or_jpg_huff(const long freq_in[257], uint8_t bits[17], uint8_t val[256], int *nval)
{
  long freq[257];
  int codesize[257], others[257];
  memcpy(freq, freq_in, sizeof(freq));
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
  cnt[i]--;                      // the reserved symbol's code
  bits[0] = 0;
  for(int k = 1; k <= 16; k++) bits[k] = (uint8_t)cnt[k];
  int p = 0;
  for(int len = 1; len <= 32; len++)
    for(int s = 0; s < 256; s++) if(codesize[s] == len) val[p++] = (uint8_t)s;
  *nval = p;
}

typedef struct { uint8_t *d; size_t n, cap; uint32_t acc; int cnt; } obuf_t;

static void put8(obuf_t *o, int b)
{
  if(o->n == o->cap) { o->cap = o->cap ? 2 * o->cap : 65536; o->d = realloc(o->d, o->cap); }
  o->d[o->n++] = (uint8_t)b;
}
static void put16(obuf_t *o, int v) { put8(o, v >> 8); put8(o, v & 255); }
static void putbits(obuf_t *o, uint32_t code, int len)
{
  for(int k = len - 1; k >= 0; k--)
  {
    o->acc = (o->acc << 1) | ((code >> k) & 1);
    if(++o->cnt == 8)
    {
      put8(o, (int)o->acc);
      if(o->acc == 0xff) put8(o, 0);
      o->acc = 0;
      o->cnt = 0;
    }
  }
}

static int nbits(int v) { int n = 0; v = v < 0 ? -v : v; while(v) { n++; v >>= 1; } return n; }

typedef struct { uint16_t code[256]; uint8_t len[256]; } hcode_t;

static void hcodes(const uint8_t bits[17], const uint8_t *val, hcode_t *h)
{
  memset(h, 0, sizeof(*h));
  int code = 0, k = 0;
  for(int len = 1; len <= 16; len++)
  {
    for(int j = 0; j < bits[len]; j++, k++) { h->code[val[k]] = (uint16_t)code++; h->len[val[k]] = (uint8_t)len; }
    code <<= 1;
  }
}

// all blocks of the scan: (component, block x, block y) in MCU order
static int nblocks(int W, int H, int hs, int vs) { return ((W + 8 * hs - 1) / (8 * hs)) * ((H + 8 * vs - 1) / (8 * vs)) * (hs * vs + 2); }

// libjpeg's input smoothing (smoothing_factor sf, 0..100) of a plane of
// ycc samples pw x ph (replicated past the image, as the planes here are):
// a full-size component mixes each sample with its 8 neighbours,
// (m (65536 - 512 sf) + sum8 64 sf + 32768) >> 16; a 2x2-subsampled one
// forms each output from its 4 members, their 8 edge neighbours (twice) and
// 4 corner neighbours, (sum4 (16384 - 80 sf) + (2 edge + corner) 16 sf +
// 32768) >> 16.  Samples outside the image repeat the nearest edge sample.
static int plane_at(const uint8_t *pl, int pw, int ph, int x, int y)
{
  x = x < 0 ? 0 : (x >= pw ? pw - 1 : x);
  y = y < 0 ? 0 : (y >= ph ? ph - 1 : y);
  return pl[(size_t)y * pw + x];
}

static void smooth_full(const uint8_t *in, uint8_t *out, int pw, int ph, int sf)
{
  const int32_t ms = 65536 - 512 * sf, ns = 64 * sf;
  for(int y = 0; y < ph; y++) for(int x = 0; x < pw; x++)
  {
    int32_t n = 0;
    for(int j = -1; j <= 1; j++) for(int i = -1; i <= 1; i++)
      if(i || j) n += plane_at(in, pw, ph, x + i, y + j);
    out[(size_t)y * pw + x] = (uint8_t)((plane_at(in, pw, ph, x, y) * ms + n * ns + 32768) >> 16);
  }
}

static int smooth_2x2(const uint8_t *in, int pw, int ph, int x, int y, int sf)
{ // output (x, y) of members (2x..2x+1, 2y..2y+1)
  const int X = 2 * x, Y = 2 * y;
  int32_t m = 0, e = 0, c = 0;
  for(int j = 0; j < 2; j++) for(int i = 0; i < 2; i++) m += plane_at(in, pw, ph, X + i, Y + j);
  for(int i = 0; i < 2; i++) e += plane_at(in, pw, ph, X + i, Y - 1) + plane_at(in, pw, ph, X + i, Y + 2);
  for(int j = 0; j < 2; j++) e += plane_at(in, pw, ph, X - 1, Y + j) + plane_at(in, pw, ph, X + 2, Y + j);
  c = plane_at(in, pw, ph, X - 1, Y - 1) + plane_at(in, pw, ph, X + 2, Y - 1) + plane_at(in, pw, ph, X - 1, Y + 2)
    + plane_at(in, pw, ph, X + 2, Y + 2);
  return (int)((m * (16384 - 80 * sf) + (2 * e + c) * (16 * sf) + 32768) >> 16);
}

int or_jpg_encode_ex(const uint8_t *rgba, int W, int H, int stride, int quality, int hs, int vs, int dct, int smooth,
                     uint8_t **out, size_t *out_len);

int or_jpg_encode(const uint8_t *rgba, int W, int H, int stride, int quality, int hs, int vs, uint8_t **out, size_t *out_len)
{
  return or_jpg_encode_ex(rgba, W, H, stride, quality, hs, vs, 0, 0, out, out_len);
}

// dct: 0 islow, 1 ifast, 2 float (libjpeg's JDCT_ISLOW, JDCT_IFAST,
// JDCT_FLOAT); smooth: libjpeg's smoothing_factor (full-size components and
// 2x2 chroma; 2x1 chroma is not smoothed, as in libjpeg)
int or_jpg_encode_ex(const uint8_t *rgba, int W, int H, int stride, int quality, int hs, int vs, int dct, int smooth,
                     uint8_t **out, size_t *out_len)
{
  if(dct < 0 || dct > 2 || smooth < 0 || smooth > 100) return -1;
  const int mw = 8 * hs, mh = 8 * vs;
  const int mx = (W + mw - 1) / mw, my = (H + mh - 1) / mh;
  const int PW = mx * mw, PH = my * mh;         // padded luma extent
  const int CW = PW / hs, CH = PH / vs;         // chroma planes
  uint8_t *Y = malloc((size_t)PW * PH), *Cb = malloc((size_t)PW * PH), *Cr = malloc((size_t)PW * PH);
  uint8_t *cb = malloc((size_t)CW * CH), *cr = malloc((size_t)CW * CH);
  for(int y = 0; y < PH; y++) for(int x = 0; x < PW; x++)
  { // edge replication to whole MCUs
    const int sx = x < W ? x : W - 1, sy = y < H ? y : H - 1;
    const uint8_t *p = rgba + (size_t)sy * stride + 4 * sx;
    int a, b, c;
    or_jpg_ycc(p[0], p[1], p[2], &a, &b, &c);
    Y[(size_t)y * PW + x] = (uint8_t)a; Cb[(size_t)y * PW + x] = (uint8_t)b; Cr[(size_t)y * PW + x] = (uint8_t)c;
  }
  if(smooth)
  { // the full-size planes: luma always, chroma at 4:4:4
    uint8_t *t = malloc((size_t)PW * PH);
    smooth_full(Y, t, PW, PH, smooth);
    memcpy(Y, t, (size_t)PW * PH);
    if(hs * vs == 1)
    {
      smooth_full(Cb, t, PW, PH, smooth);
      memcpy(Cb, t, (size_t)PW * PH);
      smooth_full(Cr, t, PW, PH, smooth);
      memcpy(Cr, t, (size_t)PW * PH);
    }
    free(t);
  }
  for(int y = 0; y < CH; y++)
  {
    int bias = hs == 2 ? (vs == 2 ? 1 : 0) : 0;
    for(int x = 0; x < CW; x++)
    {
      if(smooth && hs == 2 && vs == 2)
      {
        cb[(size_t)y * CW + x] = (uint8_t)smooth_2x2(Cb, PW, PH, x, y, smooth);
        cr[(size_t)y * CW + x] = (uint8_t)smooth_2x2(Cr, PW, PH, x, y, smooth);
        continue;
      }
      int s1 = 0, s2 = 0;
      for(int j = 0; j < vs; j++) for(int i = 0; i < hs; i++)
      {
        s1 += Cb[(size_t)(vs * y + j) * PW + hs * x + i];
        s2 += Cr[(size_t)(vs * y + j) * PW + hs * x + i];
      }
      if(hs * vs == 1) { cb[(size_t)y * CW + x] = (uint8_t)s1; cr[(size_t)y * CW + x] = (uint8_t)s2; continue; }
      const int sh = hs * vs == 4 ? 2 : 1;
      cb[(size_t)y * CW + x] = (uint8_t)((s1 + bias) >> sh);
      cr[(size_t)y * CW + x] = (uint8_t)((s2 + bias) >> sh);
      bias ^= hs * vs == 4 ? 3 : 1;
    }
  }
  int ql[64], qc[64];
  or_jpg_qtables(quality, ql, qc);
  // coefficients of every block in MCU order
  const int nb = nblocks(W, H, hs, vs);
  int16_t *coef = malloc(sizeof(int16_t) * 64 * (size_t)nb);
  uint8_t *comp = malloc((size_t)nb);
  int b = 0;
  for(int m = 0; m < mx * my; m++)
  {
    const int ux = m % mx, uy = m / mx;
    for(int c = 0; c < 3; c++)
    {
      const int bw = c == 0 ? hs : 1, bh = c == 0 ? vs : 1;
      const uint8_t *pl = c == 0 ? Y : (c == 1 ? cb : cr);
      const int pw = c == 0 ? PW : CW;
      const int *q = c == 0 ? ql : qc;
      // blocks of the component's own extent; the others of the last MCUs
      // are dummies: zero with the DC of the block to the left (right edge)
      // or of the last block of the row above (bottom rows)
      const int wib = c == 0 ? (W + 7) / 8 : ((W + hs - 1) / hs + 7) / 8;
      const int hib = c == 0 ? (H + 7) / 8 : ((H + vs - 1) / vs + 7) / 8;
      for(int j = 0; j < bh; j++) for(int i = 0; i < bw; i++, b++)
      {
        comp[b] = (uint8_t)c;
        const int bx = ux * bw + i, by = uy * bh + j;
        int16_t *o16 = coef + (size_t)b * 64;
        if(by >= hib || bx >= wib)
        {
          memset(o16, 0, 64 * sizeof(int16_t));
          o16[0] = by >= hib ? coef[(size_t)(b - i - 1) * 64] : coef[(size_t)(b - 1) * 64];
          continue;
        }
        int32_t d[64];
        const int x0 = bx * 8, y0 = by * 8;
        for(int k = 0; k < 64; k++) d[k] = (int32_t)pl[(size_t)(y0 + k / 8) * pw + x0 + k % 8] - 128;
        if(dct == 2)
        {
          float f[64];
          for(int k = 0; k < 64; k++) f[k] = (float)d[k];
          or_jpg_fdct_float(f);
          for(int k = 0; k < 64; k++) o16[k] = (int16_t)or_jpg_quant_float(f[zz[k]], or_jpg_divisor_float(q[zz[k]], zz[k]));
        }
        else if(dct == 1)
        {
          or_jpg_fdct_ifast(d);
          for(int k = 0; k < 64; k++) o16[k] = (int16_t)or_jpg_quant(d[zz[k]], or_jpg_divisor_ifast(q[zz[k]], zz[k]));
        }
        else
        {
          or_jpg_fdct(d);
          for(int k = 0; k < 64; k++) o16[k] = (int16_t)or_jpg_quant(d[zz[k]], q[zz[k]] << 3);
        }
      }
    }
  }
  // symbol statistics, then tables
  long freq[4][257];
  memset(freq, 0, sizeof(freq));
  int last[3] = { 0, 0, 0 };
  for(int k = 0; k < nb; k++)
  {
    const int16_t *z = coef + (size_t)k * 64;
    const int t = comp[k] ? 1 : 0;
    freq[2 * t][nbits(z[0] - last[comp[k]])]++;
    last[comp[k]] = z[0];
    int run = 0;
    for(int i = 1; i < 64; i++)
    {
      if(!z[i]) { run++; continue; }
      while(run > 15) { freq[2 * t + 1][0xf0]++; run -= 16; }
      freq[2 * t + 1][(run << 4) | nbits(z[i])]++;
      run = 0;
    }
    if(run) freq[2 * t + 1][0]++;
  }
  uint8_t bits[4][17], val[4][256];
  int nv[4];
  hcode_t hc[4];
  for(int t = 0; t < 4; t++) { or_jpg_huff(freq[t], bits[t], val[t], nv + t); hcodes(bits[t], val[t], hc + t); }
  // file
  obuf_t o = { 0 };
  put16(&o, 0xffd8);
  put16(&o, 0xffe0); put16(&o, 16);                                   // JFIF APP0, 300 dpi
  put8(&o, 'J'); put8(&o, 'F'); put8(&o, 'I'); put8(&o, 'F'); put8(&o, 0);
  put8(&o, 1); put8(&o, 1); put8(&o, 1); put16(&o, 300); put16(&o, 300); put8(&o, 0); put8(&o, 0);
  for(int t = 0; t < 2; t++)
  {
    put16(&o, 0xffdb); put16(&o, 67); put8(&o, t);
    for(int k = 0; k < 64; k++) put8(&o, (t ? qc : ql)[zz[k]]);
  }
  put16(&o, 0xffc0); put16(&o, 17); put8(&o, 8); put16(&o, H); put16(&o, W); put8(&o, 3);
  put8(&o, 1); put8(&o, (hs << 4) | vs); put8(&o, 0);
  put8(&o, 2); put8(&o, 0x11); put8(&o, 1);
  put8(&o, 3); put8(&o, 0x11); put8(&o, 1);
  for(int t = 0; t < 4; t++)
  {
    put16(&o, 0xffc4); put16(&o, 2 + 1 + 16 + nv[t]); put8(&o, ((t & 1) << 4) | (t >> 1));
    for(int k = 1; k <= 16; k++) put8(&o, bits[t][k]);
    for(int k = 0; k < nv[t]; k++) put8(&o, val[t][k]);
  }
  put16(&o, 0xffda); put16(&o, 12); put8(&o, 3);
  put8(&o, 1); put8(&o, 0x00); put8(&o, 2); put8(&o, 0x11); put8(&o, 3); put8(&o, 0x11);
  put8(&o, 0); put8(&o, 63); put8(&o, 0);
  last[0] = last[1] = last[2] = 0;
  for(int k = 0; k < nb; k++)
  {
    const int16_t *z = coef + (size_t)k * 64;
    const int t = comp[k] ? 1 : 0;
    const hcode_t *dc = hc + 2 * t, *ac = hc + 2 * t + 1;
    int diff = z[0] - last[comp[k]];
    last[comp[k]] = z[0];
    int n = nbits(diff);
    putbits(&o, dc->code[n], dc->len[n]);
    if(n) putbits(&o, (uint32_t)(diff < 0 ? diff - 1 : diff) & ((1u << n) - 1), n);
    int run = 0;
    for(int i = 1; i < 64; i++)
    {
      if(!z[i]) { run++; continue; }
      while(run > 15) { putbits(&o, ac->code[0xf0], ac->len[0xf0]); run -= 16; }
      n = nbits(z[i]);
      const int s = (run << 4) | n;
      putbits(&o, ac->code[s], ac->len[s]);
      putbits(&o, (uint32_t)(z[i] < 0 ? z[i] - 1 : z[i]) & ((1u << n) - 1), n);
      run = 0;
    }
    if(run) putbits(&o, ac->code[0], ac->len[0]);
  }
  if(o.cnt) putbits(&o, 0x7f, 8 - o.cnt);                           // pad with ones
  put16(&o, 0xffd9);
  free(Y); free(Cb); free(Cr); free(cb); free(cr); free(coef); free(comp);
  *out = o.d;
  *out_len = o.n;
  return 0;
}

void or_jpg_free(uint8_t *p) { free(p); }
