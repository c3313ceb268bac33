// ops_hilite.h -- per-texel arithmetic of the hilite module's four kernels
// (wavelet highlight inpainting on the mosaic, src/pipe/modules/hilite):
//   half      half.comp:20-71      mosaic block -> rgba (clipped sites skipped)
//   reduce    reduce.comp:19-60    clipping-aware 5x5 binomial reduce + edge
//   assemble  assemble.comp:23-118 expand the coarse level, rescale it to the
//                                  fine level's unclipped channels, blend
//   doub      doub.comp:20-121     back onto the mosaic sites of each block
// FP32 arithmetic in the reference's operation order; the mosaic is read
// as f16 (denoise output) or R16_UNORM (raw input) through MosAt.
#pragma once
#include "common.h"

struct hl_args
{
  float wb[3];       // white balance (unit weights when |wb|^2 <= 1e-3)
  float white;       // clipping threshold as given
  float white_pos;   // the same, 1 when not positive (reduce.comp:27-28)
  float ds;          // desat^2 (reduce.comp:48)
  float soft;
};

// the arguments of every hilite node from its push constants' white balance
// and its params white, desat, soft (reduce.comp:23-28, assemble.comp)
inline void hl_args_init(hl_args &a, const float *wb4, const float *par)
{
  float wb[3] = { wb4[0], wb4[1], wb4[2] };
  if(!(wb[0] * wb[0] + wb[1] * wb[1] + wb[2] * wb[2] > 1e-3f)) wb[0] = wb[1] = wb[2] = 1.0f;
  for(int k = 0; k < 3; k++) a.wb[k] = wb[k];
  a.white = par[0];
  a.white_pos = par[0] > 0.0f ? par[0] : 1.0f;
  a.ds = par[1] * par[1];
  a.soft = par[2];
}

// binomial taps 1 4 6 4 1 / 16 and the edge taps 1 2 0 -2 -1
DEV float hl_w(int i) { return i == 2 ? 0.375f : ((i & 1) ? 0.25f : 0.0625f); }
DEV float hl_sw(int i) { return i == 0 ? 1.0f : (i == 1 ? 2.0f : (i == 2 ? 0.0f : (i == 3 ? -2.0f : -1.0f))); }

// X-Trans: the 3x3 block at (3x,3y) read by texelFetch (zero outside),
// c[3i+j] = texel (3x+i, 3y+j)
template<class MosAt>
DEV void hl_block3(MosAt at, int x, int y, float c[9])
{
#pragma unroll
  for(int i = 0; i < 3; i++)
#pragma unroll
    for(int j = 0; j < 3; j++) c[3 * i + j] = at(3 * x + i, 3 * y + j);
}

// Bayer: textureGather at the centre of the block at (2x,2y), mirrored
// addresses: x = (0,1), y = (1,1), z = (1,0), w = (0,0)
template<class MosAt>
DEV float4 hl_gather2(MosAt at, int x, int y)
{
  return make_float4(at(2 * x, 2 * y + 1), at(2 * x + 1, 2 * y + 1), at(2 * x + 1, 2 * y), at(2 * x, 2 * y));
}

template<int XT, class MosAt>
DEV float4 hl_half_px(MosAt at, int x, int y, float white)
{
  float4 r;
  if(XT)
  {
    float c[9];
    hl_block3(at, x, y, c);
    if(c[1] >= white) c[1] = c[7];
    if(c[7] >= white) c[7] = c[1];
    if(c[3] >= white) c[3] = c[5];
    if(c[5] >= white) c[5] = c[3];
    const float maxg = fmaxf(fmaxf(fmaxf(c[0], c[2]), c[4]), fmaxf(c[6], c[8]));
    if(maxg >= white) c[0] = c[2] = c[4] = c[6] = c[8] = 1.0f;
    const float p = (c[1] + c[7]) * 0.5f, q = (c[3] + c[5]) * 0.5f;
    const bool odd = ((x + y) & 1) != 0;
    r.x = odd ? p : q;
    r.z = odd ? q : p;
    r.y = (c[0] + c[2] + c[4] + c[6] + c[8]) / 5.0f;
  }
  else
  {
    float4 g = hl_gather2(at, x, y);
    if(g.x >= white) g.x = g.z;
    if(g.z >= white) g.z = g.x;
    r.x = g.w;
    r.y = (g.x + g.z) / 2.0f;
    r.z = g.y;
  }
  r.w = 1.0f;
  return r;
}

// per fine texel of a reduce: its luminance (edge term) and, when no
// channel reaches white, the desaturated colour that enters the blur;
// valid = 0 marks a clipped texel, whose colour is then 0 (reduce.comp:37-54)
struct hl_tap
{
  float r, g, b, l, valid;
};

DEV hl_tap hl_prep(float4 t, const hl_args &a)
{
  hl_tap o;
  float r = t.x, g = t.y, b = t.z;
  o.l = lum2020(r, g, b);
  const float rw = r * a.wb[0], gw = g * a.wb[1], bw = b * a.wb[2];
  const float cmax = fmaxf(rw, fmaxf(gw, bw)), cmin = fminf(rw, fminf(gw, bw));
  const float sat = (cmax - cmin) / fmaxf(1e-3f, cmax);
  if(r < a.white_pos && g < a.white_pos && b < a.white_pos)
  {
    const float s = smoothstepf(0.2f, 1.0f, fmaxf(r, fmaxf(g, b)) / a.white_pos);
    float tt = smoothstepf(0.15f, 0.9f, sat);
    tt = clampf(5.0f * a.ds * a.ds * s * tt, 0.0f, 1.0f);
    const float m = cmax + cmin;
    o.r = mixf(r, m / a.wb[0] * 0.5f, tt);
    o.g = mixf(g, m / a.wb[1] * 0.5f, tt);
    o.b = mixf(b, m / a.wb[2] * 0.5f, tt);
    o.valid = 1.0f;
  }
  else o.r = o.g = o.b = o.valid = 0.0f;
  return o;
}

// the 5x5 sums of a reduce over prepared taps, jj outer and ii inner as the
// reference loops; Tap(ii,jj) returns the hl_tap at fine offset (ii-2,jj-2)
template<class Tap>
DEV float4 hl_reduce_sum(Tap tap)
{
  float ex = 0.0f, ey = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, wgt = 0.0f;
#pragma unroll
  for(int jj = 0; jj < 5; jj++)
#pragma unroll
    for(int ii = 0; ii < 5; ii++)
    {
      const hl_tap t = tap(ii, jj);
      ex += (hl_w(jj) * hl_sw(ii)) * t.l;
      ey += (hl_w(ii) * hl_sw(jj)) * t.l;
      // only valid taps count: a clipped tap has colour +0 and valid 0, so
      // adding its terms leaves every sum unchanged (they start at +0 and
      // can never become -0), which spares the branch per tap
      const float u = hl_w(ii) * hl_w(jj);
      cr += t.r * u;
      cg += t.g * u;
      cb += t.b * u;
      wgt += t.valid * u;
    }
  float4 o;
  if(wgt == 0.0f) o.x = o.y = o.z = 1.0f;
  else { o.x = cr / wgt; o.y = cg / wgt; o.z = cb / wgt; }
  o.w = fsqrt(ex * ex + ey * ey);
  return o;
}

// expansion of the coarse level at fine texel (ox,oy): the binomial taps of
// the texel's parity, ii outer, jj inner; Tap(i,j) mirrored coarse rgba
template<class Tap>
DEV float3 hl_expand(Tap tap, int ox, int oy)
{
  const int ix = ox >> 1, iy = oy >> 1;
  const int dx = ox & 1, dy = oy & 1;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
  for(int ii = -1; ii <= 1; ii++)
  {
    if(dx && ii < 0) continue;
    const float wx = dx ? hl_w(2 * ii + 1) : hl_w(2 * ii + 2);
#pragma unroll
    for(int jj = -1; jj <= 1; jj++)
    {
      if(dy && jj < 0) continue;
      const float wy = dy ? hl_w(2 * jj + 1) : hl_w(2 * jj + 2);
      const float4 t = tap(ix + ii, iy + jj);
      cx += t.x * wy * wx;
      cy += t.y * wy * wx;
      cz += t.z * wy * wx;
    }
  }
  // the taps of either parity sum to 0.5 per axis, so the weight sum the
  // reference divides by is exactly 0.25 (dyadic products and sums) and
  // the division an exact scaling by 4
  return make_float3(cx * 4.0f, cy * 4.0f, cz * 4.0f);
}

// colour-ratio scale of the reconstruction: the channels weighted by how
// much they dominate the blurry colour (assemble.comp:89-96, doub.comp)
DEV float hl_scale(float3 up, float vr, float vg, float vb)
{
  const float sr = vr / fmaxf(0.001f, up.x);
  const float sg = vg / fmaxf(0.001f, up.y);
  const float sb = vb / fmaxf(0.001f, up.z);
  const float wr = fexp(up.x - fmaxf(up.y, up.z));
  const float wg = fexp(up.y - fmaxf(up.x, up.z));
  const float wb = fexp(up.z - fmaxf(up.x, up.y));
  return (sr * wr + sg * wg + sb * wb) / (wr + wg + wb);
}

DEV float4 hl_assemble_px(float3 up, float4 fine, const hl_args &a)
{
  const float scale = hl_scale(up, fine.x, fine.y, fine.z);
  const float rx = clampf(up.x * scale, -65535.0f, 65535.0f);
  const float ry = clampf(up.y * scale, -65535.0f, 65535.0f);
  const float rz = clampf(up.z * scale, -65535.0f, 65535.0f);
  float t = a.soft;
  if(fine.x >= a.white || fine.y >= a.white || fine.z >= a.white) t = 1.0f;
  float fa = fine.w;
  if(__builtin_isnan(fa)) fa = 0.0f;
  t = clampf(mixf(t, 1.0f, fsqrt(fmaxf(0.0f, fa))), 0.0f, 1.0f);
  return make_float4(mixf(fine.x, rx, t), mixf(fine.y, ry, t), mixf(fine.z, rz, t), 1.0f);
}

// doub for Bayer: the gathered block g and the coarse colour up -> the block
// with clipped sites filled (returned in gather order x y z w)
DEV float4 hl_doub_bayer(float4 g, float3 up, float white)
{
  const float ming = fminf(g.x, g.z);
  const float scale = hl_scale(up, g.w, ming, g.y);
  const float softw = 0.97f * white;
  const float maxrgb = fmaxf(fmaxf(g.x, g.y), fmaxf(g.z, g.w));
  if(maxrgb > softw)
  { // mix(c, scale * upsm.gbgr, t)
    const float t = smoothstepf(softw, white, maxrgb);
    g.x = mixf(g.x, scale * up.y, t);
    g.y = mixf(g.y, scale * up.z, t);
    g.z = mixf(g.z, scale * up.y, t);
    g.w = mixf(g.w, scale * up.x, t);
  }
  return g;
}

// doub for X-Trans: c[9] as hl_block3, up already swapped by block parity
DEV void hl_doub_xtrans(float c[9], float3 up, float white)
{
  const float minr = fminf(c[1], c[7]), minb = fminf(c[3], c[5]);
  const float ming = fminf(fminf(fminf(c[0], c[2]), c[4]), fminf(c[6], c[8]));
  const float scale = hl_scale(up, minr, ming, minb);
  const float maxr = fmaxf(c[1], c[7]), maxb = fmaxf(c[3], c[5]);
  const float maxg = fmaxf(fmaxf(fmaxf(c[0], c[2]), c[4]), fmaxf(c[6], c[8]));
  const float softw = 0.97f * white;
  const float maxrgb = fmaxf(maxr, fmaxf(maxg, maxb));
  if(maxrgb > softw)
  {
    float t = smoothstepf(softw, white, maxrgb);
    c[1] = mixf(c[1], up.x * scale, t);
    c[7] = mixf(c[7], up.x * scale, t);
    c[3] = mixf(c[3], up.z * scale, t);
    c[5] = mixf(c[5], up.z * scale, t);
    t = smoothstepf(softw, white, ming);
    c[0] = mixf(c[0], up.y * scale, t);
    c[2] = mixf(c[2], up.y * scale, t);
    c[4] = mixf(c[4], up.y * scale, t);
    c[6] = mixf(c[6], up.y * scale, t);
    c[8] = mixf(c[8], up.y * scale, t);
  }
}

// ---- one texel / tile of each node (per-node kernels and the batch) -------
// The mosaic format is a template parameter (U16: R16_UNORM, else f16) and
// addresses are formed without branches (mirror_near, clamped loads with a
// select for texelFetch's zero outside): every coordinate read here lies
// within two texels of the image, where mirror_near equals mirror, and the
// loads of a texel are in flight together.

template<int U16>
DEV float hl_mos_centre(const dimg &in, int px, int py)
{
  return ld_mt<U16>(in, mirror_near(px, in.wd), mirror_near(py, in.ht));
}

template<int U16>
DEV float hl_mos_fetch(const dimg &in, int px, int py)
{
  const float v = ld_mt<U16>(in, min(max(px, 0), in.wd - 1), min(max(py, 0), in.ht - 1));
  return (px >= 0 && py >= 0 && px < in.wd && py < in.ht) ? v : 0.0f;
}

DEV float4 hl_centre_h4(const dimg &b, int x, int y) { return ld_h4(b, mirror_near(x, b.wd), mirror_near(y, b.ht)); }

template<int XT, int U16>
DEV float4 hl_half_at(const dimg &in, int x, int y, float white)
{
  if(XT) return hl_half_px<1>([&](int px, int py) { return hl_mos_fetch<U16>(in, px, py); }, x, y, white);
  return hl_half_px<0>([&](int px, int py) { return hl_mos_centre<U16>(in, px, py); }, x, y, white);
}

// reduce over BX x BY coarse texels from (X0,Y0): their fine footprint
// (2BX+3) x (2BY+3) is loaded once (mirrored addresses) and each fine texel
// prepared once in LDS -- the reference evaluates the desaturation of every
// fine texel for each of the up to 9 coarse texels whose 5x5 window covers
// it.  Every thread of the BX x BY work group calls this (it synchronises).
#define HLR_W (2 * BX + 3)
#define HLR_H (2 * BY + 3)
struct hl_lds
{
  float r[HLR_H * HLR_W], g[HLR_H * HLR_W], b[HLR_H * HLR_W], l[HLR_H * HLR_W], v[HLR_H * HLR_W];
};

DEV void hl_reduce_tile(const dimg &in, const dimg &out, const hl_args &a, int X0, int Y0, hl_lds &s)
{
  const int tid = threadIdx.y * BX + threadIdx.x;
  constexpr int N = HLR_W * HLR_H, S = BX * BY, U = 4;   // four loads in flight per pass
  for(int i0 = tid; i0 < N; i0 += U * S)
  {
    float4 t[U];
#pragma unroll
    for(int u = 0; u < U; u++)
    {
      const int i = min(i0 + u * S, N - 1);
      const int lx = i % HLR_W, ly = i / HLR_W;
      t[u] = hl_centre_h4(in, 2 * X0 - 2 + lx, 2 * Y0 - 2 + ly);
    }
#pragma unroll
    for(int u = 0; u < U; u++)
    {
      const int i = i0 + u * S;
      if(i >= N) break;
      const hl_tap p = hl_prep(t[u], a);
      s.r[i] = p.r; s.g[i] = p.g; s.b[i] = p.b; s.l[i] = p.l; s.v[i] = p.valid;
    }
  }
  __syncthreads();
  const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const int base = 2 * threadIdx.y * HLR_W + 2 * threadIdx.x;
  st_h4(out, x, y, hl_reduce_sum([&](int ii, int jj) {
    const int k = base + jj * HLR_W + ii;
    return hl_tap{ s.r[k], s.g[k], s.b[k], s.l[k], s.v[k] };
  }));
}

// (x,y) inside the fine level
DEV float4 hl_assemble_at(const dimg &fine, const dimg &coarse, const hl_args &a, int x, int y)
{
  const float3 up = hl_expand([&](int px, int py) { return hl_centre_h4(coarse, px, py); }, x, y);
  return hl_assemble_px(up, ld_h4(fine, x, y), a);
}

// one mosaic site of the output: f16 or R16_UNORM store, dropped outside
DEV void hl_store(const dimg &o, int x, int y, float v)
{
  if(x >= o.wd || y >= o.ht) return;
  if(o.fmt == DIMG_UI16) st_u16(o, x, y, v);
  else st_h1(o, x, y, v);
}

// doub of mosaic block (x,y).  doub.comp:24-25 bounds the block index by
// the output's extent, not the node's: the threads of the last 8x8 work
// groups past the node extent run too, reading zero coarse texels
// (texelFetch) and writing the sites that fall inside the output
template<int XT, int U16>
DEV void hl_doub_at(const dimg &in, const dimg &coarse, const dimg &out, int x, int y, float white)
{
  float4 u = ld_h4(coarse, min(x, coarse.wd - 1), min(y, coarse.ht - 1));
  if(x >= coarse.wd || y >= coarse.ht) u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if(XT)
  {
    const bool swap = ((x + y) & 1) == 0;
    const float3 up = swap ? make_float3(u.z, u.y, u.x) : make_float3(u.x, u.y, u.z);
    float c[9];
    hl_block3([&](int px, int py) { return hl_mos_fetch<U16>(in, px, py); }, x, y, c);
    hl_doub_xtrans(c, up, white);
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
      for(int j = 0; j < 3; j++) hl_store(out, 3 * x + i, 3 * y + j, c[3 * i + j]);
  }
  else
  {
    const float4 g = hl_doub_bayer(hl_gather2([&](int px, int py) { return hl_mos_centre<U16>(in, px, py); }, x, y),
                                   make_float3(u.x, u.y, u.z), white);
    hl_store(out, 2 * x, 2 * y, g.w);
    hl_store(out, 2 * x + 1, 2 * y, g.z);
    hl_store(out, 2 * x, 2 * y + 1, g.x);
    hl_store(out, 2 * x + 1, 2 * y + 1, g.y);
  }
}
