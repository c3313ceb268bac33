// k_denoise.hip -- the denoise module's nodes (src/pipe/modules/denoise):
// noop, half, downcov, down (x3), assemble, doub.
//
// Layouts in HBM: the CFA is ui16 W x H; the half-resolution buffers are
// rgba f16 (8 B/texel) on the block grid Q = (W/b, H/b), b = 2 (Bayer) or
// 3 (X-Trans); doub writes one f16 channel at W x H.  The down/downcov
// outputs are stored "polyphase sorted" (texel x goes to x/2 + (x&1)*(n+1)/2)
// exactly as the reference stores them, because the next level filters the
// sorted image and assemble walks the same permutation back.
//
// downcov and down are 5x5 stencils: a 64x4 work group stages its tile
// plus a 2 texel mirrored halo in LDS once and every texel reads its
// neighbourhood from there.
#include "launch.h"

struct dn_args
{
  float    wb[4], black[4], white[4];
  float    inv_range[4];     // 1/(white-black)
  int      crop[4];
  float    noise_a, noise_b;
  int      level;
  int      block;
  uint32_t filters;
  int      gainmap;
  float    strength, luma, detail, edges[4];
  float    edge_gain[3];     // exp2(12*edges.rgb + edges.a) of noise_sigma
  float    level_scale;      // 0.7^level / block factor (down.comp:85)
};

// push-constant layouts (std140, as in the GLSL push_t blocks)
struct pc_rgbc { float wb[4], black[4], white[4]; int32_t crop[4]; };

static void read_params(dn_args &a, const node_call &c)
{
  const float *p = (const float *)c.params;
  a.strength = p[0];
  a.luma = p[1];
  a.detail = p[2];
  for(int k = 0; k < 4; k++) a.edges[k] = p[4 + k];
  for(int k = 0; k < 3; k++) a.edge_gain[k] = exp2f(12.0f * a.edges[k] + a.edges[3]);
}

static void read_rgbc(dn_args &a, const void *pc)
{
  const pc_rgbc *q = (const pc_rgbc *)pc;
  for(int k = 0; k < 4; k++)
  {
    a.wb[k] = q->wb[k];
    a.black[k] = q->black[k];
    a.white[k] = q->white[k];
    a.inv_range[k] = 1.0f / (q->white[k] - q->black[k]);
    a.crop[k] = q->crop[k];
  }
}

// noise.glsl:1-11: sqrt(a + max(0,(v-black)/(white-black)) b), scaled by the
// per-channel edge gains and clamped to [1e-3, 1e3]
DEV void noise_sigma(float o[3], const dn_args &a, float black, float inv_range, float val)
{
  const float s = fsqrt(a.noise_a + fmaxf(0.0f, (val - black) * inv_range) * a.noise_b);
  o[0] = clampf(a.edge_gain[0] * s, 1e-3f, 1e3f);
  o[1] = clampf(a.edge_gain[1] * s, 1e-3f, 1e3f);
  o[2] = clampf(a.edge_gain[2] * s, 1e-3f, 1e3f);
}

// ---- noop (noop.comp:36-57, strength <= 0) -------------------------------

__global__ void __launch_bounds__(256) k_dn_noop(dimg in, dimg out, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  // ui16 single channel reads as (v, 0, 0)
  const float v = fetch_u16(in, x + a.crop[0], y + a.crop[1]);
  const float r = fmaxf(0.0f, (v - a.black[0]) / (a.white[0] - a.black[0]));
  const float g = fmaxf(0.0f, (0.0f - a.black[1]) / (a.white[1] - a.black[1]));
  const float b = fmaxf(0.0f, (0.0f - a.black[2]) / (a.white[2] - a.black[2]));
  st_h4(out, x, y, make_float4(r, g, b, 1.0f));
}

// ---- half (half.comp:24-74) ------------------------------------------------

__global__ void __launch_bounds__(256) k_dn_half(dimg in, dimg out, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float white = a.white[1];
  float4 rgba;
  if(a.filters == 9)
  {
    const int bx = a.crop[0] + 3 * x, by = a.crop[1] + 3 * y;
    const float c0 = fetch_u16(in, bx + 0, by + 0), c1 = fetch_u16(in, bx + 0, by + 1), c2 = fetch_u16(in, bx + 0, by + 2);
    const float c3 = fetch_u16(in, bx + 1, by + 0), c4 = fetch_u16(in, bx + 1, by + 1), c5 = fetch_u16(in, bx + 1, by + 2);
    const float c6 = fetch_u16(in, bx + 2, by + 0), c7 = fetch_u16(in, bx + 2, by + 1), c8 = fetch_u16(in, bx + 2, by + 2);
    const float colx = (c1 + c7) * 0.5f, coly = (c3 + c5) * 0.5f;
    if(((x + y + a.crop[0] + a.crop[1]) & 1) > 0) { rgba.x = colx; rgba.z = coly; }
    else                                          { rgba.z = colx; rgba.x = coly; }
    rgba.y = (c0 + c2 + c4 + c6 + c8) * 0.2f;
    rgba.w = 1.0f;
  }
  else
  { // textureGather of the 2x2 block at (crop + 2 ipos)
    const int i0 = a.crop[0] + 2 * x, j0 = a.crop[1] + 2 * y;
    const int x0 = mirror(i0, in.wd), x1 = mirror(i0 + 1, in.wd);
    const int y0 = mirror(j0, in.ht), y1 = mirror(j0 + 1, in.ht);
    float gx = ld_u16(in, x0, y1);         // (i0,j1)
    const float gy = ld_u16(in, x1, y1);   // (i1,j1)
    float gz = ld_u16(in, x1, y0);         // (i1,j0)
    const float gw = ld_u16(in, x0, y0);   // (i0,j0)
    if(gx >= white) gx = gz;
    if(gz >= white) gz = gx;
    rgba = make_float4(gw, (gx + gz) * 0.5f, gy, 1.0f);
  }
  st_h4(out, x, y, rgba);
}

// ---- LDS tile of an rgba f16 image with a 2 texel mirrored halo -----------

#define TH 2
#define TW (BX + 2 * TH)
#define TT (BY + 2 * TH)

DEV void load_tile4(float4 *tile, const dimg &in, int ox, int oy)
{
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < TW * TT; i += BX * BY)
  {
    const int ty = i / TW, tx = i - ty * TW;
    tile[i] = ld_h4(in, mirror(ox + tx, in.wd), mirror(oy + ty, in.ht));
  }
}

// ---- downcov (downcov.comp:41-63 + cov.glsl:21-138) ---------------------

__global__ void __launch_bounds__(256) k_dn_downcov(dimg in, dimg out, dimg cov, dn_args a)
{
  __shared__ float4 tile[TW * TT];
  __shared__ float lum[TW * TT], rlum[TW * TT];
  const int X0 = blockIdx.x * BX, Y0 = blockIdx.y * BY;
  load_tile4(tile, in, X0 + a.crop[0] - TH, Y0 + a.crop[1] - TH);
  __syncthreads();
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < TW * TT; i += BX * BY)
  {
    const float4 v = tile[i];
    const float l = lum2020(v.x, v.y, v.z);
    lum[i] = l;
    rlum[i] = 1.0f / l;    // exact: the statistics feed the ill-conditioned eigen choice
  }
  __syncthreads();
  const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const int c0 = threadIdx.y * TW + threadIdx.x;   // tile index of offset (-2,-2)

  float mw0 = 0, mw1 = 0, mb0 = 0, mb1 = 0, smw = 0, smb = 0;
#pragma unroll
  for(int j = 0; j < 5; j++)
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const float px = lum[c0 + j * TW + i], rp = rlum[c0 + j * TW + i];
      mw0 += (float)(i - 2) * px; mw1 += (float)(j - 2) * px;
      smw += px;
      mb0 += (float)(i - 2) / px; mb1 += (float)(j - 2) / px;
      smb += rp;
    }
  mw0 /= smw; mw1 /= smw;
  mb0 /= smb; mb1 /= smb;
  float Sw[4] = {0, 0, 0, 0}, Sb[4] = {0, 0, 0, 0}, sw = 0, sb = 0, mean_b = 0;
#pragma unroll
  for(int j = 0; j < 5; j++)
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const float px = lum[c0 + j * TW + i], rp = rlum[c0 + j * TW + i];
      mean_b += px / 25.0f;
      float p2 = px * px;
      float dx = (float)(i - 2) - mw0, dy = (float)(j - 2) - mw1;
      Sw[0] += p2 * dx * dx; Sw[1] += p2 * dx * dy; Sw[2] += p2 * dy * dx; Sw[3] += p2 * dy * dy;
      sw += p2;
      dx = (float)(i - 2) - mb0; dy = (float)(j - 2) - mb1;
      p2 = 1.0f / (px * px);
      Sb[0] += p2 * dx * dx; Sb[1] += p2 * dx * dy; Sb[2] += p2 * dy * dx; Sb[3] += p2 * dy * dy;
      sb += p2;
    }
#pragma unroll
  for(int k = 0; k < 4; k++) { Sw[k] /= sw; Sb[k] /= sb; }
  float S[4];
  const bool usew = det2(Sw) < det2(Sb);
#pragma unroll
  for(int k = 0; k < 4; k++) S[k] = usew ? Sw[k] : Sb[k];
  float ev[2], e0[2], e1[2];
  evd2x2(ev, e0, e1, S);
  ev[1] *= 0.05f;
  ev[0] = clampf(ev[0], 0.01f, 25.0f);
  ev[1] = clampf(ev[1], 0.01f, 25.0f);
  if(cov.p) st_h4(cov, x + a.crop[0], y + a.crop[1], make_float4(ev[0], ev[1], e0[0], e0[1]));
  const float ri0 = -0.5f * frcp(ev[0]), ri1 = -0.5f * frcp(ev[1]);
  float r0 = 0, r1 = 0, r2 = 0, wt = 0;
  const float hot = 2.0f * mean_b;
#pragma unroll
  for(int j = 0; j < 5; j++)
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const float4 v = tile[c0 + j * TW + i];
      if(v.x > hot) continue;   // hot pixel
      const float x0 = (float)(i - 2) * e0[0] + (float)(j - 2) * e0[1];
      const float x1 = (float)(i - 2) * e1[0] + (float)(j - 2) * e1[1];
      const float w = fmaxf(1e-9f, fexp(x0 * x0 * ri0 + x1 * x1 * ri1));
      r0 += w * v.x; r1 += w * v.y; r2 += w * v.z;
      wt += w;
    }
  const float d = frcp(fmaxf(wt, 1e-8f));
  float edge = clampf(75.0f * fmaxf(0.0f, ev[1] - 0.09f), 0.0f, 1.0f);
  edge = smoothstepf(0.4f, 0.75f, edge);
  edge = clampf(0.02f + edge, 0.0f, 1.0f);
  st_h4(out, sorted1(x, out.wd), sorted1(y, out.ht), make_float4(r0 * d, r1 * d, r2 * d, edge));
}

// ---- down, levels 1..3 (down.comp:59-108) --------------------------------

DEV float gamma08(float f) { return f < 0.0f ? f : fpow(f, 0.8f); }

// bilinear from the tile at base texel (bx,by) with weights (fa,fc)
DEV float4 tile_bilin(const float4 *tile, int b, float fa, float fc)
{
  const float4 p00 = tile[b], p10 = tile[b + 1], p01 = tile[b + TW], p11 = tile[b + TW + 1];
  float4 r;
  r.x = (p00.x * (1.f - fa) + p10.x * fa) * (1.f - fc) + (p01.x * (1.f - fa) + p11.x * fa) * fc;
  r.y = (p00.y * (1.f - fa) + p10.y * fa) * (1.f - fc) + (p01.y * (1.f - fa) + p11.y * fa) * fc;
  r.z = (p00.z * (1.f - fa) + p10.z * fa) * (1.f - fc) + (p01.z * (1.f - fa) + p11.z * fa) * fc;
  return r;
}

__global__ void __launch_bounds__(256) k_dn_down(dimg in, dimg out, dn_args a)
{
  __shared__ float4 tile[TW * TT];
  const int X0 = blockIdx.x * BX, Y0 = blockIdx.y * BY;
  load_tile4(tile, in, X0 + a.crop[0] - TH, Y0 + a.crop[1] - TH);
  __syncthreads();
  const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const int c = (threadIdx.y + TH) * TW + threadIdx.x + TH;   // tile index of the centre
  const float t = 0.2f;
  const float4 cc = tile[c];
  const float rgb[3] = { cc.x, cc.y, cc.z };
  float sum[3] = { t * rgb[0], t * rgb[1], t * rgb[2] };
  float wgt[3] = { t, t, t };
  float sig[3];
  noise_sigma(sig, a, a.black[1], a.inv_range[1], rgb[0]);
  float wc[3], g0[3];
#pragma unroll
  for(int k = 0; k < 3; k++) { wc[k] = frcp(a.level_scale * sig[k]); g0[k] = gamma08(rgb[k]); }
  // four flower taps at (+1.2,+.4) (-1.2,-.4) (+.4,-1.2) (-.4,+1.2)
  const int   ox[4] = { 1, -2, 0, -1 }, oy[4] = { 0, -1, -2, 1 };
  const float fa[4] = { 0.2f, 0.8f, 0.4f, 0.6f }, fc[4] = { 0.4f, 0.6f, 0.8f, 0.2f };
#pragma unroll
  for(int o = 0; o < 4; o++)
  {
    const float4 cl = tile_bilin(tile, c + oy[o] * TW + ox[o], fa[o], fc[o]);
    const float col[3] = { cl.x, cl.y, cl.z };
#pragma unroll
    for(int k = 0; k < 3; k++)
    {
      const float e = clampf(1.0f - 0.5f * (wc[k] * fabsf(gamma08(col[k]) - g0[k])), 0.0f, 1.0f);
      const float w = e * ((1.0f - t) * 0.25f);
      sum[k] += w * col[k];
      wgt[k] += w;
    }
  }
  st_h4(out, sorted1(x, out.wd), sorted1(y, out.ht),
        make_float4(sum[0] * frcp(wgt[0]), sum[1] * frcp(wgt[1]), sum[2] * frcp(wgt[2]), 1.0f));
}

// ---- assemble (assemble.comp:43-165) -------------------------------------

struct asm_mats { float rgb_to_yuv[9], yuv_to_rgb[9]; };

__global__ void __launch_bounds__(256) k_dn_assemble(dimg s0, dimg s1, dimg s2, dimg s3, dimg s4, dimg out,
                                                     dn_args a, asm_mats m)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 o0 = fetch_h4(s0, x + a.crop[0], y + a.crop[1]);
  const int szx = s1.wd + 1, szy = s1.ht + 1;
  int ex = x, ey = y;
  ex = ex / 2 + (ex & 1) * (szx / 2); ey = ey / 2 + (ey & 1) * (szy / 2);
  const float4 d1 = fetch_h4(s1, ex, ey);
  ex = ex / 2 + (ex & 1) * (szx / 2); ey = ey / 2 + (ey & 1) * (szy / 2);
  const float4 d2 = fetch_h4(s2, ex, ey);
  ex = ex / 2 + (ex & 1) * (szx / 2); ey = ey / 2 + (ey & 1) * (szy / 2);
  const float4 d3 = fetch_h4(s3, ex, ey);
  ex = ex / 2 + (ex & 1) * (szx / 2); ey = ey / 2 + (ey & 1) * (szy / 2);
  const float4 d4 = fetch_h4(s4, ex, ey);
  const float down[5][3] = { { o0.x, o0.y, o0.z }, { d1.x, d1.y, d1.z }, { d2.x, d2.y, d2.z },
                             { d3.x, d3.y, d3.z }, { d4.x, d4.y, d4.z } };
  float sig[3];
  noise_sigma(sig, a, a.black[1], a.inv_range[1], fmaxf(down[2][0], 0.0f));
  const float block = a.filters == 0u ? 1.0f : (a.filters == 9u ? 2.23607f : 1.414213f);
  const float bl[4] = { 0.7000f / block, 0.4900f / block, 0.3430f / block, 0.2401f / block };
  float rs[3];
#pragma unroll
  for(int k = 0; k < 3; k++) rs[k] = frcp(sig[k]);
  float d[4][3], len[4];
#pragma unroll
  for(int l = 0; l < 4; l++)
  {
    const float rb = 1.0f / bl[l];
#pragma unroll
    for(int k = 0; k < 3; k++) d[l][k] = (down[l][k] - down[l + 1][k]) * (rs[k] * rb);
    len[l] = fsqrt(d[l][0] * d[l][0] + d[l][1] * d[l][1] + d[l][2] * d[l][2]);
  }
  const float slope = ((len[3] - len[0]) * (1.0f / 3.0f) + (len[2] - len[1]) + (len[1] - len[0])
                     + (len[3] - len[2]) + (len[2] - len[0]) * 0.5f + (len[3] - len[1]) * 0.5f) * (1.0f / 6.0f);
  float test = fmaxf(0.0f, -slope);
  test = fmaxf(0.0f, 1.0f - test);
  test *= test; test *= test; test *= test; test *= test;   // ^16
  test = clampf(1.5f * test, 0.0f, 1.0f);
  const float thrs0 = a.strength * a.strength * a.strength * a.strength;
  float d4v[3] = { down[4][0], down[4][1], down[4][2] };
#pragma unroll
  for(int l = 3; l >= 0; l--)
  {
    const float thrs = fabsf(d[l][0]) > 10.0f ? 10000.0f : thrs0;
    const float rt = frcp(2.0f * thrs);
#pragma unroll
    for(int k = 0; k < 3; k++)
    {
      const float ad = fabsf(d[l][k]);
      const float tt = fminf(1.0f, ad * rt);
      d4v[k] += sig[k] * bl[l] * signf_(d[l][k]) * mixf(fmaxf(ad - thrs, 0.0f), ad, tt);
    }
  }
  float t0[3], t1[3];
#pragma unroll
  for(int k = 0; k < 3; k++)
  {
    t0[k] = (d4v[k] - a.black[k]) * a.inv_range[k] * a.wb[k];
    t1[k] = (down[0][k] - a.black[k]) * a.inv_range[k] * a.wb[k];
  }
  float yuv[3], yuvo0;
#pragma unroll
  for(int j = 0; j < 3; j++)
    yuv[j] = m.rgb_to_yuv[3 * j] * t0[0] + m.rgb_to_yuv[3 * j + 1] * t0[1] + m.rgb_to_yuv[3 * j + 2] * t0[2];
  yuvo0 = m.rgb_to_yuv[0] * t1[0] + m.rgb_to_yuv[1] * t1[1] + m.rgb_to_yuv[2] * t1[2];
  yuv[0] = mixf(yuvo0, yuv[0], a.luma);
  float rgb[3];
#pragma unroll
  for(int j = 0; j < 3; j++)
  {
    rgb[j] = m.yuv_to_rgb[3 * j] * yuv[0] + m.yuv_to_rgb[3 * j + 1] * yuv[1] + m.yuv_to_rgb[3 * j + 2] * yuv[2];
    rgb[j] = rgb[j] * frcp(a.wb[j]) * (a.white[j] - a.black[j]) + a.black[j];
  }
  st_h4(out, x, y, make_float4(rgb[0], rgb[1], rgb[2], test));
}

// ---- doub (doub.comp:35-116) ---------------------------------------------

__global__ void __launch_bounds__(256) k_dn_doub(dimg in, dimg crs0, dimg crs1b, dimg out, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float sx = ((float)x + 0.5f) / (float)out.wd, sy = ((float)y + 0.5f) / (float)out.ht;
  float4 upsm = bilin_h4(crs0, sx * (float)crs0.wd, sy * (float)crs0.ht);
  const float4 down = bilin_h4(crs1b, sx * (float)crs1b.wd, sy * (float)crs1b.ht);
  int ch = 1;
  float crs = upsm.y, crs1 = down.y;
  float T = 0.5f * a.strength * upsm.w, blendw = a.luma;
  if(a.filters == 9)
  {
    const int col = xtrans_colour(x, y);
    if(col != 1)
    {
      ch = col;
      crs = col == 2 ? upsm.z : upsm.x;
      crs1 = col == 2 ? down.z : down.x;
      blendw = 1.0f;
      T *= frcp(fmaxf(1e-4f, upsm.w));
    }
  }
  else
  {
    const int ox = x & 1, oy = y & 1;
    if(ox == 0 && oy == 0)      { ch = 0; crs = upsm.x; crs1 = down.x; blendw = 1.0f; }
    else if(ox == 1 && oy == 1) { ch = 2; crs = upsm.z; crs1 = down.z; blendw = 1.0f; }
  }
  const float black = a.black[ch], white = a.white[ch], inv_range = a.inv_range[ch];
  float sig[3];
  noise_sigma(sig, a, black, inv_range, crs);
  float val = centre_u16(in, x + a.crop[0], y + a.crop[1]);
  blendw = 0.5f * (blendw + 1.0f);
  if(val < white)
  {
    const float wav = (val - crs1) * frcp(fmaxf(sig[0] + sig[2], 1e-8f));
    const float tt = fminf(1.0f, wav * frcp(fmaxf(2.0f * T, 1e-8f)));
    float uw = fminf(1.0f, 1.0f * upsm.w);
    uw *= uw; uw *= uw;    // ^4
    uw = 1.0f - (1.0f - uw) * a.detail;
    const float aw = fabsf(wav);
    val = mixf(val, fmaxf(0.0f, crs + sig[1] * signf_(wav) * mixf(fmaxf(0.0f, aw - T), aw, tt)), uw * blendw);
  }
  val = fmaxf(0.0f, (val - black) * inv_range);
  st_h1(out, x, y, val);
}

// ---- host launchers --------------------------------------------------------

int launch_denoise_noop(const node_call &c)
{
  if(int e = check_blob(c, 18 * 4, 9 * 4, "denoise/noop")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "denoise/noop input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "denoise/noop output")) return e;
  const int32_t *pi = (const int32_t *)c.pc;
  const float *pf = (const float *)c.pc;
  dn_args a = {};
  for(int k = 0; k < 4; k++) { a.crop[k] = pi[k]; a.black[k] = pf[4 + k]; a.white[k] = pf[8 + k]; }
  a.filters = (uint32_t)pi[16];
  a.gainmap = pi[17];
  read_params(a, c);
  if(a.gainmap) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "denoise/noop: dng gain maps are not supported");
  hipLaunchKernelGGL(k_dn_noop, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("denoise/noop");
}

int launch_denoise_half(const node_call &c)
{
  if(int e = check_blob(c, 17 * 4, 0, "denoise/half")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "denoise/half input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "denoise/half output")) return e;
  dn_args a = {};
  read_rgbc(a, c.pc);
  a.filters = ((const uint32_t *)c.pc)[16];
  hipLaunchKernelGGL(k_dn_half, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("denoise/half");
}

static int decode_down(dn_args &a, const node_call &c)
{
  if(int e = check_blob(c, 20 * 4, 8 * 4, "denoise/down")) return e;
  read_rgbc(a, c.pc);
  const float *pf = (const float *)c.pc;
  const int32_t *pi = (const int32_t *)c.pc;
  a.noise_a = pf[16];
  a.noise_b = pf[17];
  a.level = pi[18];
  a.block = pi[19];
  read_params(a, c);
  const float block = (a.block == 3) ? 2.23607f : (a.block == 2 ? 1.414213f : 1.0f);
  a.level_scale = powf(0.7f, (float)a.level) / block;
  return 0;
}

int launch_denoise_downcov(const node_call &c)
{
  dn_args a = {};
  if(int e = decode_down(a, c)) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "denoise/downcov input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "denoise/downcov output")) return e;
  dimg cov = {};
  if(c.nconn > 2 && c.conn[2].ptr)
  { // the covariance store is dead in the reference graph; written only when bound
    if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "denoise/downcov cov")) return e;
    if(c.conn[2].wd < c.conn[1].wd + (uint32_t)a.crop[0] || c.conn[2].ht < c.conn[1].ht + (uint32_t)a.crop[1])
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "denoise/downcov: cov buffer smaller than output + crop");
    cov = to_dimg(c.conn[2]);
  }
  hipLaunchKernelGGL(k_dn_downcov, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), cov, a);
  return launch_post("denoise/downcov");
}

int launch_denoise_down(const node_call &c)
{
  dn_args a = {};
  if(int e = decode_down(a, c)) return e;
  if(a.level < 0) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "denoise/down: negative level (response branch) not supported");
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "denoise/down input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "denoise/down output")) return e;
  hipLaunchKernelGGL(k_dn_down, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("denoise/down");
}

static void matmul3(double o[9], const double a[9], const double b[9])
{
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
  {
    double s = 0;
    for(int k = 0; k < 3; k++) s += a[3 * j + k] * b[3 * k + i];
    o[3 * j + i] = s;
  }
}

int launch_denoise_assemble(const node_call &c)
{
  if(int e = check_blob(c, 19 * 4, 8 * 4, "denoise/assemble")) return e;
  for(int i = 0; i < 6; i++)
    if(int e = check_conn(c, i, VKDT_FMT_F16, 4, 1, "denoise/assemble")) return e;
  dn_args a = {};
  read_rgbc(a, c.pc);
  const float *pf = (const float *)c.pc;
  a.noise_a = pf[16];
  a.noise_b = pf[17];
  a.filters = ((const uint32_t *)c.pc)[18];
  read_params(a, c);
  // matrices.h: rec709_to_yuv * rec2020_to_rec709 and rec709_to_rec2020 * yuv_to_rec709
  static const double m709yuv[9]  = { 0.299, 0.587, 0.114, -0.147, -0.289, 0.436, 0.615, -0.515, -0.100 };
  static const double m2020709[9] = { 1.66022677, -0.58754761, -0.07283825, -0.12455334, 1.13292605, -0.00834963,
                                      -0.01815514, -0.10060303, 1.11899817 };
  static const double m7092020[9] = { 0.62750375, 0.32927542, 0.04330266, 0.06910828, 0.91951916, 0.0113596,
                                      0.01639406, 0.08801125, 0.89538035 };
  static const double myuv709[9]  = { 1.0, -3.94570707e-05, 1.13982797, 1.0, -3.94610164e-01, -5.80500316e-01,
                                      1.0, 2.03199968, -4.81376263e-04 };
  double t1[9], t2[9];
  matmul3(t1, m709yuv, m2020709);
  matmul3(t2, m7092020, myuv709);
  asm_mats m;
  for(int k = 0; k < 9; k++) { m.rgb_to_yuv[k] = (float)t1[k]; m.yuv_to_rgb[k] = (float)t2[k]; }
  hipLaunchKernelGGL(k_dn_assemble, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]),
                     to_dimg(c.conn[4]), to_dimg(c.conn[5]), a, m);
  return launch_post("denoise/assemble");
}

int launch_denoise_doub(const node_call &c)
{
  if(int e = check_blob(c, 24 * 4, 9 * 4, "denoise/doub")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "denoise/doub orig")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "denoise/doub crs0")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "denoise/doub crs1")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_F16, 1, 1, "denoise/doub output")) return e;
  dn_args a = {};
  read_rgbc(a, c.pc);
  const float *pf = (const float *)c.pc;
  const int32_t *pi = (const int32_t *)c.pc;
  a.filters = (uint32_t)pi[16];
  a.noise_a = pf[17];
  a.noise_b = pf[18];
  a.gainmap = pi[19];
  read_params(a, c);
  if(a.gainmap) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "denoise/doub: dng gain maps are not supported");
  hipLaunchKernelGGL(k_dn_doub, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), a);
  return launch_post("denoise/doub");
}
