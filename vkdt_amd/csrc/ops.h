// ops.h -- per-texel and per-tile device operations of the raw-development
// nodes.  The plain per-node kernels (k_denoise/k_demosaic/k_colour/k_llap)
// and the batched fused kernels (k_batch) are thin wrappers around these,
// so both paths evaluate the same arithmetic.
//
// Per-tile operations are work-group level (they stage LDS and call
// __syncthreads): every thread of the group must call them with the same
// tile origin.  A work group is BX x BY = 64 x 4 threads; a tile of R rows
// per thread covers 64 x 4R texels.
#pragma once
#include "common.h"

// ==========================================================================
// denoise (src/pipe/modules/denoise)
// ==========================================================================

struct dn_args
{
  float    wb[4], black[4], white[4];
  float    inv_range[4];     // 1/(white-black)
  int      crop[4];
  float    noise_a, noise_b;
  int      level;
  int      block;
  uint32_t filters;
  int      gainmap;          // apply the Bayer gain map (push and module parameter both 1)
  float    map_os[4];        // gain-map origin (x, y) and inverse extent (x, y)
  float    strength, luma, detail, edges[4];
  float    edge_gain[3];     // exp2(12*edges.rgb + edges.a) of noise_sigma
  float    level_scale;      // 0.7^level / block factor (down.comp:85)
  float    asm_bl[4];        // assemble's level scales 0.7^(l+1) / block factor (dn_asm_scales)
  float    asm_rb[4];        // and their reciprocals
};

struct asm_mats { float rgb_to_yuv[9], yuv_to_rgb[9]; };

// noop.comp:45-55, doub.comp:106-114: the gain of CFA site (x,y) at image
// position (px,py) in [0,1]^2: the rgba gain grid (one channel per site of
// the 2x2 block) sampled by texture() at (p * map_os.zw - map_os.xy)
// clamped to [0,1]
DEV float dn_gain_at(const dimg &gm, const dn_args &a, float px, float py, int x, int y)
{
  px = clampf(px * a.map_os[2] - a.map_os[0], 0.0f, 1.0f);
  py = clampf(py * a.map_os[3] - a.map_os[1], 0.0f, 1.0f);
  const float4 g = bilin_f4(gm, px * (float)gm.wd, py * (float)gm.ht);
  const int site = (x & 1) + 2 * (y & 1);
  return site == 0 ? g.x : (site == 1 ? g.y : (site == 2 ? g.z : g.w));
}

// host: the module parameters strength, luma, detail, edges[4]
// (denoise/params, float words 0..2 and 4..7)
static inline void dn_params_init(dn_args &a, const float *p)
{
  a.strength = p[0];
  a.luma = p[1];
  a.detail = p[2];
  for(int k = 0; k < 4; k++) a.edges[k] = p[4 + k];
  for(int k = 0; k < 3; k++) a.edge_gain[k] = exp2f(12.0f * a.edges[k] + a.edges[3]);
}

// host: down.comp:85, 0.7^level divided by the block's diagonal factor
static inline void dn_level_scale(dn_args &a)
{
  const float block = (a.block == 3) ? 2.23607f : (a.block == 2 ? 1.414213f : 1.0f);
  a.level_scale = powf(0.7f, (float)a.level) / block;
}

// host: assemble.comp's colour matrices (matrices.h: rec709_to_yuv *
// rec2020_to_rec709 and rec709_to_rec2020 * yuv_to_rec709), products in double
// assemble.comp's level scales bl = 0.7^(l+1) / block (block factor by CFA:
// 1, sqrt 2, sqrt 5) and 1 / bl: the same IEEE float divisions the kernel
// would take for every texel, taken once on the host
static inline void dn_asm_scales(dn_args &a)
{
  const float block = a.filters == 0u ? 1.0f : (a.filters == 9u ? 2.23607f : 1.414213f);
  const float c[4] = { 0.7000f, 0.4900f, 0.3430f, 0.2401f };
  for(int l = 0; l < 4; l++)
  {
    a.asm_bl[l] = c[l] / block;
    a.asm_rb[l] = 1.0f / a.asm_bl[l];
  }
}

static inline void asm_mats_init(asm_mats &m)
{
  static const double m709yuv[9]  = { 0.299, 0.587, 0.114, -0.147, -0.289, 0.436, 0.615, -0.515, -0.100 };
  static const double m2020709[9] = { 1.66022677, -0.58754761, -0.07283825, -0.12455334, 1.13292605, -0.00834963,
                                      -0.01815514, -0.10060303, 1.11899817 };
  static const double m7092020[9] = { 0.62750375, 0.32927542, 0.04330266, 0.06910828, 0.91951916, 0.0113596,
                                      0.01639406, 0.08801125, 0.89538035 };
  static const double myuv709[9]  = { 1.0, -3.94570707e-05, 1.13982797, 1.0, -3.94610164e-01, -5.80500316e-01,
                                      1.0, 2.03199968, -4.81376263e-04 };
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
  {
    double s1 = 0, s2 = 0;
    for(int k = 0; k < 3; k++)
    {
      s1 += m709yuv[3 * j + k] * m2020709[3 * k + i];
      s2 += m7092020[3 * j + k] * myuv709[3 * k + i];
    }
    m.rgb_to_yuv[3 * j + i] = (float)s1;
    m.yuv_to_rgb[3 * j + i] = (float)s2;
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

// half.comp:24-74: one block of the mosaic -> rgba
DEV float4 dn_half_px(const dimg &in, const dn_args &a, int x, int y)
{
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
  return rgba;
}

// LDS tile of an rgba f16 image: a work-group tile of 64 x 4R texels
// (thread (tx,ty) owns rows ty + 4k, k < R) plus a 2 texel mirrored halo
#define TH 2
#define TW (BX + 2 * TH)
#define DN_TT(R) (BY * (R) + 2 * TH)
#define DN_TN(R) (TW * DN_TT(R))

template<int R>
DEV void load_tile4(float4 *tile, const dimg &in, int ox, int oy)
{
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < DN_TN(R); i += BX * BY)
  {
    const int ty = i / TW, tx = i - ty * TW;
    tile[i] = ld_h4(in, mirror(ox + tx, in.wd), mirror(oy + ty, in.ht));
  }
}

// downcov.comp:41-63 + cov.glsl:21-138 for the tile at (X0,Y0)
// per tile texel: luminance l, 1/l, 1/(l l) and l/25, each computed once
// with the operation the statistics use.  (i-2)/l equals (i-2)*(1/l)
// exactly for the numerators 0, +-1, +-2 (a sign change or a power of two
// times a correctly rounded reciprocal), so the loops multiply.
template<int R>
struct dn_cov_smem { float4 tile[DN_TN(R)]; float lum[DN_TN(R)], rlum[DN_TN(R)], rlum2[DN_TN(R)], lum25[DN_TN(R)]; };

// The structure tensors of the 5x5 window at tile index c0 (offset (-2,-2)),
// the choice between them and its eigen decomposition; mean_b is the
// window's mean luminance.
template<int R>
DEV void dn_cov_tensor(const dn_cov_smem<R> &s, int c0, float ev[2], float e0[2], float e1[2], float &mean_b)
{
  float mw0 = 0, mw1 = 0, mb0 = 0, mb1 = 0, smw = 0, smb = 0;
#pragma unroll
  for(int j = 0; j < 5; j++)
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const float px = s.lum[c0 + j * TW + i], rp = s.rlum[c0 + j * TW + i];
      mw0 += (float)(i - 2) * px; mw1 += (float)(j - 2) * px;
      smw += px;
      mb0 += (float)(i - 2) * rp; mb1 += (float)(j - 2) * rp;
      smb += rp;
    }
  mw0 /= smw; mw1 /= smw;
  mb0 /= smb; mb1 /= smb;
  float Sw[4] = {0, 0, 0, 0}, Sb[4] = {0, 0, 0, 0}, sw = 0, sb = 0;
  mean_b = 0;
#pragma unroll
  for(int j = 0; j < 5; j++)
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const int t = c0 + j * TW + i;
      const float px = s.lum[t];
      mean_b += s.lum25[t];
      float p2 = px * px;
      float dx = (float)(i - 2) - mw0, dy = (float)(j - 2) - mw1;
      Sw[0] += p2 * dx * dx; Sw[1] += p2 * dx * dy; Sw[2] += p2 * dy * dx; Sw[3] += p2 * dy * dy;
      sw += p2;
      dx = (float)(i - 2) - mb0; dy = (float)(j - 2) - mb1;
      p2 = s.rlum2[t];
      Sb[0] += p2 * dx * dx; Sb[1] += p2 * dx * dy; Sb[2] += p2 * dy * dx; Sb[3] += p2 * dy * dy;
      sb += p2;
    }
#pragma unroll
  for(int q = 0; q < 4; q++) { Sw[q] /= sw; Sb[q] /= sb; }
  float S[4];
  const bool usew = det2(Sw) < det2(Sb);
#pragma unroll
  for(int q = 0; q < 4; q++) S[q] = usew ? Sw[q] : Sb[q];
  evd2x2(ev, e0, e1, S);
}

template<int R>
DEV void dn_downcov_tile(const dimg &in, const dimg &out, const dimg &cov, const dn_args &a, int X0, int Y0,
                         dn_cov_smem<R> &s)
{
  load_tile4<R>(s.tile, in, X0 + a.crop[0] - TH, Y0 + a.crop[1] - TH);
  __syncthreads();
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < DN_TN(R); i += BX * BY)
  {
    const float4 v = s.tile[i];
    const float l = lum2020(v.x, v.y, v.z);
    s.lum[i] = l;
    s.rlum[i] = 1.0f / l;    // IEEE: the statistics feed the ill-conditioned eigen choice
    s.rlum2[i] = 1.0f / (l * l);
    s.lum25[i] = l / 25.0f;
  }
  __syncthreads();
  const int x = X0 + threadIdx.x;
#pragma unroll 1
  for(int k = 0; k < R; k++)
  {
    const int y = Y0 + threadIdx.y + BY * k;
    if(x >= out.wd || y >= out.ht) continue;
    const int c0 = (threadIdx.y + BY * k) * TW + threadIdx.x;   // tile index of offset (-2,-2)
    float ev[2], e0[2], e1[2], mean_b;
    dn_cov_tensor<R>(s, c0, ev, e0, e1, mean_b);
    ev[1] *= 0.05f;
    ev[0] = clampf(ev[0], 0.01f, 25.0f);
    ev[1] = clampf(ev[1], 0.01f, 25.0f);
    if(cov.p) st_h4(cov, x + a.crop[0], y + a.crop[1], make_float4(ev[0], ev[1], e0[0], e0[1]));
    const float ri0 = -0.5f * frcp(ev[0]), ri1 = -0.5f * frcp(ev[1]);
    // the anisotropic weight of offset o = (i-2, j-2) equals that of -o bit
    // for bit (negating both offsets negates each rounded product and their
    // rounded sum exactly, and the sums are squared), so the exponentials
    // are taken for the half plane j > 2 or (j = 2, i >= 2) and mirrored
    float wgt[5][5];
#pragma unroll
    for(int j = 2; j < 5; j++)
#pragma unroll
      for(int i = 0; i < 5; i++)
      {
        if(j == 2 && i < 2) continue;
        const float x0 = (float)(i - 2) * e0[0] + (float)(j - 2) * e0[1];
        const float x1 = (float)(i - 2) * e1[0] + (float)(j - 2) * e1[1];
        const float w = fmaxf(1e-9f, fexp(x0 * x0 * ri0 + x1 * x1 * ri1));
        wgt[j][i] = w;
        wgt[4 - j][4 - i] = w;
      }
    float r0 = 0, r1 = 0, r2 = 0, wt = 0;
    const float hot = 2.0f * mean_b;
#pragma unroll
    for(int j = 0; j < 5; j++)
#pragma unroll
      for(int i = 0; i < 5; i++)
      {
        const float4 v = s.tile[c0 + j * TW + i];
        if(v.x > hot) continue;   // hot pixel
        const float w = wgt[j][i];
        r0 += w * v.x; r1 += w * v.y; r2 += w * v.z;
        wt += w;
      }
    const float d = frcp(fmaxf(wt, 1e-8f));
    float edge = clampf(75.0f * fmaxf(0.0f, ev[1] - 0.09f), 0.0f, 1.0f);
    edge = smoothstepf(0.4f, 0.75f, edge);
    edge = clampf(0.02f + edge, 0.0f, 1.0f);
    st_h4(out, sorted1(x, out.wd), sorted1(y, out.ht), make_float4(r0 * d, r1 * d, r2 * d, edge));
  }
}

DEV float gamma08(float f) { return f < 0.0f ? f : fpow(f, 0.8f); }

// bilinear from the tile at base texel b with weights (fa,fc).  The second
// product of each lerp is fused into its sum (an explicit FMA, so every
// caller rounds the same way): one rounding fewer than texture()'s written
// form and three instructions instead of nine per channel
DEV float bilin_fma(float p00, float p10, float p01, float p11, float fa, float fc)
{
  const float r0 = __builtin_fmaf(p10, fa, p00 * (1.f - fa));
  const float r1 = __builtin_fmaf(p11, fa, p01 * (1.f - fa));
  return __builtin_fmaf(r1, fc, r0 * (1.f - fc));
}
DEV float4 tile_bilin(const float4 *tile, int b, float fa, float fc)
{
  const float4 p00 = tile[b], p10 = tile[b + 1], p01 = tile[b + TW], p11 = tile[b + TW + 1];
  float4 r;
  r.x = bilin_fma(p00.x, p10.x, p01.x, p11.x, fa, fc);
  r.y = bilin_fma(p00.y, p10.y, p01.y, p11.y, fa, fc);
  r.z = bilin_fma(p00.z, p10.z, p01.z, p11.z, fa, fc);
  r.w = 1.0f;
  return r;
}

// down.comp:59-108 (level >= 0) for the tile at (X0,Y0).  The taps' lerps,
// the edge-stopping weight 1 - (wc/2)|d| and the weighted sums use explicit
// FMAs: the weights are continuous in their inputs (no branch or threshold
// follows), so the one rounding fewer moves the f16 output by at most a
// last-place step at rounding boundaries
template<int R>
DEV void dn_down_tile(const dimg &in, const dimg &out, const dn_args &a, int X0, int Y0, float4 *tile)
{
  load_tile4<R>(tile, in, X0 + a.crop[0] - TH, Y0 + a.crop[1] - TH);
  __syncthreads();
  const int x = X0 + threadIdx.x;
#pragma unroll 1
  for(int k = 0; k < R; k++)
  {
    const int y = Y0 + threadIdx.y + BY * k;
    if(x >= out.wd || y >= out.ht) continue;
    const int c = (threadIdx.y + BY * k + TH) * TW + threadIdx.x + TH;   // tile index of the centre
    const float t = 0.2f;
    const float4 cc = tile[c];
    const float rgb[3] = { cc.x, cc.y, cc.z };
    float sum[3] = { t * rgb[0], t * rgb[1], t * rgb[2] };
    float wgt[3] = { t, t, t };
    float sig[3];
    noise_sigma(sig, a, a.black[1], a.inv_range[1], rgb[0]);
    float hwc[3], g0[3];   // half the edge-stopping scale
#pragma unroll
    for(int q = 0; q < 3; q++) { hwc[q] = 0.5f * frcp(a.level_scale * sig[q]); g0[q] = gamma08(rgb[q]); }
    // four flower taps at (+1.2,+.4) (-1.2,-.4) (+.4,-1.2) (-.4,+1.2)
    const int   ox[4] = { 1, -2, 0, -1 }, oy[4] = { 0, -1, -2, 1 };
    const float fa[4] = { 0.2f, 0.8f, 0.4f, 0.6f }, fc[4] = { 0.4f, 0.6f, 0.8f, 0.2f };
#pragma unroll
    for(int o = 0; o < 4; o++)
    {
      const float4 cl = tile_bilin(tile, c + oy[o] * TW + ox[o], fa[o], fc[o]);
      const float col[3] = { cl.x, cl.y, cl.z };
#pragma unroll
      for(int q = 0; q < 3; q++)
      {
        const float e = clampf(__builtin_fmaf(-hwc[q], fabsf(gamma08(col[q]) - g0[q]), 1.0f), 0.0f, 1.0f);
        const float w = e * ((1.0f - t) * 0.25f);
        sum[q] = __builtin_fmaf(w, col[q], sum[q]);
        wgt[q] += w;
      }
    }
    st_h4(out, sorted1(x, out.wd), sorted1(y, out.ht),
          make_float4(sum[0] * frcp(wgt[0]), sum[1] * frcp(wgt[1]), sum[2] * frcp(wgt[2]), 1.0f));
  }
}

// assemble.comp:43-165
DEV float4 dn_assemble_px(const dimg &s0, const dimg &s1, const dimg &s2, const dimg &s3, const dimg &s4,
                          const dn_args &a, const asm_mats &m, int x, int y)
{
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
  float rs[3];
#pragma unroll
  for(int k = 0; k < 3; k++) rs[k] = frcp(sig[k]);
  float d[4][3], len[4];
#pragma unroll
  for(int l = 0; l < 4; l++)
  {
    const float rb = a.asm_rb[l];
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
      d4v[k] += sig[k] * a.asm_bl[l] * signf_(d[l][k]) * mixf(fmaxf(ad - thrs, 0.0f), ad, tt);
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
  return make_float4(rgb[0], rgb[1], rgb[2], test);
}

// doub.comp:80-116: the texel's shrinkage against the interpolated coarse
// levels (channel ch of the upsampled level crs and of the finer one crs1,
// alpha upw of the upsampled level, threshold T) and its normalisation
DEV float dn_doub_tail(const dn_args &a, int ch, float crs, float crs1, float upw, float T, float blendw, float val)
{
  const float black = a.black[ch], white = a.white[ch], inv_range = a.inv_range[ch];
  float sig[3];
  noise_sigma(sig, a, black, inv_range, crs);
  blendw = 0.5f * (blendw + 1.0f);
  if(val < white)
  {
    const float wav = (val - crs1) * frcp(fmaxf(sig[0] + sig[2], 1e-8f));
    const float tt = fminf(1.0f, wav * frcp(fmaxf(2.0f * T, 1e-8f)));
    float uw = fminf(1.0f, 1.0f * upw);
    uw *= uw; uw *= uw;    // ^4
    uw = 1.0f - (1.0f - uw) * a.detail;
    const float aw = fabsf(wav);
    val = mixf(val, fmaxf(0.0f, crs + sig[1] * signf_(wav) * mixf(fmaxf(0.0f, aw - T), aw, tt)), uw * blendw);
  }
  return fmaxf(0.0f, (val - black) * inv_range);
}

// doub.comp:35-116: one mosaic texel at full resolution
DEV float dn_doub_px(const dimg &in, const dimg &crs0, const dimg &crs1b, int owd, int oht, const dn_args &a,
                     int x, int y)
{
  const float sx = ((float)x + 0.5f) / (float)owd, sy = ((float)y + 0.5f) / (float)oht;
  const float4 upsm = bilin_h4(crs0, sx * (float)crs0.wd, sy * (float)crs0.ht);
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
  return dn_doub_tail(a, ch, crs, crs1, upsm.w, T, blendw, centre_u16(in, x + a.crop[0], y + a.crop[1]));
}

// one axis of a bilinear texture() lookup: the two mirrored taps and the
// weight of the second (bilin_h4's arithmetic)
struct bl_axis { int i0, i1; float f; };
DEV bl_axis bl_axis_at(float s, int n)
{
  const float fs = s - 0.5f, fl = floorf(fs);
  const int i = (int)fl;
  bl_axis r;
  r.i0 = mirror(i, n);
  r.i1 = mirror(i + 1, n);
  r.f = fs - fl;
  return r;
}

// channel CH of an rgba f16 texel held as two words
template<int CH> DEV float h4_chan(uint2 u)
{
  const uint32_t w = CH < 2 ? u.x : u.y;
  const __half2 h = *(const __half2 *)&w;
  return (CH & 1) ? __high2float(h) : __low2float(h);
}

// bilinear channel CH (x) and, with ALPHA, channel 3 (y) of an rgba f16
// image: bilin_h4's arithmetic for those channels alone
template<int CH, bool ALPHA>
DEV float2 bilin_h4_chan(const dimg &b, const bl_axis &ax, const bl_axis &ay)
{
  const __half *p = (const __half *)b.p;
  const long r0 = (long)ay.i0 * b.pitch, r1 = (long)ay.i1 * b.pitch;
  const uint2 u00 = *(const uint2 *)(p + 4 * (r0 + ax.i0)), u10 = *(const uint2 *)(p + 4 * (r0 + ax.i1));
  const uint2 u01 = *(const uint2 *)(p + 4 * (r1 + ax.i0)), u11 = *(const uint2 *)(p + 4 * (r1 + ax.i1));
  const float a = ax.f, c = ay.f;
  float2 r;
  r.x = (h4_chan<CH>(u00) * (1.f - a) + h4_chan<CH>(u10) * a) * (1.f - c)
      + (h4_chan<CH>(u01) * (1.f - a) + h4_chan<CH>(u11) * a) * c;
  r.y = ALPHA ? (h4_chan<3>(u00) * (1.f - a) + h4_chan<3>(u10) * a) * (1.f - c)
              + (h4_chan<3>(u01) * (1.f - a) + h4_chan<3>(u11) * a) * c : 0.0f;
  return r;
}

// a Bayer texel of colour CH (0 red, 1 green, 2 blue) from its taps
template<int CH>
DEV float dn_doub_bayer_texel(const dimg &crs0, const dimg &crs1b, const dn_args &a, const bl_axis &ux,
                              const bl_axis &uy, const bl_axis &dx, const bl_axis &dy, float val)
{
  const float2 up = bilin_h4_chan<CH, true>(crs0, ux, uy);
  const float dn = bilin_h4_chan<CH, false>(crs1b, dx, dy).x;
  return dn_doub_tail(a, CH, up.x, dn, up.y, 0.5f * a.strength * up.y, CH == 1 ? a.luma : 1.0f, val);
}

// dn_doub_px on Bayer mosaics for the 2x2 block at even (x, y), v in row
// order: the block's texels share their columns' and rows' taps, and each
// interpolates only the channel its colour reads and alpha -- per channel
// the same arithmetic, so the same values bit for bit
DEV void dn_doub_quad(const dimg &in, const dimg &crs0, const dimg &crs1b, int owd, int oht, const dn_args &a,
                      int x, int y, float v[4])
{
  bl_axis ux[2], uy[2], dx[2], dy[2];
#pragma unroll
  for(int k = 0; k < 2; k++)
  {
    const float sx = ((float)(x + k) + 0.5f) / (float)owd, sy = ((float)(y + k) + 0.5f) / (float)oht;
    ux[k] = bl_axis_at(sx * (float)crs0.wd, crs0.wd);
    dx[k] = bl_axis_at(sx * (float)crs1b.wd, crs1b.wd);
    uy[k] = bl_axis_at(sy * (float)crs0.ht, crs0.ht);
    dy[k] = bl_axis_at(sy * (float)crs1b.ht, crs1b.ht);
  }
  const int cx = x + a.crop[0], cy = y + a.crop[1];
  v[0] = dn_doub_bayer_texel<0>(crs0, crs1b, a, ux[0], uy[0], dx[0], dy[0], centre_u16(in, cx, cy));
  v[1] = dn_doub_bayer_texel<1>(crs0, crs1b, a, ux[1], uy[0], dx[1], dy[0], centre_u16(in, cx + 1, cy));
  v[2] = dn_doub_bayer_texel<1>(crs0, crs1b, a, ux[0], uy[1], dx[0], dy[1], centre_u16(in, cx, cy + 1));
  v[3] = dn_doub_bayer_texel<2>(crs0, crs1b, a, ux[1], uy[1], dx[1], dy[1], centre_u16(in, cx + 1, cy + 1));
}

// store the 2x2 block at even (x, y) of an f16 plane, v in row order: each
// row as one 4-byte store where both texels lie inside and rows start 4-byte
// aligned, texel by texel otherwise
DEV void st_h1_quad(const dimg &o, int x, int y, const float v[4])
{
  const bool pair = x + 1 < o.wd && ((o.pitch | (int)((uintptr_t)o.p >> 1)) & 1) == 0;
#pragma unroll
  for(int r = 0; r < 2; r++)
  {
    if(y + r >= o.ht) break;
    if(pair) *(__half2 *)((__half *)o.p + (long)(y + r) * o.pitch + x) = __halves2half2(f2h(v[2 * r]), f2h(v[2 * r + 1]));
    else
    {
      st_h1(o, x, y + r, v[2 * r]);
      if(x + 1 < o.wd) st_h1(o, x + 1, y + r, v[2 * r + 1]);
    }
  }
}

// ==========================================================================
// demosaic (src/pipe/modules/demosaic)
// ==========================================================================

// gauss.comp:19-126.  The covariance statistics decide between two tensors
// and feed a 2x2 eigen solve whose discriminant vanishes on flat or clipped
// blocks: they keep IEEE division so the decision follows the written
// arithmetic.
//   mos(px, py): the mosaic at mirrored (px, py)
template<int XT, typename MosAt>
DEV float4 dm_gauss_at(MosAt mos, int x, int y)
{
  constexpr int n0 = XT ? 0 : -1, n1 = 3, bs = XT ? 3 : 2;
  constexpr int NS = XT ? 5 : 8;
  // pi/px equals pi*(1/px) exactly for the offsets -1..2 (as in dn_downcov_tile)
  float px[NS], rp[NS];
  int pi[NS], pj[NS];
  int n = 0;
#pragma unroll
  for(int j = n0; j < n1; j++)
#pragma unroll
    for(int i = n0; i < n1; i++)
    {
      if(XT ? (((j + i) & 1) == 1) : (((j + i) & 1) != 1)) continue;
      px[n] = mos(bs * x + i, bs * y + j);
      rp[n] = 1.0f / px[n];
      pi[n] = i; pj[n] = j;
      n++;
    }
  float mw0 = 0, mw1 = 0, mb0 = 0, mb1 = 0, smw = 0, smb = 0;
#pragma unroll
  for(int k = 0; k < NS; k++)
  {
    mw0 += (float)pi[k] * px[k]; mw1 += (float)pj[k] * px[k];
    smw += px[k];
    mb0 += (float)pi[k] * rp[k]; mb1 += (float)pj[k] * rp[k];
    smb += rp[k];
  }
  mw0 /= smw; mw1 /= smw;
  mb0 /= smb; mb1 /= smb;
  float Sw[4] = {0, 0, 0, 0}, Sb[4] = {0, 0, 0, 0}, sw = 0, sb = 0;
#pragma unroll
  for(int k = 0; k < NS; k++)
  {
    float p2 = px[k] * px[k];
    float dx = (float)pi[k] - mw0, dy = (float)pj[k] - mw1;
    Sw[0] += p2 * dx * dx; Sw[1] += p2 * dx * dy; Sw[2] += p2 * dy * dx; Sw[3] += p2 * dy * dy;
    sw += p2;
    dx = (float)pi[k] - mb0; dy = (float)pj[k] - mb1;
    p2 = 1.0f / p2;
    Sb[0] += p2 * dx * dx; Sb[1] += p2 * dx * dy; Sb[2] += p2 * dy * dx; Sb[3] += p2 * dy * dy;
    sb += p2;
  }
#pragma unroll
  for(int k = 0; k < 4; k++) { Sw[k] /= sw; Sb[k] /= sb; }
  const bool usew = det2(Sw) < det2(Sb);
  float S[4];
#pragma unroll
  for(int k = 0; k < 4; k++) S[k] = usew ? Sw[k] : Sb[k];
  float ev[2], e0[2], e1[2];
  evd2x2(ev, e0, e1, S);
  if(!XT)
  {
    ev[0] *= 0.2f; ev[1] *= 0.2f;
    if(fabsf(e0[0]) > fabsf(e0[1])) { e0[0] = 1.0f; e0[1] = 0.0f; }
    else                            { e0[0] = 0.0f; e0[1] = 1.0f; }
  }
  else
  {
    ev[0] *= 0.4f; ev[1] *= 0.4f;
    if     (fabsf(e0[0]) > 2.0f * fabsf(e0[1])) { e0[0] = 1.0f; e0[1] = 0.0f; }
    else if(fabsf(e0[1]) > 2.0f * fabsf(e0[0])) { e0[0] = 0.0f; e0[1] = 1.0f; }
    else { ev[0] = 0.1f; ev[1] = 0.1f; }
  }
  return make_float4(ev[0], ev[1], e0[0], e0[1]);
}

template<int XT, int U16 = 0>
DEV float4 dm_gauss_px(const dimg &orig, int x, int y)
{
  return dm_gauss_at<XT>([&](int px, int py) { return centre_mt<U16>(orig, px, py); }, x, y);
}

// the taps of block texel (x, y) lie inside the mosaic: mirroring is the
// identity there
template<int XT>
DEV bool dm_gauss_inside(const dimg &orig, int x, int y)
{
  constexpr int n0 = XT ? 0 : -1, bs = XT ? 3 : 2;
  return bs * x + n0 >= 0 && bs * y + n0 >= 0 && bs * x + 2 < orig.wd && bs * y + 2 < orig.ht;
}

// gaussian weight of offset (i,j) under cov = (eval, major axis):
// exp(-0.5 |E o|^2_eval) clamped to [lo, 1], eval passed as -0.5/eval
DEV float gauss_w(float4 cov, float r0, float r1, int i, int j, float lo)
{
  const float ofx = cov.z * (float)i + cov.w * (float)j;
  const float ofy = -cov.w * (float)i + cov.z * (float)j;
  return clampf(fexp(ofx * ofx * r0 + ofy * ofy * r1), lo, 1.0f);
}

// gauss_w of every offset |i|,|j| <= R for one covariance.  The arguments
// of (i,j) and (-i,-j) are equal bit for bit (the offsets enter through
// products whose sign flips exactly and are then squared), so each
// exponential is taken once, for the half plane j > 0 or (j = 0, i > 0),
// and mirrored; every tap then reads its weight whatever the CFA colour.
// With PAR >= 0 only the offsets whose i + j has that parity are weighted
// (a Bayer site's taps of one colour class all share it; (i,j) and (-i,-j)
// do too), the rest are left unset and must not be read.
template<int R, int PAR = -1>
struct sym_weights
{
  float e[2 * R + 1][2 * R + 1];
  __device__ __forceinline__ sym_weights(float4 cov, float r0, float r1, float lo)
  {
#pragma unroll
    for(int j = 0; j <= R; j++)
#pragma unroll
      for(int i = -R; i <= R; i++)
      {
        if(j == 0 && i <= 0) continue;
        if(PAR >= 0 && ((i + j) & 1) != PAR) continue;
        const float w = gauss_w(cov, r0, r1, i, j, lo);
        e[j + R][i + R] = w;
        e[R - j][R - i] = w;
      }
    if(PAR <= 0) e[R][R] = gauss_w(cov, r0, r1, 0, 0, lo);
  }
  __device__ __forceinline__ float operator()(int i, int j) const { return e[j + R][i + R]; }
};

// splat.comp:16-140: green at (x,y).  Only the green splats reach the
// stored channel, so only green sites are weighted (same sum and order).
//   mos(px, py): the mosaic at (px, py), 0 outside (texelFetch);
//   W, H its extent
//   PAR: -1, or on Bayer mosaics the parity of i + j of the site's green taps
//        (1 - ((x + y) & 1)): then only those taps' weights are evaluated
template<int XT, int PAR = -1, typename MosAt>
DEV float dm_splat_at(MosAt mos, int W, int H, const dimg &ga, int x, int y)
{
  const float4 cov = XT ? fetch_h4(ga, x / 3, y / 3) : fetch_h4(ga, (x + 1) / 2, (y + 1) / 2);
  const float r0 = -0.5f * frcp(clampf(cov.x, 0.01f, 25.0f)), r1 = -0.5f * frcp(clampf(cov.y, 0.01f, 25.0f));
  constexpr int r = XT ? 2 : 1;
  const sym_weights<r, PAR> tapw(cov, r0, r1, 1e-4f);
  float g = 0.0f, w = 0.0f;
#pragma unroll
  for(int j = -r; j <= r; j++)
#pragma unroll
    for(int i = -r; i <= r; i++)
    {
      if(PAR >= 0 && ((i + j) & 1) != PAR) continue;   // not green
      const bool green = XT ? xtrans_colour(x + i + 6, y + j + 6) == 1 : (((x + i + y + j) & 1) == 1);
      if(!green) continue;
      int px = x + i, py = y + j;
      if(px < 0) px += 6;
      if(py < 0) py += 6;
      if(px >= W) px -= 6;
      if(py >= H) py -= 6;
      const float col = mos(px, py);
      const float wt = (i == 0 && j == 0) ? 666.0f : tapw(i, j);
      g += col * wt;
      w += wt;
    }
  return g * frcp(fmaxf(1e-8f, w));
}

template<int XT, int U16 = 0, int PAR = -1>
DEV float dm_splat_px(const dimg &in, const dimg &ga, int x, int y)
{
  return dm_splat_at<XT, PAR>([&](int px, int py) { return fetch_mt<U16>(in, px, py); }, in.wd, in.ht, ga, x, y);
}

// fix.comp:21-135: rgb at (x,y) from the mosaic, the green plane (through
// `green(px,py)`, mirrored coordinates) and the block covariance.  Green
// sites copy the centre green, red and blue sites are ratio-weighted.
//   mos(px, py): the mosaic at mirrored (px, py); R the fix radius
//   PAR: -1, or on Bayer mosaics the parity of i + j of the site's red and
//        blue taps ((x + y) & 1): then only those taps' weights are evaluated
template<int XT, int R, int PAR = -1, typename MosAt, typename GreenAt>
DEV float4 dm_fix_at(MosAt mos, const dimg &cv, int x, int y, GreenAt green)
{
  float sr = 0.0f, sb = 0.0f, wr = 0.0f, wb = 0.0f;
  const float gc = green(x, y);
  const float4 cov = XT ? fetch_h4(cv, (x + 1) / 3, (y + 1) / 3) : fetch_h4(cv, (x + 1) / 2, (y + 1) / 2);
  const float hi = XT ? 10.0f : 49.0f, sc = XT ? 3.0f : 2.0f;
  const float r0 = -0.5f * frcp(sc * clampf(cov.x, 1.0f, hi)), r1 = -0.5f * frcp(sc * clampf(cov.y, 1.0f, hi));
  const float gcn = 1e-4f + gc;
  const sym_weights<R, PAR> tapw(cov, r0, r1, 1e-3f);
  bool any_green = false;
#pragma unroll
  for(int j = -R; j <= R; j++)
#pragma unroll
    for(int i = -R; i <= R; i++)
    {
      if(PAR >= 0 && ((i + j) & 1) != PAR) { any_green = true; continue; }   // a green tap
      const int px = x + i, py = y + j;   // may be negative: truncating division as in GLSL
      int k;
      if(XT)
      {
        const int blue_top = ((px / 3 + py / 3) & 1) > 0;
        const int mx = px - (px / 3) * 3, my = py - (py / 3) * 3;
        k = (((mx + my) & 1) == 0) ? 1 : ((blue_top ^ (my == 1)) ? 2 : 0);
      }
      else k = (((px + py) & 1) == 1) ? 1 : ((px & 1) == 0 ? 0 : 2);
      if(k == 1) { any_green = true; continue; }
      const float gh = green(px, py);
      const float col = mos(px, py);
      const float wt = tapw(i, j);
      const float v = col * gcn * frcp(1e-4f + gh) * wt;
      if(k == 0) { sr += v; wr += wt; }
      else       { sb += v; wb += wt; }
    }
  const float g = any_green ? gc : 0.0f;
  return make_float4(sr * frcp(fmaxf(1e-8f, wr)), g, sb * frcp(fmaxf(1e-8f, wb)), 1.0f);
}

// the same for any radius r, one weight per tap
template<int XT, typename MosAt, typename GreenAt>
DEV float4 dm_fix_loop(MosAt mos, const dimg &cv, int r, int x, int y, GreenAt green)
{
  float sr = 0.0f, sb = 0.0f, wr = 0.0f, wb = 0.0f;
  const float gc = green(x, y);
  const float4 cov = XT ? fetch_h4(cv, (x + 1) / 3, (y + 1) / 3) : fetch_h4(cv, (x + 1) / 2, (y + 1) / 2);
  const float hi = XT ? 10.0f : 49.0f, sc = XT ? 3.0f : 2.0f;
  const float r0 = -0.5f * frcp(sc * clampf(cov.x, 1.0f, hi)), r1 = -0.5f * frcp(sc * clampf(cov.y, 1.0f, hi));
  const float gcn = 1e-4f + gc;
  bool any_green = false;
  for(int j = -r; j <= r; j++)
    for(int i = -r; i <= r; i++)
    {
      const int px = x + i, py = y + j;   // may be negative: truncating division as in GLSL
      int k;
      if(XT)
      {
        const int blue_top = ((px / 3 + py / 3) & 1) > 0;
        const int mx = px - (px / 3) * 3, my = py - (py / 3) * 3;
        k = (((mx + my) & 1) == 0) ? 1 : ((blue_top ^ (my == 1)) ? 2 : 0);
      }
      else k = (((px + py) & 1) == 1) ? 1 : ((px & 1) == 0 ? 0 : 2);
      if(k == 1) { any_green = true; continue; }
      const float gh = green(px, py);
      const float col = mos(px, py);
      const float wt = gauss_w(cov, r0, r1, i, j, 1e-3f);
      const float v = col * gcn * frcp(1e-4f + gh) * wt;
      if(k == 0) { sr += v; wr += wt; }
      else       { sb += v; wb += wt; }
    }
  const float g = any_green ? gc : 0.0f;
  return make_float4(sr * frcp(fmaxf(1e-8f, wr)), g, sb * frcp(fmaxf(1e-8f, wb)), 1.0f);
}

// the fix radius r is uniform per image (1..2 Bayer, 2..3 X-Trans); the
// default radius (demosaic colour 0) runs unrolled with symmetric weights
template<int XT, int PAR = -1, typename MosAt, typename GreenAt>
DEV float4 dm_fix_r(MosAt mos, const dimg &cv, int r, int x, int y, GreenAt green)
{
  constexpr int R0 = XT ? 2 : 1;
  if(r == R0) return dm_fix_at<XT, R0, PAR>(mos, cv, x, y, green);
  return dm_fix_loop<XT>(mos, cv, r, x, y, green);
}

template<int XT, int U16 = 0, int PAR = -1, typename GreenAt>
DEV float4 dm_fix_px(const dimg &in, const dimg &cv, int r, int x, int y, GreenAt green)
{
  return dm_fix_r<XT, PAR>([&](int px, int py) { return centre_mt<U16>(in, px, py); }, cv, r, x, y, green);
}

// ==========================================================================
// colour (colour/main-impl.glsl:200-341, matrix path) and filmcurv
// (filmcurv/main.comp, modes 1 and 3)
// ==========================================================================

struct co_args
{
  float M[9];        // camera to rec2020, row major (identity for primaries 2)
  float toc[9];      // cat16 M * rec2020_to_xyz
  float fromc[9];    // xyz_to_rec2020 * cat16 M^-1
  float ratio[3];    // cl_dst / cl_src of the white-balance adaptation
  float expo;        // 2^exposure
};

DEV void mv3(float o[3], const float *M, float a, float b, float c)
{
  o[0] = M[0] * a + M[1] * b + M[2] * c;
  o[1] = M[3] * a + M[4] * b + M[5] * c;
  o[2] = M[6] * a + M[7] * b + M[8] * c;
}

DEV float4 co_px(float4 t, const co_args &a)
{
  float rgb[3], cl[3];
  mv3(rgb, a.M, t.x, t.y, t.z);
  mv3(cl, a.toc, rgb[0], rgb[1], rgb[2]);
  cl[0] *= a.ratio[0]; cl[1] *= a.ratio[1]; cl[2] *= a.ratio[2];
  mv3(rgb, a.fromc, cl[0], cl[1], cl[2]);
  return make_float4(clampf(rgb[0] * a.expo, -65535.0f, 65535.0f), clampf(rgb[1] * a.expo, -65535.0f, 65535.0f),
                     clampf(rgb[2] * a.expo, -65535.0f, 65535.0f), 1.0f);
}

// shared.glsl:371-387
DEV void adjust_colour_dng(float c0[3], float c1[3])
{
  float t;
  bool fx = false, fy = false, fz = false;
  if(c0[2] > c0[1]) { t = c0[2]; c0[2] = c0[1]; c0[1] = t; t = c1[2]; c1[2] = c1[1]; c1[1] = t; fx = true; }
  if(c0[1] > c0[0]) { t = c0[0]; c0[0] = c0[1]; c0[1] = t; t = c1[0]; c1[0] = c1[1]; c1[1] = t; fy = true; }
  if(c0[2] > c0[1]) { t = c0[2]; c0[2] = c0[1]; c0[1] = t; t = c1[2]; c1[2] = c1[1]; c1[1] = t; fz = true; }
  c1[1] = mixf(c1[2], c1[0], (c0[1] - c0[2] + 1e-6f) * frcp(c0[0] - c0[2] + 1e-6f));
  if(fz) { t = c1[2]; c1[2] = c1[1]; c1[1] = t; }
  if(fy) { t = c1[0]; c1[0] = c1[1]; c1[1] = t; }
  if(fx) { t = c1[2]; c1[2] = c1[1]; c1[1] = t; }
}

// filmcurv: fc_args, the per-channel and colour modes, fc_args_init
#include "ops_filmcurv.h"

// host: colour's committed params (colour/params, 242 words) to the matrix
// path's arguments.  Returns an error text for the paths not implemented
// (clut, rbf, saturation, gamut clip, non-linear trc, other primaries).
static inline const char *co_args_init(co_args &a, const void *params)
{
  const float *p = (const float *)params;
  const uint32_t *u = (const uint32_t *)params;
  if(u[16] > 0 || p[226] != 1.0f || p[231] > 0.0f || u[225] != 0 || u[230] != 0)
    return "only the matrix path (no clut/pick/rbf/sat/clip, linear trc) is supported";
  if(u[229] != 0 && u[229] != 2) return "primaries not supported (camera matrix and rec2020 are)";
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
    a.M[3 * j + i] = u[229] == 0 ? p[4 + 4 * i + j] : (i == j ? 1.0f : 0.0f);
  static const double cat16[9]  = { 0.401288, 0.650173, -0.051461, -0.250268, 1.204414, 0.045854, -0.002079, 0.048952, 0.953127 };
  static const double cat16i[9] = { 1.86206786, -1.01125463, 0.14918677, 0.38752654, 0.62144744, -0.00897398, -0.01584150, -0.03412294, 1.04996444 };
  static const double r2x[9] = { 0.636958048301290991, 0.144616903586208406, 0.168880975164172054, 0.26270021201126692,
                                 0.677998071518871148, 0.0593017164698619384, 4.9999999999999999e-17, 0.0280726930490874452,
                                 1.06098505771079066 };
  static const double x2r[9] = { 1.71665119, -0.35567078, -0.25336628, -0.66668435, 1.61648124, 0.01576855, 0.01763986,
                                 -0.04277061, 0.94210312 };
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
  {
    double s1 = 0, s2 = 0;
    for(int k = 0; k < 3; k++)
    {
      s1 += cat16[3 * j + k] * r2x[3 * k + i];
      s2 += x2r[3 * j + k] * cat16i[3 * k + i];
    }
    a.toc[3 * j + i] = (float)s1;
    a.fromc[3 * j + i] = (float)s2;
  }
  // cat16 source white (1,1,1), destination white mul.rgb
  for(int j = 0; j < 3; j++)
  {
    const float src = a.toc[3 * j] + a.toc[3 * j + 1] + a.toc[3 * j + 2];
    const float dst = a.toc[3 * j] * p[0] + a.toc[3 * j + 1] * p[1] + a.toc[3 * j + 2] * p[2];
    a.ratio[j] = dst / src;
  }
  a.expo = p[3];
  return 0;
}

// binary16 round trip: what a store to an f16 buffer and the next node's
// load do to a value
DEV float r16(float v) { return __half2float(__float2half_rn(v)); }
DEV float4 r16(float4 v) { return make_float4(r16(v.x), r16(v.y), r16(v.z), r16(v.w)); }

// ==========================================================================
// llap (src/pipe/modules/llap)
// ==========================================================================

#define LL_MAX_GAMMA 32

struct ll_args
{
  float sigma, shadows, highlights, clarity;
  float inv2s;               // 1/(2 sigma)
  float c3;                  // 3/(2 sigma^2)
  int   num_gamma;
  int   first;
  float gam[LL_MAX_GAMMA];   // gamma_from_i(i) = i/(num_gamma-1), correctly rounded
};

// curve.comp:39-63
DEV float ll_curve(float x, float g, const ll_args &a)
{
  const float c = x - g;
  float val;
  const bool pos = c > 0.0f;
  const float ssigma = pos ? a.sigma : -a.sigma;
  const float shadhi = pos ? a.shadows : a.highlights;
  if(fabsf(c) > 2.0f * a.sigma) val = g + ssigma + shadhi * (c - ssigma);
  else
  {
    const float t = clampf(c * (pos ? a.inv2s : -a.inv2s), 0.0f, 1.0f);
    const float t2 = t * t;
    const float mt = 1.0f - t;
    val = g + ssigma * 2.0f * mt * t + t2 * (ssigma + ssigma * shadhi);
  }
  val += a.clarity * c * fexp(-c * c * a.c3);
  return val;
}

// llap.glsl:16-21: first curve index above v (1..num_gamma-1)
DEV int gamma_hi_from_v(float v, const ll_args &a)
{
  int hi = 1;
  for(; hi < a.num_gamma - 1 && a.gam[hi] <= v; hi++);
  return hi;
}

// luminance of the clamped rgb (curve.comp:72)
DEV float ll_lum(float4 t)
{
  return lum2020(clampf(t.x, -1000.0f, 1000.0f), clampf(t.y, -1000.0f, 1000.0f), clampf(t.z, -1000.0f, 1000.0f));
}

// reduce.comp:17-35 (sample_semisoft at 2*o+.5): [1 2 1]/4 x [1 2 1]/4 on
// texel 2*o, mirrored, for `planes` planes of one output texel
DEV void ll_reduce_px(const dimg &in, const dimg &out, int planes, int x, int y)
{
  const int xm = mirror(2 * x - 1, in.wd), x0 = mirror(2 * x, in.wd), xp = mirror(2 * x + 1, in.wd);
  const int ym = mirror(2 * y - 1, in.ht), y0 = mirror(2 * y, in.ht), yp = mirror(2 * y + 1, in.ht);
  const __half *p = (const __half *)in.p;
  const long rm = (long)ym * in.pitch, r0 = (long)y0 * in.pitch, rp = (long)yp * in.pitch;
  for(int z = 0; z < planes; z++)
  {
    const __half *q = p + z * in.pstride;
    const float a = h2f(q[rm + xm]) + 2.0f * h2f(q[rm + x0]) + h2f(q[rm + xp]);
    const float b = h2f(q[r0 + xm]) + 2.0f * h2f(q[r0 + x0]) + h2f(q[r0 + xp]);
    const float c = h2f(q[rp + xm]) + 2.0f * h2f(q[rp + x0]) + h2f(q[rp + xp]);
    st_h1(out, x, y, (a + 2.0f * b + c) * (1.0f / 16.0f), z);
  }
}

// four outputs (x0..x0+3, y) of ll_reduce_px.  When the footprint
// 2x0-1 .. 2x0+7 lies inside the row and rows and planes are aligned
// (`vec` bits 0-1: 2 for 16-byte input rows, 1 for 8-byte ones, 3 for
// 4-byte ones), the texels 2x0 .. 2x0+7 of each input row arrive in one
// 16-byte load (two 8-byte, four 4-byte loads), and the four results leave
// in one 8-byte store where output rows
// are 8-byte aligned (`vec` bit 2); the sums are ll_reduce_px's, term for
// term.
DEV void ll_reduce_px4(const dimg &in, const dimg &out, int planes, int x0, int y, int vec)
{
  if(!(vec & 3) || 2 * x0 + 8 > in.wd)
  {
    for(int k = 0; k < 4; k++) if(x0 + k < out.wd) ll_reduce_px(in, out, planes, x0 + k, y);
    return;
  }
  const bool vout = (vec & 4) && x0 + 4 <= out.wd;
  const int xl = mirror(2 * x0 - 1, in.wd);
  const long rw[3] = { (long)mirror(2 * y - 1, in.ht) * in.pitch, (long)mirror(2 * y, in.ht) * in.pitch,
                       (long)mirror(2 * y + 1, in.ht) * in.pitch };
  const __half *p = (const __half *)in.p;
  __half *po = (__half *)out.p + (long)y * out.pitch + x0;
  for(int z = 0; z < planes; z++)
  {
    const __half *q = p + z * in.pstride;
    float h[3][4];
#pragma unroll
    for(int r = 0; r < 3; r++)
    {
      float c[9];
      c[0] = h2f(q[rw[r] + xl]);
      union { uint4 u; uint2 h[2]; __half e[8]; } v;
      if((vec & 3) == 2) v.u = *(const uint4 *)(q + rw[r] + 2 * x0);
      else if((vec & 3) == 1)
      {
        v.h[0] = *(const uint2 *)(q + rw[r] + 2 * x0);
        v.h[1] = *(const uint2 *)(q + rw[r] + 2 * x0 + 4);
      }
      else
      {
        const uint32_t *w = (const uint32_t *)(q + rw[r] + 2 * x0);
        v.u = make_uint4(w[0], w[1], w[2], w[3]);
      }
#pragma unroll
      for(int i = 0; i < 8; i++) c[i + 1] = h2f(v.e[i]);
#pragma unroll
      for(int k = 0; k < 4; k++) h[r][k] = c[2 * k] + 2.0f * c[2 * k + 1] + c[2 * k + 2];
    }
    float o[4];
#pragma unroll
    for(int k = 0; k < 4; k++) o[k] = (h[0][k] + 2.0f * h[1][k] + h[2][k]) * (1.0f / 16.0f);
    __half *qo = po + z * out.pstride;
    if(vout)
    {
      union { uint2 u; __half e[4]; } w;
#pragma unroll
      for(int k = 0; k < 4; k++) w.e[k] = f2h(o[k]);
      *(uint2 *)qo = w.u;
    }
    else
      for(int k = 0; k < 4; k++) if(x0 + k < out.wd) qo[k] = f2h(o[k]);
  }
}

// assemble.comp:169-203 for a fine tile of 64 x 4R texels; thread (tx,ty)
// owns the texels (X0+tx, Y0+ty+4k), k < R.  Every fine texel expands
// (sample_soft at o/2+.5 = (1/2 1/2 1 1/2 1/2)/3 even, (1 1/2 1/2 1)/3 odd,
// per axis) the coarse reconstruction and the two curve planes its
// luminance selects.  The tile's coarse footprint is staged in LDS only for
// the planes some texel of the tile selects (the block's [min hi-1, max hi]
// range), plus the coarse reconstruction in the last slot.
//   first        the coarsest level: expand the luminance plane of l1
//                instead of a coarse reconstruction
//   v[k]         the fine luminance plane at the thread's texel k
//   plane(k, p)  the fine curve plane p there; called for every k, also for
//                texels outside (owd, oht), so it must clamp its address
// Texels outside (owd, oht) get no result.  Loads are issued unconditionally
// (clamped addresses) so that the rows' loads and the staging's are in
// flight together instead of one memory latency each.
#define AS_TX 36                       // coarse tile width for a 64 wide fine tile (+-2 halo)
#define AS_TY(R) (2 * (R) + 4)         // coarse tile height for a 4R high fine tile
#define AS_TN(R) (AS_TX * AS_TY(R))

// The expand filter's five taps along one axis: weights 1/2 1/2 1 1/2 1/2 at
// even texels and 0 1 1/2 1/2 1 at odd ones, summed from zero in tap order
// (llap.glsl upsample).  Scaling by two commutes with rounding, so twice that
// sum is exactly (((t0 + t1) + 2 t2) + t3) + t4 at even texels and
// ((2 t1 + t2) + t3) + 2 t4 at odd ones (an all-zero sum may differ in the
// sign of zero): ll_expand5 returns it, one multiply, three FMAs (each product
// by 1 or 2 is exact) and an add instead of five multiplies and five adds.
struct ll_expand_w { float z0, m1, m2, m4; };
DEV ll_expand_w ll_expand_weights(int odd)
{
  ll_expand_w w;
  w.z0 = odd ? 0.0f : 1.0f;
  w.m1 = odd ? 2.0f : 1.0f;
  w.m2 = odd ? 1.0f : 2.0f;
  w.m4 = odd ? 2.0f : 1.0f;
  return w;
}
DEV float ll_expand5(const ll_expand_w &w, float t0, float t1, float t2, float t3, float t4)
{
  return __builtin_fmaf(t4, w.m4, __builtin_fmaf(t2, w.m2, __builtin_fmaf(t1, w.m1, t0 * w.z0)) + t3);
}

DEV int wave_min(int v) { for(int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o)); return v; }
DEV int wave_max(int v) { for(int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o)); return v; }

template<int R, typename PlaneAt>
DEV void ll_assemble_tile(const dimg &coarse, const dimg &l1, int owd, int oht, const ll_args &a, int first,
                          int X0, int Y0, float *tile, int *lims, const float v[R], PlaneAt plane, float res[R])
{
  constexpr int CN = AS_TN(R);
  const int ng = a.num_gamma;
  const int x = X0 + threadIdx.x;
  if(threadIdx.x == 0 && threadIdx.y == 0) { lims[0] = LL_MAX_GAMMA; lims[1] = 0; }
  int hi[R];
  float pv0[R], pv1[R];
  int lo_t = LL_MAX_GAMMA, hi_t = 0;
#pragma unroll
  for(int k = 0; k < R; k++)
  {
    const int y = Y0 + threadIdx.y + 4 * k;
    hi[k] = gamma_hi_from_v(v[k], a);
    // issued before the staging so their latency overlaps it
    pv0[k] = plane(k, hi[k] - 1);
    pv1[k] = plane(k, hi[k]);
    if(x < owd && y < oht)
    {
      lo_t = min(lo_t, hi[k] - 1);
      hi_t = max(hi_t, hi[k]);
    }
  }
  lo_t = wave_min(lo_t);
  hi_t = wave_max(hi_t);
  __syncthreads();
  if(threadIdx.x == 0) { atomicMin(&lims[0], lo_t); atomicMax(&lims[1], hi_t); }
  __syncthreads();
  const int p0 = lims[0], np = lims[1] - p0 + 1;
  const int cx0 = X0 / 2 - 2, cy0 = Y0 / 2 - 2;
  const int tid = threadIdx.y * BX + threadIdx.x;
  // slots 0..np-1: planes p0.. of l1; slot np: the coarse reconstruction
  // (l1's plane ng on the first assemble).  Four loads per thread and pass.
  const int n = (np + 1) * CN;
  constexpr int S = BX * BY, U = 4;
  for(int i0 = tid; i0 < n; i0 += U * S)
  {
    float c[U];
#pragma unroll
    for(int u = 0; u < U; u++)
    {
      const int idx = min(i0 + u * S, n - 1);
      const int sl = idx / CN, r = idx - sl * CN;
      const int ty = r / AS_TX, tx = r - ty * AS_TX;
      const int gx = mirror(cx0 + tx, l1.wd), gy = mirror(cy0 + ty, l1.ht);
      const bool fc = sl >= np && first <= 0;
      const __half *src = fc ? (const __half *)coarse.p : (const __half *)l1.p;
      const long off = fc ? (long)gy * coarse.pitch + gx
                          : (long)(sl < np ? p0 + sl : ng) * l1.pstride + (long)gy * l1.pitch + gx;
      c[u] = h2f(src[off]);
    }
#pragma unroll
    for(int u = 0; u < U; u++)
      if(i0 + u * S < n) tile[i0 + u * S] = c[u];
  }
  __syncthreads();
  const ll_expand_w wx = ll_expand_weights(x & 1);
  const int lx = x / 2 - cx0 - 2;   // tile column of offset -2
#pragma unroll
  for(int k = 0; k < R; k++)
  {
    const int y = Y0 + threadIdx.y + 4 * k;
    if(x >= owd || y >= oht) continue;
    const ll_expand_w wy = ll_expand_weights(y & 1);
    const int ly = y / 2 - cy0 - 2;
    const int h = hi[k];
    const float al = clampf((v[k] - a.gam[h - 1]) / (a.gam[h] - a.gam[h - 1]), 0.0f, 1.0f);
    const float *tc = tile + np * CN, *t0 = tile + (h - 1 - p0) * CN, *t1 = tile + (h - p0) * CN;
    // rows: twice the horizontal sums; columns: twice again, so the
    // reference's weighted sums are a quarter of these, exactly
    float sc[5], s0[5], s1[5];
#pragma unroll
    for(int j = 0; j < 5; j++)
    {
      const int o = (ly + j) * AS_TX + lx;
      sc[j] = ll_expand5(wx, tc[o], tc[o + 1], tc[o + 2], tc[o + 3], tc[o + 4]);
      s0[j] = ll_expand5(wx, t0[o], t0[o + 1], t0[o + 2], t0[o + 3], t0[o + 4]);
      s1[j] = ll_expand5(wx, t1[o], t1[o + 1], t1[o + 2], t1[o + 3], t1[o + 4]);
    }
    const float rs = 0.25f * ll_expand5(wy, sc[0], sc[1], sc[2], sc[3], sc[4]);
    const float e0 = 0.25f * ll_expand5(wy, s0[0], s0[1], s0[2], s0[3], s0[4]);
    const float e1 = 0.25f * ll_expand5(wy, s1[0], s1[1], s1[2], s1[3], s1[4]);
    const float s = 1.0f / 9.0f;
    const float lv0 = pv0[k] - e0 * s;
    const float lv1 = pv1[k] - e1 * s;
    res[k] = rs * s + lv0 * (1.0f - al) + lv1 * al;
  }
}

// colour.comp:221-240: black rolloff and luminance transfer
DEV float4 ll_colour_px(float yv, float4 t)
{
  const float yo = fmaxf(lum2020(t.x, t.y, t.z), 1e-8f);
  if(yv < yo) yv = yo * fexp(yv - yo);
  const float r = yv * frcp(yo);
  return make_float4(fmaxf(0.0f, t.x * r), fmaxf(0.0f, t.y * r), fmaxf(0.0f, t.z * r), 1.0f);
}

// host-side argument setup shared by both paths
static inline void ll_args_init(ll_args &a, const float *p, int ng)
{
  a.sigma = p[0];
  a.shadows = p[1];
  a.highlights = p[2];
  a.clarity = p[3];
  a.inv2s = 1.0f / (2.0f * a.sigma);
  a.c3 = 1.0f / (2.0f * a.sigma * a.sigma / 3.0f);
  a.num_gamma = ng;
  a.first = 0;
  for(int i = 0; i < LL_MAX_GAMMA; i++) a.gam[i] = i < ng ? (float)i / ((float)ng - 1.0f) : 2.0f;
}
