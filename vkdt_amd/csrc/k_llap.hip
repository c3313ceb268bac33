// k_llap.hip -- local laplacian pyramid nodes (src/pipe/modules/llap):
// curve, reduce, assemble, colour.
//
// Pyramid buffers are arrays of single-channel f16 planes: plane i < G-1 is
// the input luminance remapped by curve i, plane G-1 (= num_gamma) is the
// luminance itself.  Planes of one level are contiguous, plane_stride apart.
#include "launch.h"

struct ll_args
{
  float sigma, shadows, highlights, clarity;
  int   num_gamma;
  int   first;
};

// curve.comp:39-63
DEV float ll_curve(float x, float g, float sigma, float shadows, float highlights, float clarity)
{
  const float c = x - g;
  float val;
  const float ssigma = c > 0.0f ? sigma : -sigma;
  const float shadhi = c > 0.0f ? shadows : highlights;
  if(fabsf(c) > 2.0f * sigma) val = g + ssigma + shadhi * (c - ssigma);
  else
  {
    const float t = clampf(c / (2.0f * ssigma), 0.0f, 1.0f);
    const float t2 = t * t;
    const float mt = 1.0f - t;
    val = g + ssigma * 2.0f * mt * t + t2 * (ssigma + ssigma * shadhi);
  }
  val += clarity * c * expf(-c * c / (2.0f * sigma * sigma / 3.0f));
  return val;
}

// llap.glsl:3-21 with linear spacing over num_gamma samples
DEV float gamma_from_i(int i, int ng) { return (float)i / ((float)ng - 1.0f); }
DEV int gamma_hi_from_v(float v, int ng)
{
  int hi = 1;
  for(; hi < ng - 1 && gamma_from_i(hi, ng) <= v; hi++);
  return hi;
}

__global__ void __launch_bounds__(256) k_ll_curve(dimg in, dimg out, ll_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= in.wd || y >= in.ht) return;
  const float4 t = fetch_h4(in, x, y);
  const float yl = lum2020(clampf(t.x, -1000.0f, 1000.0f), clampf(t.y, -1000.0f, 1000.0f), clampf(t.z, -1000.0f, 1000.0f));
  for(int i = 0; i < a.num_gamma; i++)
    st_h1(out, x, y, ll_curve(yl, gamma_from_i(i, a.num_gamma), a.sigma, a.shadows, a.highlights, a.clarity), i);
  st_h1(out, x, y, yl, a.num_gamma);
}

// reduce.comp:17-35: sample_semisoft at 2*opos+.5, i.e. [1 2 1]/4 x [1 2 1]/4
// centred on texel 2*opos with mirrored borders
__global__ void __launch_bounds__(256) k_ll_reduce(dimg in, dimg out)
{
  const int x = tid_x(), y = tid_y(), z = blockIdx.z;
  if(x >= out.wd || y >= out.ht) return;
  const float s = 2.0f * x + 0.5f, t = 2.0f * y + 0.5f;
  const float r = (bilin_h1(in, s - 0.5f, t - 0.5f, z) + bilin_h1(in, s + 0.5f, t - 0.5f, z)
                 + bilin_h1(in, s - 0.5f, t + 0.5f, z) + bilin_h1(in, s + 0.5f, t + 0.5f, z)) / 4.0f;
  st_h1(out, x, y, r, z);
}

// assemble.comp:134-138 gauss_expand = sample_soft at opos/2+.5
DEV float expand(const dimg &img, int pl, int ox, int oy)
{
  const float sc = (float)ox * 0.5f + 0.5f, tc = (float)oy * 0.5f + 0.5f;
  float r = 0.0f;
#pragma unroll
  for(int j = -1; j <= 1; j++)
#pragma unroll
    for(int i = -1; i <= 1; i++)
      r += bilin_h1(img, sc + 1.5f * i, tc + 1.5f * j, pl);
  return r / 9.0f;
}

// assemble.comp:169-203
__global__ void __launch_bounds__(256) k_ll_assemble(dimg coarse, dimg l0, dimg l1, dimg out, ll_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const int ng = a.num_gamma;
  const float res = a.first > 0 ? expand(l1, ng, x, y) : expand(coarse, 0, x, y);
  const float v = fetch_h1(l0, x, y, ng);
  const int hi = gamma_hi_from_v(v, ng);
  const int lo = hi - 1;
  const float glo = gamma_from_i(lo, ng), ghi = gamma_from_i(hi, ng);
  const float al = clampf((v - glo) / (ghi - glo), 0.0f, 1.0f);
  const float lv0 = fetch_h1(l0, x, y, hi - 1) - expand(l1, hi - 1, x, y);
  const float lv1 = fetch_h1(l0, x, y, hi) - expand(l1, hi, x, y);
  st_h1(out, x, y, res + lv0 * (1.0f - al) + lv1 * al);
}

// colour.comp:221-240; output rgba f16 or f32 (f32 when it feeds the export sink)
template<int F32>
__global__ void __launch_bounds__(256) k_ll_colour(dimg lum, dimg org, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float yv = fetch_h1(lum, x, y);
  const float4 t = fetch_h4(org, x, y);
  const float yo = fmaxf(lum2020(t.x, t.y, t.z), 1e-8f);
  if(yv < yo) yv = yo * expf(1.0f * (yv - yo));
  const float4 o = make_float4(fmaxf(0.0f, t.x * yv / yo), fmaxf(0.0f, t.y * yv / yo), fmaxf(0.0f, t.z * yv / yo), 1.0f);
  if(F32) st_f4(out, x, y, o);
  else    st_h4(out, x, y, o);
}

// ---- host launchers --------------------------------------------------------

static void read_params(ll_args &a, const node_call &c)
{
  const float *p = (const float *)c.params;
  a.sigma = p[0];
  a.shadows = p[1];
  a.highlights = p[2];
  a.clarity = p[3];
}

int launch_llap_curve(const node_call &c)
{
  if(int e = check_blob(c, 4, 4 * 4, "llap/curve")) return e;
  ll_args a = {};
  read_params(a, c);
  a.num_gamma = ((const int32_t *)c.pc)[0];
  if(a.num_gamma < 2 || a.num_gamma > 64) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/curve: num_gamma %d", a.num_gamma);
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "llap/curve input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, a.num_gamma + 1, "llap/curve output")) return e;
  if(c.conn[1].wd < c.conn[0].wd || c.conn[1].ht < c.conn[0].ht)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/curve: output smaller than input");
  hipLaunchKernelGGL(k_ll_curve, grid2d(c.conn[0].wd, c.conn[0].ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("llap/curve");
}

int launch_llap_reduce(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, c.dp, "llap/reduce inhi")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, c.dp, "llap/reduce outlo")) return e;
  hipLaunchKernelGGL(k_ll_reduce, grid2d(c.wd, c.ht, c.dp), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]));
  return launch_post("llap/reduce");
}

int launch_llap_assemble(const node_call &c)
{
  if(int e = check_blob(c, 8, 0, "llap/assemble")) return e;
  ll_args a = {};
  a.num_gamma = ((const int32_t *)c.pc)[0];
  a.first = ((const int32_t *)c.pc)[1];
  if(a.num_gamma < 2 || a.num_gamma > 64) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/assemble: num_gamma %d", a.num_gamma);
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "llap/assemble coarse")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, a.num_gamma + 1, "llap/assemble currlo")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 1, a.num_gamma + 1, "llap/assemble currhi")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_F16, 1, 1, "llap/assemble fine")) return e;
  if(c.conn[1].wd < c.conn[3].wd || c.conn[1].ht < c.conn[3].ht)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/assemble: currlo smaller than output");
  hipLaunchKernelGGL(k_ll_assemble, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), a);
  return launch_post("llap/assemble");
}

int launch_llap_colour(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "llap/colour lum")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "llap/colour input")) return e;
  const int f32 = c.nconn > 2 && c.conn[2].format == VKDT_FMT_F32;
  if(int e = check_conn(c, 2, f32 ? VKDT_FMT_F32 : VKDT_FMT_F16, 4, 1, "llap/colour output")) return e;
  if(f32)
    hipLaunchKernelGGL(k_ll_colour<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  else
    hipLaunchKernelGGL(k_ll_colour<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("llap/colour");
}
