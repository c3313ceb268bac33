// k_llap.hip -- local laplacian pyramid nodes (src/pipe/modules/llap):
// curve, reduce, assemble, colour.
//
// Pyramid buffers are arrays of single-channel f16 planes: plane i < G-1 is
// the input luminance remapped by curve i, plane G-1 (= num_gamma) is the
// luminance itself.  Planes of one level are contiguous, plane_stride apart.
//
// The reference samples with bilinear taps at half-texel positions
// (sample_semisoft / sample_soft, shared.glsl:99-148); at those positions
// the taps reduce to fixed separable weights, which is what these kernels
// evaluate: reduce = [1 2 1]/4 x [1 2 1]/4 centred on texel 2*o, expand =
// (1/2 1/2 1 1/2 1/2)/3 (even o) or (1 1/2 1/2 1)/3 (odd o) per axis around
// texel o/2, all with mirrored-repeat addressing.
#include "launch.h"

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

__global__ void __launch_bounds__(256) k_ll_curve(dimg in, dimg out, ll_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= in.wd || y >= in.ht) return;
  const float4 t = fetch_h4(in, x, y);
  const float yl = lum2020(clampf(t.x, -1000.0f, 1000.0f), clampf(t.y, -1000.0f, 1000.0f), clampf(t.z, -1000.0f, 1000.0f));
  for(int i = 0; i < a.num_gamma; i++)
    st_h1(out, x, y, ll_curve(yl, a.gam[i], a), i);
  st_h1(out, x, y, yl, a.num_gamma);
}

// reduce.comp:17-35, all planes of one texel per thread
__global__ void __launch_bounds__(256) k_ll_reduce(dimg in, dimg out, int planes)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
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

// assemble.comp:169-203.  The coarse level of the work group's fine tile is
// staged in LDS for all curve planes and the coarse reconstruction, then
// every fine texel expands the planes it selects from there.
#define AS_TX 36   // coarse tile width for a 64 wide fine tile (+-2 halo)
#define AS_TY 6    // coarse tile height for a 4 high fine tile
#define AS_TN (AS_TX * AS_TY)

__global__ void __launch_bounds__(256) k_ll_assemble(dimg coarse, dimg l0, dimg l1, dimg out, ll_args a)
{
  extern __shared__ float tile[];             // (num_gamma + 1) planes x AS_TN, coarse slot last
  const int ng = a.num_gamma;
  const int X0 = blockIdx.x * BX, Y0 = blockIdx.y * BY;
  const int cx0 = X0 / 2 - 2, cy0 = Y0 / 2 - 2;
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int idx = tid; idx < (ng + 1) * AS_TN; idx += BX * BY)
  {
    const int p = idx / AS_TN, r = idx - p * AS_TN;
    const int ty = r / AS_TX, tx = r - ty * AS_TX;
    const int gx = mirror(cx0 + tx, l1.wd), gy = mirror(cy0 + ty, l1.ht);
    float v;
    if(p < ng) v = ld_h1(l1, gx, gy, p);
    else if(a.first > 0) v = ld_h1(l1, gx, gy, ng);
    else v = ld_h1(coarse, gx, gy, 0);
    tile[idx] = v;
  }
  __syncthreads();
  const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  // per-axis weights over coarse offsets -2..2 (sum 3 per axis, 9 total)
  const float we[5] = { 0.5f, 0.5f, 1.0f, 0.5f, 0.5f }, wo[5] = { 0.0f, 1.0f, 0.5f, 0.5f, 1.0f };
  const float *wx = (x & 1) ? wo : we, *wy = (y & 1) ? wo : we;
  const int lx = x / 2 - cx0 - 2, ly = y / 2 - cy0 - 2;   // tile index of offset -2
  const float v = fetch_h1(l0, x, y, ng);
  const int hi = gamma_hi_from_v(v, a);
  const float al = clampf((v - a.gam[hi - 1]) / (a.gam[hi] - a.gam[hi - 1]), 0.0f, 1.0f);
  const float *tc = tile + ng * AS_TN, *t0 = tile + (hi - 1) * AS_TN, *t1 = tile + hi * AS_TN;
  float res = 0.0f, e0 = 0.0f, e1 = 0.0f;
#pragma unroll
  for(int j = 0; j < 5; j++)
  {
    float rc = 0.0f, r0 = 0.0f, r1 = 0.0f;
    const int o = (ly + j) * AS_TX + lx;
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      rc += wx[i] * tc[o + i];
      r0 += wx[i] * t0[o + i];
      r1 += wx[i] * t1[o + i];
    }
    res += wy[j] * rc;
    e0 += wy[j] * r0;
    e1 += wy[j] * r1;
  }
  const float s = 1.0f / 9.0f;
  const float lv0 = fetch_h1(l0, x, y, hi - 1) - e0 * s;
  const float lv1 = fetch_h1(l0, x, y, hi) - e1 * s;
  st_h1(out, x, y, res * s + lv0 * (1.0f - al) + lv1 * al);
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
  if(yv < yo) yv = yo * fexp(yv - yo);
  const float r = yv * frcp(yo);
  const float4 o = make_float4(fmaxf(0.0f, t.x * r), fmaxf(0.0f, t.y * r), fmaxf(0.0f, t.z * r), 1.0f);
  if(F32) st_f4(out, x, y, o);
  else    st_h4(out, x, y, o);
}

// ---- host launchers --------------------------------------------------------

static int read_args(ll_args &a, const node_call &c, int ng)
{
  const float *p = c.params ? (const float *)c.params : nullptr;
  if(!p || c.params_size < 16) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap: params missing");
  a.sigma = p[0];
  a.shadows = p[1];
  a.highlights = p[2];
  a.clarity = p[3];
  a.inv2s = 1.0f / (2.0f * a.sigma);
  a.c3 = 1.0f / (2.0f * a.sigma * a.sigma / 3.0f);
  if(ng < 2 || ng > LL_MAX_GAMMA) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap: num_gamma %d outside [2,%d]", ng, LL_MAX_GAMMA);
  a.num_gamma = ng;
  for(int i = 0; i < LL_MAX_GAMMA; i++) a.gam[i] = i < ng ? (float)i / ((float)ng - 1.0f) : 2.0f;
  return VKDT_HIP_OK;
}

int launch_llap_curve(const node_call &c)
{
  if(int e = check_blob(c, 4, 4 * 4, "llap/curve")) return e;
  ll_args a = {};
  if(int e = read_args(a, c, ((const int32_t *)c.pc)[0])) return e;
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
  hipLaunchKernelGGL(k_ll_reduce, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), (int)c.dp);
  return launch_post("llap/reduce");
}

int launch_llap_assemble(const node_call &c)
{
  if(int e = check_blob(c, 8, 16, "llap/assemble")) return e;
  ll_args a = {};
  if(int e = read_args(a, c, ((const int32_t *)c.pc)[0])) return e;
  a.first = ((const int32_t *)c.pc)[1];
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "llap/assemble coarse")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, a.num_gamma + 1, "llap/assemble currlo")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 1, a.num_gamma + 1, "llap/assemble currhi")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_F16, 1, 1, "llap/assemble fine")) return e;
  if(c.conn[1].wd < c.conn[3].wd || c.conn[1].ht < c.conn[3].ht)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/assemble: currlo smaller than output");
  if(!a.first && (c.conn[0].wd != c.conn[2].wd || c.conn[0].ht != c.conn[2].ht))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap/assemble: coarse and currhi extents differ");
  const size_t lds = sizeof(float) * (a.num_gamma + 1) * AS_TN;
  hipLaunchKernelGGL(k_ll_assemble, grid2d(c.wd, c.ht, 1), dim3(BX, BY), lds, c.stream,
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
