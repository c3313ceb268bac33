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
#include "ops.h"

// four adjacent texels per thread; with 8-byte aligned output rows and
// planes (`vec`) each plane's four values leave in one store
__global__ void __launch_bounds__(256) k_ll_curve(dimg in, dimg out, ll_args a, int vec)
{
  const int x0 = 4 * tid_x(), y = tid_y();
  if(x0 >= in.wd || y >= in.ht) return;
  if(!vec || x0 + 4 > in.wd)
  {
    for(int k = 0; k < 4 && x0 + k < in.wd; k++)
    {
      const float yl = ll_lum(ld_h4(in, x0 + k, y));
      for(int i = 0; i < a.num_gamma; i++) st_h1(out, x0 + k, y, ll_curve(yl, a.gam[i], a), i);
      st_h1(out, x0 + k, y, yl, a.num_gamma);
    }
    return;
  }
  float yl[4];
#pragma unroll
  for(int k = 0; k < 4; k++) yl[k] = ll_lum(ld_h4(in, x0 + k, y));
  __half *po = (__half *)out.p + (long)y * out.pitch + x0;
  union { uint2 u; __half e[4]; } w;
  for(int i = 0; i < a.num_gamma; i++)
  {
#pragma unroll
    for(int k = 0; k < 4; k++) w.e[k] = f2h(ll_curve(yl[k], a.gam[i], a));
    *(uint2 *)(po + i * out.pstride) = w.u;
  }
#pragma unroll
  for(int k = 0; k < 4; k++) w.e[k] = f2h(yl[k]);
  *(uint2 *)(po + a.num_gamma * out.pstride) = w.u;
}

// eight adjacent texels per thread where the input rows are 16-byte aligned
// and the output rows and planes too: four 16-byte loads, and one 16-byte
// store a plane; the values of k_ll_curve
__global__ void __launch_bounds__(256) k_ll_curve8(dimg in, dimg out, ll_args a)
{
  const int x0 = 8 * tid_x(), y = tid_y();
  if(x0 >= in.wd || y >= in.ht) return;
  if(x0 + 8 > in.wd)
  {
    for(int k = 0; k < 8 && x0 + k < in.wd; k++)
    {
      const float yl = ll_lum(ld_h4(in, x0 + k, y));
      for(int i = 0; i < a.num_gamma; i++) st_h1(out, x0 + k, y, ll_curve(yl, a.gam[i], a), i);
      st_h1(out, x0 + k, y, yl, a.num_gamma);
    }
    return;
  }
  float yl[8];
  const uint4 *pi = (const uint4 *)((const __half *)in.p + 4 * ((long)y * in.pitch + x0));
#pragma unroll
  for(int k = 0; k < 4; k++)
  {
    const uint4 u = pi[k];
    const __half2 a0 = *(const __half2 *)&u.x, c0 = *(const __half2 *)&u.y;
    const __half2 a1 = *(const __half2 *)&u.z, c1 = *(const __half2 *)&u.w;
    yl[2 * k] = ll_lum(make_float4(__low2float(a0), __high2float(a0), __low2float(c0), __high2float(c0)));
    yl[2 * k + 1] = ll_lum(make_float4(__low2float(a1), __high2float(a1), __low2float(c1), __high2float(c1)));
  }
  __half *po = (__half *)out.p + (long)y * out.pitch + x0;
  union { uint4 u; __half e[8]; } w;
  for(int i = 0; i < a.num_gamma; i++)
  {
#pragma unroll
    for(int k = 0; k < 8; k++) w.e[k] = f2h(ll_curve(yl[k], a.gam[i], a));
    *(uint4 *)(po + i * out.pstride) = w.u;
  }
#pragma unroll
  for(int k = 0; k < 8; k++) w.e[k] = f2h(yl[k]);
  *(uint4 *)(po + a.num_gamma * out.pstride) = w.u;
}

// four adjacent output texels of one plane (blockIdx.z) per thread: the
// planes' loads are independent, so they go to separate waves rather than
// one thread's serial loop
__global__ void __launch_bounds__(256) k_ll_reduce(dimg in, dimg out, int vec)
{
  const int x = 4 * tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  dimg i1 = in, o1 = out;
  i1.p = (__half *)in.p + (long)blockIdx.z * in.pstride;
  o1.p = (__half *)out.p + (long)blockIdx.z * out.pstride;
  ll_reduce_px4(i1, o1, 1, x, y, vec);
}

#define AS_R 4   // fine rows per thread: 64 x 16 fine texels per work group

__global__ void __launch_bounds__(256) k_ll_assemble(dimg coarse, dimg l0, dimg l1, dimg out, ll_args a)
{
  extern __shared__ float tile[];             // up to (num_gamma + 1) slots of AS_TN(AS_R)
  __shared__ int lims[2];
  const int2 t = xcd_block();
  const int X0 = t.x * BX, Y0 = t.y * (BY * AS_R);
  const int x = X0 + threadIdx.x;
  float v[AS_R], res[AS_R];
#pragma unroll
  for(int k = 0; k < AS_R; k++)
  {
    const int y = Y0 + threadIdx.y + 4 * k;
    v[k] = ld_h1(l0, min(x, out.wd - 1), min(y, out.ht - 1), a.num_gamma);   // clamped: no result outside
  }
  const int xc = min(x, out.wd - 1);
  ll_assemble_tile<AS_R>(coarse, l1, out.wd, out.ht, a, a.first, X0, Y0, tile, lims, v,
                         [&](int k, int p) { return ld_h1(l0, xc, min(Y0 + (int)threadIdx.y + 4 * k, out.ht - 1), p); },
                         res);
#pragma unroll
  for(int k = 0; k < AS_R; k++)
  {
    const int y = Y0 + threadIdx.y + 4 * k;
    if(x < out.wd && y < out.ht) st_h1(out, x, y, res[k]);
  }
}

// The coarse levels, at most LL_SMALL texels: one thread a texel reads its
// coarse footprint straight from memory (L2 holds these levels) instead of
// staging the tile's planes in LDS, which costs the tiny levels several
// dependent memory round trips; the values, weights and the order of the
// sums are those of ll_assemble_tile, so the result is the same bit for bit.
#define LL_SMALL (1 << 19)
__global__ void __launch_bounds__(256) k_ll_assemble_small(dimg coarse, dimg l0, dimg l1, dimg out, ll_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float v = ld_h1(l0, x, y, a.num_gamma);
  const int h = gamma_hi_from_v(v, a);
  const float pv0 = ld_h1(l0, x, y, h - 1), pv1 = ld_h1(l0, x, y, h);
  const ll_expand_w wx = ll_expand_weights(x & 1), wy = ll_expand_weights(y & 1);
  const float al = clampf((v - a.gam[h - 1]) / (a.gam[h] - a.gam[h - 1]), 0.0f, 1.0f);
  float sc[5], s0[5], s1[5];
#pragma unroll
  for(int j = 0; j < 5; j++)
  {
    const int gy = mirror(y / 2 - 2 + j, l1.ht);
    float c[5], r0[5], r1[5];
#pragma unroll
    for(int i = 0; i < 5; i++)
    {
      const int gx = mirror(x / 2 - 2 + i, l1.wd);
      c[i] = a.first <= 0 ? ld_h1(coarse, gx, gy) : ld_h1(l1, gx, gy, a.num_gamma);
      r0[i] = ld_h1(l1, gx, gy, h - 1);
      r1[i] = ld_h1(l1, gx, gy, h);
    }
    sc[j] = ll_expand5(wx, c[0], c[1], c[2], c[3], c[4]);
    s0[j] = ll_expand5(wx, r0[0], r0[1], r0[2], r0[3], r0[4]);
    s1[j] = ll_expand5(wx, r1[0], r1[1], r1[2], r1[3], r1[4]);
  }
  const float rs = 0.25f * ll_expand5(wy, sc[0], sc[1], sc[2], sc[3], sc[4]);
  const float e0 = 0.25f * ll_expand5(wy, s0[0], s0[1], s0[2], s0[3], s0[4]);
  const float e1 = 0.25f * ll_expand5(wy, s1[0], s1[1], s1[2], s1[3], s1[4]);
  const float s = 1.0f / 9.0f;
  const float lv0 = pv0 - e0 * s;
  const float lv1 = pv1 - e1 * s;
  st_h1(out, x, y, rs * s + lv0 * (1.0f - al) + lv1 * al);
}

// colour.comp:221-240; output rgba f16 or f32 (f32 when it feeds the export sink)
template<int F32>
__global__ void __launch_bounds__(256) k_ll_colour(dimg lum, dimg org, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 o = ll_colour_px(fetch_h1(lum, x, y), fetch_h4(org, x, y));
  if(F32) st_f4(out, x, y, o);
  else    st_h4(out, x, y, o);
}

// two texels a thread where lum, org and out have the output's extent and
// even, aligned rows: a 4-byte luminance pair, a 16-byte rgba f16 pair, and
// two 16-byte (f32) or one 16-byte (f16) store; the arithmetic of k_ll_colour
template<int F32>
__global__ void __launch_bounds__(256) k_ll_colour2(dimg lum, dimg org, dimg out)
{
  const int x = 2 * tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  if(x + 1 >= out.wd)
  {
    const float4 o = ll_colour_px(ld_h1(lum, x, y), ld_h4(org, x, y));
    if(F32) st_f4(out, x, y, o);
    else    st_h4(out, x, y, o);
    return;
  }
  const __half2 l = *(const __half2 *)((const __half *)lum.p + (long)y * lum.pitch + x);
  const uint4 u = *(const uint4 *)((const __half *)org.p + 4 * ((long)y * org.pitch + x));
  const __half2 a = *(const __half2 *)&u.x, c = *(const __half2 *)&u.y;
  const __half2 d = *(const __half2 *)&u.z, e = *(const __half2 *)&u.w;
  const float4 o0 = ll_colour_px(__low2float(l), make_float4(__low2float(a), __high2float(a), __low2float(c), __high2float(c)));
  const float4 o1 = ll_colour_px(__high2float(l), make_float4(__low2float(d), __high2float(d), __low2float(e), __high2float(e)));
  if(F32)
  {
    float4 *q = (float4 *)((float *)out.p + 4 * ((long)y * out.pitch + x));
    q[0] = o0;
    q[1] = o1;
  }
  else
  {
    st_h4(out, x, y, o0);
    st_h4(out, x + 1, y, o1);
  }
}

// ---- host launchers --------------------------------------------------------

static int read_args(ll_args &a, const node_call &c, int ng)
{
  const float *p = c.params ? (const float *)c.params : nullptr;
  if(!p || c.params_size < 16) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap: params missing");
  if(ng < 2 || ng > LL_MAX_GAMMA) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "llap: num_gamma %d outside [2,%d]", ng, LL_MAX_GAMMA);
  ll_args_init(a, p, ng);
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
  const dimg out = to_dimg(c.conn[1]), in = to_dimg(c.conn[0]);
  if(out.pitch % 8 == 0 && out.pstride % 8 == 0 && (uintptr_t)out.p % 16 == 0 && in.pitch % 2 == 0 &&
     (uintptr_t)in.p % 16 == 0)
  {
    hipLaunchKernelGGL(k_ll_curve8, grid2d((c.conn[0].wd + 7) / 8, c.conn[0].ht, 1), dim3(BX, BY), 0, c.stream, in,
                       out, a);
    return launch_post("llap/curve");
  }
  const int vec = out.pitch % 4 == 0 && out.pstride % 4 == 0 && (uintptr_t)out.p % 8 == 0;
  hipLaunchKernelGGL(k_ll_curve, grid2d((c.conn[0].wd + 3) / 4, c.conn[0].ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), out, a, vec);
  return launch_post("llap/curve");
}

int launch_llap_reduce(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, c.dp, "llap/reduce inhi")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, c.dp, "llap/reduce outlo")) return e;
  const dimg in = to_dimg(c.conn[0]), out = to_dimg(c.conn[1]);
  // 16-byte (2), 8-byte (1) or 4-byte (3) input rows, and 8-byte output
  // groups (4), at every row and plane start
  const int vout = out.pitch % 4 == 0 && out.pstride % 4 == 0 && (uintptr_t)out.p % 8 == 0;
  const int vin = (in.pitch % 8 == 0 && in.pstride % 8 == 0 && (uintptr_t)in.p % 16 == 0) ? 2
                : (in.pitch % 4 == 0 && in.pstride % 4 == 0 && (uintptr_t)in.p % 8 == 0) ? 1
                : (in.pitch % 2 == 0 && in.pstride % 2 == 0 && (uintptr_t)in.p % 4 == 0) ? 3 : 0;
  const int vec = vin | (vout ? 4 : 0);
  hipLaunchKernelGGL(k_ll_reduce, grid2d((c.wd + 3) / 4, c.ht, c.dp), dim3(BX, BY), 0, c.stream, in, out, vec);
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
  if((uint64_t)c.wd * c.ht <= LL_SMALL)
  {
    hipLaunchKernelGGL(k_ll_assemble_small, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), a);
    return launch_post("llap/assemble");
  }
  const size_t lds = sizeof(float) * (a.num_gamma + 1) * AS_TN(AS_R);
  const dim3 grid((c.wd + BX - 1) / BX, (c.ht + BY * AS_R - 1) / (BY * AS_R), 1);
  hipLaunchKernelGGL(k_ll_assemble, grid, dim3(BX, BY), lds, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), a);
  return launch_post("llap/assemble");
}

int launch_llap_colour(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "llap/colour lum")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "llap/colour input")) return e;
  const int f32 = c.nconn > 2 && c.conn[2].format == VKDT_FMT_F32;
  if(int e = check_conn(c, 2, f32 ? VKDT_FMT_F32 : VKDT_FMT_F16, 4, 1, "llap/colour output")) return e;
  const vkdt_buf_t &l = c.conn[0], &o = c.conn[1], &q = c.conn[2];
  const bool pairs = l.wd == q.wd && l.ht == q.ht && o.wd == q.wd && o.ht == q.ht && c.wd == q.wd && c.ht == q.ht &&
                     l.pitch % 2 == 0 && o.pitch % 2 == 0 && q.pitch % 2 == 0 && (uintptr_t)l.ptr % 4 == 0 &&
                     (uintptr_t)o.ptr % 16 == 0 && (uintptr_t)q.ptr % 16 == 0;
  if(pairs)
    hipLaunchKernelGGL(f32 ? k_ll_colour2<1> : k_ll_colour2<0>, grid2d((c.wd + 1) / 2, c.ht, 1), dim3(BX, BY), 0,
                       c.stream, to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  else if(f32)
    hipLaunchKernelGGL(k_ll_colour<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  else
    hipLaunchKernelGGL(k_ll_colour<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("llap/colour");
}
