// k_hilite.hip -- the hilite module's nodes (highlight inpainting on the
// mosaic, src/pipe/modules/hilite/main.c:6-86):
//   half      mosaic W x H (f16 or ui16) -> rgba f16 W/b x H/b
//   reduce    rgba f16 level l -> level l+1 (ceil-half extent), x up to 14
//   assemble  fine + coarse rgba f16 -> rgba f16 at the fine extent
//   doub      mosaic + rgba f16 W/b x H/b -> mosaic W x H (f16 or ui16)
// Push constants: wb[4] (float bits) and filters; params: white, desat, soft.
#include "launch.h"
#include "ops_hilite.h"

template<int XT>
__global__ void __launch_bounds__(256) k_hl_half(dimg in, dimg out, float white)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float4 r;
  if(XT) r = hl_half_px<1>([&](int px, int py) { return fetch_m(in, px, py); }, x, y, white);
  else   r = hl_half_px<0>([&](int px, int py) { return centre_m(in, px, py); }, x, y, white);
  st_h4(out, x, y, r);
}

// reduce: a work group covers BX x BY coarse texels.  Their fine
// footprint (2BX+3) x (2BY+3) is loaded once (mirrored addresses) and each
// fine texel prepared once in LDS -- the reference evaluates the
// desaturation of every fine texel for each of the up to 9 coarse texels
// whose 5x5 window covers it.
#define HLR_W (2 * BX + 3)
#define HLR_H (2 * BY + 3)
__global__ void __launch_bounds__(256) k_hl_reduce(dimg in, dimg out, hl_args a)
{
  __shared__ float s_r[HLR_H * HLR_W], s_g[HLR_H * HLR_W], s_b[HLR_H * HLR_W];
  __shared__ float s_l[HLR_H * HLR_W], s_v[HLR_H * HLR_W];
  const int2 blk = xcd_block();
  const int X0 = blk.x * BX, Y0 = blk.y * BY;
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < HLR_W * HLR_H; i += BX * BY)
  {
    const int lx = i % HLR_W, ly = i / HLR_W;
    const hl_tap t = hl_prep(centre_h4(in, 2 * X0 - 2 + lx, 2 * Y0 - 2 + ly), a);
    s_r[i] = t.r; s_g[i] = t.g; s_b[i] = t.b; s_l[i] = t.l; s_v[i] = t.valid;
  }
  __syncthreads();
  const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const int base = 2 * threadIdx.y * HLR_W + 2 * threadIdx.x;
  st_h4(out, x, y, hl_reduce_sum([&](int ii, int jj) {
    const int k = base + jj * HLR_W + ii;
    return hl_tap{ s_r[k], s_g[k], s_b[k], s_l[k], s_v[k] };
  }));
}

__global__ void __launch_bounds__(256) k_hl_assemble(dimg fine, dimg coarse, dimg out, hl_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float3 up = hl_expand([&](int px, int py) { return centre_h4(coarse, px, py); }, x, y);
  st_h4(out, x, y, hl_assemble_px(up, fetch_h4(fine, x, y), a));
}

// one mosaic site of the output: f16 or R16_UNORM store, dropped outside
DEV void hl_store(const dimg &o, int x, int y, float v)
{
  if(x >= o.wd || y >= o.ht) return;
  if(o.fmt == DIMG_UI16) st_u16(o, x, y, v);
  else st_h1(o, x, y, v);
}

template<int XT>
__global__ void __launch_bounds__(256) k_hl_doub(dimg in, dimg coarse, dimg out, int nx, int ny, float white)
{ // one thread per block of the mosaic.  doub.comp:24-25 bounds the block
  // index by the output's extent, not the node's: the threads of the last
  // 8x8 work groups past the node extent run too, reading zero coarse texels
  // (texelFetch) and writing the sites that fall inside the output
  const int x = tid_x(), y = tid_y();
  if(x >= nx || y >= ny || x >= out.wd || y >= out.ht) return;
  const float4 u = fetch_h4(coarse, x, y);
  if(XT)
  {
    const bool swap = ((x + y) & 1) == 0;
    const float3 up = swap ? make_float3(u.z, u.y, u.x) : make_float3(u.x, u.y, u.z);
    float c[9];
    hl_block3([&](int px, int py) { return fetch_m(in, px, py); }, x, y, c);
    hl_doub_xtrans(c, up, white);
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
      for(int j = 0; j < 3; j++) hl_store(out, 3 * x + i, 3 * y + j, c[3 * i + j]);
  }
  else
  {
    const float4 g = hl_doub_bayer(hl_gather2([&](int px, int py) { return centre_m(in, px, py); }, x, y),
                                   make_float3(u.x, u.y, u.z), white);
    hl_store(out, 2 * x, 2 * y, g.w);
    hl_store(out, 2 * x + 1, 2 * y, g.z);
    hl_store(out, 2 * x, 2 * y + 1, g.x);
    hl_store(out, 2 * x + 1, 2 * y + 1, g.y);
  }
}

// ---- host launchers --------------------------------------------------------

// decoded push constants + params (the params blob has no committed form)
static int decode(hl_args &a, uint32_t &filters, const node_call &c, const char *what)
{
  if(int e = check_blob(c, 5 * 4, 3 * 4, what)) return e;
  const float *pf = (const float *)c.pc;
  const float *par = (const float *)c.params;
  filters = ((const uint32_t *)c.pc)[4];
  float wb[3] = { pf[0], pf[1], pf[2] };
  if(!(wb[0] * wb[0] + wb[1] * wb[1] + wb[2] * wb[2] > 1e-3f)) wb[0] = wb[1] = wb[2] = 1.0f;
  for(int k = 0; k < 3; k++) a.wb[k] = wb[k];
  a.white = par[0];
  a.white_pos = par[0] > 0.0f ? par[0] : 1.0f;
  a.ds = par[1] * par[1];
  a.soft = par[2];
  return 0;
}

int launch_hilite_half(const node_call &c)
{
  hl_args a;
  uint32_t filters;
  if(int e = decode(a, filters, c, "hilite/half")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "hilite/half input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "hilite/half output")) return e;
  const dimg in = to_dimg(c.conn[0]), out = to_dimg(c.conn[1]);
  if(filters == 9)
    hipLaunchKernelGGL(k_hl_half<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream, in, out, a.white);
  else
    hipLaunchKernelGGL(k_hl_half<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream, in, out, a.white);
  return launch_post("hilite/half");
}

int launch_hilite_reduce(const node_call &c)
{
  hl_args a;
  uint32_t filters;
  if(int e = decode(a, filters, c, "hilite/reduce")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "hilite/reduce input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "hilite/reduce output")) return e;
  hipLaunchKernelGGL(k_hl_reduce, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("hilite/reduce");
}

int launch_hilite_assemble(const node_call &c)
{
  hl_args a;
  uint32_t filters;
  if(int e = decode(a, filters, c, "hilite/assemble")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "hilite/assemble fine")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "hilite/assemble coarse")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "hilite/assemble output")) return e;
  hipLaunchKernelGGL(k_hl_assemble, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), a);
  return launch_post("hilite/assemble");
}

int launch_hilite_doub(const node_call &c)
{
  hl_args a;
  uint32_t filters;
  if(int e = decode(a, filters, c, "hilite/doub")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "hilite/doub input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "hilite/doub coarse")) return e;
  if(int e = check_conn2(c, 2, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "hilite/doub output")) return e;
  // the reference dispatches ceil(wd/8) x ceil(ht/8) groups of 8x8
  const int nx = (int)((c.wd + 7) / 8 * 8), ny = (int)((c.ht + 7) / 8 * 8);
  const dimg in = to_dimg(c.conn[0]), crs = to_dimg(c.conn[1]), out = to_dimg(c.conn[2]);
  if(filters == 9)
    hipLaunchKernelGGL(k_hl_doub<1>, grid2d(nx, ny, 1), dim3(BX, BY), 0, c.stream, in, crs, out, nx, ny, a.white);
  else
    hipLaunchKernelGGL(k_hl_doub<0>, grid2d(nx, ny, 1), dim3(BX, BY), 0, c.stream, in, crs, out, nx, ny, a.white);
  return launch_post("hilite/doub");
}
