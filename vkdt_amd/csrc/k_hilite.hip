// k_hilite.hip -- the hilite module's nodes (highlight inpainting on the
// mosaic, src/pipe/modules/hilite/main.c:6-86):
//   half      mosaic W x H (f16 or ui16) -> rgba f16 W/b x H/b
//   reduce    rgba f16 level l -> level l+1 (ceil-half extent), x up to 14
//   assemble  fine + coarse rgba f16 -> rgba f16 at the fine extent
//   doub      mosaic + rgba f16 W/b x H/b -> mosaic W x H (f16 or ui16)
// Push constants: wb[4] (float bits) and filters; params: white, desat, soft.
#include "launch.h"
#include "ops_hilite.h"

template<int XT, int U16>
__global__ void __launch_bounds__(256) k_hl_half(dimg in, dimg out, float white)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  st_h4(out, x, y, hl_half_at<XT, U16>(in, x, y, white));
}

// reduce: a work group covers BX x BY coarse texels (hl_reduce_tile)
__global__ void __launch_bounds__(256) k_hl_reduce(dimg in, dimg out, hl_args a)
{
  __shared__ hl_lds s;
  const int2 blk = xcd_block();
  hl_reduce_tile(in, out, a, blk.x * BX, blk.y * BY, s);
}

__global__ void __launch_bounds__(256) k_hl_assemble(dimg fine, dimg coarse, dimg out, hl_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  st_h4(out, x, y, hl_assemble_at(fine, coarse, a, x, y));
}

template<int XT, int U16>
__global__ void __launch_bounds__(256) k_hl_doub(dimg in, dimg coarse, dimg out, int nx, int ny, float white)
{ // one thread per block of the mosaic (hl_doub_at)
  const int x = tid_x(), y = tid_y();
  if(x >= nx || y >= ny || x >= out.wd || y >= out.ht) return;
  hl_doub_at<XT, U16>(in, coarse, out, x, y, white);
}

// ---- host launchers --------------------------------------------------------

// decoded push constants + params (the params blob has no committed form)
static int decode(hl_args &a, uint32_t &filters, const node_call &c, const char *what)
{
  if(int e = check_blob(c, 5 * 4, 3 * 4, what)) return e;
  filters = ((const uint32_t *)c.pc)[4];
  hl_args_init(a, (const float *)c.pc, (const float *)c.params);
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
  const int u16 = in.fmt == DIMG_UI16;
  auto k = filters == 9 ? (u16 ? k_hl_half<1, 1> : k_hl_half<1, 0>) : (u16 ? k_hl_half<0, 1> : k_hl_half<0, 0>);
  hipLaunchKernelGGL(k, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream, in, out, a.white);
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
  const int u16 = in.fmt == DIMG_UI16;
  auto k = filters == 9 ? (u16 ? k_hl_doub<1, 1> : k_hl_doub<1, 0>) : (u16 ? k_hl_doub<0, 1> : k_hl_doub<0, 0>);
  hipLaunchKernelGGL(k, grid2d(nx, ny, 1), dim3(BX, BY), 0, c.stream, in, crs, out, nx, ny, a.white);
  return launch_post("hilite/doub");
}
