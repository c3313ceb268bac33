// k_colour.hip -- colour (matrix path) and filmcurv per-pixel nodes.
//   colour/main   src/pipe/modules/colour/main-impl.glsl:200-341
//   filmcurv/main src/pipe/modules/filmcurv/main.comp:68-167 (colour modes 0..5)
// Both read and write rgba f16 W x H: two texels a thread with one 16-byte
// load and store where the buffers allow, else one texel with 8 bytes.
#include "launch.h"
#include "ops.h"

__global__ void __launch_bounds__(256) k_co_main(dimg in, dimg out, co_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  st_h4(out, x, y, co_px(centre_h4(in, x, y), a));
}

// one instantiation for modes 1 and 3 (the default path) and one for the
// colour modes, so that the per-channel kernel keeps its registers
template<bool COLOUR>
__global__ void __launch_bounds__(256) k_fc_main(dimg in, dimg out, fc_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 t = fetch_h4(in, x, y);
  st_h4(out, x, y, COLOUR ? fc_px(t, a) : fc_px_basic(t, a));
}

// Two texels a thread, one 16-byte load and store of the pair, where rows
// start 16-byte aligned, the extents of input and output agree and the pair
// lies inside the row (`in` and `out` even pitch, aligned bases).  The same
// per-texel arithmetic as above.
DEV void ld_h4x2(const dimg &b, int x, int y, float4 &p, float4 &q)
{
  const uint4 u = *(const uint4 *)((const __half *)b.p + 4 * ((long)y * b.pitch + x));
  const __half2 a = *(const __half2 *)&u.x, c = *(const __half2 *)&u.y;
  const __half2 d = *(const __half2 *)&u.z, e = *(const __half2 *)&u.w;
  p = make_float4(__low2float(a), __high2float(a), __low2float(c), __high2float(c));
  q = make_float4(__low2float(d), __high2float(d), __low2float(e), __high2float(e));
}

DEV void st_h4x2(const dimg &b, int x, int y, float4 p, float4 q)
{
  const __half2 a = __floats2half2_rn(p.x, p.y), c = __floats2half2_rn(p.z, p.w);
  const __half2 d = __floats2half2_rn(q.x, q.y), e = __floats2half2_rn(q.z, q.w);
  uint4 u;
  u.x = *(const uint32_t *)&a;
  u.y = *(const uint32_t *)&c;
  u.z = *(const uint32_t *)&d;
  u.w = *(const uint32_t *)&e;
  *(uint4 *)((__half *)b.p + 4 * ((long)y * b.pitch + x)) = u;
}

// a 2x2 block of texels per thread, each row's pair in one 16-byte load and
// store: both rows' loads are in flight before either pair is converted
template<typename Px>
DEV void px_pair2(const dimg &in, const dimg &out, Px px)
{
  const int x = 2 * tid_x(), y = 2 * tid_y();
  if(x >= out.wd || y >= out.ht) return;
  if(x + 1 < out.wd && y + 1 < out.ht)
  {
    float4 p0, q0, p1, q1;
    ld_h4x2(in, x, y, p0, q0);
    ld_h4x2(in, x, y + 1, p1, q1);
    st_h4x2(out, x, y, px(p0), px(q0));
    st_h4x2(out, x, y + 1, px(p1), px(q1));
    return;
  }
  for(int r = 0; r < 2 && y + r < out.ht; r++)
  {
    if(x + 1 < out.wd)
    {
      float4 p, q;
      ld_h4x2(in, x, y + r, p, q);
      st_h4x2(out, x, y + r, px(p), px(q));
    }
    else st_h4(out, x, y + r, px(ld_h4(in, x, y + r)));
  }
}

static bool pairs_ok(const vkdt_buf_t &i, const vkdt_buf_t &o)
{
  return i.wd == o.wd && i.ht == o.ht && i.pitch % 2 == 0 && o.pitch % 2 == 0 && (uintptr_t)i.ptr % 16 == 0 &&
         (uintptr_t)o.ptr % 16 == 0;
}

__global__ void __launch_bounds__(256) k_co_main2(dimg in, dimg out, co_args a)
{
  px_pair2(in, out, [&](float4 t) { return co_px(t, a); });
}

template<bool COLOUR>
__global__ void __launch_bounds__(256) k_fc_main2(dimg in, dimg out, fc_args a)
{
  px_pair2(in, out, [&](float4 t) { return COLOUR ? fc_px(t, a) : fc_px_basic(t, a); });
}

// ---- host launchers --------------------------------------------------------

int launch_colour_main(const node_call &c)
{
  if(int e = check_blob(c, 0, 242 * 4, "colour/main")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "colour/main input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "colour/main output")) return e;
  const int32_t *pc = (const int32_t *)c.pc;
  const int have_clut = (pc && c.pc_size >= 12) ? pc[0] : 0;
  const int have_pick = (pc && c.pc_size >= 12) ? pc[1] : 0;
  const int have_abney = (pc && c.pc_size >= 12) ? pc[2] : 0;
  if(have_clut || have_pick || have_abney)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "colour/main: only the matrix path (no clut/pick/abney) is supported");
  co_args a;
  if(const char *err = co_args_init(a, c.params)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "colour/main: %s", err);
  if(pairs_ok(c.conn[0], c.conn[1]) && c.wd == c.conn[1].wd && c.ht == c.conn[1].ht)
    hipLaunchKernelGGL(k_co_main2, grid2d((c.wd + 1) / 2, (c.ht + 1) / 2, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  else
    hipLaunchKernelGGL(k_co_main, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("colour/main");
}

int launch_filmcurv_main(const node_call &c)
{
  if(int e = check_blob(c, 0, 4 * 4, "filmcurv/main")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "filmcurv/main input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "filmcurv/main output")) return e;
  fc_args a;
  if(const char *err = fc_args_init(a, c.params, c.params_size, false)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "filmcurv/main: %s", err);
  const bool basic = a.mode == 1 || a.mode == 3;
  if(pairs_ok(c.conn[0], c.conn[1]) && c.wd == c.conn[1].wd && c.ht == c.conn[1].ht)
    hipLaunchKernelGGL(basic ? k_fc_main2<false> : k_fc_main2<true>, grid2d((c.wd + 1) / 2, (c.ht + 1) / 2, 1), dim3(BX, BY),
                       0, c.stream, to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  else
    hipLaunchKernelGGL(basic ? k_fc_main<false> : k_fc_main<true>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("filmcurv/main");
}
