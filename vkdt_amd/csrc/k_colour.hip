// k_colour.hip -- colour (matrix path) and filmcurv per-pixel nodes.
//   colour/main   src/pipe/modules/colour/main-impl.glsl:200-341
//   filmcurv/main src/pipe/modules/filmcurv/main.comp:68-167 (colour modes 0..5)
// Both read and write rgba f16 W x H with one 8-byte load and store per texel.
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
  hipLaunchKernelGGL((a.mode == 1 || a.mode == 3) ? k_fc_main<false> : k_fc_main<true>, grid2d(c.wd, c.ht, 1),
                     dim3(BX, BY), 0, c.stream, to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("filmcurv/main");
}
