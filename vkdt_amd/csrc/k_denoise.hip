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
#include "ops.h"

// push-constant layouts (std140, as in the GLSL push_t blocks)
struct pc_rgbc { float wb[4], black[4], white[4]; int32_t crop[4]; };

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

static void read_params(dn_args &a, const node_call &c) { dn_params_init(a, (const float *)c.params); }

// ---- noop (noop.comp:36-57, strength <= 0) -------------------------------

__global__ void __launch_bounds__(256) k_dn_noop(dimg in, dimg out, dimg gm, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  // ui16 single channel reads as (v, 0, 0)
  const float v = fetch_u16(in, x + a.crop[0], y + a.crop[1]);
  float r = fmaxf(0.0f, (v - a.black[0]) / (a.white[0] - a.black[0]));
  float g = fmaxf(0.0f, (0.0f - a.black[1]) / (a.white[1] - a.black[1]));
  float b = fmaxf(0.0f, (0.0f - a.black[2]) / (a.white[2] - a.black[2]));
  if(a.gainmap)
  { // position of the texel centre in the input, the uncropped raw
    const float px = (0.5f + (float)x + (float)a.crop[0]) / (float)in.wd;
    const float py = (0.5f + (float)y + (float)a.crop[1]) / (float)in.ht;
    const float gain = dn_gain_at(gm, a, px, py, x, y);
    r *= gain; g *= gain; b *= gain;
  }
  if(out.chan == 1) { st_h1(out, x, y, r); return; }   // the mosaic's one channel
  st_h4(out, x, y, make_float4(r, g, b, 1.0f));
}

__global__ void __launch_bounds__(256) k_dn_half(dimg in, dimg out, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  st_h4(out, x, y, dn_half_px(in, a, x, y));
}

#define COV_R 2    // downcov: 64 x 8 texels per work group
#define DOWN_R 4   // down: 64 x 16 texels per work group

__global__ void __launch_bounds__(256) k_dn_downcov(dimg in, dimg out, dimg cov, dn_args a)
{
  __shared__ dn_cov_smem<COV_R> s;
  const int2 t = xcd_block();
  dn_downcov_tile<COV_R>(in, out, cov, a, t.x * BX, t.y * BY * COV_R, s);
}

__global__ void __launch_bounds__(256) k_dn_down(dimg in, dimg out, dn_args a)
{
  __shared__ float4 tile[DN_TN(DOWN_R)];
  const int2 t = xcd_block();
  dn_down_tile<DOWN_R>(in, out, a, t.x * BX, t.y * BY * DOWN_R, tile);
}

__global__ void __launch_bounds__(256) k_dn_assemble(dimg s0, dimg s1, dimg s2, dimg s3, dimg s4, dimg out,
                                                     dn_args a, asm_mats m)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  st_h4(out, x, y, dn_assemble_px(s0, s1, s2, s3, s4, a, m, x, y));
}

__global__ void __launch_bounds__(256) k_dn_doub(dimg in, dimg crs0, dimg crs1b, dimg out, dimg gm, dn_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float v = dn_doub_px(in, crs0, crs1b, out.wd, out.ht, a, x, y);
  if(a.gainmap)
  { // the block's centre on the half-resolution grid of the output
    const float px = (0.5f + (float)((x + a.crop[0]) / 2)) / (float)(out.wd / 2);
    const float py = (0.5f + (float)((y + a.crop[1]) / 2)) / (float)(out.ht / 2);
    v *= dn_gain_at(gm, a, px, py, x, y);
  }
  st_h1(out, x, y, v);
}

// Bayer mosaics: one 2x2 block per thread (dn_doub_quad)
__global__ void __launch_bounds__(256) k_dn_doub_bayer(dimg in, dimg crs0, dimg crs1b, dimg out, dimg gm, dn_args a)
{
  const int x = 2 * tid_x(), y = 2 * tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float v[4];
  dn_doub_quad(in, crs0, crs1b, out.wd, out.ht, a, x, y, v);
  if(a.gainmap)
#pragma unroll
    for(int k = 0; k < 4; k++)
    {
      const int xx = x + (k & 1), yy = y + (k >> 1);
      const float px = (0.5f + (float)((xx + a.crop[0]) / 2)) / (float)(out.wd / 2);
      const float py = (0.5f + (float)((yy + a.crop[1]) / 2)) / (float)(out.ht / 2);
      v[k] *= dn_gain_at(gm, a, px, py, xx, yy);
    }
  st_h1_quad(out, x, y, v);
}

// the gain map binding of noop (connector 2) and doub (connector 4): the
// map applies when the push constant and the module's `gainmap` parameter
// (uniform word 8) are both 1 on a Bayer mosaic (filters neither 0 nor 9);
// it is then an rgba f32 grid of at least one texel
static int read_gainmap(dn_args &a, const node_call &c, int conn, int gflag, const float *map_os, const char *what)
{
  const int param = ((const int32_t *)c.params)[8];
  a.gainmap = gflag == 1 && param == 1 && a.filters != 0 && a.filters != 9;
  for(int k = 0; k < 4; k++) a.map_os[k] = map_os[k];
  if(!a.gainmap) return 0;
  if(c.nconn <= conn) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: gain map binding missing", what);
  return check_conn(c, conn, VKDT_FMT_F32, 4, 1, what);
}

// ---- host launchers --------------------------------------------------------

int launch_denoise_noop(const node_call &c)
{
  if(int e = check_blob(c, 18 * 4, 9 * 4, "denoise/noop")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "denoise/noop input")) return e;
  const uint32_t och = c.nconn > 1 && c.conn[1].chan == 1 ? 1 : 4;   // the mosaic channel, or rgba
  if(int e = check_conn(c, 1, VKDT_FMT_F16, och, 1, "denoise/noop output")) return e;
  const int32_t *pi = (const int32_t *)c.pc;
  const float *pf = (const float *)c.pc;
  dn_args a = {};
  for(int k = 0; k < 4; k++) { a.crop[k] = pi[k]; a.black[k] = pf[4 + k]; a.white[k] = pf[8 + k]; }
  a.filters = (uint32_t)pi[16];
  read_params(a, c);
  if(int e = read_gainmap(a, c, 2, pi[17], pf + 12, "denoise/noop gainmap")) return e;
  hipLaunchKernelGGL(k_dn_noop, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a.gainmap ? to_dimg(c.conn[2]) : dimg{}, a);
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
  dn_level_scale(a);
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
  hipLaunchKernelGGL(k_dn_downcov, grid2d(c.wd, (c.ht + COV_R - 1) / COV_R, 1), dim3(BX, BY), 0, c.stream,
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
  hipLaunchKernelGGL(k_dn_down, grid2d(c.wd, (c.ht + DOWN_R - 1) / DOWN_R, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("denoise/down");
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
  dn_asm_scales(a);
  asm_mats m;
  asm_mats_init(m);
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
  read_params(a, c);
  if(int e = read_gainmap(a, c, 4, pi[19], pf + 20, "denoise/doub gainmap")) return e;
  if(a.filters == 9)
    hipLaunchKernelGGL(k_dn_doub, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]),
                       a.gainmap ? to_dimg(c.conn[4]) : dimg{}, a);
  else
    hipLaunchKernelGGL(k_dn_doub_bayer, grid2d((c.wd + 1) / 2, (c.ht + 1) / 2, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]),
                       a.gainmap ? to_dimg(c.conn[4]) : dimg{}, a);
  return launch_post("denoise/doub");
}
