// k_demosaic.hip -- the demosaic module's gaussian-splat nodes
// (src/pipe/modules/demosaic): down, gauss, splat, fix.
//
// Inputs: the denoised mosaic, one f16 channel W x H.  gauss writes an rgba
// f16 covariance (eigenvalues, major axis) per CFA block; splat writes the
// interpolated green as one f16 channel; fix writes rgba f16 W x H.
//
// splat only stores the green sum of its splats and fix only weights the
// red and blue sites (green sites copy the centre green), so both kernels
// evaluate gaussian weights for exactly the taps whose sums reach the
// output: the same sums in the same order as the reference's full loops.
#include "launch.h"
#include "ops.h"

struct dm_args
{
  uint32_t filters;
  int      fixup;
};

// ---- down (down.comp:18-47; never sampled by gauss) ----------------------

__global__ void __launch_bounds__(256) k_dm_down(dimg in, dimg out, dm_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float lum;
  if(a.filters == 9)
  {
    float s = 0.0f;
    // texelFetch order of the reference: (0,0) (0,1) (0,2) (1,0) ...
    for(int i = 0; i < 3; i++) for(int j = 0; j < 3; j++) s += fetch_m(in, 3 * x + i, 3 * y + j);
    lum = s / 9.0f;
  }
  else
  {
    const float c0 = fetch_m(in, 2 * x, 2 * y), c1 = fetch_m(in, 2 * x, 2 * y + 1);
    const float c2 = fetch_m(in, 2 * x + 1, 2 * y), c3 = fetch_m(in, 2 * x + 1, 2 * y + 1);
    lum = (c0 + c1 + c2 + c3) * 0.25f;
  }
  st_h1(out, x, y, lum);
}

// Waves whose every tap lies inside the image read without the mirrored
// addressing (the identity there): the same values, fewer instructions.
template<int XT, int U16>
__global__ void __launch_bounds__(256) k_dm_gauss(dimg orig, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  if(__all(dm_gauss_inside<XT>(orig, x, y)))
    st_h4(out, x, y, dm_gauss_at<XT>([&](int px, int py) { return ld_mt<U16>(orig, px, py); }, x, y));
  else
    st_h4(out, x, y, dm_gauss_px<XT, U16>(orig, x, y));
}

// texel of this thread: Bayer tiles are handed out colour-uniform per wave
template<int XT>
DEV int2 cfa_xy()
{
  const int2 t = xcd_block();
  if(XT) return make_int2(t.x * BX + threadIdx.x, t.y * BY + threadIdx.y);
  return bayer_wave_xy(t.x * BX, t.y * BY);
}

template<int XT, int U16, int PAR>
DEV float dm_splat_node(const dimg &in, const dimg &ga, int x, int y)
{
  constexpr int R = XT ? 2 : 1;
  if(__all(x >= R && y >= R && x + R < in.wd && y + R < in.ht))
    return dm_splat_at<XT, PAR>([&](int px, int py) { return ld_mt<U16>(in, px, py); }, in.wd, in.ht, ga, x, y);
  return dm_splat_px<XT, U16, PAR>(in, ga, x, y);
}

// Bayer waves are colour-uniform (cfa_xy), so the branch on the site's
// parity is uniform too, and each wave weights only its colour's taps
template<int XT, int U16>
__global__ void __launch_bounds__(256) k_dm_splat(dimg in, dimg ga, dimg out)
{
  const int2 p = cfa_xy<XT>();
  const int x = p.x, y = p.y;
  if(x >= out.wd || y >= out.ht) return;
  float v;
  if(XT) v = dm_splat_node<XT, U16, -1>(in, ga, x, y);
  else if((x + y) & 1) v = dm_splat_node<XT, U16, 0>(in, ga, x, y);
  else v = dm_splat_node<XT, U16, 1>(in, ga, x, y);
  st_h1(out, x, y, v);
}

template<int XT, int U16, int PAR>
DEV float4 dm_fix_node(const dimg &in, const dimg &gr, const dimg &cv, int r, int x, int y)
{
  if(__all(x >= r && y >= r && x + r < min(in.wd, gr.wd) && y + r < min(in.ht, gr.ht)))
    return dm_fix_r<XT, PAR>([&](int px, int py) { return ld_mt<U16>(in, px, py); }, cv, r, x, y,
                             [&](int px, int py) { return ld_h1(gr, px, py); });
  return dm_fix_px<XT, U16, PAR>(in, cv, r, x, y, [&](int px, int py) { return centre_h1(gr, px, py); });
}

template<int XT, int U16>
__global__ void __launch_bounds__(256) k_dm_fix(dimg in, dimg gr, dimg cv, dimg out, int r)
{
  const int2 p = cfa_xy<XT>();
  const int x = p.x, y = p.y;
  if(x >= out.wd || y >= out.ht) return;
  float4 v;
  if(XT) v = dm_fix_node<XT, U16, -1>(in, gr, cv, r, x, y);
  else if((x + y) & 1) v = dm_fix_node<XT, U16, 1>(in, gr, cv, r, x, y);
  else v = dm_fix_node<XT, U16, 0>(in, gr, cv, r, x, y);
  st_h4(out, x, y, v);
}

// ---- halfsize (halfsize.comp:19-50): one rgba texel per CFA block --------

template<int XT>
__global__ void __launch_bounds__(256) k_dm_halfsize(dimg in, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float4 o;
  if(XT)
  { // c[3i+j] = texel (3x+i, 3y+j), texelFetch (zero outside)
    float c[9];
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
      for(int j = 0; j < 3; j++) c[3 * i + j] = fetch_m(in, 3 * x + i, 3 * y + j);
    o.y = (c[0] + c[2] + c[4] + c[6] + c[8]) * 1.0f / 5.0f;
    const float p = (c[1] + c[7]) * 0.5f, q = (c[3] + c[5]) * 0.5f;
    const bool odd = ((x + y) & 1) != 0;
    o.x = odd ? p : q;
    o.z = odd ? q : p;
    o.w = 1.0f;   // left unset by the reference (undefined alpha)
  }
  else
  { // textureGather at the block centre: r = (0,0), g = mean of (0,1),(1,0), b = (1,1)
    const float gx = centre_m(in, 2 * x, 2 * y + 1), gy = centre_m(in, 2 * x + 1, 2 * y + 1);
    const float gz = centre_m(in, 2 * x + 1, 2 * y), gw = centre_m(in, 2 * x, 2 * y);
    o = make_float4(gw, (gx + gz) / 2.0f, gy, 1.0f);
  }
  st_h4(out, x, y, o);
}

// ---- host launchers --------------------------------------------------------

// the <XT, U16> instantiation of a demosaic kernel: CFA and mosaic format
#define DM_PICK(kern, xt, u16) \
  ((xt) ? ((u16) ? kern<1, 1> : kern<1, 0>) : ((u16) ? kern<0, 1> : kern<0, 0>))

static int decode(dm_args &a, const node_call &c, const char *what)
{
  if(int e = check_blob(c, 5 * 4, 0, what)) return e;
  a.filters = ((const uint32_t *)c.pc)[4];
  a.fixup = (c.params && c.params_size >= 4) ? ((const int32_t *)c.params)[0] : 0;
  return 0;
}

int launch_demosaic_down(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/down")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/down input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, 1, "demosaic/down output")) return e;
  hipLaunchKernelGGL(k_dm_down, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("demosaic/down");
}

int launch_demosaic_gauss(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/gauss")) return e;
  if(int e = check_conn2(c, 1, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/gauss orig")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "demosaic/gauss output")) return e;
  const int u16 = c.conn[1].format == VKDT_FMT_UI16;
  hipLaunchKernelGGL(DM_PICK(k_dm_gauss, a.filters == 9, u16), grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("demosaic/gauss");
}

int launch_demosaic_splat(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/splat")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/splat input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "demosaic/splat gauss")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 1, 1, "demosaic/splat output")) return e;
  // the +-6 border wrap of eval_gauss stays inside images of at least 6x6
  if(c.conn[0].wd < 6 || c.conn[0].ht < 6)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "demosaic/splat: input smaller than 6x6");
  const int u16 = c.conn[0].format == VKDT_FMT_UI16;
  hipLaunchKernelGGL(DM_PICK(k_dm_splat, a.filters == 9, u16), grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("demosaic/splat");
}

int launch_demosaic_fix(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/fix")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/fix input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, 1, "demosaic/fix green")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "demosaic/fix cov")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_F16, 4, 1, "demosaic/fix output")) return e;
  int r;
  if(a.filters == 9) { r = a.fixup + 2; r = r < 2 ? 2 : (r > 3 ? 3 : r); }
  else               { r = a.fixup + 1; r = r < 1 ? 1 : (r > 2 ? 2 : r); }
  const int u16 = c.conn[0].format == VKDT_FMT_UI16;
  hipLaunchKernelGGL(DM_PICK(k_dm_fix, a.filters == 9, u16), grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), r);
  return launch_post("demosaic/fix");
}

int launch_demosaic_halfsize(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/halfsize")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/halfsize input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "demosaic/halfsize output")) return e;
  if(a.filters == 9)
    hipLaunchKernelGGL(k_dm_halfsize<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]));
  else
    hipLaunchKernelGGL(k_dm_halfsize<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]));
  return launch_post("demosaic/halfsize");
}
