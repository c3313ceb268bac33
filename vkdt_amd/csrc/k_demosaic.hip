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
    for(int i = 0; i < 3; i++) for(int j = 0; j < 3; j++) s += fetch_h1(in, 3 * x + i, 3 * y + j);
    lum = s / 9.0f;
  }
  else
  {
    const float c0 = fetch_h1(in, 2 * x, 2 * y), c1 = fetch_h1(in, 2 * x, 2 * y + 1);
    const float c2 = fetch_h1(in, 2 * x + 1, 2 * y), c3 = fetch_h1(in, 2 * x + 1, 2 * y + 1);
    lum = (c0 + c1 + c2 + c3) * 0.25f;
  }
  st_h1(out, x, y, lum);
}

// ---- gauss (gauss.comp:19-126) ---------------------------------------------
// The covariance statistics decide between two tensors and feed a 2x2
// eigen solve whose discriminant vanishes on flat or clipped blocks: they
// keep IEEE division so that decision follows the written arithmetic.

template<int XT>
__global__ void __launch_bounds__(256) k_dm_gauss(dimg orig, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  constexpr int n0 = XT ? 0 : -1, n1 = 3, bs = XT ? 3 : 2;
  constexpr int NS = XT ? 5 : 8;
  float px[NS], rp[NS];
  int pi[NS], pj[NS];
  int n = 0;
#pragma unroll
  for(int j = n0; j < n1; j++)
#pragma unroll
    for(int i = n0; i < n1; i++)
    {
      if(XT ? (((j + i) & 1) == 1) : (((j + i) & 1) != 1)) continue;
      px[n] = centre_h1(orig, bs * x + i, bs * y + j);
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
    mb0 += (float)pi[k] / px[k]; mb1 += (float)pj[k] / px[k];
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
  st_h4(out, x, y, make_float4(ev[0], ev[1], e0[0], e0[1]));
}

// ---- splat (splat.comp:16-140) ---------------------------------------------

// gaussian weight of offset (i,j) under cov = (eval, major axis):
// exp(-0.5 |E o|^2_eval) clamped to [lo, 1], eval given as -0.5/eval
DEV float gauss_w(float4 cov, float r0, float r1, int i, int j, float lo)
{
  const float ofx = cov.z * (float)i + cov.w * (float)j;
  const float ofy = -cov.w * (float)i + cov.z * (float)j;
  return clampf(fexp(ofx * ofx * r0 + ofy * ofy * r1), lo, 1.0f);
}

template<int XT>
__global__ void __launch_bounds__(256) k_dm_splat(dimg in, dimg ga, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 cov = XT ? fetch_h4(ga, x / 3, y / 3) : fetch_h4(ga, (x + 1) / 2, (y + 1) / 2);
  const float r0 = -0.5f * frcp(clampf(cov.x, 0.01f, 25.0f)), r1 = -0.5f * frcp(clampf(cov.y, 0.01f, 25.0f));
  constexpr int r = XT ? 2 : 1;
  float g = 0.0f, w = 0.0f;
#pragma unroll
  for(int j = -r; j <= r; j++)
#pragma unroll
    for(int i = -r; i <= r; i++)
    {
      // only green sites contribute to the stored channel
      const bool green = XT ? xtrans_colour(x + i + 6, y + j + 6) == 1 : (((x + i + y + j) & 1) == 1);
      if(!green) continue;
      int px = x + i, py = y + j;
      if(px < 0) px += 6;
      if(py < 0) py += 6;
      if(px >= in.wd) px -= 6;
      if(py >= in.ht) py -= 6;
      const float col = fetch_h1(in, px, py);
      const float wt = (i == 0 && j == 0) ? 666.0f : gauss_w(cov, r0, r1, i, j, 1e-4f);
      g += col * wt;
      w += wt;
    }
  st_h1(out, x, y, g * frcp(fmaxf(1e-8f, w)));
}

// ---- fix (fix.comp:21-135) -------------------------------------------------

template<int XT>
__global__ void __launch_bounds__(256) k_dm_fix(dimg in, dimg gr, dimg cv, dimg out, int r)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  float sr = 0.0f, sb = 0.0f, wr = 0.0f, wb = 0.0f;
  const float gc = centre_h1(gr, x, y);
  float4 cov = XT ? fetch_h4(cv, (x + 1) / 3, (y + 1) / 3) : fetch_h4(cv, (x + 1) / 2, (y + 1) / 2);
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
      const float gh = centre_h1(gr, px, py);
      const float col = centre_h1(in, px, py);
      const float wt = gauss_w(cov, r0, r1, i, j, 1e-3f);
      const float v = col * gcn * frcp(1e-4f + gh) * wt;
      if(k == 0) { sr += v; wr += wt; }
      else       { sb += v; wb += wt; }
    }
  const float g = any_green ? gc : 0.0f;
  st_h4(out, x, y, make_float4(sr * frcp(fmaxf(1e-8f, wr)), g, sb * frcp(fmaxf(1e-8f, wb)), 1.0f));
}

// ---- host launchers --------------------------------------------------------

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
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "demosaic/down input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, 1, "demosaic/down output")) return e;
  hipLaunchKernelGGL(k_dm_down, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("demosaic/down");
}

int launch_demosaic_gauss(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/gauss")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, 1, "demosaic/gauss orig")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "demosaic/gauss output")) return e;
  if(a.filters == 9)
    hipLaunchKernelGGL(k_dm_gauss<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  else
    hipLaunchKernelGGL(k_dm_gauss<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("demosaic/gauss");
}

int launch_demosaic_splat(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/splat")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "demosaic/splat input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "demosaic/splat gauss")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 1, 1, "demosaic/splat output")) return e;
  // the +-6 border wrap of eval_gauss stays inside images of at least 6x6
  if(c.conn[0].wd < 6 || c.conn[0].ht < 6)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "demosaic/splat: input smaller than 6x6");
  if(a.filters == 9)
    hipLaunchKernelGGL(k_dm_splat<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  else
    hipLaunchKernelGGL(k_dm_splat<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]));
  return launch_post("demosaic/splat");
}

int launch_demosaic_fix(const node_call &c)
{
  dm_args a;
  if(int e = decode(a, c, "demosaic/fix")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 1, 1, "demosaic/fix input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 1, 1, "demosaic/fix green")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F16, 4, 1, "demosaic/fix cov")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_F16, 4, 1, "demosaic/fix output")) return e;
  int r;
  if(a.filters == 9) { r = a.fixup + 2; r = r < 2 ? 2 : (r > 3 ? 3 : r); }
  else               { r = a.fixup + 1; r = r < 1 ? 1 : (r > 2 ? 2 : r); }
  if(a.filters == 9)
    hipLaunchKernelGGL(k_dm_fix<1>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), r);
  else
    hipLaunchKernelGGL(k_dm_fix<0>, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                       to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), r);
  return launch_post("demosaic/fix");
}
