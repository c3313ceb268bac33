// k_colour.hip -- colour (matrix path) and filmcurv per-pixel nodes.
//   colour/main   src/pipe/modules/colour/main-impl.glsl:200-341
//   filmcurv/main src/pipe/modules/filmcurv/main.comp:153-252 (modes 1, 3)
// Both read and write rgba f16 W x H with one 8-byte load and store per texel.
#include "launch.h"

struct co_args
{
  float M[9];        // camera to rec2020, row major (identity for primaries 2)
  float toc[9];      // cat16 M * rec2020_to_xyz
  float fromc[9];    // xyz_to_rec2020 * cat16 M^-1
  float ratio[3];    // cl_dst / cl_src of the white-balance adaptation
  float expo;        // 2^exposure
};

DEV void mv3(float o[3], const float *M, float a, float b, float c)
{
  o[0] = M[0] * a + M[1] * b + M[2] * c;
  o[1] = M[3] * a + M[4] * b + M[5] * c;
  o[2] = M[6] * a + M[7] * b + M[8] * c;
}

__global__ void __launch_bounds__(256) k_co_main(dimg in, dimg out, co_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 t = centre_h4(in, x, y);
  float rgb[3], cl[3];
  mv3(rgb, a.M, t.x, t.y, t.z);
  mv3(cl, a.toc, rgb[0], rgb[1], rgb[2]);
  cl[0] *= a.ratio[0]; cl[1] *= a.ratio[1]; cl[2] *= a.ratio[2];
  mv3(rgb, a.fromc, cl[0], cl[1], cl[2]);
  st_h4(out, x, y, make_float4(clampf(rgb[0] * a.expo, -65535.0f, 65535.0f),
                               clampf(rgb[1] * a.expo, -65535.0f, 65535.0f),
                               clampf(rgb[2] * a.expo, -65535.0f, 65535.0f), 1.0f));
}

// ---- filmcurv -----------------------------------------------------------

struct fc_args { float il, k, bias; int mode; };

DEV float weibull(float x, float il, float k)
{ // filmcurv/params.glsl:16-20
  return 1.0f - fexp(-fpow(fmaxf(x, 1e-7f) * il, k));
}

// shared.glsl:371-387
DEV void adjust_colour_dng(float c0[3], float c1[3])
{
  float t;
  bool fx = false, fy = false, fz = false;
  if(c0[2] > c0[1]) { t = c0[2]; c0[2] = c0[1]; c0[1] = t; t = c1[2]; c1[2] = c1[1]; c1[1] = t; fx = true; }
  if(c0[1] > c0[0]) { t = c0[0]; c0[0] = c0[1]; c0[1] = t; t = c1[0]; c1[0] = c1[1]; c1[1] = t; fy = true; }
  if(c0[2] > c0[1]) { t = c0[2]; c0[2] = c0[1]; c0[1] = t; t = c1[2]; c1[2] = c1[1]; c1[1] = t; fz = true; }
  c1[1] = mixf(c1[2], c1[0], (c0[1] - c0[2] + 1e-6f) * frcp(c0[0] - c0[2] + 1e-6f));
  if(fz) { t = c1[2]; c1[2] = c1[1]; c1[1] = t; }
  if(fy) { t = c1[0]; c1[0] = c1[1]; c1[1] = t; }
  if(fx) { t = c1[2]; c1[2] = c1[1]; c1[1] = t; }
}

__global__ void __launch_bounds__(256) k_fc_main(dimg in, dimg out, fc_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 t = fetch_h4(in, x, y);
  float c0[3] = { t.x + a.bias, t.y + a.bias, t.z + a.bias };
  float c1[3] = { weibull(c0[0], a.il, a.k), weibull(c0[1], a.il, a.k), weibull(c0[2], a.il, a.k) };
  if(a.mode == 3) adjust_colour_dng(c0, c1);
  st_h4(out, x, y, make_float4(c1[0], c1[1], c1[2], 1.0f));
}

// ---- host launchers --------------------------------------------------------

static void matmul3(double o[9], const double a[9], const double b[9])
{
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
  {
    double s = 0;
    for(int k = 0; k < 3; k++) s += a[3 * j + k] * b[3 * k + i];
    o[3 * j + i] = s;
  }
}

int launch_colour_main(const node_call &c)
{
  if(int e = check_blob(c, 0, 242 * 4, "colour/main")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "colour/main input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "colour/main output")) return e;
  const float *p = (const float *)c.params;
  const uint32_t *u = (const uint32_t *)c.params;
  const int32_t *pc = (const int32_t *)c.pc;
  const int have_clut = (pc && c.pc_size >= 12) ? pc[0] : 0;
  const int have_pick = (pc && c.pc_size >= 12) ? pc[1] : 0;
  const int have_abney = (pc && c.pc_size >= 12) ? pc[2] : 0;
  // only the matrix path of the north-star graph is implemented; anything
  // else is refused rather than silently ignored
  if(have_clut || have_pick || have_abney || u[16] > 0 || p[226] != 1.0f || p[231] > 0.0f || u[225] != 0 || u[230] != 0)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "colour/main: only the matrix path (no clut/pick/rbf/sat/clip, linear trc) is supported");
  if(u[229] != 0 && u[229] != 2)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "colour/main: primaries %u not supported", u[229]);
  co_args a;
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++)
    a.M[3 * j + i] = u[229] == 0 ? p[4 + 4 * i + j] : (i == j ? 1.0f : 0.0f);
  static const double cat16[9]  = { 0.401288, 0.650173, -0.051461, -0.250268, 1.204414, 0.045854, -0.002079, 0.048952, 0.953127 };
  static const double cat16i[9] = { 1.86206786, -1.01125463, 0.14918677, 0.38752654, 0.62144744, -0.00897398, -0.01584150, -0.03412294, 1.04996444 };
  static const double r2x[9] = { 0.636958048301290991, 0.144616903586208406, 0.168880975164172054, 0.26270021201126692,
                                 0.677998071518871148, 0.0593017164698619384, 4.9999999999999999e-17, 0.0280726930490874452,
                                 1.06098505771079066 };
  static const double x2r[9] = { 1.71665119, -0.35567078, -0.25336628, -0.66668435, 1.61648124, 0.01576855, 0.01763986,
                                 -0.04277061, 0.94210312 };
  double A[9], B[9];
  matmul3(A, cat16, r2x);
  matmul3(B, x2r, cat16i);
  for(int k = 0; k < 9; k++) { a.toc[k] = (float)A[k]; a.fromc[k] = (float)B[k]; }
  // cat16 source white (1,1,1), destination white mul.rgb
  float src[3], dst[3];
  for(int j = 0; j < 3; j++)
  {
    src[j] = a.toc[3 * j] + a.toc[3 * j + 1] + a.toc[3 * j + 2];
    dst[j] = a.toc[3 * j] * p[0] + a.toc[3 * j + 1] * p[1] + a.toc[3 * j + 2] * p[2];
    a.ratio[j] = dst[j] / src[j];
  }
  a.expo = p[3];
  hipLaunchKernelGGL(k_co_main, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("colour/main");
}

int launch_filmcurv_main(const node_call &c)
{
  if(int e = check_blob(c, 0, 4 * 4, "filmcurv/main")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "filmcurv/main input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "filmcurv/main output")) return e;
  const float *p = (const float *)c.params;
  fc_args a;
  a.il = fmaxf(5e-3f, p[0]);
  a.k = fmaxf(1e-4f, p[1]);
  a.bias = p[2];
  a.mode = ((const int32_t *)c.params)[3];
  if(a.mode != 1 && a.mode != 3)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "filmcurv/main: colour mode %d not supported (1 and 3 are)", a.mode);
  hipLaunchKernelGGL(k_fc_main, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("filmcurv/main");
}
