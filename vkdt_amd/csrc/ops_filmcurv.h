// ops_filmcurv.h -- filmcurv/main (src/pipe/modules/filmcurv/main.comp:68-167):
// the weibull tone curve per channel (mode 1), the hsl hack (3, with
// adjust_colour_dng from ops.h), and the colour modes darktable UCS (0),
// Munsell hue constancy (2), AgX (4) and the oklab lum ratio (5), with the
// colour spaces they use (src/pipe/modules/colourspaces.glsl,
// shared/dtucs.glsl, shared/munsell.glsl).
//
// GLSL matrices are column-major; the row-major tables here list the same
// coefficients by rows, so that o[i] = sum_j M[3i+j] v[j] (mv3).  The
// colour-space conversions use the accurate library functions (powf,
// atan2f, cosf, sinf); the tone curve itself uses the hardware exp2/log2
// as in the per-channel mode.  Included by ops.h after mv3, mixf, clampf,
// smoothstepf and adjust_colour_dng.
#pragma once
#include "common.h"
#include "munsell.h"

#define FC_PI 3.14159265358979323846f

struct fc_args
{
  float il, k, bias;
  int   mode;
  float chroma, rolloff, red, yellow, blue, shadows;   // lum ratio (mode 5)
};

// filmcurv/params: light, contrast, bias, colour, chroma, rolloff, red,
// yellow, blue, shadows
DEV float weibull(float x, float il, float k)
{ // filmcurv/params.glsl:16-20
  return 1.0f - fexp(-fpow(fmaxf(x, 1e-7f) * il, k));
}

DEV float weibull_pdf(float x, float il, float k)
{ // filmcurv/params.glsl:25-29
  x = fmaxf(x, 1e-7f);
  return k * il * powf(x * il, k - 1.0f) * expf(-powf(x * il, k));
}

DEV float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
DEV float glsl_fract(float x) { return x - floorf(x); }

// matrices.h: rec2020 <-> XYZ (row-major as makemat lists them)
__constant__ static const float fc_r2x[9] = { 0.636958048301290991f, 0.144616903586208406f, 0.168880975164172054f,
                                              0.26270021201126692f, 0.677998071518871148f, 0.0593017164698619384f,
                                              4.9999999999999999e-17f, 0.0280726930490874452f, 1.06098505771079066f };
__constant__ static const float fc_x2r[9] = { 1.71665119f, -0.35567078f, -0.25336628f, -0.66668435f, 1.61648124f,
                                              0.01576855f, 0.01763986f, -0.04277061f, 0.94210312f };

// ---- oklab (colourspaces.glsl:20-50) -----------------------------------------

__constant__ static const float ok_m1[9] = { 0.61668844f, 0.36015907f, 0.02304329f, 0.2651402f, 0.63585648f,
                                             0.09903023f, 0.10015065f, 0.20400432f, 0.69632468f };
__constant__ static const float ok_m2[9] = { 0.21045426f, 0.79361779f, -0.00407205f, 1.9779985f, -2.42859221f,
                                             0.45059371f, 0.02590404f, 0.78277177f, -0.80867577f };
__constant__ static const float ok_m2i[9] = { 1.0f, 0.39633779f, 0.21580376f, 1.00000001f, -0.10556134f, -0.06385417f,
                                              1.00000005f, -0.08948418f, -1.29148554f };
__constant__ static const float ok_m1i[9] = { 2.14014041f, -1.24635595f, 0.10643173f, -0.88483245f, 2.16317272f,
                                              -0.27836159f, -0.04857906f, -0.45449091f, 1.50235629f };

DEV void rec2020_to_oklab(float lab[3], const float c[3])
{
  float lms[3];
  mv3(lms, ok_m1, c[0], c[1], c[2]);
  for(int i = 0; i < 3; i++) lms[i] = powf(fmaxf(0.0f, lms[i]), 1.0f / 3.0f);
  mv3(lab, ok_m2, lms[0], lms[1], lms[2]);
}

DEV void oklab_to_rec2020(float c[3], const float lab[3])
{
  float lms[3];
  mv3(lms, ok_m2i, lab[0], lab[1], lab[2]);
  for(int i = 0; i < 3; i++) lms[i] = lms[i] * lms[i] * lms[i];
  mv3(c, ok_m1i, lms[0], lms[1], lms[2]);
}

// "hsv" as oklab hue/chroma/lightness (colourspaces.glsl:64-76)
DEV void rgb2hsv(float h[3], const float c[3])
{
  float lab[3];
  rec2020_to_oklab(lab, c);
  h[0] = glsl_fract(1.0f + atan2f(lab[2], lab[1]) / (2.0f * FC_PI));
  h[1] = sqrtf(lab[1] * lab[1] + lab[2] * lab[2]);
  h[2] = lab[0];
}

DEV void hsv2rgb(float c[3], const float h[3])
{
  const float lab[3] = { h[2], h[1] * cosf(2.0f * FC_PI * h[0]), h[1] * sinf(2.0f * FC_PI * h[0]) };
  if(lab[0] <= 0.0f) { c[0] = c[1] = c[2] = 0.0f; return; }
  oklab_to_rec2020(c, lab);
}

DEV float lerp_chromaticity_angle(float h1, float h2, float t)
{ // filmcurv/main.comp:17-26
  const float delta = h2 - h1;
  if(delta > 0.5f) h2 -= 1.0f;
  else if(delta < -0.5f) h2 += 1.0f;
  return glsl_mod(h1 + t * (h2 - h1), 1.0f);
}

// ---- darktable UCS 22 (shared/dtucs.glsl:11-78) ----------------------------

DEV float Y_to_dt_UCS_L_star(float Y)
{
  const float yh = powf(Y, 0.631651345306265f);
  return 2.098883786377f * yh / (yh + 1.12426773749357f);
}

DEV float dt_UCS_L_star_to_Y(float L)
{
  return powf(1.12426773749357f * L / (2.098883786377f - L), 1.5831518565279648f);
}

DEV void xy_to_dt_UCS_UV(float uv[2], float x, float y)
{
  // M1 * (x, y, 1), M1 given by columns in the reference
  const float u = -0.783941002840055f * x + 0.277512987809202f * y + 0.153836578598858f;
  const float v = 0.745273540913283f * x - 0.205375866083878f * y - 0.165478376301988f;
  const float d = 0.318707282433486f * x + 2.16743692732158f * y + 0.291320554395942f;
  const float us = u / d, vs = v / d;
  const float U = 1.39656225667f * us / (fabsf(us) + 1.49217352929f);
  const float V = 1.4513954287f * vs / (fabsf(vs) + 1.52488637914f);
  uv[0] = -1.124983854323892f * U - 0.980483721769325f * V;
  uv[1] = 1.86323315098672f * U + 1.971853092390862f * V;
}

DEV void xyY_to_dt_UCS_JCH(float jch[3], const float xyY[3], float L_white)
{
  float uv[2];
  xy_to_dt_UCS_UV(uv, xyY[0], xyY[1]);
  const float L = Y_to_dt_UCS_L_star(xyY[2]);
  const float M2 = uv[0] * uv[0] + uv[1] * uv[1];
  jch[0] = L / L_white;
  jch[1] = 15.932993652962535f * powf(L, 0.6523997524738018f) * powf(M2, 0.6007557017508491f) / L_white;
  jch[2] = atan2f(uv[1], uv[0]);
}

DEV void dt_UCS_JCH_to_xyY(float xyY[3], const float jch[3], float L_white)
{
  const float L = jch[0] * L_white;
  float M = powf(jch[1] * L_white / (15.932993652962535f * powf(L, 0.6523997524738018f)), 0.8322850678616855f);
  M = clampf(M, 0.0f, 0.05f);
  const float us0 = M * cosf(jch[2]), vs0 = M * sinf(jch[2]);
  const float us = -5.037522385190711f * us0 - 2.504856328185843f * vs0;
  const float vs = 4.760029407436461f * us0 + 2.874012963239247f * vs0;
  const float U = -1.49217352929f * us / (fabsf(us) - 1.39656225667f);
  const float V = -1.52488637914f * vs / (fabsf(vs) - 1.4513954287f);
  const float xd = 0.167171472114775f * U + 0.141299802443708f * V - 0.00801531300850582f;
  const float yd = -0.150959086409163f * U - 0.155185060382272f * V - 0.00843312433578007f;
  const float zd = 0.940254742367256f * U + 1.0f * V - 0.0256325967652889f;
  xyY[0] = xd / zd;
  xyY[1] = yd / zd;
  xyY[2] = dt_UCS_L_star_to_Y(L);
}

// ---- Munsell chart (shared/munsell.glsl:8-137) ------------------------------

__constant__ static const uint32_t fc_munsell[MUNSELL_HDIM * MUNSELL_CDIM] = { MUNSELL_XY_TABLE };

DEV float2 munsell_lookup(int h, int c)
{
  h = (h % MUNSELL_HDIM + MUNSELL_HDIM) % MUNSELL_HDIM;
  c = c < 0 ? 0 : (c > MUNSELL_CDIM - 1 ? MUNSELL_CDIM - 1 : c);
  const uint32_t u = fc_munsell[MUNSELL_CDIM * h + c];
  const __half lo = __ushort_as_half((unsigned short)(u & 0xffffu)), hi = __ushort_as_half((unsigned short)(u >> 16));
  return make_float2(__half2float(lo), __half2float(hi));
}

DEV float munsell_hue_angle(float x, float y)
{ // angle around the white, growing with the chart's hue index
  return glsl_mod(2.0f * FC_PI - 2.52f - atan2f(y - 0.32902f, x - 0.31271f), 2.0f * FC_PI);
}

DEV float munsell_side(float2 v0, float2 v1, float2 p)
{
  v1.x -= v0.x; v1.y -= v0.y;
  p.x -= v0.x; p.y -= v0.y;
  return v1.x * p.y - v1.y * p.x;
}

DEV float2 munsell_to_xy(float2 mhc)
{
  const float hm = mhc.x * MUNSELL_HDIM, cm = fmaxf(mhc.y, 0.0f) * MUNSELL_CDIM;
  const int hi = (int)hm, ci = (int)cm;
  const float hu = hm - hi, cu = cm - ci;
  const float2 r3 = munsell_lookup(hi, ci + 1), r2 = munsell_lookup(hi + 1, ci + 1);
  const float2 r0 = munsell_lookup(hi, ci), r1 = munsell_lookup(hi + 1, ci);
  if(hu >= cu)
    return make_float2((1.0f - hu) * r0.x + (hu - cu) * r1.x + cu * r2.x, (1.0f - hu) * r0.y + (hu - cu) * r1.y + cu * r2.y);
  return make_float2(hu * r2.x + (cu - hu) * r3.x + (1.0f - cu) * r0.x, hu * r2.y + (cu - hu) * r3.y + (1.0f - cu) * r0.y);
}

DEV float2 munsell_from_xy(float2 xy)
{
  int hm = 0, hM = MUNSELL_HDIM, cm = 0, cM = MUNSELL_CDIM - 1;
  const float theta = munsell_hue_angle(xy.x, xy.y);
  const float dx = xy.x - 0.31271f, dy = xy.y - 0.32902f;
  const float rad2 = dx * dx + dy * dy;
  for(int i = 0; i < 10; i++)
  { // bisect hue and chroma towards the cell holding xy
    const int h = (hm + hM) / 2, c = (cm + cM) / 2;
    const float2 r = munsell_lookup(h, c);
    const float th = munsell_hue_angle(r.x, r.y);
    const float ex = r.x - 0.31271f, ey = r.y - 0.32902f;
    const float r2 = ex * ex + ey * ey;
    if(th <= theta) hm = h;
    else            hM = h;
    if(r2 <= rad2) cm = c;
    else           cM = c;
    if(hM <= hm + 1 && cM <= cm + 1) break;
  }
  for(int i = 0; i < 10; i++)
  { // walk the grid to the quad that contains xy
    const float2 r3 = munsell_lookup(hm, cm + 1), r2 = munsell_lookup(hm + 1, cm + 1);
    const float2 r0 = munsell_lookup(hm, cm), r1 = munsell_lookup(hm + 1, cm);
    const float s0 = munsell_side(r0, r1, xy), s1 = munsell_side(r1, r2, xy);
    const float s2 = munsell_side(r2, r3, xy), s3 = munsell_side(r3, r0, xy);
    if(s0 < 0 && cm > 0) cm--;
    else if(s0 < 0 && cm == 0) hm = ((hm + MUNSELL_HDIM / 2) % MUNSELL_HDIM + MUNSELL_HDIM) % MUNSELL_HDIM;
    else if(s2 < 0 && cm < MUNSELL_CDIM - 2) cm++;
    if(s1 < 0) hm++;
    else if(s3 < 0) hm--;
    if(s0 >= 0 && s1 >= 0 && s3 >= 0 && (s2 >= 0 || cm >= MUNSELL_CDIM - 2))
    { // barycentric coordinates in one of the quad's two triangles
      const float t0 = munsell_side(r0, r1, r2), t1 = munsell_side(r2, r3, r0);
      float u0, u1, u2, u3;
      if(cm > 0 && s0 + s1 <= t0) { u2 = s0 / t0; u0 = s1 / t0; u1 = 1.0f - u0 - u2; u3 = 0.0f; }
      else                        { u2 = s3 / t1; u0 = s2 / t1; u3 = 1.0f - u0 - u2; u1 = 0.0f; }
      const float h = u0 * hm + u1 * (hm + 1.0f) + u2 * (hm + 1.0f) + u3 * hm;
      const float c = u0 * cm + u1 * cm + u2 * (cm + 1.0f) + u3 * (cm + 1.0f);
      return make_float2(h / MUNSELL_HDIM, fmaxf(0.0f, c / MUNSELL_CDIM));
    }
  }
  return make_float2(1.0f, 1.0f);
}

// ---- the modes ----------------------------------------------------------------

DEV void fc_weibull3(float o[3], const float c[3], const fc_args &a)
{
  for(int i = 0; i < 3; i++) o[i] = weibull(c[i], a.il, a.k);
}

DEV void fc_mode_ucs(float c1[3], const float c0[3], const fc_args &a)
{ // main.comp:72-89
  fc_weibull3(c1, c0, a);
  float xyz0[3], xyz1[3];
  mv3(xyz0, fc_r2x, c0[0], c0[1], c0[2]);
  mv3(xyz1, fc_r2x, c1[0], c1[1], c1[2]);
  const float n0 = fmaxf(1e-4f, xyz0[0] + xyz0[1] + xyz0[2]), n1 = fmaxf(1e-4f, xyz1[0] + xyz1[1] + xyz1[2]);
  const float xyY0[3] = { xyz0[0] / n0, xyz0[1] / n0, xyz0[1] };
  float xyY1[3] = { xyz1[0] / n1, xyz1[1] / n1, xyz0[1] };
  float jch0[3], jch1[3];
  xyY_to_dt_UCS_JCH(jch0, xyY0, 1.0f);
  xyY_to_dt_UCS_JCH(jch1, xyY1, 1.0f);
  jch1[2] = jch0[2];
  dt_UCS_JCH_to_xyY(xyY1, jch1, 1.0f);
  const float s = xyz1[1] / fmaxf(1e-4f, xyY1[1]);
  const float x = xyY1[0] * s, y = xyY1[1] * s, z = (1.0f - xyY1[0] - xyY1[1]) * s;
  mv3(c1, fc_x2r, x, y, z);
}

DEV void rec2020_to_xyY(float o[3], const float c[3])
{ // colourspaces.glsl:2-7
  float xyz[3];
  mv3(xyz, fc_r2x, c[0], c[1], c[2]);
  const float n = xyz[0] + xyz[1] + xyz[2];
  o[0] = xyz[0] / n;
  o[1] = xyz[1] / n;
  o[2] = xyz[1];
}

DEV void xyY_to_rec2020(float c[3], const float xyY[3])
{ // colourspaces.glsl:15-19
  const float x = xyY[0] * xyY[2] / xyY[1], y = xyY[1] * xyY[2] / xyY[1], z = (1.0f - xyY[0] - xyY[1]) * xyY[2] / xyY[1];
  mv3(c, fc_x2r, x, y, z);
}

DEV void fc_mode_munsell(float c1[3], const float c0[3], const fc_args &a)
{ // main.comp:90-99
  fc_weibull3(c1, c0, a);
  float xyY0[3], xyY1[3];
  rec2020_to_xyY(xyY0, c0);
  const float2 m0 = munsell_from_xy(make_float2(xyY0[0], xyY0[1]));
  rec2020_to_xyY(xyY1, c1);
  const float2 m1 = munsell_from_xy(make_float2(xyY1[0], xyY1[1]));
  const float2 xy = munsell_to_xy(make_float2(m0.x, m1.y));
  xyY1[0] = xy.x;
  xyY1[1] = xy.y;
  xyY_to_rec2020(c1, xyY1);
}

__constant__ static const float agx_m[9] = { 0.856627153315983f, 0.0951212405381588f, 0.0482516061458583f,
                                             0.137318972929847f, 0.761241990602591f, 0.101439036467562f,
                                             0.11189821299995f, 0.0767994186031903f, 0.811302368396859f };
__constant__ static const float agx_mi[9] = { 1.1271005818144368f, -0.11060664309660323f, -0.016493938717834573f,
                                              -0.1413297634984383f, 1.157823702216272f, -0.016493938717834257f,
                                              -0.14132976349843826f, -0.11060664309660294f, 1.2519364065950405f };

DEV void fc_mode_agx(float c1[3], const float c0[3], const fc_args &a)
{ // main.comp:28-55
  float c[3], h0[3], h1[3], w[3];
  mv3(c, agx_m, c0[0], c0[1], c0[2]);
  rgb2hsv(h0, c);
  fc_weibull3(w, c, a);
  rgb2hsv(h1, w);
  h1[0] = lerp_chromaticity_angle(h0[0], h1[0], 0.4f);
  hsv2rgb(c, h1);
  mv3(c1, agx_mi, c[0], c[1], c[2]);
}

DEV float hue_bump(float h, float h0, float w)
{ // main.comp:58-63
  const float d = fabsf(glsl_mod(h - h0 + FC_PI, 2.0f * FC_PI) - FC_PI);
  return d < w ? 0.5f + 0.5f * cosf(FC_PI * d / w) : 0.0f;
}

DEV void fc_mode_lum_ratio(float c1[3], const float c0[3], const fc_args &a)
{ // main.comp:109-170
  float lab0[3];
  rec2020_to_oklab(lab0, c0);
  const float L0 = fmaxf(lab0[0], 1e-7f);
  const float lum0 = fmaxf(0.2627f * c0[0] + 0.6780f * c0[1] + 0.0593f * c0[2], 1e-7f);
  float lum1 = weibull(lum0, a.il, a.k);
  lum1 = lum1 + a.rolloff * lum1 * lum1 * (1.0f - lum1);
  if(a.shadows != 0.0f)
  {
    const float toe_gamma = 1.0f - 0.5f * a.shadows;
    const float lum_toe = powf(fmaxf(lum1, 1e-7f), toe_gamma);
    lum1 = mixf(lum1, lum_toe, smoothstepf(0.3f, 0.0f, lum1));
  }
  const float L1 = L0 * powf(lum1 / lum0, 1.0f / 3.0f);
  float w[3], lab_pc[3];
  fc_weibull3(w, c0, a);
  rec2020_to_oklab(lab_pc, w);
  float C1 = sqrtf(lab_pc[1] * lab_pc[1] + lab_pc[2] * lab_pc[2]);
  const float tame = 1.0f - 0.6f * a.rolloff * smoothstepf(0.15f, 0.5f, lum1);
  C1 *= mixf(tame, 1.0f, clampf(a.chroma - 1.0f, 0.0f, 1.0f));
  float h = atan2f(lab_pc[2], lab_pc[1]);
  const float h_target = 0.96f;
  const float h_dist = fabsf(glsl_mod(h - h_target + FC_PI, 2.0f * FC_PI) - FC_PI);
  if(h_dist < 0.7f)
  {
    const float away = fabsf(lum1 - 0.35f);
    const float converge = 0.3f * smoothstepf(0.0f, 0.3f, away);
    h = lerp_chromaticity_angle(h, h_target, converge);
  }
  const float c = a.chroma - 1.0f;
  const float deriv = weibull_pdf(lum0, a.il, a.k);
  const float hi = powf(fmaxf(1.0f, 1.0f / fmaxf(deriv, 0.15f)), c * 0.1f);
  const float sw = smoothstepf(0.3f, 0.0f, lum1);
  const float lo = 1.0f + c * 0.15f * sw;
  float cr = a.chroma * hi * lo;
  cr *= 1.0f + a.red * hue_bump(h, 0.7f, 1.0f) + a.yellow * hue_bump(h, 1.76f, 1.0f) + a.blue * hue_bump(h, -1.76f, 1.0f);
  const float lab1[3] = { L1, C1 * cr * cosf(h), C1 * cr * sinf(h) };
  oklab_to_rec2020(c1, lab1);
}

// modes 1 and 3 (the batched develop runs these only)
DEV float4 fc_px_basic(float4 t, const fc_args &a)
{
  float c0[3] = { t.x + a.bias, t.y + a.bias, t.z + a.bias };
  float c1[3];
  fc_weibull3(c1, c0, a);
  if(a.mode == 3) adjust_colour_dng(c0, c1);
  return make_float4(c1[0], c1[1], c1[2], 1.0f);
}

// every mode (main.comp:74-166)
DEV float4 fc_px(float4 t, const fc_args &a)
{
  if(a.mode == 1 || a.mode == 3) return fc_px_basic(t, a);
  const float c0[3] = { t.x + a.bias, t.y + a.bias, t.z + a.bias };
  float c1[3];
  if(a.mode == 0)      fc_mode_ucs(c1, c0, a);
  else if(a.mode == 2) fc_mode_munsell(c1, c0, a);
  else if(a.mode == 4) fc_mode_agx(c1, c0, a);
  else                 fc_mode_lum_ratio(c1, c0, a);
  return make_float4(c1[0], c1[1], c1[2], 1.0f);
}

// host: filmcurv's params to the kernel's arguments.  `bytes` is the blob's
// size: words past it take the parameter defaults (chroma 1, others 0).
// `modes13` limits the colour modes to 1 and 3.
static inline const char *fc_args_init(fc_args &a, const void *params, size_t bytes, bool modes13)
{
  const float *p = (const float *)params;
  const size_t nw = bytes / 4;
  a.il = fmaxf(5e-3f, p[0]);
  a.k = fmaxf(1e-4f, p[1]);
  a.bias = p[2];
  a.mode = ((const int32_t *)params)[3];
  const float dflt[6] = { 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
  float *dst[6] = { &a.chroma, &a.rolloff, &a.red, &a.yellow, &a.blue, &a.shadows };
  for(int i = 0; i < 6; i++) *dst[i] = (size_t)(4 + i) < nw ? p[4 + i] : dflt[i];
  if(modes13 && a.mode != 1 && a.mode != 3) return "colour mode not supported here (1 and 3 are)";
  if(a.mode < 0 || a.mode > 5) return "colour mode out of range (0..5)";
  return 0;
}
