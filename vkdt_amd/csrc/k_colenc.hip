// k_colenc.hip -- colenc/main (src/pipe/modules/colenc/main.comp:15-79):
// linear rec2020 rgba -> the export's primaries (a 3x3 matrix) and transfer
// function.  Reads f16 or f32 rgba and writes either (the output module's
// format); alpha is written 1.
#include "launch.h"

// rec2020 -> primaries p (modules/matrices.h), rows of the matrix; the
// published constants of the reference, so exports match its colours
struct enc_mat { float m[9]; };
static const enc_mat k_mats[11] = {
  {{ 1, 0, 0, 0, 1, 0, 0, 0, 1 }},                                        // 0 custom: unchanged
  {{ 1.66022677f, -0.58754761f, -0.07283825f, -0.12455334f, 1.13292605f, -0.00834963f,
     -0.01815514f, -0.10060303f, 1.11899817f }},                           // 1 rec709
  {{ 1, 0, 0, 0, 1, 0, 0, 0, 1 }},                                        // 2 rec2020
  {{ 1.15194302f, -0.09753232f, -0.05448118f, -0.12454585f, 1.13290963f, -0.00837122f,
     -0.02253539f, -0.04979918f, 1.07275365f }},                           // 3 adobe rgb
  {{ 1.34353337f, -0.28218904f, -0.06142427f, -0.06530851f, 1.07578268f, -0.01048453f,
     0.00282971f, -0.01961215f, 1.01717851f }},                            // 4 display p3
  {{ 0.636958048301290991f, 0.144616903586208406f, 0.168880975164172054f, 0.26270021201126692f,
     0.677998071518871148f, 0.0593017164698619384f, 4.9999999999999999e-17f, 0.0280726930490874452f,
     1.06098505771079066f }},                                              // 5 xyz
  {{ 6.68685575e-01f, 1.51817679e-01f, 1.77189677e-01f, 4.49002044e-02f, 8.62145497e-01f,
     1.01922441e-01f, -2.66851927e-09f, 2.78271109e-02f, 1.05170358f }},  // 6 aces ap0
  {{ 9.62918591e-01f, 1.16137050e-02f, 2.55863361e-02f, 4.16800770e-04f, 9.99378426e-01f,
     -8.82457347e-05f, 5.31123331e-03f, 2.18655328e-02f, 9.75907920e-01f }}, // 7 aces ap1
  {{ 1, 0, 0, 0, 1, 0, 0, 0, 1 }},                                        // 8 arri 3: not an output space
  {{ 1, 0, 0, 0, 1, 0, 0, 0, 1 }},                                        // 9 arri 4: not an output space
  {{ 0.853263f, 0.079695f, 0.067042f, 0.029375f, 0.809195f, 0.161430f, 0.051575f, 0.208097f,
     0.740329f }},                                                         // 10 red wide gamut
};

struct enc_args
{
  enc_mat M;
  int     mat;   // 0: no matrix
  int     trc;
};

DEV float4 ld_rgba(const dimg &b, int x, int y)
{
  if(b.fmt == DIMG_F32) return *(const float4 *)((const float *)b.p + 4 * ((long)y * b.pitch + x));
  return ld_h4(b, x, y);
}

DEV float enc_trc(float v, int trc)
{
  if(trc == 1)
  { // rec709
    const float a = 1.09929682680944f, b = 0.018053968510807f;
    return v > b ? fpow(v, 1.0f / 2.2f) * a - (a - 1.0f) : v * 4.5f;
  }
  if(trc == 2) return v > 0.0031308f ? fpow(v, 1.0f / 2.4f) * 1.055f - 0.055f : v * 12.92f;   // sRGB
  if(trc == 3)
  { // PQ (the cancellation-free arrangement of the reference)
    const float c3 = 2392.0f / 128.0f, c2 = 2413.0f / 128.0f, c1 = c3 - c2 + 1.0f;
    const float m1 = 1305.0f / 8192.0f, m2 = 2523.0f / 32.0f;
    const float p = fpow(fmaxf(0.0f, v), m1);
    const float num = (c1 - 1.0f) + (c2 - c3) * p, den = 1.0f + c3 * p;
    return fpow(1.0f + num / den, m2);
  }
  if(trc == 4) return fpow(v, 1.0f / 2.6f);                                   // DCI
  if(trc == 5)
  { // HLG
    const float a = 0.17883277f, b = 1.0f - 4.0f * a, c = 0.5f - a * logf(4.0f * a);
    return v > 1.0f / 12.0f ? a * logf(12.0f * v - b) + c : sqrtf(3.0f * v);
  }
  if(trc == 6) return fpow(v, 1.0f / 2.2f);                                   // gamma 2.2
  return v;                                                                   // linear
}

__global__ void __launch_bounds__(256) k_colenc(dimg in, dimg out, enc_args a)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float4 t = x < in.wd && y < in.ht ? ld_rgba(in, x, y) : make_float4(0.f, 0.f, 0.f, 0.f);
  float r = t.x, g = t.y, b = t.z;
  if(a.mat)
  {
    const float *m = a.M.m;
    const float r2 = m[0] * r + m[1] * g + m[2] * b;
    const float g2 = m[3] * r + m[4] * g + m[5] * b;
    const float b2 = m[6] * r + m[7] * g + m[8] * b;
    r = r2; g = g2; b = b2;
  }
  const float4 o = make_float4(enc_trc(r, a.trc), enc_trc(g, a.trc), enc_trc(b, a.trc), 1.0f);
  if(out.fmt == DIMG_F32) st_f4(out, x, y, o);
  else if(out.fmt == VKDT_FMT_UI8)
  { // R8G8B8A8_UNORM: round(clamp(v, 0, 1) * 255), ties to even
    const float v[4] = { o.x, o.y, o.z, o.w };
    uint32_t u = 0;
    for(int k = 0; k < 4; k++)
    {
      const float c = v[k] > 0.0f ? (v[k] < 1.0f ? v[k] : 1.0f) : 0.0f;   // NaN -> 0
      u |= (uint32_t)__builtin_rintf(c * 255.0f) << (8 * k);
    }
    ((uint32_t *)out.p)[(long)y * out.pitch + x] = u;
  }
  else st_h4(out, x, y, o);
}

int launch_colenc_main(const node_call &c)
{
  if(int e = check_blob(c, 0, 2 * 4, "colenc/main")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_F32, 4, 1, "colenc/main input")) return e;
  if(c.nconn > 1 && c.conn[1].format == VKDT_FMT_UI8)
  { if(int e = check_conn(c, 1, VKDT_FMT_UI8, 4, 1, "colenc/main output")) return e; }
  else if(int e = check_conn2(c, 1, VKDT_FMT_F16, VKDT_FMT_F32, 4, 1, "colenc/main output")) return e;
  const int32_t *p = (const int32_t *)c.params;
  enc_args a;
  // primaries the reference encodes (colenc/main.comp:26-34); others stay rec2020
  const int prim = p[0];
  const int known = prim == 1 || prim == 3 || prim == 4 || prim == 5 || prim == 6 || prim == 7 || prim == 10;
  a.mat = known;
  a.M = k_mats[known ? prim : 0];
  a.trc = p[1];
  hipLaunchKernelGGL(k_colenc, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), a);
  return launch_post("colenc/main");
}
