// k_shared.hip -- shared/resample (src/pipe/modules/shared/resample.comp:23-35):
// rgba f16 resampled to the output extent with the Catmull-Rom kernel as
// nine bilinear taps (shared.glsl:46-95: the middle two of each axis' four
// cubic taps merged into one bilinear tap between them).  Texture
// coordinates go through the same normalised form as the reference
// (texel / size, then size * coordinate in the sampler).
#include "launch.h"

__global__ void __launch_bounds__(256) k_sh_resample(dimg in, dimg out)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  const float W = (float)in.wd, H = (float)in.ht;
  // uv = (opos + 0.5) / output size; sample position in texels of the input
  const float u = ((float)x + 0.5f) / (float)out.wd, v = ((float)y + 0.5f) / (float)out.ht;
  const float sx = u * W, sy = v * H;
  const float t1x = floorf(sx - 0.5f) + 0.5f, t1y = floorf(sy - 0.5f) + 0.5f;
  const float fx = sx - t1x, fy = sy - t1y;
  const float w0x = fx * (-0.5f + fx * (1.0f - 0.5f * fx)), w0y = fy * (-0.5f + fy * (1.0f - 0.5f * fy));
  const float w1x = 1.0f + fx * fx * (-2.5f + 1.5f * fx), w1y = 1.0f + fy * fy * (-2.5f + 1.5f * fy);
  const float w2x = fx * (0.5f + fx * (2.0f - 1.5f * fx)), w2y = fy * (0.5f + fy * (2.0f - 1.5f * fy));
  const float w3x = fx * fx * (-0.5f + 0.5f * fx), w3y = fy * fy * (-0.5f + 0.5f * fy);
  const float w12x = w1x + w2x, w12y = w1y + w2y;
  const float o12x = w2x / (w1x + w2x), o12y = w2y / (w1y + w2y);
  // normalised tap coordinates, and back to texel space as the sampler does
  const float px[3] = { (t1x - 1.0f) / W * W, (t1x + o12x) / W * W, (t1x + 2.0f) / W * W };
  const float py[3] = { (t1y - 1.0f) / H * H, (t1y + o12y) / H * H, (t1y + 2.0f) / H * H };
  const float wx[3] = { w0x, w12x, w3x }, wy[3] = { w0y, w12y, w3y };
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for(int j = 0; j < 3; j++)
#pragma unroll
    for(int i = 0; i < 3; i++)
    {
      const float4 t = bilin_h4(in, px[i], py[j]);   // tap * wx * wy, left to right
      r.x += t.x * wx[i] * wy[j];
      r.y += t.y * wx[i] * wy[j];
      r.z += t.z * wx[i] * wy[j];
      r.w += t.w * wx[i] * wy[j];
    }
  st_h4(out, x, y, r);
}

int launch_shared_resample(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "shared/resample input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "shared/resample output")) return e;
  if(c.dp > 1) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "shared/resample: arrays (dp %u) not supported", c.dp);
  hipLaunchKernelGGL(k_sh_resample, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]));
  return launch_post("shared/resample");
}
