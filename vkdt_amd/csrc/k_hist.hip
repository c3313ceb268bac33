// k_hist.hip -- the display/exposure reductions of the hist and autoexp
// modules (src/pipe/modules/hist, src/pipe/modules/autoexp):
//   hist/collect     rgba f16 -> waveform histogram: per output column the
//                    counts of r, g and b per output row, packed 10 bits a
//                    channel into one u32 (collect.comp:13-25)
//   hist/map         counts -> rgba f16, pow(count/255, 0.3) (map.comp:13-24)
//   autoexp/hist     256-bin log2 luminance histogram of a radially thinned
//                    pixel set, plus the count, into u32 (hist.comp:18-60)
//   autoexp/lum      mean bin -> luminance, smoothed over frames into the
//                    protected f32 buffer (lum.comp:27-56)
//   autoexp/exp      rgb * clamp(0.18 / luminance, 1e-3, 1e3) (exp.comp:22-37)
// The counts are integer atomics as in the reference (LDS per work group for
// autoexp/hist), so they are exact; the float steps keep the shaders'
// operation order, unfused.
#include "launch.h"

namespace {

// ---- hist ------------------------------------------------------------------

// float -> row index as uint() then clamp(0, n-1) does for in-range values;
// below 0 (and NaN) the conversion is undefined in GLSL: row 0 here
DEV uint32_t hist_row(float v, int n)
{
  if(!(v > 0.0f)) return 0u;
  const float top = (float)(n - 1);
  return v >= top ? (uint32_t)(n - 1) : (uint32_t)v;
}

__global__ void __launch_bounds__(BX *BY) k_hist_collect(dimg in, uint32_t *out, int ow, int oh, int opitch)
{
  const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
  if(x >= in.wd || y >= in.ht) return;
  const float4 rgb = ld_h4(in, x, y);
  const float scale = (float)ow / (float)in.wd;
  const int op = min(max((int)(scale * (float)x), 0), ow - 1);
  const float fh = (float)oh;
  const uint32_t yr = hist_row((1.0f - rgb.x) * fh + 0.5f, oh);
  const uint32_t yg = hist_row((1.0f - rgb.y) * fh + 0.5f, oh);
  const uint32_t yb = hist_row((1.0f - rgb.z) * fh + 0.5f, oh);
  atomicAdd(out + (long)yr * opitch + op, 1u);
  atomicAdd(out + (long)yg * opitch + op, 1024u);
  atomicAdd(out + (long)yb * opitch + op, 1048576u);
}

__global__ void __launch_bounds__(BX *BY) k_hist_map(const uint32_t *cnt, int ipitch, dimg out)
{
  const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const uint32_t c = cnt[(long)y * ipitch + x];
  const float r = powf(1.0f / 255.0f * (float)(c & 1023u), 0.3f);
  const float g = powf(1.0f / 255.0f * (float)((c >> 10) & 1023u), 0.3f);
  const float b = powf(1.0f / 255.0f * (float)((c >> 20) & 1023u), 0.3f);
  st_h4(out, x, y, make_float4(r, g, b, 1.0f));
}

// ---- autoexp ---------------------------------------------------------------

#define AE_BINS 256

DEV float4 ae_ld(const dimg &b, int x, int y)
{
  if(b.fmt == DIMG_F32) return *(const float4 *)((const float *)b.p + 4 * ((long)y * b.pitch + x));
  return ld_h4(b, x, y);
}

// hist.comp:18-25; a luminance that is not >= 2^-15 (NaN included) is bin 0,
// an infinite one bin 255
DEV uint32_t ae_bin(float r, float g, float b)
{
  const float l = 2.62700212e-01f * r + 6.77998072e-01f * g + 5.93017165e-02f * b;
  if(!(l >= 3.0517578125e-05f)) return 0u;
  const float normalized = (log2f(l) - (-15.0f)) / 20.0f;
  const float nb = normalized * (float)(AE_BINS - 2) + 1.0f;   // saturating uint(), infinite l included
  return nb >= (float)(AE_BINS - 1) ? AE_BINS - 1 : (uint32_t)nb;
}

// the radial thinning of hist.comp:37-42: every pixel near the centre,
// every 13th column and 7th row towards the border
DEV bool ae_mask(int x, int y, int wd, int ht)
{
  const float dx = (float)(wd / 2) - (float)x, dy = (float)(ht / 2) - (float)y;
  const float d = sqrtf(dx * dx + dy * dy);
  const float e0 = 0.1f * (float)wd / 2.0f, e1 = (float)(wd / 2);
  float t = (d - e0) / (e1 - e0);
  t = t > 0.0f ? (t < 1.0f ? t : 1.0f) : 0.0f;
  const float s = t * t * (3.0f - 2.0f * t);
  const uint32_t mx = max((uint32_t)(s * 13.0f), 1u), my = max((uint32_t)(s * 7.0f), 1u);
  return ((uint32_t)x % mx) == 0u && ((uint32_t)y % my) == 0u;
}

// 16x16 work groups over the image, the reference's dispatch
__global__ void __launch_bounds__(256) k_ae_hist(dimg in, uint32_t *hist)
{
  __shared__ uint32_t local_hist[AE_BINS];
  __shared__ uint32_t count;
  const int tid = threadIdx.y * 16 + threadIdx.x;
  const int x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
  local_hist[tid] = 0u;
  if(tid == 0) count = 0u;
  __syncthreads();
  if(x < in.wd && y < in.ht && ae_mask(x, y, in.wd, in.ht))
  {
    const float4 c = ae_ld(in, x, y);
    const uint32_t bin = ae_bin(c.x, c.y, c.z);
    if(bin > 0u)
    {
      atomicAdd(local_hist + bin, 1u);
      atomicAdd(&count, 1u);
    }
  }
  __syncthreads();
  if(local_hist[tid]) atomicAdd(hist + tid, local_hist[tid]);
  if(tid == 0 && count) atomicAdd(hist + AE_BINS, count);
}

struct ae_params { float vup, vdown, timediff; int32_t reset; uint32_t frame; };

__global__ void __launch_bounds__(256) k_ae_lum(const uint32_t *hist, float *lum, ae_params p)
{
  __shared__ uint32_t local_hist[AE_BINS];
  __shared__ uint32_t count;
  const int tid = threadIdx.x;
  const uint32_t local_bin = hist[tid];
  local_hist[tid] = local_bin * (uint32_t)tid;
  if(tid == 0) count = hist[AE_BINS];
  __syncthreads();
  for(int split = AE_BINS >> 1; split > 0; split >>= 1)
  {
    if(tid < split) local_hist[tid] += local_hist[tid + split];
    __syncthreads();
  }
  if(tid == 0)
  {
    const float num_bright = fmaxf((float)count - (float)local_bin, 1.0f);
    const float average_bin = ((float)local_hist[0] / num_bright) - 1.0f;
    const float average_l = exp2f(((average_bin / (float)(AE_BINS - 2)) * 20.0f) + (-15.0f));
    float ml = average_l;
    float last = (p.reset == 1 || p.frame == 0u) ? ml : lum[0];
    if(isnan(last) || isinf(last)) last = ml;
    const float tau = -p.timediff * (ml > last ? p.vup : p.vdown);
    ml = last + (ml - last) * (1.0f - expf(tau));
    lum[0] = ml;
  }
}

__global__ void __launch_bounds__(BX *BY) k_ae_exp(dimg in, const float *lum, dimg out)
{
  const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
  if(x >= out.wd || y >= out.ht) return;
  const float e = fminf(fmaxf(0.18f / lum[0], 1e-3f), 1e3f);
  const float4 c = (x < in.wd && y < in.ht) ? ae_ld(in, x, y) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 o = make_float4(c.x * e, c.y * e, c.z * e, 1.0f);
  if(out.fmt == DIMG_F32) st_f4(out, x, y, o);
  else st_h4(out, x, y, o);
}

} // namespace

int launch_hist_collect(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_F16, 4, 1, "hist/collect input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_UI32, 1, 1, "hist/collect output")) return e;
  const vkdt_buf_t &o = c.conn[1];
  if(o.wd == 0 || o.ht == 0) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "hist/collect: empty histogram");
  const dimg in = to_dimg(c.conn[0]);
  hipLaunchKernelGGL(k_hist_collect, grid2d(in.wd, in.ht, 1), dim3(BX, BY), 0, c.stream, in, (uint32_t *)o.ptr,
                     (int)o.wd, (int)o.ht, (int)o.pitch);
  return launch_post("hist/collect");
}

int launch_hist_map(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_UI32, 1, 1, "hist/map input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "hist/map output")) return e;
  const vkdt_buf_t &h = c.conn[0];
  const dimg out = to_dimg(c.conn[1]);
  if(h.wd < (uint32_t)out.wd || h.ht < (uint32_t)out.ht)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "hist/map: counts %ux%u for a %dx%d output", h.wd, h.ht, out.wd, out.ht);
  hipLaunchKernelGGL(k_hist_map, grid2d(out.wd, out.ht, 1), dim3(BX, BY), 0, c.stream, (const uint32_t *)h.ptr,
                     (int)h.pitch, out);
  return launch_post("hist/map");
}

static int ae_input(const node_call &c, const char *what)
{
  return check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_F32, 4, 1, what);
}

static int ae_counts(const node_call &c, int idx, const char *what)
{
  if(int e = check_conn(c, idx, VKDT_FMT_UI32, 1, 1, what)) return e;
  if((size_t)c.conn[idx].wd * c.conn[idx].ht < AE_BINS + 1)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: %u counters, need %d", what, c.conn[idx].wd * c.conn[idx].ht,
                          AE_BINS + 1);
  return VKDT_HIP_OK;
}

int launch_autoexp_hist(const node_call &c)
{
  if(int e = ae_input(c, "autoexp/hist input")) return e;
  if(int e = ae_counts(c, 1, "autoexp/hist output")) return e;
  // node extent ceil(wd/16)*8 x ceil(ht/16)*8 dispatched in groups of 8x8
  // invocations' worth, each a 16x16 work group (main.c:12)
  const dim3 grid((c.wd + 7) / 8, (c.ht + 7) / 8);
  hipLaunchKernelGGL(k_ae_hist, grid, dim3(16, 16), 0, c.stream, to_dimg(c.conn[0]), (uint32_t *)c.conn[1].ptr);
  return launch_post("autoexp/hist");
}

int launch_autoexp_lum(const node_call &c)
{
  if(int e = check_blob(c, 0, sizeof(ae_params), "autoexp/lum")) return e;
  if(int e = ae_counts(c, 1, "autoexp/lum histogram")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F32, 1, 1, "autoexp/lum output")) return e;
  const ae_params p = *(const ae_params *)c.params;
  hipLaunchKernelGGL(k_ae_lum, dim3(1), dim3(AE_BINS), 0, c.stream, (const uint32_t *)c.conn[1].ptr,
                     (float *)c.conn[2].ptr, p);
  return launch_post("autoexp/lum");
}

int launch_autoexp_exp(const node_call &c)
{
  if(int e = ae_input(c, "autoexp/exp input")) return e;
  if(int e = ae_counts(c, 1, "autoexp/exp histogram")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_F32, 1, 1, "autoexp/exp luminance")) return e;
  if(int e = check_conn2(c, 3, VKDT_FMT_F16, VKDT_FMT_F32, 4, 1, "autoexp/exp output")) return e;
  const dimg out = to_dimg(c.conn[3]);
  hipLaunchKernelGGL(k_ae_exp, grid2d(out.wd, out.ht, 1), dim3(BX, BY), 0, c.stream, to_dimg(c.conn[0]),
                     (const float *)c.conn[2].ptr, out);
  return launch_post("autoexp/exp");
}
