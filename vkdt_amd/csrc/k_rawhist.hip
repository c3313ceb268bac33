// k_rawhist.hip -- the rawhist module's nodes (raw noise statistics for noise
// profiling, src/pipe/modules/rawhist):
//   collect  ui16 mosaic -> accumulators (bins x 5, FP64): per bin of
//            bins*log2(mean+1) the count, the sum of block means and the sum
//            of block second moments of each CFA block's green sites
//            (collect.comp:19-72)
//   map      accumulators -> rgba f16 plot of count, variance and the
//            current noise model (map.comp:20-52)
// The reference adds into R32F with float atomics; here every work group
// bins its blocks into an LDS histogram (integer counts, FP64 moments) and
// adds its non-empty bins into the FP64 accumulators once, so the
// per-frame moments are FP64 sums of the FP32 per-block values.
#include "launch.h"

#define RH_MAX_BINS 2048
#define RH_BATCH 4

// the module's committed uniform (rawhist/main.c:19-29): noise model
// c + dx * x, filters, black and white in raw units
struct rh_params
{
  float    c, dx;
  uint32_t filters, black, white;
};

DEV float rh_norm(float v, float black, float white)
{
  return clampf((v - black) / (white - black), 0.0f, 1.0f);
}

// one invocation of collect.comp at block (bx,by): false when it adds nothing
DEV bool rh_block(const dimg &in, int bx, int by, uint32_t filters, float black, float white, int bins,
                  int &bin, float &mean, float &mom2)
{
  mean = 0.0f;
  mom2 = 0.0f;
  float x = 0.0f;
  if(filters == 9u)
  { // the five greens of the 3x3 block: corners and centre
    const int px = 3 * bx, py = 3 * by;
    if(px >= in.wd || py >= in.ht) return false;
    const float v0 = rh_norm(fetch_u16(in, px, py), black, white);
    const float v1 = rh_norm(fetch_u16(in, px + 2, py), black, white);
    const float v2 = rh_norm(fetch_u16(in, px + 1, py + 1), black, white);
    const float v3 = rh_norm(fetch_u16(in, px, py + 2), black, white);
    const float v4 = rh_norm(fetch_u16(in, px + 2, py + 2), black, white);
    mean = (v0 + v1 + v2 + v3 + v4) / 5.0f;
    mom2 = (v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3 + v4 * v4) / 5.0f;
    x = (float)bins * log2f(mean + 1.0f);
  }
  else if(filters != 0u)
  { // the two greens of the RGGB block
    const int px = 2 * bx, py = 2 * by;
    if(px >= in.wd || py >= in.ht) return false;
    const float v0 = rh_norm(fetch_u16(in, px + 1, py), black, white);
    const float v1 = rh_norm(fetch_u16(in, px, py + 1), black, white);
    mean = (v0 + v1) / 2.0f;
    mom2 = (v0 * v0 + v1 * v1) / 2.0f;
    x = (float)bins * log2f(mean + 1.0f);
  }
  bin = (int)x;                 // ivec2(x, 0): truncation; x >= 0
  return bin < bins;            // imageAtomicAdd outside the image adds nothing
}

__global__ void __launch_bounds__(256) k_rh_collect(dimg in, double *hist, int bins, int nx, int ny,
                                                    uint32_t filters, float black, float white)
{
  __shared__ uint32_t s_cnt[RH_MAX_BINS];
  __shared__ double s_m1[RH_MAX_BINS], s_m2[RH_MAX_BINS];
  for(int i = threadIdx.x; i < bins; i += blockDim.x)
  {
    s_cnt[i] = 0u;
    s_m1[i] = 0.0;
    s_m2[i] = 0.0;
  }
  __syncthreads();
  // each work group takes one contiguous run of blocks (a band of block
  // rows), so its histogram spans the brightness of one band and its flush
  // below touches few bins; RH_BATCH blocks per thread and pass have their
  // loads in flight before the first LDS atomic
  const int n = nx * ny, per = (n + gridDim.x - 1) / gridDim.x;
  const int beg = blockIdx.x * per, end = min(n, beg + per), stride = blockDim.x;
  for(int t0 = beg + threadIdx.x; t0 < end; t0 += RH_BATCH * stride)
  {
    int bin[RH_BATCH];
    float mean[RH_BATCH], mom2[RH_BATCH];
    bool ok[RH_BATCH];
#pragma unroll
    for(int k = 0; k < RH_BATCH; k++)
    {
      const int t = t0 + k * stride;
      ok[k] = t < end && rh_block(in, t % nx, t / nx, filters, black, white, bins, bin[k], mean[k], mom2[k]);
    }
#pragma unroll
    for(int k = 0; k < RH_BATCH; k++)
    {
      if(!ok[k]) continue;
      atomicAdd(&s_cnt[bin[k]], 1u);
      unsafeAtomicAdd(&s_m1[bin[k]], (double)mean[k]);
      unsafeAtomicAdd(&s_m2[bin[k]], (double)mom2[k]);
    }
  }
  __syncthreads();
  for(int i = threadIdx.x; i < bins; i += blockDim.x)
  {
    if(!s_cnt[i]) continue;
    unsafeAtomicAdd(hist + i, (double)s_cnt[i]);
    unsafeAtomicAdd(hist + bins + i, s_m1[i]);
    unsafeAtomicAdd(hist + 2 * bins + i, s_m2[i]);
  }
}

__global__ void __launch_bounds__(256) k_rh_map(const double *hist, int hist_wd, int hist_pitch, dimg out,
                                                float fit_c, float fit_dx)
{
  const int x = tid_x(), y = tid_y();
  if(x >= out.wd || y >= out.ht) return;
  // texelFetch of the accumulators (as float, the reference's R32F), zero outside
  const float cnt = x < hist_wd ? (float)hist[x] : 0.0f;
  const float mean = x < hist_wd ? (float)hist[hist_pitch + x] : 0.0f;
  const float mom2 = x < hist_wd ? (float)hist[2 * hist_pitch + x] : 0.0f;
  const float var = mom2 / cnt - mean / cnt * mean / cnt;
  const float yy = (float)(out.ht - y - 1) / (float)out.ht;
  const float c1 = 20.0f, c2 = 0.05f;
  const float lin = mean / cnt;
  const float fit = fit_c + fit_dx * lin;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if(fabsf(yy - sqrtf(var) * c1) < 0.015f)
  {
    r = 0.5f * (r + 1.0f);
    g = 0.5f * (g + 0.2f);
    b = 0.5f * (b + 0.2f);
  }
  if(yy < log2f(cnt) * c2)
  {
    r += 0.5f;
    g += 0.5f;
    b += 0.5f;
  }
  const float df = fabsf(yy - sqrtf(fit) * c1);
  if(df < 0.005f)
  {
    const float s = smoothstepf(0.0f, 1.0f, 1.0f - df / 0.005f);
    r += s;
    g += s;
    b += s;
  }
  st_h4(out, x, y, make_float4(r, g, b, 1.0f));
}

// ---- host launchers --------------------------------------------------------

int launch_rawhist_collect(const node_call &c)
{
  if(int e = check_blob(c, 0, sizeof(rh_params), "rawhist/collect")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "rawhist/collect input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_ATOM, 1, 1, "rawhist/collect output")) return e;
  const rh_params *p = (const rh_params *)c.params;
  const vkdt_buf_t &o = c.conn[1];
  if(o.ht < 3 || o.wd > RH_MAX_BINS || o.pitch != o.wd)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "rawhist/collect: accumulators %ux%u (pitch %u), need <= %d x >= 3",
                          o.wd, o.ht, o.pitch, RH_MAX_BINS);
  // every invocation of the ceil(wd/8) x ceil(ht/8) groups of 8x8 runs
  const int nx = (int)((c.wd + 7) / 8 * 8), ny = (int)((c.ht + 7) / 8 * 8);
  const long n = (long)nx * ny;
  if(n >= (1l << 30)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "rawhist/collect: %ld blocks", n);
  const int groups = (int)(n / 256 + 1 < 1024 ? n / 256 + 1 : 1024);
  // collect.comp:28-29: levels in raw units over 65536
  const float black = (float)p->black / 65536.0f, white = (float)p->white / 65536.0f;
  hipLaunchKernelGGL(k_rh_collect, dim3(groups), dim3(256), 0, c.stream, to_dimg(c.conn[0]), (double *)o.ptr,
                     (int)o.wd, nx, ny, p->filters, black, white);
  return launch_post("rawhist/collect");
}

int launch_rawhist_map(const node_call &c)
{
  if(int e = check_blob(c, 0, sizeof(rh_params), "rawhist/map")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_ATOM, 1, 1, "rawhist/map input")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_F16, 4, 1, "rawhist/map output")) return e;
  const rh_params *p = (const rh_params *)c.params;
  const vkdt_buf_t &h = c.conn[0];
  if(h.ht < 3) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "rawhist/map: accumulators need 3 rows");
  hipLaunchKernelGGL(k_rh_map, grid2d(c.wd, c.ht, 1), dim3(BX, BY), 0, c.stream, (const double *)h.ptr, (int)h.wd,
                     (int)h.pitch, to_dimg(c.conn[1]), p->c, p->dx);
  return launch_post("rawhist/map");
}
