// k_rcd.hip -- RCD demosaicing of RGGB mosaics (ratio corrected
// demosaicing; demosaic method 1, src/pipe/modules/demosaic):
//   rcd_conv  per site: vertical/horizontal discriminator vh (W x H), the
//             diagonal discriminator pq at red/blue sites and the low-pass
//             green lp at green sites (both W/2 x H)   (rcd_conv.comp:11-49)
//   rcd_fill  per 64x32 tile (58x26 written, 3-site border): green at
//             red/blue, red/blue at blue/red, red and blue at green, in LDS
//             as f16, one thread per pair of sites   (rcd_fill.comp:24-161)
// The tile stages read LDS at linear offsets of the tile (a column offset
// past the tile edge lands in the neighbouring row, as the reference's index
// arithmetic does); offsets outside the tile's 64x32 array read 0.  Those
// reads only reach sites in the 3-site border that is not written out, or
// values derived from them.
#include "launch.h"

#define RCD_TW 64
#define RCD_TH 32
#define RCD_B 3
#define RCD_N (RCD_TW * RCD_TH)

// texture() at a texel centre of a single-channel buffer (mirrored)
DEV float rcd_tex(const dimg &b, int x, int y) { return centre_m(b, x, y); }

__global__ void __launch_bounds__(256) k_rcd_conv(dimg cfa, dimg vh, dimg pq, dimg lp)
{
  const int x = tid_x(), y = tid_y();
  if(x >= cfa.wd || y >= cfa.ht) return;
  auto c = [&](int dx, int dy) { return rcd_tex(cfa, x + dx, y + dy); };
  float vx = 0.0f, vy = 0.0f;
#pragma unroll
  for(int i = -1; i <= 1; i++)
  {
    vx += c(i, -3) - 3.0f * c(i, -2) - c(i, -1) + 6.0f * c(i, 0) - c(i, 1) - 3.0f * c(i, 2) + c(i, 3);
    vy += c(-3, i) - 3.0f * c(-2, i) - c(-1, i) + 6.0f * c(0, i) - c(1, i) - 3.0f * c(2, i) + c(3, i);
  }
  vx *= vx;
  vy *= vy;
  st_h1(vh, x, y, vx / (1e-5f + vx + vy));
  const int hx = x / 2;
  if(((x + y) & 1) == 0)
  { // red/blue site: diagonal discriminator
    float px = 1e-5f, py = 1e-5f;
#pragma unroll
    for(int i = -1; i <= 1; i++)
    {
      px += c(-3 + i, -3 + i) - c(1 + i, -1 + i) - c(1 + i, 1 + i) + c(3 + i, 3 + i)
            - 3.0f * (c(-2 + i, -2 + i) + c(2 + i, 2 + i)) + 6.0f * c(i, i);
      py += c(3 + i, -3 - i) - c(1 + i, -1 - i) - c(-1 + i, 1 - i) + c(-3 + i, 3 - i)
            - 3.0f * (c(2 + i, -2 - i) + c(-2 + i, 2 - i)) + 6.0f * c(i, -i);
    }
    px *= px;
    py *= py;
    if(hx < pq.wd && y < pq.ht) st_h1(pq, hx, y, px / (px + py));
  }
  else
  { // green site: 3x3 [1 2 1] low pass of the neighbouring red/blue column
    const int off = (x & 1) ? -1 : 1;
    const float w[3] = { 0.5f, 1.0f, 0.5f };
    float l = 0.0f;
#pragma unroll
    for(int j = -1; j <= 1; j++)
#pragma unroll
      for(int i = -1; i <= 1; i++) l += w[j + 1] * w[i + 1] * c(i + off, j);
    if(hx < lp.wd && y < lp.ht) st_h1(lp, hx, y, fmaxf(1e-6f, l));
  }
}

// RGGB site colour: 0 red, 1 green, 2 blue (rcd_fill.comp:22)
DEV int rcd_col(int x, int y) { return ((x + y) & 1) ? 1 : ((y & 1) ? 2 : 0); }

// the tile arrays: LDS f16, reads outside the array give 0
struct rcd_tile
{
  __half r[RCD_N], g[RCD_N], b[RCD_N];
};
DEV float rcd_ld(const __half *a, int i) { return (i >= 0 && i < RCD_N) ? __half2float(a[i]) : 0.0f; }

// discriminator blend: the neighbourhood mean where it is more decided
DEV float rcd_discr(float c, float n) { return fabsf(0.5f - c) < fabsf(0.5f - n) ? n : c; }

__global__ void __launch_bounds__(1024) k_rcd_fill(dimg cfa, dimg vh, dimg pq, dimg lp, dimg out, float4 wb)
{
  __shared__ rcd_tile s;
  int px = blockIdx.x * (RCD_TW - 2 * RCD_B), py = blockIdx.y * (RCD_TH - 2 * RCD_B);
  if(px >= out.wd || py >= out.ht) return;          // uniform over the work group
  const int lx = 2 * threadIdx.x, ly = threadIdx.y;   // this thread's pair (lx, lx+1) in the tile
  px += lx - RCD_B;
  py += ly - RCD_B;
  const float eps = 1e-5f;
  // the mosaic (texelFetch, zero outside) and the discriminators at offsets of tc
  auto cf = [&](int x, int y) { return fetch_m(cfa, x, y); };
  // pq and lp live on the half-width grid: site x of row y is texel
  // (x + (y odd ? 0 : 1)) / 2, integer division as the reference's ivec math
  auto half_tex = [&](const dimg &b, int x, int y) { return rcd_tex(b, (x + ((y & 1) ? 0 : 1)) / 2, y); };
  { // fill: white balanced mosaic into the channel of each site
    const int base = RCD_TW * ly + lx;
#pragma unroll
    for(int k = 0; k < 2; k++)
    {
      const int c = rcd_col(px + k, py);
      const float v = clampf(cf(px + k, py), 0.0f, 65535.0f);
      s.r[base + k] = __float2half_rn(c == 0 ? wb.x * v : 0.0f);
      s.g[base + k] = __float2half_rn(c == 1 ? wb.y * v : 0.0f);
      s.b[base + k] = __float2half_rn(c == 2 ? wb.z * v : 0.0f);
    }
  }
  __syncthreads();
  // the red/blue site of the pair, its index and position
  const int rbk = rcd_col(px, py) == 1 ? 1 : 0;
  const int rbx = px + rbk, rby = py, rbi = RCD_TW * ly + lx + rbk;
  { // green at the red/blue site (rcd_fill.comp:56-84)
    auto vhs = [&](int dx, int dy) { return rcd_tex(vh, rbx + dx, rby + dy); };
    auto lps = [&](int dx, int dy) { return half_tex(lp, rbx + dx, rby + dy); };
    auto c = [&](int dx, int dy) { return cf(rbx + dx, rby + dy); };
    const float vhc = vhs(0, 0);
    const float vhn = 0.25f * (vhs(-1, -1) + vhs(1, -1) + vhs(-1, 1) + vhs(1, 1));
    const float d = rcd_discr(vhc, vhn);
    const float N_grad = eps + fabsf(c(0, -1) - c(0, 1)) + fabsf(c(0, 0) - c(0, -2))
                       + fabsf(c(0, -1) - c(0, -3)) + fabsf(c(0, -2) - c(0, -4));
    const float S_grad = eps + fabsf(c(0, -1) - c(0, 1)) + fabsf(c(0, 0) - c(0, 2))
                       + fabsf(c(0, 1) - c(0, 3)) + fabsf(c(0, 2) - c(0, 4));
    const float W_grad = eps + fabsf(c(-1, 0) - c(1, 0)) + fabsf(c(0, 0) - c(-2, 0))
                       + fabsf(c(-1, 0) - c(-3, 0)) + fabsf(c(-2, 0) - c(-4, 0));
    const float E_grad = eps + fabsf(c(-1, 0) - c(1, 0)) + fabsf(c(0, 0) - c(2, 0))
                       + fabsf(c(1, 0) - c(3, 0)) + fabsf(c(2, 0) - c(4, 0));
    const float l0 = lps(0, 0);
    const float N_est = c(0, -1) * 2.0f * l0 / (eps + l0 + lps(0, -2));
    const float S_est = c(0, 1) * 2.0f * l0 / (eps + l0 + lps(0, 2));
    const float W_est = c(-1, 0) * 2.0f * l0 / (eps + l0 + lps(-2, 0));
    const float E_est = c(1, 0) * 2.0f * l0 / (eps + l0 + lps(2, 0));
    const float v_est = clampf((S_grad * N_est + N_grad * S_est) / (N_grad + S_grad), 0.0f, 65534.0f);
    const float h_est = clampf((W_grad * E_est + E_grad * W_est) / (E_grad + W_grad), 0.0f, 65534.0f);
    s.g[rbi] = __float2half_rn(mixf(v_est, h_est, d));
  }
  __syncthreads();
  { // the other of red/blue at the red/blue site, along the diagonals (:86-115)
    const bool red = rcd_col(rbx, rby) == 0;
    auto pqs = [&](int dx, int dy) { return half_tex(pq, rbx + dx, rby + dy); };
    const float pqc = pqs(0, 0);
    const float pqn = 0.25f * (pqs(-1, -1) + pqs(1, -1) + pqs(-1, 1) + pqs(1, 1));
    const float d = rcd_discr(pqc, pqn);
    const __half *oc = red ? s.b : s.r;
    auto sg = [&](int dx, int dy) { return rcd_ld(s.g, rbi + RCD_TW * dy + dx); };
    auto sc = [&](int dx, int dy) { return rcd_ld(oc, rbi + RCD_TW * dy + dx); };
    const float NW_grad = eps + fabsf(sc(-1, -1) - sc(1, 1)) + fabsf(sc(-1, -1) - sc(-3, -3)) + fabsf(sg(0, 0) - sg(-2, -2));
    const float NE_grad = eps + fabsf(sc(1, -1) - sc(-1, 1)) + fabsf(sc(1, -1) - sc(3, -3)) + fabsf(sg(0, 0) - sg(2, -2));
    const float SW_grad = eps + fabsf(sc(1, -1) - sc(-1, 1)) + fabsf(sc(-1, 1) - sc(-3, 3)) + fabsf(sg(0, 0) - sg(-2, 2));
    const float SE_grad = eps + fabsf(sc(-1, -1) - sc(1, 1)) + fabsf(sc(1, 1) - sc(3, 3)) + fabsf(sg(0, 0) - sg(2, 2));
    const float NW_est = sc(-1, -1) - sg(-1, -1);
    const float NE_est = sc(1, -1) - sg(1, -1);
    const float SW_est = sc(-1, 1) - sg(-1, 1);
    const float SE_est = sc(1, 1) - sg(1, 1);
    const float p_est = (NW_grad * SE_est + SE_grad * NW_est) / (NW_grad + SE_grad);
    const float q_est = (NE_grad * SW_est + SW_grad * NE_est) / (NE_grad + SW_grad);
    const __half v = __float2half_rn(clampf(sg(0, 0) + mixf(p_est, q_est, d), 0.0f, 65535.0f));
    if(red) s.b[rbi] = v;
    else s.r[rbi] = v;
  }
  __syncthreads();
  { // red and blue at the green site (:117-146)
    const int gk = 1 - rbk;
    const int gx = px + gk, gy = py, gi = RCD_TW * ly + lx + gk;
    auto vhs = [&](int dx, int dy) { return rcd_tex(vh, gx + dx, gy + dy); };
    const float vhc = vhs(0, 0);
    const float vhn = 0.25f * (vhs(-1, -1) + vhs(1, -1) + vhs(-1, 1) + vhs(1, 1));
    const float d = rcd_discr(vhc, vhn);
    auto sg = [&](int dx, int dy) { return rcd_ld(s.g, gi + RCD_TW * dy + dx); };
    const float N1 = eps + fabsf(sg(0, 0) - sg(0, -2));
    const float S1 = eps + fabsf(sg(0, 0) - sg(0, 2));
    const float W1 = eps + fabsf(sg(0, 0) - sg(-2, 0));
    const float E1 = eps + fabsf(sg(0, 0) - sg(2, 0));
    float res[2];
#pragma unroll
    for(int c = 0; c < 2; c++)
    {
      const __half *ch = c == 0 ? s.r : s.b;
      auto sc = [&](int dx, int dy) { return rcd_ld(ch, gi + RCD_TW * dy + dx); };
      const float SNabs = fabsf(sc(0, -1) - sc(0, 1));
      const float EWabs = fabsf(sc(-1, 0) - sc(1, 0));
      const float N_grad = N1 + SNabs + fabsf(sc(0, -1) - sc(0, -3));
      const float S_grad = S1 + SNabs + fabsf(sc(0, 1) - sc(0, 3));
      const float W_grad = W1 + EWabs + fabsf(sc(-1, 0) - sc(-3, 0));
      const float E_grad = E1 + EWabs + fabsf(sc(1, 0) - sc(3, 0));
      const float N_est = sc(0, -1) - sg(0, -1);
      const float S_est = sc(0, 1) - sg(0, 1);
      const float W_est = sc(-1, 0) - sg(-1, 0);
      const float E_est = sc(1, 0) - sg(1, 0);
      const float v_est = (N_grad * S_est + S_grad * N_est) / (N_grad + S_grad);
      const float h_est = (E_grad * W_est + W_grad * E_est) / (E_grad + W_grad);
      res[c] = clampf(sg(0, 0) + mixf(v_est, h_est, d), 0.0f, 65535.0f);
    }
    // both channels are computed from the tile before either is stored: the
    // red pass's store at this site is not read by the blue pass (it reads
    // the four neighbours and the sites three away)
    s.r[gi] = __float2half_rn(res[0]);
    s.b[gi] = __float2half_rn(res[1]);
  }
  __syncthreads();
#pragma unroll
  for(int k = 0; k < 2; k++)
  { // write the inner 58x26 sites: the channels over the white balance
    const int tx = lx + k, x = px + k, i = RCD_TW * ly + tx;
    if(tx >= RCD_B && ly >= RCD_B && tx < RCD_TW - RCD_B && ly < RCD_TH - RCD_B && x < out.wd && py < out.ht)
      st_h4(out, x, py, make_float4(__half2float(s.r[i]) / wb.x, __half2float(s.g[i]) / wb.y,
                                    __half2float(s.b[i]) / wb.z, 1.0f));
  }
}

// ---- host launchers --------------------------------------------------------

int launch_demosaic_rcd_conv(const node_call &c)
{
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/rcd_conv cfa")) return e;
  for(int k = 1; k < 4; k++)
    if(int e = check_conn(c, k, VKDT_FMT_F16, 1, 1, "demosaic/rcd_conv output")) return e;
  hipLaunchKernelGGL(k_rcd_conv, grid2d(c.conn[0].wd, c.conn[0].ht, 1), dim3(BX, BY), 0, c.stream,
                     to_dimg(c.conn[0]), to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]));
  return launch_post("demosaic/rcd_conv");
}

int launch_demosaic_rcd_fill(const node_call &c)
{
  if(int e = check_blob(c, 5 * 4, 0, "demosaic/rcd_fill")) return e;
  if(int e = check_conn2(c, 0, VKDT_FMT_F16, VKDT_FMT_UI16, 1, 1, "demosaic/rcd_fill cfa")) return e;
  for(int k = 1; k < 4; k++)
    if(int e = check_conn(c, k, VKDT_FMT_F16, 1, 1, "demosaic/rcd_fill discriminator")) return e;
  if(int e = check_conn(c, 4, VKDT_FMT_F16, 4, 1, "demosaic/rcd_fill output")) return e;
  const float *w = (const float *)c.pc;
  // the node's extent counts 8 invocations per tile: one 32x32 work group each
  const dim3 grid((c.wd + 7) / 8, (c.ht + 7) / 8);
  hipLaunchKernelGGL(k_rcd_fill, grid, dim3(RCD_TW / 2, RCD_TH), 0, c.stream, to_dimg(c.conn[0]),
                     to_dimg(c.conn[1]), to_dimg(c.conn[2]), to_dimg(c.conn[3]), to_dimg(c.conn[4]),
                     make_float4(w[0], w[1], w[2], w[3]));
  return launch_post("demosaic/rcd_fill");
}
