// k_batch.hip -- the raw-development graph of many images per launch.
//
// A batch holds stills or frames of any sizes and CFAs.  Every stage is one
// launch over the tiles of all images that have it: a flat tile list with a
// per-image job table (binary searched by tile index), run either as one
// tile per work group or by persistent work groups that pull tiles from
// per-XCD queues (XCD k starts on the k-th eighth of the list, then steals
// from the next eighths).  Stages:
//
//   group 0  half      denoise/half                   CFA ui16 -> Q rgba f16
//   group 1  wavelet   denoise/downcov, down x3       Q -> Q (four launches)
//   group 2  dnasm     denoise/assemble               5x Q -> Q
//   group 3  doub      denoise/doub                   CFA + 2x Q -> mosaic f16
//   group 4  hilite    hilite/half, reduce 1..L, assemble L-1..0, doub
//                      (images whose graph has hilite between denoise and
//                      demosaic)                      mosaic -> mosaic f16
//   group 5  gauss     demosaic/gauss                 mosaic -> block covariance
//   group 6  develop   demosaic/splat + demosaic/fix + colour + filmcurv,
//                      fused: the green of a tile and its halo lives in LDS,
//                      fix/colour/filmcurv run in registers -> rgba f16
//   group 7  pyr1      llap/curve + llap/reduce to level 1, fused: the
//                      full-resolution curve planes exist only in LDS
//   group 8  pyramid   llap/reduce levels 2.., llap/assemble levels nl-2..1
//   group 9  final     llap/assemble level 0 + llap/colour, fused: level 0
//                      planes are re-evaluated from the filmcurv output
//
// Wherever the per-node graph stores an intermediate as f16 and the next
// node reads it back, the fused stages round the value to binary16 at the
// same point (r16), so each image's output equals its per-node graph run.
// Per 24 MP Bayer still this skips the splat/fix/colour intermediates and
// the 11 full-resolution curve planes (about 96 B per CFA pixel).
#include <vector>
#include <algorithm>
#include <mutex>
#include "launch.h"
#include "ops.h"
#include "ops_hilite.h"

#define BT_LEVELS 12
#define HL_LEVELS 15  // hilite: the block grid and at most 14 coarser levels (hilite/main.c)
#define COV_R 2     // downcov: 64 x 8 texels per tile
#define DOWN_R 4    // down: 64 x 16 texels per tile
#define DEV_R 4     // develop: 64 x 16 texels per tile
#define ASM_R 4     // assemble / final: 64 x 16 texels per tile

struct bt_img
{
  dimg cfa, out;
  dimg qh, qd[4], qa;      // denoise: half, downcov + down 1..3, assemble (Q rgba f16)
  dimg mos, cov, film;     // demosaic input (f16), gauss output (rgba f16), filmcurv output (rgba f16)
  dimg dmos;               // denoise/doub output: mos itself unless hilite runs in between
  dimg hq[HL_LEVELS];      // hilite: half output (level 0) and reduce outputs (levels 1..L)
  dimg ha[HL_LEVELS];      // hilite: assemble outputs (levels 0..L-1); ha[L] aliases hq[L]
  hl_args hl;
  int  hl_nx, hl_ny;       // hilite/doub: blocks run (8-aligned node extent, bounded by the mosaic)
  dimg lp[BT_LEVELS];      // llap planes of level l >= 1, num_gamma + 1 planes
  dimg la[BT_LEVELS];      // llap reconstruction of level l, 1 <= l <= nl-2
  int  nl, xt, dm_r;
  dn_args dn_half, dn_down[4], dn_asm, dn_doub;
  co_args co;
  fc_args fc;
  ll_args ll;
};

struct bt_job { int img, tiles_x, start, level; };

struct bt_launch
{
  const bt_img *img;
  const bt_job *job;
  int njob, total;
  int *ctr;          // NUM_XCD tile counters (persistent launches)
  int persist;       // 0: one tile per work group, k: k persistent work groups per CU
  int rows;          // per-pixel stages: rows per thread (tile 64 x 4 rows)
};

// ---- tile scheduling --------------------------------------------------------

// values every lane of the work group holds alike (job and image records):
// kept in scalar registers
DEV int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
DEV long rfl(long v)
{
  return (long)(((uint64_t)(uint32_t)rfl((int)((uint64_t)v >> 32)) << 32) | (uint32_t)rfl((int)(uint64_t)v));
}
DEV dimg rfl(const dimg &d)
{
  dimg r;
  r.p = (void *)rfl((long)d.p);
  r.wd = rfl(d.wd); r.ht = rfl(d.ht); r.pitch = rfl(d.pitch); r.chan = rfl(d.chan);
  r.pstride = rfl(d.pstride); r.fmt = rfl(d.fmt);
  return r;
}
DEV bt_job rfl(const bt_job &j) { return bt_job{ rfl(j.img), rfl(j.tiles_x), rfl(j.start), rfl(j.level) }; }

DEV int bt_find(const bt_job *job, int n, int t)
{
  int lo = 0, hi = n - 1;
  while(lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if(job[mid].start <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// next tile for this work group: its home queue (work group id mod XCDs)
// first, then the following queues; `tried` counts exhausted queues
DEV int bt_next(const bt_launch &L, int &tried)
{
  const int home = blockIdx.x % NUM_XCD;
  for(; tried < NUM_XCD; tried++)
  {
    const int k = (home + tried) % NUM_XCD;
    const int s = (int)((long long)L.total * k / NUM_XCD), e = (int)((long long)L.total * (k + 1) / NUM_XCD);
    if(s >= e) continue;
    const int t = s + atomicAdd(L.ctr + k, 1);
    if(t < e) return t;
  }
  return L.total;
}

// Run the launch's tiles.  setup(job) reads an image's arguments (LDS
// copies, buffer views in registers) and runs once per work group and
// image; fn(job, local tile) does one tile.  Both may use __syncthreads
// (every thread calls them alike).  Persistent work groups keep the image
// of their last tile and redo setup only when the queue moves them on.
template<typename S, typename F>
DEV void bt_tiles(const bt_launch &L, S setup, F fn)
{
  if(!L.persist)
  { // one tile per work group, XCD k on the k-th eighth of the list
    const int n = gridDim.x, b = blockIdx.x;
    const int q = n / NUM_XCD, r = n - q * NUM_XCD, k = b % NUM_XCD;
    const int t = k * q + min(k, r) + b / NUM_XCD;
    if(t >= L.total) return;
    const bt_job jb = rfl(L.job[bt_find(L.job, L.njob, t)]);
    setup(jb);
    __syncthreads();
    fn(jb, t - jb.start);
    return;
  }
  __shared__ int s_t;
  int tried = 0, js = 0, je = 0;   // tile range [js, je) of the current image
  bt_job jb = { 0, 1, 0, 0 };
  for(;;)
  {
    __syncthreads();   // the previous tile is done with s_t and its LDS
    if(threadIdx.x == 0 && threadIdx.y == 0) s_t = bt_next(L, tried);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_t);
    if(t >= L.total) break;
    if(t < js || t >= je)
    {
      const int j = bt_find(L.job, L.njob, t);
      jb = rfl(L.job[j]);
      js = jb.start;
      je = rfl(j + 1 < L.njob ? L.job[j + 1].start : L.total);
      setup(jb);
      __syncthreads();
    }
    fn(jb, t - js);
  }
}

__global__ void kb_reset(int *ctr, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) ctr[i] = 0;
}

// Per-image arguments live in global memory, which the stages' stores may
// alias as far as the compiler knows: read after a store they would be
// re-fetched with vector loads on every use.  setup() therefore copies a
// stage's argument block into LDS and its buffer views into registers.
template<typename T>
DEV void lds_copy(T &dst, const T &src)
{
  static_assert(sizeof(T) % 4 == 0, "word copy");
  const int tid = threadIdx.y * BX + threadIdx.x;
  for(int i = tid; i < (int)(sizeof(T) / 4); i += BX * BY) ((int *)&dst)[i] = ((const int *)&src)[i];
}

// per-pixel stages: tiles of 64 x 4R texels, R = L.rows (1..8)
#define PIX_TILE(L, jb, lt, X0, Y0) \
  const int X0 = ((lt) % (jb).tiles_x) * BX, Y0 = ((lt) / (jb).tiles_x) * (BY * (L).rows)

// ---- denoise -----------------------------------------------------------------

__global__ void __launch_bounds__(256) kb_half(bt_launch L)
{
  __shared__ dn_args a;
  dimg cfa, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.dn_half);
    cfa = rfl(im.cfa); o = rfl(im.qh);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < o.wd && y < o.ht) st_h4(o, x, y, dn_half_px(cfa, a, x, y));
    }
  });
}

__global__ void __launch_bounds__(256) kb_downcov(bt_launch L)
{
  __shared__ dn_cov_smem<COV_R> s;
  __shared__ dn_args a;
  dimg in, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.dn_down[0]);
    in = rfl(im.qh); o = rfl(im.qd[0]);
  }, [&](const bt_job &jb, int lt) {
    const dimg nocov = {};
    dn_downcov_tile<COV_R>(in, o, nocov, a, (lt % jb.tiles_x) * BX, (lt / jb.tiles_x) * (BY * COV_R), s);
  });
}

// at most 72 VGPRs (7 waves per SIMD), measured faster than the compiler's choice
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) kb_down(bt_launch L)
{
  __shared__ float4 tile[DN_TN(DOWN_R)];
  __shared__ dn_args a;
  dimg in, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.dn_down[jb.level]);
    in = rfl(im.qd[jb.level - 1]); o = rfl(im.qd[jb.level]);
  }, [&](const bt_job &jb, int lt) {
    dn_down_tile<DOWN_R>(in, o, a, (lt % jb.tiles_x) * BX, (lt / jb.tiles_x) * (BY * DOWN_R), tile);
  });
}

// at most 72 VGPRs (7 waves per SIMD), measured faster than the compiler's choice
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) kb_dnasm(bt_launch L, asm_mats m)
{
  __shared__ dn_args a;
  dimg s0, s1, s2, s3, s4, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.dn_asm);
    s0 = rfl(im.qh); s1 = rfl(im.qd[0]); s2 = rfl(im.qd[1]); s3 = rfl(im.qd[2]); s4 = rfl(im.qd[3]); o = rfl(im.qa);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < o.wd && y < o.ht) st_h4(o, x, y, dn_assemble_px(s0, s1, s2, s3, s4, a, m, x, y));
    }
  });
}

__global__ void __launch_bounds__(256) kb_doub(bt_launch L)
{
  __shared__ dn_args a;
  dimg cfa, crs0, crs1, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.dn_doub);
    cfa = rfl(im.cfa); crs0 = rfl(im.qa); crs1 = rfl(im.qh); o = rfl(im.dmos);
  }, [&](const bt_job &jb, int lt) {
    if(a.filters != 9u)
    { // Bayer: 2x2 blocks per thread, tiles of 2 BX x 2 BY rows texels
      const int x = ((lt % jb.tiles_x) * BX + threadIdx.x) * 2, Y0 = (lt / jb.tiles_x) * (BY * L.rows);
#pragma unroll 1
      for(int k = 0; k < L.rows; k++)
      {
        const int y = (Y0 + threadIdx.y + BY * k) * 2;
        if(x >= o.wd || y >= o.ht) continue;
        float v[4];
        dn_doub_quad(cfa, crs0, crs1, o.wd, o.ht, a, x, y, v);
        st_h1_quad(o, x, y, v);
      }
      return;
    }
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < o.wd && y < o.ht) st_h1(o, x, y, dn_doub_px(cfa, crs0, crs1, o.wd, o.ht, a, x, y));
    }
  });
}

// ---- hilite -----------------------------------------------------------------

__global__ void __launch_bounds__(256) kb_hl_half(bt_launch L)
{
  __shared__ hl_args a;
  dimg in, o;
  int xt = 0;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.hl);
    in = rfl(im.dmos); o = rfl(im.hq[0]); xt = rfl(im.xt);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < o.wd && y < o.ht) st_h4(o, x, y, xt ? hl_half_at<1, 0>(in, x, y, a.white) : hl_half_at<0, 0>(in, x, y, a.white));
    }
  });
}

// reduce to level jb.level: tiles of BX x BY coarse texels
__global__ void __launch_bounds__(256) kb_hl_reduce(bt_launch L)
{
  __shared__ hl_lds t;
  __shared__ hl_args a;
  dimg in, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.hl);
    in = rfl(im.hq[jb.level - 1]); o = rfl(im.hq[jb.level]);
  }, [&](const bt_job &jb, int lt) {
    hl_reduce_tile(in, o, a, (lt % jb.tiles_x) * BX, (lt / jb.tiles_x) * BY, t);
  });
}

// assemble of level jb.level from its fine level and the next coarser one
__global__ void __launch_bounds__(256) kb_hl_asm(bt_launch L)
{
  __shared__ hl_args a;
  dimg fine, crs, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.hl);
    fine = rfl(im.hq[jb.level]); crs = rfl(im.ha[jb.level + 1]); o = rfl(im.ha[jb.level]);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < o.wd && y < o.ht) st_h4(o, x, y, hl_assemble_at(fine, crs, a, x, y));
    }
  });
}

// back onto the mosaic: one thread per block
// at most 64 VGPRs (8 waves per SIMD), measured faster than the compiler's choice
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) kb_hl_doub(bt_launch L)
{
  __shared__ hl_args a;
  dimg in, crs, o;
  int xt = 0, nx = 0, ny = 0;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.hl);
    in = rfl(im.dmos); crs = rfl(im.ha[0]); o = rfl(im.mos);
    xt = rfl(im.xt); nx = rfl(im.hl_nx); ny = rfl(im.hl_ny);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x >= nx || y >= ny) continue;
      if(xt) hl_doub_at<1, 0>(in, crs, o, x, y, a.white);   // the denoised mosaic is f16
      else   hl_doub_at<0, 0>(in, crs, o, x, y, a.white);
    }
  });
}

// ---- demosaic + colour + filmcurv --------------------------------------------

__global__ void __launch_bounds__(256) kb_gauss(bt_launch L)
{
  dimg in, o;
  int xt = 0;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    in = rfl(im.mos); o = rfl(im.cov); xt = rfl(im.xt);
  }, [&](const bt_job &jb, int lt) {
    PIX_TILE(L, jb, lt, X0, Y0);
    const int x = X0 + threadIdx.x;
#pragma unroll 1
    for(int k = 0; k < L.rows; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x >= o.wd || y >= o.ht) continue;
      // waves whose taps all lie inside the mosaic read it without mirroring
      auto direct = [&](int px, int py) { return ld_mt<0>(in, px, py); };
      if(xt) st_h4(o, x, y, __all(dm_gauss_inside<1>(in, x, y)) ? dm_gauss_at<1>(direct, x, y) : dm_gauss_px<1>(in, x, y));
      else   st_h4(o, x, y, __all(dm_gauss_inside<0>(in, x, y)) ? dm_gauss_at<0>(direct, x, y) : dm_gauss_px<0>(in, x, y));
    }
  });
}

// green of the tile plus an r-texel halo at mirrored coordinates (what the
// per-node fix reads from the stored splat output), then fix, colour and
// filmcurv per texel with the f16 rounding of each stored node output.
// Tiles whose footprint -- the fix radius r plus the splat radius rs around
// the tile -- lies inside the image first stage that mosaic window in LDS
// (no mirroring or border wrap can occur there); border tiles read the
// mosaic through the bounds-checked global path.  Both see the same values.
#define DEV_M 5   // r + rs <= 3 + 2

// (row, column) of the elements i0, i0 + step, ... of a row-major tile w
// wide, stepped without a division per element
struct tile_walk
{
  int ty, tx, dq, dr, w;
  __device__ tile_walk(int i0, int step, int w_) : w(w_)
  {
    ty = i0 / w; tx = i0 - ty * w;
    dq = step / w; dr = step - dq * w;
  }
  __device__ void next()
  {
    tx += dr; ty += dq;
    if(tx >= w) { tx -= w; ty++; }
  }
};

template<int XT>
DEV void develop_tile(const dimg &mos, const dimg &cov, const dimg &film, const co_args &co, const fc_args &fc,
                      int r, int X0, int Y0, float *gt, float *mt)
{
  constexpr int rs = XT ? 2 : 1;
  const int gw = BX + 2 * r, gh = BY * DEV_R + 2 * r;
  const int W = film.wd, H = film.ht;
  const int m = r + rs, mw = BX + 2 * m, mh = BY * DEV_R + 2 * m;
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int x = X0 + threadIdx.x;
  auto green = [&](int px, int py) { return gt[(py - Y0 + r) * gw + (px - X0 + r)]; };
  if(X0 - m >= 0 && Y0 - m >= 0 && X0 + BX + m <= W && Y0 + BY * DEV_R + m <= H)
  {
    tile_walk wm(tid, BX * BY, mw);
    for(int i = tid; i < mw * mh; i += BX * BY, wm.next())
      mt[i] = ld_h1(mos, X0 - m + wm.tx, Y0 - m + wm.ty);
    __syncthreads();
    auto mosl = [&](int px, int py) { return mt[(py - Y0 + m) * mw + (px - X0 + m)]; };
    if(XT)
    {
      tile_walk wg(tid, BX * BY, gw);
      for(int i = tid; i < gw * gh; i += BX * BY, wg.next())
        gt[i] = r16(dm_splat_at<XT>(mosl, W, H, cov, X0 - r + wg.tx, Y0 - r + wg.ty));
    }
    else
    { // Bayer: the sites with tx + ty even first, then the odd ones, so that
      // consecutive lanes share one splat pattern (X0, Y0 are even and gw is
      // even, so the parity of tx + ty is that of the texel's x + y)
      const int hw = gw / 2, ht2 = gh * hw;
      int c = tid >= ht2;
      tile_walk wf(tid - c * ht2, BX * BY, hw);
      for(int e = tid; e < gw * gh; e += BX * BY, wf.next())
      {
        if(!c && e >= ht2) { c = 1; wf = tile_walk(e - ht2, BX * BY, hw); }
        const int ty = wf.ty, tx = 2 * wf.tx + ((ty + c) & 1);
        // c is the site's parity: only its green taps' weights are evaluated
        gt[ty * gw + tx] = r16(c ? dm_splat_at<XT, 0>(mosl, W, H, cov, X0 - r + tx, Y0 - r + ty)
                                 : dm_splat_at<XT, 1>(mosl, W, H, cov, X0 - r + tx, Y0 - r + ty));
      }
    }
    __syncthreads();
#pragma unroll 1
    for(int k = 0; k < DEV_R; k++)
    { // Bayer: each wave one CFA colour (bayer_wave_xy) in each 64x4 band
      const int2 q = XT ? make_int2(x, Y0 + threadIdx.y + BY * k) : bayer_wave_xy(X0, Y0 + BY * k);
      const int xq = q.x, y = q.y;
      const float4 f = XT ? dm_fix_r<XT>(mosl, cov, r, xq, y, green)
                          : (((xq + y) & 1) ? dm_fix_r<XT, 1>(mosl, cov, r, xq, y, green)
                                            : dm_fix_r<XT, 0>(mosl, cov, r, xq, y, green));
      const float4 c = r16(co_px(r16(f), co));
      st_h4(film, xq, y, fc_px_basic(c, fc));
    }
    return;
  }
  tile_walk wg(tid, BX * BY, gw);
  for(int i = tid; i < gw * gh; i += BX * BY, wg.next())
    gt[i] = r16(dm_splat_px<XT>(mos, cov, mirror(X0 - r + wg.tx, W), mirror(Y0 - r + wg.ty, H)));
  __syncthreads();
#pragma unroll 1
  for(int k = 0; k < DEV_R; k++)
  {
    const int y = Y0 + threadIdx.y + BY * k;
    if(x >= W || y >= H) continue;
    const float4 f = dm_fix_px<XT>(mos, cov, r, x, y, green);
    const float4 c = r16(co_px(r16(f), co));
    st_h4(film, x, y, fc_px_basic(c, fc));
  }
}

// at most 80 VGPRs, 6 waves per SIMD (X-Trans needed 96, 5 waves): the
// staging and splat phases' memory latency is hidden by the extra waves
template<int XT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) kb_develop(bt_launch L)
{
  __shared__ float gt[(BX + 6) * (BY * DEV_R + 6)];                 // green, halo r <= 3
  __shared__ float mt[(BX + 2 * DEV_M) * (BY * DEV_R + 2 * DEV_M)];  // mosaic, halo r + rs
  __shared__ co_args co;
  __shared__ fc_args fc;
  dimg mos, cov, film;
  int r = 1;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(co, im.co);
    lds_copy(fc, im.fc);
    mos = rfl(im.mos); cov = rfl(im.cov); film = rfl(im.film); r = rfl(im.dm_r);
  }, [&](const bt_job &jb, int lt) {
    const int X0 = (lt % jb.tiles_x) * BX, Y0 = (lt / jb.tiles_x) * (BY * DEV_R);
    develop_tile<XT>(mos, cov, film, co, fc, r, X0, Y0, gt, mt);
  });
}

// ---- llap ------------------------------------------------------------------

// level-1 tile (64 x 4) of curve + reduce: the footprint (2X0-1 .. 2X0+128,
// 2Y0-1 .. 2Y0+8, mirrored) of every level-0 plane is evaluated once into
// LDS as binary16 -- the values the curve node stores -- and reduced from
// there in reduce.comp's order
#define P1_W (2 * BX + 2)
#define P1_H (2 * BY + 1)
#define P1_N (P1_W * P1_H)

__global__ void __launch_bounds__(256) kb_pyr1(bt_launch L)
{
  extern __shared__ __half pl[];   // (num_gamma + 1) x P1_N
  __shared__ ll_args a;
  dimg f, o;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.ll);
    f = rfl(im.film); o = rfl(im.lp[1]);
  }, [&](const bt_job &jb, int lt) {
    const int ng = a.num_gamma;
    const int X0 = (lt % jb.tiles_x) * BX, Y0 = (lt / jb.tiles_x) * BY;
    const int tid = threadIdx.y * BX + threadIdx.x;
    for(int i = tid; i < P1_N; i += BX * BY)
    {
      const int ty = i / P1_W, tx = i - ty * P1_W;
      const float yl = ll_lum(ld_h4(f, mirror(2 * X0 - 1 + tx, f.wd), mirror(2 * Y0 - 1 + ty, f.ht)));
      for(int p = 0; p < ng; p++) pl[p * P1_N + i] = f2h(ll_curve(yl, a.gam[p], a));
      pl[ng * P1_N + i] = f2h(yl);
    }
    __syncthreads();
    const int x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
    if(x >= o.wd || y >= o.ht) return;
    const int c = 2 * threadIdx.y * P1_W + 2 * threadIdx.x;   // local (2x-1, 2y-1)
    for(int p = 0; p <= ng; p++)
    {
      const __half *q = pl + p * P1_N + c;
      const float s0 = h2f(q[0]) + 2.0f * h2f(q[1]) + h2f(q[2]);
      const float s1 = h2f(q[P1_W]) + 2.0f * h2f(q[P1_W + 1]) + h2f(q[P1_W + 2]);
      const float s2 = h2f(q[2 * P1_W]) + 2.0f * h2f(q[2 * P1_W + 1]) + h2f(q[2 * P1_W + 2]);
      st_h1(o, x, y, (s0 + 2.0f * s1 + s2) * (1.0f / 16.0f), p);
    }
  });
}

// reduce level l from l-1, four outputs per thread (256 x 4 per tile)
__global__ void __launch_bounds__(256) kb_reduce(bt_launch L)
{
  dimg in, o;
  int planes = 1;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    in = rfl(im.lp[jb.level - 1]); o = rfl(im.lp[jb.level]); planes = rfl(im.ll.num_gamma + 1);
  }, [&](const bt_job &jb, int lt) {
    const int x0 = ((lt % jb.tiles_x) * BX + threadIdx.x) * 4, y = (lt / jb.tiles_x) * BY + threadIdx.y;
    if(x0 < o.wd && y < o.ht) ll_reduce_px4(in, o, planes, x0, y, 2 | 4);
  });
}

// assemble level l (1 <= l <= nl-2) from level l+1
__global__ void __launch_bounds__(256) kb_asm(bt_launch L)
{
  extern __shared__ float tile[];
  __shared__ int lims[2];
  __shared__ ll_args a;
  dimg fine, l1, coarse, o;
  int first = 0;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    const int l = jb.level;
    lds_copy(a, im.ll);
    first = rfl((int)(l == im.nl - 2));
    fine = rfl(im.lp[l]); l1 = rfl(im.lp[l + 1]); coarse = rfl(first ? im.lp[l + 1] : im.la[l + 1]); o = rfl(im.la[l]);
  }, [&](const bt_job &jb, int lt) {
    const int X0 = (lt % jb.tiles_x) * BX, Y0 = (lt / jb.tiles_x) * (BY * ASM_R);
    const int x = X0 + threadIdx.x, ng = a.num_gamma;
    float v[ASM_R], res[ASM_R];
#pragma unroll
    for(int k = 0; k < ASM_R; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      v[k] = ld_h1(fine, min(x, fine.wd - 1), min(y, fine.ht - 1), ng);   // clamped: no result outside
    }
    const int xc = min(x, fine.wd - 1);
    ll_assemble_tile<ASM_R>(coarse, l1, fine.wd, fine.ht, a, first, X0, Y0, tile, lims, v,
                            [&](int k, int p) {
                              return ld_h1(fine, xc, min(Y0 + (int)threadIdx.y + BY * k, fine.ht - 1), p);
                            }, res);
#pragma unroll
    for(int k = 0; k < ASM_R; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < fine.wd && y < fine.ht) st_h1(o, x, y, res[k]);
    }
  });
}

// assemble level 0 + llap/colour: level 0 planes from the filmcurv output.
// At most 80 VGPRs (6 waves per SIMD instead of 5): its dependent phases
// (rows, plane range, staging) need the extra waves to hide memory latency
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) kb_final(bt_launch L)
{
  extern __shared__ float tile[];
  __shared__ int lims[2];
  __shared__ ll_args a;
  dimg f, l1, coarse, o;
  int first = 0;
  bt_tiles(L, [&](const bt_job &jb) {
    const bt_img &im = L.img[jb.img];
    lds_copy(a, im.ll);
    first = rfl((int)(im.nl == 2));
    f = rfl(im.film); l1 = rfl(im.lp[1]); coarse = rfl(first ? im.lp[1] : im.la[1]); o = rfl(im.out);
  }, [&](const bt_job &jb, int lt) {
    const int X0 = (lt % jb.tiles_x) * BX, Y0 = (lt / jb.tiles_x) * (BY * ASM_R);
    const int x = X0 + threadIdx.x;
    float4 t[ASM_R];
    float yl[ASM_R], v[ASM_R], res[ASM_R];
#pragma unroll
    for(int k = 0; k < ASM_R; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      t[k] = ld_h4(f, min(x, f.wd - 1), min(y, f.ht - 1));   // clamped: no result outside
      yl[k] = ll_lum(t[k]);
      v[k] = r16(yl[k]);
    }
    ll_assemble_tile<ASM_R>(coarse, l1, f.wd, f.ht, a, first, X0, Y0, tile, lims, v,
                            [&](int k, int p) { return r16(ll_curve(yl[k], a.gam[p], a)); }, res);
#pragma unroll
    for(int k = 0; k < ASM_R; k++)
    {
      const int y = Y0 + threadIdx.y + BY * k;
      if(x < f.wd && y < f.ht) st_f4(o, x, y, ll_colour_px(r16(res[k]), t[k]));
    }
  });
}

// ---- host ---------------------------------------------------------------------

namespace {

enum kind_t { K_HALF, K_DOWNCOV, K_DOWN, K_DNASM, K_DOUB, K_HL_HALF, K_HL_REDUCE, K_HL_ASM, K_HL_DOUB, K_GAUSS,
              K_DEVELOP, K_DEVELOP_XT, K_PYR1, K_REDUCE, K_ASM, K_FINAL };
const int kind_group[] = { 0, 1, 1, 2, 3, 4, 4, 4, 4, 5, 6, 6, 7, 8, 8, 9 };
const char *group_name[VKDT_BATCH_STAGES] = { "half", "wavelet", "dnasm", "doub", "hilite", "gauss", "develop",
                                              "pyr1", "pyramid", "final" };

struct launch_t
{
  int kind, level;
  int job0, njob, total;
  size_t lds;          // dynamic LDS bytes
  int ctr0;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
int    pitch8(int w) { return (w + 7) & ~7; }

// extents of one image's buffers
struct dims_t
{
  int W, H, qw, qh, cw, ch, nl, ng, block;
  int lw[BT_LEVELS], lh[BT_LEVELS];
  int hl, hn;                          // hilite on, its coarse levels L
  int hw[HL_LEVELS], hh[HL_LEVELS];    // hilite level extents (0: the block grid)
};

int image_dims(const vkdt_develop_image_t &d, dims_t &m, int idx)
{
  if(!d.filters)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: not a mosaic (filters 0)", idx);
  if(!(d.dn_params[0] > 0.0f))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: denoise strength <= 0 (noop path) not batched", idx);
  if(d.wd < 6 || d.ht < 6 || d.wd > (1u << 16) || d.ht > (1u << 16))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: extent %ux%u outside [6, 65536]", idx, d.wd, d.ht);
  if(!d.cfa_wd || !d.cfa_ht || d.cfa_pitch < d.cfa_wd)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: mosaic %ux%u pitch %u", idx, d.cfa_wd, d.cfa_ht, d.cfa_pitch);
  if(d.ll_numgamma < 2 || d.ll_numgamma > LL_MAX_GAMMA)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: num_gamma %d outside [2,%d]", idx, d.ll_numgamma, LL_MAX_GAMMA);
  if(d.ll_levels < 2 || d.ll_levels > BT_LEVELS)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: %d pyramid levels outside [2,%d]", idx, d.ll_levels, BT_LEVELS);
  m.W = d.wd; m.H = d.ht;
  m.block = d.filters == 9u ? 3 : 2;
  m.qw = m.W / m.block; m.qh = m.H / m.block;   // denoise/main.c: roi / block
  m.cw = m.W / m.block; m.ch = m.H / m.block;   // demosaic/main.c: wd / block
  if(m.qw < 1 || m.qh < 1) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: empty block grid", idx);
  m.nl = d.ll_levels;
  m.ng = d.ll_numgamma;
  m.lw[0] = m.W; m.lh[0] = m.H;
  for(int l = 1; l < m.nl; l++) { m.lw[l] = (m.lw[l - 1] - 1) / 2 + 1; m.lh[l] = (m.lh[l - 1] - 1) / 2 + 1; }
  m.hl = d.hl_on != 0;
  m.hn = 0;
  if(m.hl)
  { // hilite/main.c create_nodes: a reduce per level until the next coarser
    // extent reaches 1 or 14 levels exist
    if(m.W % m.block || m.H % m.block)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: hilite over %dx%d, not whole %dx%d blocks", idx,
                            m.W, m.H, m.block, m.block);
    m.hw[0] = m.qw; m.hh[0] = m.qh;
    for(int l = 1; l < HL_LEVELS; l++)
    {
      m.hw[l] = (m.hw[l - 1] - 1) / 2 + 1;
      m.hh[l] = (m.hh[l - 1] - 1) / 2 + 1;
      m.hn = l;
      if((m.hw[l] - 1) / 2 + 1 <= 1 || (m.hh[l] - 1) / 2 + 1 <= 1) break;
    }
    if(m.hn != d.hl_levels)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: %d hilite levels, the block grid %dx%d has %d", idx,
                            d.hl_levels, m.qw, m.qh, m.hn);
  }
  return VKDT_HIP_OK;
}

size_t image_scratch(const dims_t &m)
{
  size_t s = 0;
  const size_t q = align256((size_t)m.qw * m.qh * 8);
  s += 6 * q;                                              // qh, qd[4], qa
  s += align256((size_t)pitch8(m.W) * m.H * 2);            // mos
  s += align256((size_t)m.cw * m.ch * 8);                  // cov
  s += align256((size_t)m.W * m.H * 8);                    // film
  for(int l = 1; l < m.nl; l++)
    s += align256((size_t)(m.ng + 1) * pitch8(m.lw[l]) * m.lh[l] * 2);
  for(int l = 1; l <= m.nl - 2; l++)
    s += align256((size_t)pitch8(m.lw[l]) * m.lh[l] * 2);
  if(m.hl)
  {
    s += align256((size_t)pitch8(m.W) * m.H * 2);               // dmos
    for(int l = 0; l <= m.hn; l++) s += align256((size_t)m.hw[l] * m.hh[l] * 8);
    for(int l = 0; l < m.hn; l++) s += align256((size_t)m.hw[l] * m.hh[l] * 8);
  }
  return s;
}

dimg mk(void *p, int wd, int ht, int pitch, int chan, long pstride, int fmt = DIMG_F16)
{
  dimg d;
  d.p = p; d.wd = wd; d.ht = ht; d.pitch = pitch; d.chan = chan; d.pstride = pstride; d.fmt = fmt;
  return d;
}

} // namespace

struct vkdt_batch_t
{
  int n = 0, ncu = 256, max_nl = 0, max_ng = 0, max_hn = 0;
  std::vector<dims_t> dm;
  std::vector<bt_img> img;         // host copies of the device tables
  std::vector<bt_job> job;
  bt_img *d_img = nullptr;         // device tables (in the caller's scratch)
  bt_job *d_job = nullptr;
  int *d_ctr = nullptr;
  int nctr = 0;
  int dirty = 1;                   // tables to upload before the next run
  std::vector<launch_t> seq;
  int32_t tune[VKDT_BATCH_STAGES];
  asm_mats mats;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int timing = 0;
  hipEvent_t ev[VKDT_BATCH_STAGES + 1] = {};
  float ms[VKDT_BATCH_STAGES] = {};
};

// the device tables at the end of the scratch: image records, job lists
// (at most one job per image and launch) and the tile counters
#define BT_MAX_LAUNCH (10 + 2 * BT_LEVELS + 2 * HL_LEVELS)
static size_t table_bytes(int n)
{
  return align256(sizeof(bt_img) * n) + align256(sizeof(bt_job) * (size_t)BT_MAX_LAUNCH * n)
       + align256(sizeof(int) * NUM_XCD * BT_MAX_LAUNCH);
}

int vkdt_hip_batch_scratch_bytes(const vkdt_develop_image_t *img, int n, size_t *bytes)
{
  if(!bytes || n < 0 || (n && !img)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: bad arguments");
  size_t s = 0;
  for(int i = 0; i < n; i++)
  {
    dims_t m;
    if(int e = image_dims(img[i], m, i)) return e;
    s += image_scratch(m);
  }
  *bytes = s + table_bytes(n);
  return VKDT_HIP_OK;
}

static void dn_common(dn_args &a, const vkdt_develop_image_t &d)
{
  a = dn_args{};
  for(int k = 0; k < 4; k++)
  {
    a.wb[k] = d.wb[k];
    a.black[k] = d.black[k];
    a.white[k] = d.white[k];
    a.inv_range[k] = 1.0f / (d.white[k] - d.black[k]);
  }
  a.noise_a = d.noise_a;
  a.noise_b = d.noise_b;
  a.filters = d.filters;
  dn_params_init(a, d.dn_params);
}

static int build_image(bt_img &b, const vkdt_develop_image_t &d, const dims_t &m, char *&sp, int idx)
{
  b = bt_img{};
  b.cfa = mk(const_cast<void *>(d.cfa), d.cfa_wd, d.cfa_ht, d.cfa_pitch, 1, 0, DIMG_UI16);
  b.out = mk(d.out, m.W, m.H, m.W, 4, 0, DIMG_F32);
  const size_t q = align256((size_t)m.qw * m.qh * 8);
  b.qh = mk(sp, m.qw, m.qh, m.qw, 4, 0); sp += q;
  for(int l = 0; l < 4; l++) { b.qd[l] = mk(sp, m.qw, m.qh, m.qw, 4, 0); sp += q; }
  b.qa = mk(sp, m.qw, m.qh, m.qw, 4, 0); sp += q;
  b.mos = mk(sp, m.W, m.H, pitch8(m.W), 1, 0); sp += align256((size_t)pitch8(m.W) * m.H * 2);
  b.cov = mk(sp, m.cw, m.ch, m.cw, 4, 0); sp += align256((size_t)m.cw * m.ch * 8);
  b.film = mk(sp, m.W, m.H, m.W, 4, 0); sp += align256((size_t)m.W * m.H * 8);
  for(int l = 1; l < m.nl; l++)
  {
    const long ps = (long)pitch8(m.lw[l]) * m.lh[l];
    b.lp[l] = mk(sp, m.lw[l], m.lh[l], pitch8(m.lw[l]), 1, ps);
    sp += align256((size_t)(m.ng + 1) * ps * 2);
  }
  for(int l = 1; l <= m.nl - 2; l++)
  {
    b.la[l] = mk(sp, m.lw[l], m.lh[l], pitch8(m.lw[l]), 1, 0);
    sp += align256((size_t)pitch8(m.lw[l]) * m.lh[l] * 2);
  }
  b.dmos = b.mos;
  if(m.hl)
  {
    b.dmos = mk(sp, m.W, m.H, pitch8(m.W), 1, 0); sp += align256((size_t)pitch8(m.W) * m.H * 2);
    for(int l = 0; l <= m.hn; l++)
    {
      b.hq[l] = mk(sp, m.hw[l], m.hh[l], m.hw[l], 4, 0);
      sp += align256((size_t)m.hw[l] * m.hh[l] * 8);
    }
    for(int l = 0; l < m.hn; l++)
    {
      b.ha[l] = mk(sp, m.hw[l], m.hh[l], m.hw[l], 4, 0);
      sp += align256((size_t)m.hw[l] * m.hh[l] * 8);
    }
    b.ha[m.hn] = b.hq[m.hn];
    hl_args_init(b.hl, d.hl_wb, d.hl_params);
    // hilite/doub: ceil(wd/8) x ceil(ht/8) groups of 8x8 blocks, bounded by the mosaic's extent
    b.hl_nx = std::min((m.qw + 7) / 8 * 8, m.W);
    b.hl_ny = std::min((m.qh + 7) / 8 * 8, m.H);
  }
  b.nl = m.nl;
  b.xt = d.filters == 9u;
  b.dm_r = b.xt ? std::min(3, std::max(2, d.dm_fixup + 2)) : std::min(2, std::max(1, d.dm_fixup + 1));
  // denoise push constants per node (denoise/main.c create_nodes): half and
  // doub crop the mosaic, the wavelet levels and assemble run on the block grid
  dn_common(b.dn_half, d);
  for(int k = 0; k < 4; k++) b.dn_half.crop[k] = d.crop[k];
  for(int l = 0; l < 4; l++)
  {
    dn_common(b.dn_down[l], d);
    b.dn_down[l].level = l;
    b.dn_down[l].block = m.block;
    dn_level_scale(b.dn_down[l]);
  }
  dn_common(b.dn_asm, d);
  dn_asm_scales(b.dn_asm);
  dn_common(b.dn_doub, d);
  for(int k = 0; k < 4; k++) b.dn_doub.crop[k] = d.crop[k];
  if(const char *err = co_args_init(b.co, d.co_params))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: colour: %s", idx, err);
  if(const char *err = fc_args_init(b.fc, d.fc_params, sizeof(d.fc_params), true))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: filmcurv: %s", idx, err);
  ll_args_init(b.ll, d.ll_params, m.ng);
  return VKDT_HIP_OK;
}

static int is_pixel_stage(int kind)
{
  return kind == K_HALF || kind == K_DNASM || kind == K_DOUB || kind == K_GAUSS || kind == K_HL_HALF ||
         kind == K_HL_ASM || kind == K_HL_DOUB;
}

// launch shape of a stage group: tune = persist | rows << 8 (rows: per-pixel
// stages' rows per thread, 1..8, default 4)
static int tune_persist(int32_t t) { return t & 0xff; }
static int tune_rows(int32_t t) { const int r = (t >> 8) & 0xff; return r < 1 ? 4 : (r > 8 ? 8 : r); }

static int tiles_of(int kind, const dims_t &m, int level, int rows, int &tx)
{
  int w, h, tw = BX, th = BY;
  switch(kind)
  {
    case K_HALF: case K_DNASM: w = m.qw; h = m.qh; th = BY * rows; break;
    case K_DOWNCOV: w = m.qw; h = m.qh; th = BY * COV_R; break;
    case K_DOWN: w = m.qw; h = m.qh; th = BY * DOWN_R; break;
    case K_DOUB: w = m.block == 2 ? (m.W + 1) / 2 : m.W; h = m.block == 2 ? (m.H + 1) / 2 : m.H; th = BY * rows; break;
    case K_HL_HALF: w = m.qw; h = m.qh; th = BY * rows; break;
    case K_HL_REDUCE: w = m.hw[level]; h = m.hh[level]; break;
    case K_HL_ASM: w = m.hw[level]; h = m.hh[level]; th = BY * rows; break;
    case K_HL_DOUB: w = std::min((m.qw + 7) / 8 * 8, m.W); h = std::min((m.qh + 7) / 8 * 8, m.H); th = BY * rows; break;
    case K_GAUSS: w = m.cw; h = m.ch; th = BY * rows; break;
    case K_DEVELOP: case K_DEVELOP_XT: w = m.W; h = m.H; th = BY * DEV_R; break;
    case K_PYR1: w = m.lw[1]; h = m.lh[1]; break;
    case K_REDUCE: w = m.lw[level]; h = m.lh[level]; tw = 4 * BX; break;
    case K_ASM: w = m.lw[level]; h = m.lh[level]; th = BY * ASM_R; break;
    default: w = m.W; h = m.H; th = BY * ASM_R; break;   // K_FINAL
  }
  tx = (w + tw - 1) / tw;
  return tx * ((h + th - 1) / th);
}

// launch sequence and job tables for the current tune
static void build_launches(vkdt_batch_t *b)
{
  b->seq.clear();
  b->job.clear();
  b->nctr = 0;
  auto add = [&](int kind, int level, size_t lds) {
    const int rows = is_pixel_stage(kind) ? tune_rows(b->tune[kind_group[kind]]) : 1;
    launch_t L = { kind, level, (int)b->job.size(), 0, 0, lds, b->nctr };
    for(int i = 0; i < b->n; i++)
    {
      const dims_t &m = b->dm[i];
      if(kind == K_REDUCE && level >= m.nl) continue;
      if(kind == K_ASM && level > m.nl - 2) continue;
      if(kind == K_DEVELOP && m.block == 3) continue;      // Bayer and X-Trans develop
      if(kind == K_DEVELOP_XT && m.block != 3) continue;   // run as two launches
      if(kind >= K_HL_HALF && kind <= K_HL_DOUB && !m.hl) continue;
      if(kind == K_HL_REDUCE && level > m.hn) continue;
      if(kind == K_HL_ASM && level >= m.hn) continue;
      int tx;
      const int t = tiles_of(kind, m, level, rows, tx);
      if(t <= 0) continue;
      b->job.push_back(bt_job{ i, tx, L.total, level });
      L.total += t;
      L.njob++;
    }
    if(L.njob) { b->seq.push_back(L); b->nctr += NUM_XCD; }
  };
  const size_t asm_lds = sizeof(float) * (b->max_ng + 1) * AS_TN(ASM_R);
  add(K_HALF, 0, 0);
  add(K_DOWNCOV, 0, 0);
  for(int l = 1; l < 4; l++) add(K_DOWN, l, 0);
  add(K_DNASM, 0, 0);
  add(K_DOUB, 0, 0);
  add(K_HL_HALF, 0, 0);
  for(int l = 1; l <= b->max_hn; l++) add(K_HL_REDUCE, l, 0);
  for(int l = b->max_hn - 1; l >= 0; l--) add(K_HL_ASM, l, 0);
  add(K_HL_DOUB, 0, 0);
  add(K_GAUSS, 0, 0);
  add(K_DEVELOP, 0, 0);
  add(K_DEVELOP_XT, 0, 0);
  add(K_PYR1, 1, sizeof(__half) * (b->max_ng + 1) * P1_N);
  for(int l = 2; l < b->max_nl; l++) add(K_REDUCE, l, 0);
  for(int l = b->max_nl - 2; l >= 1; l--) add(K_ASM, l, asm_lds);
  add(K_FINAL, 0, asm_lds);
  b->dirty = 1;
}

int vkdt_hip_batch_create(vkdt_batch_t **out, const vkdt_develop_image_t *img, int n, void *scratch,
                          size_t scratch_bytes, const int32_t *tune)
{
  if(!out || n <= 0 || !img) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: no images");
  *out = nullptr;
  std::vector<dims_t> dm(n);
  size_t need = 0;
  int max_nl = 0, max_ng = 0, max_hn = 0;
  for(int i = 0; i < n; i++)
  {
    if(int e = image_dims(img[i], dm[i], i)) return e;
    if(!img[i].cfa || !img[i].out) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch image %d: no mosaic or output", i);
    need += image_scratch(dm[i]);
    max_nl = std::max(max_nl, dm[i].nl);
    max_ng = std::max(max_ng, dm[i].ng);
    max_hn = std::max(max_hn, dm[i].hn);
  }
  const size_t tables = need;
  need += table_bytes(n);
  if(!scratch || scratch_bytes < need || ((uintptr_t)scratch & 255))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: scratch %zu bytes at %p, need %zu (256-byte aligned)",
                          scratch_bytes, scratch, need);
  std::vector<bt_img> h(n);
  char *sp = (char *)scratch;
  for(int i = 0; i < n; i++)
    if(int e = build_image(h[i], img[i], dm[i], sp, i)) return e;

  vkdt_batch_t *b = new vkdt_batch_t;
  b->n = n;
  b->dm = dm;
  b->img = h;
  b->max_nl = max_nl;
  b->max_ng = max_ng;
  b->max_hn = max_hn;
  for(int s = 0; s < VKDT_BATCH_STAGES; s++) b->tune[s] = tune ? tune[s] : (4 << 8);
  asm_mats_init(b->mats);
  int dev = 0;
  if(hipGetDevice(&dev) != hipSuccess) dev = 0;
  if(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || b->ncu <= 0)
    b->ncu = 256;
  char *tp = (char *)scratch + tables;
  b->d_img = (bt_img *)tp;
  tp += align256(sizeof(bt_img) * n);
  b->d_job = (bt_job *)tp;
  tp += align256(sizeof(bt_job) * (size_t)BT_MAX_LAUNCH * n);
  b->d_ctr = (int *)tp;
  build_launches(b);
  *out = b;
  return VKDT_HIP_OK;
}

// the tables go up in stream order before the first launch that reads them
static int upload_tables(vkdt_batch_t *b, hipStream_t s)
{
  if(!b->dirty) return VKDT_HIP_OK;
  if(int e = hip_fail(hipMemcpyAsync(b->d_img, b->img.data(), sizeof(bt_img) * b->n, hipMemcpyHostToDevice, s),
                      "batch tables")) return e;
  if(int e = hip_fail(hipMemcpyAsync(b->d_job, b->job.data(), sizeof(bt_job) * b->job.size(), hipMemcpyHostToDevice, s),
                      "batch tables")) return e;
  b->dirty = 0;
  return VKDT_HIP_OK;
}

static int enqueue(vkdt_batch_t *b, hipStream_t s, int timed)
{
  hipLaunchKernelGGL(kb_reset, dim3((b->nctr + 255) / 256), dim3(256), 0, s, b->d_ctr, b->nctr);
  int group = -1;
  for(const launch_t &L : b->seq)
  {
    const int g = kind_group[L.kind];
    if(timed && g != group)
    {
      for(int k = group + 1; k <= g; k++)
        if(int e = hip_fail(hipEventRecord(b->ev[k], s), "batch timing")) return e;
      group = g;
    }
    const int persist = tune_persist(b->tune[g]);
    const int rows = is_pixel_stage(L.kind) ? tune_rows(b->tune[g]) : 1;
    bt_launch a = { b->d_img, b->d_job + L.job0, L.njob, L.total, b->d_ctr + L.ctr0, persist, rows };
    const dim3 grid(persist ? b->ncu * persist : L.total), blk(BX, BY);
    switch(L.kind)
    {
      case K_HALF:    hipLaunchKernelGGL(kb_half, grid, blk, 0, s, a); break;
      case K_DOWNCOV: hipLaunchKernelGGL(kb_downcov, grid, blk, 0, s, a); break;
      case K_DOWN:    hipLaunchKernelGGL(kb_down, grid, blk, 0, s, a); break;
      case K_DNASM:   hipLaunchKernelGGL(kb_dnasm, grid, blk, 0, s, a, b->mats); break;
      case K_DOUB:    hipLaunchKernelGGL(kb_doub, grid, blk, 0, s, a); break;
      case K_HL_HALF: hipLaunchKernelGGL(kb_hl_half, grid, blk, 0, s, a); break;
      case K_HL_REDUCE: hipLaunchKernelGGL(kb_hl_reduce, grid, blk, 0, s, a); break;
      case K_HL_ASM:  hipLaunchKernelGGL(kb_hl_asm, grid, blk, 0, s, a); break;
      case K_HL_DOUB: hipLaunchKernelGGL(kb_hl_doub, grid, blk, 0, s, a); break;
      case K_GAUSS:   hipLaunchKernelGGL(kb_gauss, grid, blk, 0, s, a); break;
      case K_DEVELOP: hipLaunchKernelGGL(kb_develop<0>, grid, blk, 0, s, a); break;
      case K_DEVELOP_XT: hipLaunchKernelGGL(kb_develop<1>, grid, blk, 0, s, a); break;
      case K_PYR1:    hipLaunchKernelGGL(kb_pyr1, grid, blk, L.lds, s, a); break;
      case K_REDUCE:  hipLaunchKernelGGL(kb_reduce, grid, blk, 0, s, a); break;
      case K_ASM:     hipLaunchKernelGGL(kb_asm, grid, blk, L.lds, s, a); break;
      default:        hipLaunchKernelGGL(kb_final, grid, blk, L.lds, s, a); break;
    }
    if(int e = hip_fail(hipGetLastError(), group_name[g])) return e;
  }
  if(timed)
    for(int k = group + 1; k <= VKDT_BATCH_STAGES; k++)
      if(int e = hip_fail(hipEventRecord(b->ev[k], s), "batch timing")) return e;
  return VKDT_HIP_OK;
}

static void drop_graph(vkdt_batch_t *b)
{
  if(b->exec) (void)hipGraphExecDestroy(b->exec);
  if(b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int vkdt_hip_batch_run(vkdt_batch_t *b, void *stream, int graph)
{
  if(!b) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: null");
  hipStream_t s = (hipStream_t)stream;
  if(int e = upload_tables(b, s)) return e;
  if(!graph || !s || b->timing)
  { // eager (timing records events between the stage groups)
    if(int e = enqueue(b, s, b->timing)) return e;
    if(b->timing)
    {
      if(int e = hip_fail(hipEventSynchronize(b->ev[VKDT_BATCH_STAGES]), "batch timing")) return e;
      for(int k = 0; k < VKDT_BATCH_STAGES; k++)
        if(int e = hip_fail(hipEventElapsedTime(&b->ms[k], b->ev[k], b->ev[k + 1]), "batch timing")) return e;
    }
    return VKDT_HIP_OK;
  }
  if(!b->exec)
  {
    if(int e = hip_fail(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "batch capture")) return e;
    const int rc = enqueue(b, s, 0);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(s, &g);
    if(rc) { if(g) (void)hipGraphDestroy(g); return rc; }
    if(int e = hip_fail(ee, "batch end capture")) return e;
    b->graph = g;
    if(int e = hip_fail(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "batch graph instantiate"))
    {
      drop_graph(b);
      return e;
    }
  }
  return hip_fail(hipGraphLaunch(b->exec, s), "batch graph launch");
}

int vkdt_hip_batch_set_timing(vkdt_batch_t *b, int on)
{
  if(!b) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: null");
  for(int k = 0; on && k <= VKDT_BATCH_STAGES; k++)
    if(!b->ev[k])
      if(int e = hip_fail(hipEventCreate(&b->ev[k]), "batch timing")) return e;
  b->timing = on != 0;
  return VKDT_HIP_OK;
}

int vkdt_hip_batch_stage_ms(vkdt_batch_t *b, float *ms)
{
  if(!b || !ms) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: null");
  for(int k = 0; k < VKDT_BATCH_STAGES; k++) ms[k] = b->ms[k];
  return VKDT_HIP_OK;
}

const char *vkdt_hip_batch_stage_name(int stage)
{
  return stage >= 0 && stage < VKDT_BATCH_STAGES ? group_name[stage] : "";
}

int vkdt_hip_batch_autotune(vkdt_batch_t *b, void *stream, int32_t *tune)
{
  if(!b) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch: null");
  // launch shape per stage group: persistence (0 = one tile per work
  // group, k = k persistent work groups per CU pulling from the tile
  // queues) and, for the per-pixel stages, rows per thread
  static const int pers[] = { 0, 1, 2, 4, 8 };
  static const int rows[] = { 1, 2, 4, 8 };
  const int nc = 12;
  float best[VKDT_BATCH_STAGES];
  int32_t pick[VKDT_BATCH_STAGES];
  for(int k = 0; k < VKDT_BATCH_STAGES; k++) { best[k] = 1e30f; pick[k] = b->tune[k]; }
  const int was = b->timing;
  if(int e = vkdt_hip_batch_set_timing(b, 1)) return e;
  drop_graph(b);
  int rc = VKDT_HIP_OK;
  for(int c = 0; c < nc && !rc; c++)
  {
    int32_t cand[VKDT_BATCH_STAGES];
    for(int k = 0; k < VKDT_BATCH_STAGES; k++)
    {
      const int pix = k == kind_group[K_HALF] || k == kind_group[K_DNASM] || k == kind_group[K_DOUB] ||
                      k == kind_group[K_GAUSS] || k == kind_group[K_HL_HALF];
      cand[k] = pix ? (pers[(c % 3) * 2] | rows[c / 3] << 8) : (pers[c % 5] | 4 << 8);
      b->tune[k] = cand[k];
    }
    build_launches(b);
    for(int rep = 0; rep < 3 && !rc; rep++)
    {
      rc = vkdt_hip_batch_run(b, stream, 0);
      if(rc || rep == 0) continue;   // first run warms
      for(int k = 0; k < VKDT_BATCH_STAGES; k++)
        if(b->ms[k] < best[k]) { best[k] = b->ms[k]; pick[k] = cand[k]; }
    }
  }
  b->timing = was;
  for(int k = 0; k < VKDT_BATCH_STAGES; k++) b->tune[k] = pick[k];
  build_launches(b);
  if(tune) for(int k = 0; k < VKDT_BATCH_STAGES; k++) tune[k] = pick[k];
  return rc;
}

void vkdt_hip_batch_destroy(vkdt_batch_t *b)
{
  if(!b) return;
  drop_graph(b);
  for(int s = 0; s <= VKDT_BATCH_STAGES; s++) if(b->ev[s]) (void)hipEventDestroy(b->ev[s]);
  delete b;
}

// ---------------------------------------------------------------------------
// A batch in the library's own device scratch, for callers that cannot
// allocate.  `scratch` is a file-scope device array of VKDT_BATCH_STATIC_BYTES;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (kept per device), never through its name.  Every
// call on a device shares that device's array: calls that overlap on one
// device race, and a call leaves its intermediates behind (the tile
// counters, the only state that must start at zero, are reset by a kernel
// at the start of each run).
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[VKDT_BATCH_STATIC_BYTES];

#define STATIC_MAX_DEV 64
static void *g_static_ptr[STATIC_MAX_DEV];
static std::mutex g_static_mx;

static int static_scratch(void **ptr)
{
  int dev = 0;
  if(int e = hip_fail(hipGetDevice(&dev), "batch static: hipGetDevice")) return e;
  if(dev < 0 || dev >= STATIC_MAX_DEV) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch static: device %d", dev);
  std::lock_guard<std::mutex> lk(g_static_mx);
  if(!g_static_ptr[dev])
  { // once per device, with that device current
    void *p = nullptr;
    if(int e = hip_fail(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "batch static: hipGetSymbolAddress")) return e;
    g_static_ptr[dev] = p;
  }
  *ptr = g_static_ptr[dev];
  return VKDT_HIP_OK;
}

int vkdt_hip_batch_develop_static(const vkdt_develop_image_t *img, int n, void *stream)
{
  size_t need = 0;
  if(int e = vkdt_hip_batch_scratch_bytes(img, n, &need)) return e;
  if(need > (size_t)VKDT_BATCH_STATIC_BYTES)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "batch static: the images need %zu bytes of scratch, the static "
                          "scratch holds %zu", need, (size_t)VKDT_BATCH_STATIC_BYTES);
  void *base = nullptr;
  if(int e = static_scratch(&base)) return e;
  vkdt_batch_t *b = nullptr;
  if(int e = vkdt_hip_batch_create(&b, img, n, base, VKDT_BATCH_STATIC_BYTES, nullptr)) return e;
  int e = vkdt_hip_batch_run(b, stream, 0);
  // the tables go up from the batch's host memory: keep it until the stream has drained
  const int es = hip_fail(hipStreamSynchronize((hipStream_t)stream), "batch static: sync");
  vkdt_hip_batch_destroy(b);
  return e ? e : es;
}
