// k_mcraw.hip -- the GPU codec of MotionCam raw video frames, the i-mcraw
// module's nodes (src/pipe/modules/i-mcraw): `prefix` turns the blocks' bit
// counts into byte offsets of their data, `decode` unpacks the blocks into
// the ui16 mosaic.
//
// A frame is cut into blocks of 64 samples.  Block i has a 4-bit count
// (two per byte of `bitcnt`, low nibble first), a 16-bit reference value and
// count-dependent data: the 64 offsets from the reference at `bits` bits
// each (count 7 stores 8 bits, 9 stores 10, 11..16 store 16), laid out for
// eight 16-bit SIMD registers of eight lanes (decode.comp:21-290).  The data
// of the blocks are concatenated in block order, so block i starts where the
// inclusive prefix sum of the block sizes up to i-1 ends (prefix.comp).
//
// Four consecutive blocks cover a 64x4 tile of the frame: block 4k+q holds
// the sites of CFA phase (q&1, q>>1) of rows 0 and 2 of tile k, 32 per row
// (decode.comp:314-337).  The samples are stored as R16 UNORM of v/65536, as
// the reference's image store does: a sample of 32769 or more comes back one
// less (the conversion rounds v * 65535/65536 to nearest, ties to even).
#include "launch.h"

#include <algorithm>

// bytes of one block's data by its count (prefix.comp:40-55)
DEV uint32_t mc_block_bytes(uint32_t cnt)
{
  if(cnt <= 6) return 8 * cnt;
  if(cnt <= 8) return 64;
  if(cnt <= 10) return 80;
  return 128;
}

DEV uint32_t mc_count(const uint8_t *bits, uint32_t i)
{
  uint32_t r = bits[i >> 1];
  if(i & 1) r >>= 4;
  return r & 15;
}

// ---- prefix: inclusive prefix sum of the block sizes ----------------------
// two launches, no waits between groups: k_mc_part_sums writes the data
// bytes of each part of 4096 blocks into the scratch connector; k_mc_prefix
// scans its part locally and adds the sum of the parts before it (at most a
// few dozen words, read by every group)

#define MC_ROWS 16
#define MC_WG 256
#define MC_PART (MC_WG * MC_ROWS)

DEV uint32_t block_inclusive_scan(uint32_t v, uint32_t *sh)
{ // Hillis-Steele over the work group; returns the inclusive sum at this thread
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for(int d = 1; d < MC_WG; d <<= 1)
  {
    const uint32_t o = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += o;
    __syncthreads();
  }
  return sh[t];
}

// this thread's 16 blocks of the part: inclusive sums of their data bytes
// (vec: the counts are 8-byte aligned for one load)
DEV uint32_t mc_thread_blocks(const uint8_t *bits, uint32_t base, uint32_t n, int vec, uint32_t loc[MC_ROWS])
{
  uint32_t acc = 0;
  if(vec && base + MC_ROWS <= n)
  {
    const uint2 q = *(const uint2 *)(bits + base / 2);
#pragma unroll
    for(int r = 0; r < MC_ROWS; r++)
    {
      acc += mc_block_bytes(((r < 8 ? q.x : q.y) >> (4 * (r & 7))) & 15u);
      loc[r] = acc;
    }
  }
  else
  {
#pragma unroll
    for(int r = 0; r < MC_ROWS; r++)
    {
      const uint32_t i = base + r;
      acc += i < n ? mc_block_bytes(mc_count(bits, i)) : 0u;
      loc[r] = acc;
    }
  }
  return acc;
}

__global__ void __launch_bounds__(MC_WG) k_mc_part_sums(const uint8_t *bits, uint32_t *part, uint32_t n, int vec)
{
  __shared__ uint32_t sh[MC_WG];
  uint32_t loc[MC_ROWS];
  const uint32_t acc = mc_thread_blocks(bits, blockIdx.x * MC_PART + threadIdx.x * MC_ROWS, n, vec, loc);
  block_inclusive_scan(acc, sh);
  if(threadIdx.x == MC_WG - 1) part[blockIdx.x] = sh[MC_WG - 1];
}

// vec: bits 8-byte and off 16-byte aligned (vector loads and stores)
__global__ void __launch_bounds__(MC_WG) k_mc_prefix(const uint8_t *bits, uint32_t *off, const uint32_t *part,
                                                     uint32_t n, int vec)
{
  __shared__ uint32_t sh[MC_WG];
  // the parts before this one
  uint32_t carry = 0;
  for(uint32_t j = threadIdx.x; j < blockIdx.x; j += MC_WG) carry += part[j];
  block_inclusive_scan(carry, sh);
  carry = sh[MC_WG - 1];
  __syncthreads();
  const uint32_t base = blockIdx.x * MC_PART + threadIdx.x * MC_ROWS;
  uint32_t loc[MC_ROWS];
  const uint32_t acc = mc_thread_blocks(bits, base, n, vec, loc);
  const uint32_t excl = block_inclusive_scan(acc, sh) - acc + carry;
  if(vec && base + MC_ROWS <= n)
  {
    uint4 *o = (uint4 *)(off + base);
#pragma unroll
    for(int r = 0; r < MC_ROWS; r += 4)
      o[r / 4] = make_uint4(loc[r] + excl, loc[r + 1] + excl, loc[r + 2] + excl, loc[r + 3] + excl);
  }
  else
  {
#pragma unroll
    for(int r = 0; r < MC_ROWS; r++)
      if(base + r < n) off[base + r] = loc[r] + excl;
  }
}

// ---- decode ------------------------------------------------------------------
// The reference's decoder reads the values of two adjacent register lanes
// 2s and 2s + 1 as a pair, from bytes `reg` and `reg + 1` of the block at the
// same bit offset.  Here a block is decoded by a whole wave, one value a
// lane: h = 0 the pair's first (byte reg), h = 1 its second (byte reg + 1).
// The decoders below are written for the pair and return the lane's value
// in both components.  Bytes past the data buffer read as 0.
struct mc_block
{
  const uint8_t *d;
  uint32_t off, len;
  int h;
  __device__ uint32_t at(int r) const { return off + (uint32_t)r < len ? d[off + r] : 0u; }
  __device__ uint2 pair(int reg, int bit, int n) const
  {
    const uint32_t m = (1u << n) - 1u, v = (at(reg + h) >> bit) & m;
    return make_uint2(v, v);
  }
};

DEV uint2 operator|(uint2 a, uint2 b) { return make_uint2(a.x | b.x, a.y | b.y); }
DEV uint2 shl(uint2 a, int s) { return make_uint2(a.x << s, a.y << s); }

// decode.comp:38-290, one function per stored width
DEV uint2 mc_decode1(const mc_block &b, int s) { return b.pair(2 * (s % 4), s / 4, 1); }
DEV uint2 mc_decode2(const mc_block &b, int s) { return b.pair(2 * (s % 8), 2 * (s / 8), 2); }

DEV uint2 mc_decode3(const mc_block &b, int s)
{ // rows of 4 lanes read 3 bits at offsets 0, 3, 6 of bytes 0..7, 8..15,
  // 16..23; the offset-6 groups take their third bit from bytes 16..23
  // (bit 6 for bytes 0..7, bit 7 for bytes 8..15)
  const int g = s / 4, q = s % 4;
  const int reg = (g < 3 ? 0 : (g < 6 ? 8 : 16)) + 2 * q;
  const int off = 3 * (g % 3);
  uint2 r = b.pair(reg, off, 3);
  if(g == 2) r = r | shl(b.pair(16 + 2 * q, 6, 1), 2);
  else if(g == 5) r = r | shl(b.pair(16 + 2 * q, 7, 1), 2);
  return r;
}

DEV uint2 mc_decode4(const mc_block &b, int s)
{
  const int bit = ((2 * s / 8) & 1) == 0 ? 0 : 4;
  const int reg = (2 * s / 16) * 8 + ((2 * s) % 8);
  return b.pair(reg, bit, 4);
}

DEV uint2 mc_decode5(const mc_block &b, int s)
{ // five low bits of bytes 0..39, the three high bits of bytes 0..23 and
  // the top bits of 24..39 complete lanes 20..31
  if(s < 20) return b.pair(2 * s, 0, 5);
  if(s < 28)
  {
    const int rl = s < 24 ? 2 * (s - 20) : 8 + 2 * (s - 24), ru = s < 24 ? 24 + 2 * (s - 20) : 32 + 2 * (s - 24);
    return b.pair(rl, 5, 3) | shl(b.pair(ru, 5, 2), 3);
  }
  const int rl = 16 + 2 * (s - 28), rm = 24 + 2 * (s - 28), rh = 32 + 2 * (s - 28);
  return b.pair(rl, 5, 3) | shl(b.pair(rm, 7, 1), 3) | shl(b.pair(rh, 7, 1), 4);
}

DEV uint2 mc_decode6(const mc_block &b, int s)
{ // six low bits of bytes 0..47, lanes 24..31 from the two high bits of
  // three bytes each
  if(s < 24) return b.pair(2 * s, 0, 6);
  const int q = s < 28 ? s - 24 : s - 28, r0 = s < 28 ? 0 : 24;
  return b.pair(r0 + 2 * q, 6, 2) | shl(b.pair(r0 + 8 + 2 * q, 6, 2), 2) | shl(b.pair(r0 + 16 + 2 * q, 6, 2), 4);
}

DEV uint2 mc_decode8(const mc_block &b, int s) { const uint32_t v = b.at(2 * s + b.h); return make_uint2(v, v); }

DEV uint2 mc_decode10(const mc_block &b, int s)
{ // eight low bits from bytes 0..31 / 40..71, two high bits from 32..39 / 72..79
  const int rl = s < 16 ? 2 * s : 40 + 2 * (s - 16);
  const int rh = (s < 16 ? 32 : 72) + 2 * (s % 4);
  const int off = 2 * ((s & 15) / 4);
  const uint32_t lo = b.at(rl + b.h);
  return make_uint2(lo, lo) | shl(b.pair(rh, off, 2), 8);
}

DEV uint2 mc_decode16(const mc_block &b, int s)
{ // little-endian 16-bit values
  const uint32_t v = b.at(4 * s + 2 * b.h) | (b.at(4 * s + 2 * b.h + 1) << 8);
  return make_uint2(v, v);
}

DEV uint2 mc_decode(const mc_block &b, uint32_t cnt, int s)
{
  switch(cnt)
  {
    case 0: return make_uint2(0u, 0u);
    case 1: return mc_decode1(b, s);
    case 2: return mc_decode2(b, s);
    case 3: return mc_decode3(b, s);
    case 4: return mc_decode4(b, s);
    case 5: return mc_decode5(b, s);
    case 6: return mc_decode6(b, s);
    case 7: case 8: return mc_decode8(b, s);
    case 9: case 10: return mc_decode10(b, s);
    default: return mc_decode16(b, s);
  }
}

struct mc_args { int ox, oy, owd, oht; uint32_t nblocks, data_len; };

// R16 UNORM of v/65536, round(v * 65535/65536) to nearest with ties to
// even, in exact arithmetic: v up to 32768 stays, larger v lose one, and
// v/65536 >= 1 saturates
DEV uint32_t mc_unorm(uint32_t v) { return v >= 65536u ? 65535u : (v <= 32768u ? v : v - 1u); }

// A wave decodes one 64x4 tile, its four blocks 4k..4k+3 in turn, so that
// every lane of the wave takes the same count's path (decode.comp:314-337):
// block 4k+q holds CFA phase (q&1, q>>1) of rows 0 and 2; its register lane
// 2s+h (lane 2s+h of the wave, s < 16 on row 0, s >= 16 on row 2) lands at
// column 4(s%16) + 2h.  So each lane ends with two adjacent columns of two
// rows and stores each row as one 4-byte store where the destination is
// aligned and inside the image.  The loads of the four blocks' counts,
// offsets and references are issued before their data loads.
__global__ void __launch_bounds__(256) k_mc_decode(dimg out, const uint8_t *__restrict__ bits,
                                                   const uint32_t *__restrict__ off, const uint16_t *__restrict__ ref,
                                                   const uint8_t *__restrict__ dat, mc_args a)
{
  const uint32_t k = blockIdx.x * 4 + threadIdx.x / 64;
  const uint32_t b0 = 4 * k;
  if(b0 >= a.nblocks) return;
  const int l = threadIdx.x % 64, s = l / 2, h = l % 2;
  const int nq = (int)min(4u, a.nblocks - b0);
  uint32_t cnt[4], start[4], rv[4];
#pragma unroll
  for(int q = 0; q < 4; q++)
  {
    const uint32_t i = b0 + q;
    const bool have = q < nq;
    cnt[q] = have ? mc_count(bits, i) : 0u;
    start[q] = have && i > 0 ? off[i - 1] : 0u;
    rv[q] = have ? ref[i] : 0u;
  }
  uint32_t v[4];
#pragma unroll
  for(int q = 0; q < 4; q++)
  {
    const mc_block b = { dat, start[q], a.data_len, h };
    v[q] = q < nq ? mc_unorm(mc_decode(b, cnt[q], s).x + rv[q]) : 0u;
  }
  const uint32_t bpr = (uint32_t)a.owd / 64;
  const int x = 64 * (int)(k % bpr) + 4 * (s % 16) + 2 * h - a.ox;
  const int y = 4 * (int)(k / bpr) + (s >= 16 ? 2 : 0) - a.oy;
  uint16_t *o = (uint16_t *)out.p;
#pragma unroll
  for(int r = 0; r < 2; r++)
  { // row y + r: phase 2r at column x, phase 2r + 1 at x + 1
    const int yy = y + r;
    if(yy < 0 || yy >= out.ht) continue;
    const long at = (long)yy * out.pitch + x;
    if(2 * r + 1 < nq && x >= 0 && x + 1 < out.wd && ((((uintptr_t)o) >> 1) + at) % 2 == 0)
      *(uint32_t *)(o + at) = v[2 * r] | (v[2 * r + 1] << 16);
    else
    {
      if(2 * r < nq && x >= 0 && x < out.wd) o[at] = (uint16_t)v[2 * r];
      if(2 * r + 1 < nq && x + 1 >= 0 && x + 1 < out.wd) o[at + 1] = (uint16_t)v[2 * r + 1];
    }
  }
}

// ---- launchers ---------------------------------------------------------------

static size_t conn_elems(const vkdt_buf_t &b) { return (size_t)b.pitch * b.ht; }

int launch_mcraw_prefix(const node_call &c)
{
  if(int e = check_conn(c, 0, VKDT_FMT_UI8, 1, 1, "i-mcraw/prefix bitcnt")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_UI32, 1, 1, "i-mcraw/prefix offset")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_UI32, 1, 1, "i-mcraw/prefix scratch")) return e;
  // the offsets of every block the offset buffer holds (its extent is the
  // block count rounded up to 64), counted in the bit counts buffer
  const uint32_t n = (uint32_t)conn_elems(c.conn[1]);
  if(conn_elems(c.conn[0]) * 2 < n)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "i-mcraw/prefix: %u blocks, %zu bit counts", n,
                          conn_elems(c.conn[0]) * 2);
  const uint32_t np = (n + MC_PART - 1) / MC_PART;
  if(conn_elems(c.conn[2]) < np)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "i-mcraw/prefix: scratch holds %zu of %u partials",
                          conn_elems(c.conn[2]), np);
  const uint8_t *bits = (const uint8_t *)c.conn[0].ptr;
  uint32_t *off = (uint32_t *)c.conn[1].ptr, *part = (uint32_t *)c.conn[2].ptr;
  const int vec = ((uintptr_t)bits % 8 == 0) && ((uintptr_t)off % 16 == 0);
  hipLaunchKernelGGL(k_mc_part_sums, dim3(np), dim3(MC_WG), 0, c.stream, bits, part, n, vec);
  hipLaunchKernelGGL(k_mc_prefix, dim3(np), dim3(MC_WG), 0, c.stream, bits, off, (const uint32_t *)part, n, vec);
  return launch_post("i-mcraw/prefix");
}

int launch_mcraw_decode(const node_call &c)
{
  if(int e = check_blob(c, 4 * 4, 0, "i-mcraw/decode")) return e;
  if(int e = check_conn(c, 0, VKDT_FMT_UI16, 1, 1, "i-mcraw/decode output")) return e;
  if(int e = check_conn(c, 1, VKDT_FMT_UI8, 1, 1, "i-mcraw/decode bitcnt")) return e;
  if(int e = check_conn(c, 2, VKDT_FMT_UI32, 1, 1, "i-mcraw/decode offset")) return e;
  if(int e = check_conn(c, 3, VKDT_FMT_UI16, 1, 1, "i-mcraw/decode refval")) return e;
  if(int e = check_conn(c, 4, VKDT_FMT_UI8, 1, 1, "i-mcraw/decode datain")) return e;
  const int32_t *pc = (const int32_t *)c.pc;
  mc_args a = { pc[0], pc[1], pc[2], pc[3], 0, 0 };
  if(a.owd <= 0 || a.oht <= 0 || a.owd % 64)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "i-mcraw/decode: frame %dx%d (width a multiple of 64)", a.owd, a.oht);
  a.nblocks = (uint32_t)((uint64_t)a.owd * a.oht / 64);
  if(conn_elems(c.conn[1]) * 2 < a.nblocks || conn_elems(c.conn[2]) < a.nblocks || conn_elems(c.conn[3]) < a.nblocks)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "i-mcraw/decode: %u blocks exceed the bit count, offset or "
                          "reference buffers", a.nblocks);
  a.data_len = (uint32_t)std::min<size_t>(conn_elems(c.conn[4]), 0xffffffffu);
  const uint32_t ntiles = (a.nblocks + 3) / 4;
  hipLaunchKernelGGL(k_mc_decode, dim3((ntiles + 3) / 4), dim3(256), 0, c.stream, to_dimg(c.conn[0]),
                     (const uint8_t *)c.conn[1].ptr, (const uint32_t *)c.conn[2].ptr, (const uint16_t *)c.conn[3].ptr,
                     (const uint8_t *)c.conn[4].ptr, a);
  return launch_post("i-mcraw/decode");
}
