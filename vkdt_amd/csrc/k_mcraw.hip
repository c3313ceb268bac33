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
// three passes without inter-group waits: 256 threads x 16 blocks per group
// scan locally and publish their total; one group scans the totals; every
// group adds its exclusive total back

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

__global__ void __launch_bounds__(MC_WG) k_mc_scan_local(const uint8_t *bits, uint32_t *off, uint32_t *part, uint32_t n)
{
  __shared__ uint32_t sh[MC_WG];
  const uint32_t base = blockIdx.x * MC_PART + threadIdx.x * MC_ROWS;
  uint32_t loc[MC_ROWS], acc = 0;
#pragma unroll
  for(int r = 0; r < MC_ROWS; r++)
  {
    const uint32_t i = base + r;
    acc += i < n ? mc_block_bytes(mc_count(bits, i)) : 0u;
    loc[r] = acc;
  }
  const uint32_t incl = block_inclusive_scan(acc, sh);
  const uint32_t excl = incl - acc;
#pragma unroll
  for(int r = 0; r < MC_ROWS; r++)
    if(base + r < n) off[base + r] = loc[r] + excl;
  if(threadIdx.x == MC_WG - 1) part[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(MC_WG) k_mc_scan_parts(uint32_t *part, uint32_t np)
{ // in place: group totals -> exclusive prefix of the totals
  __shared__ uint32_t sh[MC_WG];
  uint32_t carry = 0;
  for(uint32_t c0 = 0; c0 < np; c0 += MC_WG)
  {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < np ? part[i] : 0u;
    const uint32_t incl = block_inclusive_scan(v, sh);
    const uint32_t tot = sh[MC_WG - 1];
    if(i < np) part[i] = carry + incl - v;
    carry += tot;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(MC_WG) k_mc_scan_add(uint32_t *off, const uint32_t *part, uint32_t n)
{
  if(blockIdx.x == 0) return;
  const uint32_t add = part[blockIdx.x];
  const uint32_t base = blockIdx.x * MC_PART;
  for(uint32_t r = threadIdx.x; r < MC_PART; r += MC_WG)
    if(base + r < n) off[base + r] += add;
}

// ---- decode ------------------------------------------------------------------
// A block is decoded by 32 lanes (half a wave); lane s yields the values of
// two adjacent register lanes, the pair read from bytes `reg` and `reg + 1`
// of the block at the same bit offset.  Bytes past the data buffer read as 0.

struct mc_block
{
  const uint8_t *d;
  uint32_t off, len;
  __device__ uint32_t at(int r) const { return off + (uint32_t)r < len ? d[off + r] : 0u; }
  __device__ uint2 pair(int reg, int bit, int n) const
  {
    const uint32_t m = (1u << n) - 1u;
    return make_uint2((at(reg) >> bit) & m, (at(reg + 1) >> bit) & m);
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

DEV uint2 mc_decode8(const mc_block &b, int s) { return make_uint2(b.at(2 * s), b.at(2 * s + 1)); }

DEV uint2 mc_decode10(const mc_block &b, int s)
{ // eight low bits from bytes 0..31 / 40..71, two high bits from 32..39 / 72..79
  const int rl = s < 16 ? 2 * s : 40 + 2 * (s - 16);
  const int rh = (s < 16 ? 32 : 72) + 2 * (s % 4);
  const int off = 2 * ((s & 15) / 4);
  return make_uint2(b.at(rl), b.at(rl + 1)) | shl(b.pair(rh, off, 2), 8);
}

DEV uint2 mc_decode16(const mc_block &b, int s)
{ // little-endian 16-bit values
  return make_uint2(b.at(4 * s) | (b.at(4 * s + 1) << 8), b.at(4 * s + 2) | (b.at(4 * s + 3) << 8));
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

DEV void mc_store(const dimg &out, int x, int y, uint32_t v)
{
  if(x < 0 || y < 0 || x >= out.wd || y >= out.ht) return;
  ((uint16_t *)out.p)[(long)y * out.pitch + x] = (uint16_t)mc_unorm(v);
}

__global__ void __launch_bounds__(256) k_mc_decode(dimg out, const uint8_t *bits, const uint32_t *off,
                                                   const uint16_t *ref, const uint8_t *dat, mc_args a)
{
  const uint32_t idx = blockIdx.x * 8 + threadIdx.x / 32;   // two blocks per wave
  if(idx >= a.nblocks) return;
  const int s = threadIdx.x % 32;
  const uint32_t cnt = mc_count(bits, idx);
  const mc_block b = { dat, idx > 0 ? off[idx - 1] : 0u, a.data_len };
  uint2 res = mc_decode(b, cnt, s);
  res.x += ref[idx];
  res.y += ref[idx];
  // decode.comp:314-337: tile k = idx/4 of 64x4, phase idx&3, lanes 16..31 on row 2
  const uint32_t k = idx / 4, bpr = (uint32_t)a.owd / 64;
  int px = 64 * (int)(k % bpr), py = 4 * (int)(k / bpr);
  if(s >= 16) { px -= 64; py += 2; }
  px += (idx & 1);
  py += (idx >> 1) & 1;
  px += 4 * s;
  mc_store(out, px - a.ox, py - a.oy, res.x);
  mc_store(out, px + 2 - a.ox, py - a.oy, res.y);
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
  uint32_t *off = (uint32_t *)c.conn[1].ptr, *part = (uint32_t *)c.conn[2].ptr;
  hipLaunchKernelGGL(k_mc_scan_local, dim3(np), dim3(MC_WG), 0, c.stream, (const uint8_t *)c.conn[0].ptr, off, part, n);
  hipLaunchKernelGGL(k_mc_scan_parts, dim3(1), dim3(MC_WG), 0, c.stream, part, np);
  hipLaunchKernelGGL(k_mc_scan_add, dim3(np), dim3(MC_WG), 0, c.stream, off, (const uint32_t *)part, n);
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
  hipLaunchKernelGGL(k_mc_decode, dim3((a.nblocks + 7) / 8), dim3(256), 0, c.stream, to_dimg(c.conn[0]),
                     (const uint8_t *)c.conn[1].ptr, (const uint32_t *)c.conn[2].ptr, (const uint16_t *)c.conn[3].ptr,
                     (const uint8_t *)c.conn[4].ptr, a);
  return launch_post("i-mcraw/decode");
}
