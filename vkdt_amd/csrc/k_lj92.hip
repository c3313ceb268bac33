// k_lj92.hip -- lossless JPEG (ITU-T T.81 process 14, "LJ92") decoding of
// the tiles or strips of compressed raw files (DNG Compression = 7).
//
// The reference decodes on the host through rawler (i-raw, a Rust crate not
// in its tree) and through liblj92 (i-mlv).  Here one parser serves two
// decoders with the same arithmetic:
//   * vkdt_lj92_decode_host: one stream on the CPU (i-raw's drop-in path);
//   * vkdt_hip_lj92_decode: many streams of many images in one launch, one
//     lane per stream (the streams are independent: a batch of tiled raws
//     has thousands), so uploads stay compressed and the unpack leaves the
//     host.
// Supported: 2..16 bit samples, 1..4 interleaved components, predictors
// 1..7, point transform, one Huffman table per component; restart
// intervals are refused.  Decoded samples fill the tile row-major (sample i
// of the scan at (i mod tile_wd, i div tile_wd)); sites outside the image
// are decoded and dropped.
#include "launch.h"

#include <map>
#include <string.h>
#include <vector>

namespace {

struct lj92_hdr
{
  int bits, wd, ht, comp, pred, pt;   // frame: wd samples per line and component
  int tbl[4];                         // Huffman table per scan component
  size_t scan;                        // offset of the entropy-coded segment
  uint8_t huff_bits[4][16];           // BITS of the DC tables
  uint8_t huff_val[4][17];
  int have[4];
};

int be16(const uint8_t *p) { return (p[0] << 8) | p[1]; }

// markers up to the scan (T.81 B.2): SOI, DHT, SOF3, SOS; DRI refused
int lj92_parse(const uint8_t *d, size_t n, lj92_hdr *h)
{
  memset(h, 0, sizeof(*h));
  if(n < 4 || d[0] != 0xff || d[1] != 0xd8) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: no SOI");
  size_t i = 2;
  int sof = 0;
  while(i + 4 <= n)
  {
    if(d[i] != 0xff) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: marker expected at %zu", i);
    const int m = d[i + 1];
    if(m == 0xff) { i++; continue; }
    const size_t L = be16(d + i + 2);
    if(i + 2 + L > n) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: segment past the end");
    const uint8_t *s = d + i + 4;
    if(m == 0xc4)
    { // DHT: one or more tables
      size_t k = 0;
      while(k + 17 <= L - 2)
      {
        const int tc = s[k] >> 4, th = s[k] & 15;
        int cnt = 0;
        for(int b = 0; b < 16; b++) cnt += s[k + 1 + b];
        if(tc != 0 || th > 3 || cnt > 17 || k + 17 + cnt > L - 2)
          return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: bad DHT");
        memcpy(h->huff_bits[th], s + k + 1, 16);
        memcpy(h->huff_val[th], s + k + 17, cnt);
        h->have[th] = 1;
        k += 17 + cnt;
      }
    }
    else if(m == 0xc3)
    { // SOF3
      h->bits = s[0];
      h->ht = be16(s + 1);
      h->wd = be16(s + 3);
      h->comp = s[5];
      if(h->bits < 2 || h->bits > 16 || h->comp < 1 || h->comp > 4 || h->wd < 1 || h->ht < 1)
        return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: frame %d bits %dx%d x%d", h->bits, h->wd, h->ht, h->comp);
      sof = 1;
    }
    else if(m == 0xdd)
    {
      if(be16(s) != 0) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: restart intervals are not supported");
    }
    else if(m == 0xda)
    { // SOS
      if(!sof) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: scan before frame");
      const int ns = s[0];
      if(ns != h->comp) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: %d scan components, frame has %d", ns, h->comp);
      for(int c = 0; c < ns; c++)
      {
        h->tbl[c] = s[2 + 2 * c] >> 4;
        if(h->tbl[c] > 3 || !h->have[h->tbl[c]]) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: missing table");
      }
      h->pred = s[1 + 2 * ns];
      h->pt = s[3 + 2 * ns] & 15;
      if(h->pred < 1 || h->pred > 7) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: predictor %d", h->pred);
      h->scan = i + 2 + L;
      if(h->scan >= n) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: no entropy-coded data");
      return VKDT_HIP_OK;
    }
    else if(m == 0xd9) break;
    else if(m >= 0xc0 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: not a lossless frame (SOF%d)", m - 0xc0);
    i += 2 + L;
  }
  return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: no scan");
}

// a DC table (T.81 C.2, F.2.2.3) in the form the decoder reads: codes of up
// to LJ92_FAST bits by one lookup of the stream's next LJ92_FAST bits (entry =
// category << 8 | length, 0 where the code is longer or matches nothing),
// longer ones by the canonical bounds per length (maxcode -1 where a length
// has no codes; category = val[code + delta])
#define LJ92_FAST 9
struct __attribute__((aligned(16))) lj92_tab
{
  uint16_t fast[1 << LJ92_FAST];
  int32_t  maxcode[17];
  int32_t  delta[17];
  uint8_t  val[20];
};

int lj92_table(const uint8_t bits[16], const uint8_t *val, lj92_tab *t)
{
  memset(t, 0, sizeof(*t));
  int code = 0, k = 0;
  for(int len = 1; len <= 16; len++)
  {
    t->maxcode[len] = -1;
    if(bits[len - 1])
    {
      t->delta[len] = k - code;
      for(int j = 0; j < bits[len - 1]; j++, k++, code++)
      {
        if(code >= (1 << len) || val[k] > 16) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: bad Huffman table");
        t->val[k] = val[k];
        if(len <= LJ92_FAST)
        {
          const int lo = code << (LJ92_FAST - len), cnt = 1 << (LJ92_FAST - len);
          for(int e = 0; e < cnt; e++) t->fast[lo + e] = (uint16_t)((val[k] << 8) | len);
        }
      }
      t->maxcode[len] = code - 1;
    }
    code <<= 1;
  }
  return VKDT_HIP_OK;
}

} // namespace

// ---- the decoder, host and device alike ------------------------------------

struct lj92_job
{
  const uint8_t *data;     // entropy-coded segment (readable as whole aligned
  long     len;            //   words: the buffer is padded past its end)
  const lj92_tab *tab[4];  // per component
  int      same;           // bit c: component c uses component 0's table
  int      bits, fwd, fht, comp, pred, pt;
  uint16_t *out;           // the image
  int      img_wd, img_ht, pitch;
  int      x0, y0, tile_wd;
  uint16_t *row;           // fwd * comp samples: the previous line (global memory variant)
};

// bit reader over a segment with 0xFF00 stuffing; a marker ends the data
// (zeros are fed past it).  The stuffed bytes arrive a 32-bit word at a time
// with the next word already requested, so a lane waits on memory once per
// four bytes at most.  Words are read only up to the one holding the
// segment's last byte.
// loads and stores of global memory reached through pointers held in the
// job: named as global on the device so they do not count against LDS waits
template<class T> __host__ __device__ static inline T lj92_ld(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) T *)p;
#else
  return *p;
#endif
}

template<class T> __host__ __device__ static inline void lj92_st(T *p, T v)
{
#if defined(__HIP_DEVICE_COMPILE__)
  *(__attribute__((address_space(1))) T *)p = v;
#else
  *p = v;
#endif
}

struct lj92_bits
{
  const uint32_t *wp;      // next word to request
  const uint32_t *wlast;   // the word holding the segment's last byte
  uint32_t pre;            // requested word as loaded
  int      pre_ok;         // whether it holds segment bytes (else it counts as zero)
  uint64_t raw;            // stuffed bytes not yet taken, left-justified
  int      rawn;           // how many
  long     left;           // segment bytes from raw's first on
  uint64_t acc;            // unstuffed bits, left-justified
  int      cnt;
};

__host__ __device__ static inline uint32_t lj92_bswap(uint32_t w)
{
  return (w >> 24) | ((w >> 8) & 0xff00u) | ((w << 8) & 0xff0000u) | (w << 24);
}

// requests the next word: always a load (of the last word once past it), so
// no branch joins on its result and the lane waits only where it is used
__host__ __device__ static inline void lj92_request(lj92_bits &b)
{
  b.pre_ok = b.wp <= b.wlast;
  b.pre = lj92_ld(b.pre_ok ? b.wp : b.wlast);
  b.wp++;
}

__host__ __device__ static inline uint32_t lj92_take(const lj92_bits &b)
{
  return b.pre_ok ? lj92_bswap(b.pre) : 0u;
}

// n >= 1 segment bytes at p
__host__ __device__ static inline void lj92_init(lj92_bits &b, const uint8_t *p, long n)
{
  const int skip = (int)((uintptr_t)p & 3);
  b.wp = (const uint32_t *)(p - skip);
  b.wlast = (const uint32_t *)((uintptr_t)(p + n - 1) & ~(uintptr_t)3);
  lj92_request(b);
  b.raw = ((uint64_t)lj92_take(b) << 32) << (8 * skip);
  b.rawn = 4 - skip;
  lj92_request(b);
  b.left = n;
  b.acc = 0;
  b.cnt = 0;
}

__host__ __device__ static inline void lj92_fill(lj92_bits &b)
{
  while(b.cnt <= 56)
  {
    if(b.rawn <= 4)
    { // five bytes or more in raw: a 0xFF and the byte after it are both there
      b.raw |= (uint64_t)lj92_take(b) << (32 - 8 * b.rawn);
      b.rawn += 4;
      lj92_request(b);
    }
    uint32_t byte = 0;
    if(b.left > 0)
    {
      byte = (uint32_t)(b.raw >> 56);
      if(byte == 0xff)
      {
        if(b.left >= 2 && ((b.raw >> 48) & 0xff) == 0) { b.raw <<= 16; b.rawn -= 2; b.left -= 2; }
        else { byte = 0; b.left = 0; }   // a marker: the segment ends
      }
      else { b.raw <<= 8; b.rawn--; b.left--; }
    }
    b.acc |= (uint64_t)byte << (56 - b.cnt);
    b.cnt += 8;
  }
}

template<class R> __host__ __device__ static inline int lj92_get(R &b, int k)
{
  const int v = (int)(b.acc >> (64 - k));
  b.acc <<= k;
  b.cnt -= k;
  return v;
}

// next category from table t, -1 on a prefix no code matches
template<class R> __host__ __device__ static inline int lj92_symbol(R &b, const lj92_tab *t)
{
  const int peek = (int)(b.acc >> (64 - LJ92_FAST));
  const uint16_t e = t->fast[peek];
  if(e)
  {
    lj92_get(b, e & 0xff);
    return e >> 8;
  }
  const int p16 = (int)(b.acc >> 48);
  for(int l = LJ92_FAST + 1; l <= 16; l++)
  {
    const int code = p16 >> (16 - l);
    if(code <= t->maxcode[l])
    {
      lj92_get(b, l);
      return t->val[code + t->delta[l]];
    }
  }
  return -1;
}

typedef unsigned int lj92_u32x4 __attribute__((ext_vector_type(4)));

// every 16-bit half of w shifted left by pt (the point transform)
__host__ __device__ static inline unsigned lj92_pt2(unsigned w, int pt)
{
  return ((w << pt) & 0xffffu) | ((((w >> 16) << pt) & 0xffffu) << 16);
}

// decodes one stream into the image; row holds a scan line (fwd * comp
// samples): the previous one while a line decodes, which it then replaces.
// Each decoded line is written out after it completes, in 16-byte stores
// where the tile's lines allow: stores interleaved with the decode would
// make every wait for the next stream word wait for them too.
// With SAME every component decodes with tab0 (a copy of their common table
// in the caller's fastest memory), else with the job's table per component.
// Returns 0, or -1 on corrupt data.
template<int SAME>
__host__ __device__ static inline int lj92_run(const lj92_job &jg, const lj92_tab *tab0, uint16_t *row)
{
  const lj92_job j = jg;   // in registers: the stores below cannot alias it
  const lj92_tab *const *tab = jg.tab;   // indexed per component: read where it lies
  lj92_bits b;
  lj92_init(b, j.data, j.len);
  const int C = j.comp, W = j.fwd, L = W * C;
  const int base = 1 << (j.bits - j.pt - 1);
  int tx = 0, ty = 0;      // tile site of the next line's first sample
  long o = (long)j.y0 * j.pitch + j.x0;               // the tile's origin relative to the window (may be outside)
  const bool inside = j.x0 >= 0 && j.x0 + j.tile_wd <= j.img_wd;   // every column of the tile lands in the window
  // a scan line is a tile line that lands whole, on 16-byte boundaries
  const bool vec = L == j.tile_wd && inside && (L & 7) == 0 && (j.pitch & 7) == 0 &&
                   (((uintptr_t)(j.out + o)) & 15) == 0 && (((uintptr_t)row) & 15) == 0;
  for(int r = 0; r < j.fht; r++)
  {
    int left0 = 0, left1 = 0, left2 = 0, left3 = 0, ul0 = 0, ul1 = 0, ul2 = 0, ul3 = 0;
    for(int x = 0; x < W; x++)
      for(int c = 0; c < C; c++)
      {
        if(b.cnt < 32) lj92_fill(b);
        const int t = lj92_symbol(b, SAME ? tab0 : lj92_ld(tab + c));
        if(t < 0) return -1;
        int diff;
        if(t == 0) diff = 0;
        else if(t == 16) diff = 32768;
        else
        { // the fill above left 32 bits or more: a code of up to 16 and up to 15 more
          diff = lj92_get(b, t);
          if(diff < (1 << (t - 1))) diff -= (1 << t) - 1;
        }
        const int k = x * C + c;
        const int up = row[k];
        int p;
        if(r == 0) p = x == 0 ? base : left0;
        else if(x == 0) p = up;
        else
        {
          const int a = left0, bb = up, cc = ul0;
          switch(j.pred)
          {
            case 1: p = a; break;
            case 2: p = bb; break;
            case 3: p = cc; break;
            case 4: p = a + bb - cc; break;
            case 5: p = a + ((bb - cc) >> 1); break;
            case 6: p = bb + ((a - cc) >> 1); break;
            default: p = (a + bb) >> 1; break;
          }
        }
        const int v = (p + diff) & 0xffff;
        // left0/ul0 always hold the next sample's component: rotate by one
        if(C == 1) { left0 = v; ul0 = up; }
        else if(C == 2) { left0 = left1; left1 = v; ul0 = ul1; ul1 = up; }
        else if(C == 3) { left0 = left1; left1 = left2; left2 = v; ul0 = ul1; ul1 = ul2; ul2 = up; }
        else { left0 = left1; left1 = left2; left2 = left3; left3 = v; ul0 = ul1; ul1 = ul2; ul2 = ul3; ul3 = up; }
        row[k] = (uint16_t)v;
      }
    // the scan fills the tile row-major: this line's L samples from (tx, ty) on
    if(vec)
    {
      if(j.y0 + ty >= 0 && j.y0 + ty < j.img_ht)
        for(int k = 0; k < L; k += 8)
        {
          lj92_u32x4 w = *(const lj92_u32x4 *)(row + k);
          if(j.pt) w = lj92_u32x4{ lj92_pt2(w.x, j.pt), lj92_pt2(w.y, j.pt), lj92_pt2(w.z, j.pt), lj92_pt2(w.w, j.pt) };
          lj92_st((lj92_u32x4 *)(j.out + o + k), w);
        }
      ty++;
      o += j.pitch;
    }
    else
      for(int k = 0; k < L; k++)
      {
        if(j.y0 + ty >= 0 && j.y0 + ty < j.img_ht && (inside || (j.x0 + tx >= 0 && j.x0 + tx < j.img_wd)))
          lj92_st(j.out + o + tx, (uint16_t)(row[k] << j.pt));
        if(++tx == j.tile_wd) { tx = 0; ty++; o += j.pitch; }
      }
  }
  return 0;
}

// one lane per stream.  For lines of up to LJ92_ROW samples the lane keeps its
// line in LDS (ROW_LDS), and for streams whose components share one table
// the lane's copy of it (SAME), so that the only global loads of the decode
// are the stream's own words.
#define LJ92_ROW 512
template<int ROW_LDS, int SAME>
__global__ void __launch_bounds__(64) k_lj92(const lj92_job *jobs, int n, int *status)
{
  __shared__ lj92_tab tabs[SAME ? 64 : 1];
  __shared__ uint16_t rows[ROW_LDS ? 64 : 1][ROW_LDS ? LJ92_ROW : 1];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const lj92_job &j = jobs[i];
  lj92_tab *t = tabs + (SAME ? threadIdx.x : 0);
  if(SAME)
  {
    const lj92_u32x4 *src = (const lj92_u32x4 *)lj92_ld(j.tab);
    lj92_u32x4 *dst = (lj92_u32x4 *)t;
#pragma unroll 8
    for(int k = 0; k < (int)(sizeof(lj92_tab) / (sizeof(lj92_u32x4))); k++) dst[k] = lj92_ld(src + k);
  }
  if(lj92_run<SAME>(j, t, ROW_LDS ? rows[threadIdx.x] : lj92_ld(&j.row))) atomicAdd(status, 1);
}

// ---- many lanes per stream ----------------------------------------------------
//
// A stream is a serial chain of variable-length symbols, so one lane per
// stream leaves most of the chip idle (a 24 MP raw in 256x256 tiles is 384
// streams).  Streams whose components share one table and use predictor 1
// (the form raw converters write) are instead decoded by many lanes each,
// after the published scheme for Huffman data without restart markers
// (Weissenberger and Schmidt, "Massively Parallel Huffman Decoding on GPUs",
// ICPP 2018), which rests on prefix codes resynchronising within a few
// symbols of a wrong start:
//   unstuff   one work group per stream: the segment without its stuffed
//             zero bytes, up to the first marker, into aligned words
//   first     lane g decodes subsequence k = bits [k S, (k+1) S) from k S:
//             its symbol count and where its last symbol ends
//   sync      while any lane's start differs from its predecessor's end, it
//             decodes again from there (ping-pong ends; converges in as many
//             rounds as the longest chain of wrong starts, a few)
//   scan      symbol offsets per subsequence (one work group per stream)
//   write     every lane decodes its subsequence once more into the
//             differences; the last lane of a stream continues past the
//             segment's end on zero bits as the serial decoder does
//   col0/rows predictor 1 is two prefix sums: down the first column of
//             each component, then along each scan line.
#define LJ92_SUB 512

struct lj92_pst
{
  const uint8_t *src;      // entropy-coded segment
  uint32_t *clean;         // unstuffed bytes as words, zero past the data
  uint32_t len, cwords;    // segment bytes; words of the clean region
  uint32_t sbase, K;       // first subsequence, subsequences (len * 8 / S rounded up)
  uint32_t rbase;          // first scan line in the numbering of all streams
  uint32_t S;              // samples of the scan
  const lj92_tab *tab;
  uint16_t *diff;          // S differences
  uint64_t *col0;          // fht first-column values, four 16-bit components packed
  int      fwd, fht, comp, bits, pt;
  uint16_t *out;
  int      img_wd, img_ht, pitch, x0, y0, tile_wd;
};

struct lj92_cbits          // reader of an unstuffed segment
{
  const uint32_t *w;
  uint32_t nw, wi;         // words, next word to request
  uint32_t pre;
  int      pre_ok;
  uint64_t acc;
  int      cnt;
};

__device__ static inline void lj92_crequest(lj92_cbits &b)
{
  b.pre_ok = b.wi < b.nw;
  b.pre = lj92_ld(b.w + (b.pre_ok ? b.wi : b.nw - 1));
  b.wi++;
}

__device__ static inline uint32_t lj92_ctake(const lj92_cbits &b)
{
  return b.pre_ok ? lj92_bswap(b.pre) : 0u;
}

__device__ static inline void lj92_cinit(lj92_cbits &b, const lj92_pst &p, uint32_t pos)
{
  b.w = p.clean;
  b.nw = p.cwords;
  b.wi = pos >> 5;
  lj92_crequest(b);
  const uint64_t w0 = lj92_ctake(b);
  lj92_crequest(b);
  const uint64_t w1 = lj92_ctake(b);
  lj92_crequest(b);
  const int sh = pos & 31;
  b.acc = ((w0 << 32) | w1) << sh;
  b.cnt = 64 - sh;
}

__device__ static inline void lj92_cfill(lj92_cbits &b)
{
  if(b.cnt < 32)
  {
    b.acc |= (uint64_t)lj92_ctake(b) << (32 - b.cnt);
    b.cnt += 32;
    lj92_crequest(b);
  }
}

// one symbol at the reader: its difference, and the bits it took; a prefix no
// code matches takes one bit and yields no symbol (ok = 0), so that a decode
// from any bit position is a function of that position alone
__device__ static inline int lj92_csym(lj92_cbits &b, const lj92_tab *t, int &used, int &ok)
{
  lj92_cfill(b);
  const int c0 = b.cnt;
  const int cat = lj92_symbol(b, t);
  if(cat < 0)
  {
    lj92_get(b, 1);
    used = 1;
    ok = 0;
    return 0;
  }
  int diff;
  if(cat == 0) diff = 0;
  else if(cat == 16) diff = 32768;
  else
  {
    diff = lj92_get(b, cat);
    if(diff < (1 << (cat - 1))) diff -= (1 << cat) - 1;
  }
  used = c0 - b.cnt;
  ok = 1;
  return diff;
}

// the stream and subsequence of global subsequence g (streams sorted by sbase)
__device__ static inline int lj92_find(const lj92_pst *ps, int np, uint32_t g, bool rows)
{
  int lo = 0, hi = np - 1;
  while(lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    const uint32_t b = rows ? lj92_ld(&ps[mid].rbase) : lj92_ld(&ps[mid].sbase);
    if(b <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// symbols starting in [pos, lim): their count; returns where the last ends
__device__ static inline uint32_t lj92_count(const lj92_pst &p, uint32_t pos, uint32_t lim, uint32_t &n)
{
  n = 0;
  if(pos >= lim) return pos;
  lj92_cbits b;
  lj92_cinit(b, p, pos);
  const lj92_tab *t = p.tab;
  while(pos < lim)
  {
    int used, ok;
    lj92_csym(b, t, used, ok);
    pos += used;
    n += ok;
  }
  return pos;
}

__global__ void __launch_bounds__(256) k_lj92_unstuff(const lj92_pst *ps, uint32_t *nbits)
{
  __shared__ uint32_t s_sum[256];
  __shared__ uint32_t s_mark;
  const lj92_pst &p = ps[blockIdx.x];
  const uint8_t *src = p.src;
  uint8_t *dst = (uint8_t *)p.clean;
  const uint32_t len = p.len, t = threadIdx.x;
  uint32_t kept = 0;
  bool ended = false;
  for(uint32_t base = 0; base < len && !ended; base += 256 * 4)
  {
    const uint32_t i0 = base + 4 * t;
    uint32_t by[6];                                    // bytes i0-1 .. i0+4, zero outside
    for(int j = 0; j < 6; j++)
    {
      const long i = (long)i0 - 1 + j;
      by[j] = i >= 0 && i < (long)len ? lj92_ld(src + i) : 0u;
    }
    uint32_t mark = 0xffffffffu;                       // first marker among this lane's bytes
    for(int j = 3; j >= 0; j--)
      if(i0 + j < len && by[j + 1] == 0xff && (i0 + j + 1 >= len || by[j + 2] != 0)) mark = i0 + j;
    if(t == 0) s_mark = 0xffffffffu;
    __syncthreads();
    if(mark != 0xffffffffu) atomicMin(&s_mark, mark);
    __syncthreads();
    const uint32_t m = s_mark;
    // kept: before the marker and not a zero stuffed after 0xff
    uint32_t keep = 0;
    for(int j = 0; j < 4; j++)
    {
      const uint32_t i = i0 + j;
      if(i < len && i < m && !(by[j + 1] == 0 && by[j] == 0xff && i > 0)) keep |= 1u << j;
    }
    // exclusive scan of the kept counts over the work group
    uint32_t c = __popc(keep);
    s_sum[t] = c;
    __syncthreads();
    for(int o = 1; o < 256; o <<= 1)
    {
      const uint32_t v = t >= (uint32_t)o ? s_sum[t - o] : 0u;
      __syncthreads();
      s_sum[t] += v;
      __syncthreads();
    }
    uint32_t at = kept + s_sum[t] - c;
    for(int j = 0; j < 4; j++)
      if(keep & (1u << j)) lj92_st(dst + at++, (uint8_t)by[j + 1]);
    kept += s_sum[255];
    ended = m != 0xffffffffu;
    __syncthreads();
  }
  // zeros after the data up to the end of the clean region
  for(uint32_t i = kept + t; i < p.cwords * 4; i += 256) lj92_st(dst + i, (uint8_t)0);
  if(t == 0) nbits[blockIdx.x] = kept * 8;
}

__global__ void __launch_bounds__(256) k_lj92_first(const lj92_pst *ps, int np, uint32_t ntot, const uint32_t *nbits,
                                                    uint32_t *start, uint32_t *end, uint32_t *cnt)
{
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if(g >= ntot) return;
  const int i = lj92_find(ps, np, g, false);
  const lj92_pst &p = ps[i];
  const uint32_t k = g - p.sbase, B = nbits[i];
  const uint32_t s0 = k * LJ92_SUB, lim = min(s0 + LJ92_SUB, B);
  uint32_t n;
  start[g] = s0;
  end[g] = lj92_count(p, s0, lim, n);
  cnt[g] = n;
}

__global__ void __launch_bounds__(256) k_lj92_sync(const lj92_pst *ps, int np, uint32_t ntot, const uint32_t *nbits,
                                                   uint32_t *start, const uint32_t *end_in, uint32_t *end_out,
                                                   uint32_t *cnt, int *changed)
{
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if(g >= ntot) return;
  const int i = lj92_find(ps, np, g, false);
  const lj92_pst &p = ps[i];
  const uint32_t k = g - p.sbase;
  const uint32_t e = end_in[g];
  // past the data a subsequence holds no symbol whatever its start: left alone,
  // it does not hand a changed start down the rest of the stream one round at a time
  if(k == 0 || end_in[g - 1] == start[g] || k * LJ92_SUB >= nbits[i]) { end_out[g] = e; return; }
  const uint32_t s = end_in[g - 1], lim = min(k * LJ92_SUB + LJ92_SUB, nbits[i]);
  uint32_t n;
  start[g] = s;
  end_out[g] = lj92_count(p, s, lim, n);
  cnt[g] = n;
  atomicAdd(changed, 1);
}

__global__ void __launch_bounds__(256) k_lj92_scan(const lj92_pst *ps, const uint32_t *cnt, uint32_t *off)
{
  __shared__ uint32_t s_sum[256];
  const lj92_pst &p = ps[blockIdx.x];
  const uint32_t t = threadIdx.x;
  uint32_t carry = 0;
  for(uint32_t k0 = 0; k0 < p.K; k0 += 256)
  {
    const uint32_t k = k0 + t;
    const uint32_t c = k < p.K ? cnt[p.sbase + k] : 0u;
    s_sum[t] = c;
    __syncthreads();
    for(int o = 1; o < 256; o <<= 1)
    {
      const uint32_t v = t >= (uint32_t)o ? s_sum[t - o] : 0u;
      __syncthreads();
      s_sum[t] += v;
      __syncthreads();
    }
    if(k < p.K) off[p.sbase + k] = carry + s_sum[t] - c;
    carry += s_sum[255];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_lj92_write(const lj92_pst *ps, int np, uint32_t ntot, const uint32_t *nbits,
                                                    const uint32_t *start, const uint32_t *off, int *status)
{
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if(g >= ntot) return;
  const int i = lj92_find(ps, np, g, false);
  const lj92_pst &p = ps[i];
  const uint32_t k = g - p.sbase, B = nbits[i];
  const uint32_t kreal = (B + LJ92_SUB - 1) / LJ92_SUB;
  const bool tail = k == (kreal ? kreal - 1 : 0);     // continues to the scan's last sample
  uint32_t pos = start[g];
  const uint32_t lim = min(k * LJ92_SUB + LJ92_SUB, B);
  uint32_t o = off[g];
  if(!(pos < lim || tail) || o >= p.S) return;
  lj92_cbits b;
  lj92_cinit(b, p, pos);
  const lj92_tab *t = p.tab;
  uint16_t *diff = p.diff;
  // samples up to a multiple of four one by one, then four to a store
  uint64_t pack = 0;
  int npk = 0;
  while((pos < lim || tail) && o < p.S)
  {
    int used, ok;
    const int d = lj92_csym(b, t, used, ok);
    pos += used;
    if(!ok) { atomicAdd(status, 1); return; }      // a sample of the scan with no valid code
    if(npk == 0 && (o & 3)) lj92_st(diff + o, (uint16_t)d);
    else
    {
      pack |= (uint64_t)(uint16_t)d << (16 * npk);
      if(++npk == 4)
      {
        lj92_st((uint64_t *)(diff + o - 3), pack);
        pack = 0;
        npk = 0;
      }
    }
    o++;
  }
  for(int j = 0; j < npk; j++) lj92_st(diff + o - npk + j, (uint16_t)(pack >> (16 * j)));
}

// first-column values per scan line: base plus the running sum of the lines'
// first differences, four components in 16-bit lanes (one work group per stream)
__device__ static inline uint64_t lj92_add4(uint64_t a, uint64_t b)
{
  const uint64_t h = 0x8000800080008000ull;
  return ((a & ~h) + (b & ~h)) ^ ((a ^ b) & h);
}

__global__ void __launch_bounds__(256) k_lj92_col0(const lj92_pst *ps)
{
  __shared__ uint64_t s_sum[256];
  const lj92_pst &p = ps[blockIdx.x];
  const uint32_t t = threadIdx.x, L = p.fwd * p.comp;
  const uint64_t base = (uint64_t)(1u << (p.bits - p.pt - 1)) * 0x0001000100010001ull;
  uint64_t carry = base;
  for(int r0 = 0; r0 < p.fht; r0 += 256)
  {
    const int r = r0 + t;
    uint64_t v = 0;
    if(r < p.fht)
      for(int c = 0; c < p.comp; c++) v |= (uint64_t)lj92_ld(p.diff + (size_t)r * L + c) << (16 * c);
    s_sum[t] = v;
    __syncthreads();
    for(int o = 1; o < 256; o <<= 1)
    {
      const uint64_t w = t >= (uint32_t)o ? s_sum[t - o] : 0ull;
      __syncthreads();
      s_sum[t] = lj92_add4(s_sum[t], w);
      __syncthreads();
    }
    if(r < p.fht) lj92_st(p.col0 + r, lj92_add4(carry, s_sum[t]));
    carry = lj92_add4(carry, s_sum[255]);
    __syncthreads();
  }
}

// one lane per scan line: the running sum along the line per component, the
// samples into the tile row-major
__global__ void __launch_bounds__(256) k_lj92_rows(const lj92_pst *ps, int np, uint32_t nrows)
{
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if(g >= nrows) return;
  const int i = lj92_find(ps, np, g, true);
  const lj92_pst &p = ps[i];
  const int r = g - p.rbase, C = p.comp, L = p.fwd * C;
  const uint16_t *d = p.diff + (size_t)r * L;
  const uint64_t c0 = lj92_ld(p.col0 + r);
  uint32_t v0 = c0 & 0xffff, v1 = (c0 >> 16) & 0xffff, v2 = (c0 >> 32) & 0xffff, v3 = c0 >> 48;
  const long s0 = (long)r * L;
  int tx = (int)(s0 % p.tile_wd), ty = (int)(s0 / p.tile_wd);
  long o = (long)(p.y0 + ty) * p.pitch + p.x0;
  const bool inside = p.x0 >= 0 && p.x0 + p.tile_wd <= p.img_wd;
  for(int k = 0; k < L; k++)
  {
    const int c = k % C;
    uint32_t v;
    if(k < C) v = c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : v3;
    else
    {
      const uint32_t dd = lj92_ld(d + k);
      if(c == 0) v = v0 = (v0 + dd) & 0xffff;
      else if(c == 1) v = v1 = (v1 + dd) & 0xffff;
      else if(c == 2) v = v2 = (v2 + dd) & 0xffff;
      else v = v3 = (v3 + dd) & 0xffff;
    }
    if(p.y0 + ty >= 0 && p.y0 + ty < p.img_ht && (inside || (p.x0 + tx >= 0 && p.x0 + tx < p.img_wd)))
      lj92_st(p.out + o + tx, (uint16_t)(v << p.pt));
    if(++tx == p.tile_wd) { tx = 0; ty++; o += p.pitch; }
  }
}

// one wave per scan line where the components are 1, 2 or 4 (every lane's
// four samples start at component 0): coalesced loads of four differences a
// lane, their sums per component scanned over the wave in 16-bit lanes, the
// first column's values in place of its differences
__device__ static inline uint64_t lj92_shfl_up(uint64_t v, int d)
{
  const uint32_t lo = __shfl_up((uint32_t)v, d), hi = __shfl_up((uint32_t)(v >> 32), d);
  return ((uint64_t)hi << 32) | lo;
}

__global__ void __launch_bounds__(256) k_lj92_rows4(const lj92_pst *ps, int np, uint32_t nrows)
{
  const uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if(g >= nrows) return;                               // whole waves: g is uniform in a wave
  const int i = lj92_find(ps, np, g, true);
  const lj92_pst &p = ps[i];
  const int r = g - p.rbase, C = p.comp, L = p.fwd * C;
  const uint16_t *d = p.diff + (size_t)r * L;
  const uint64_t c0 = lj92_ld(p.col0 + r);
  const bool direct = L == p.tile_wd && p.x0 >= 0 && p.x0 + p.tile_wd <= p.img_wd;   // scan line = tile line inside the window
  const long s0 = (long)r * L;
  const int ty0 = (int)(s0 / p.tile_wd);
  const bool wide = (L & 3) == 0;                      // lines start on 8-byte boundaries of the differences
  uint64_t carry = 0;
  for(int k0 = 0; k0 < L; k0 += 256)
  {
    const int k = k0 + 4 * lane;
    uint32_t dd[4];
    if(wide && k + 3 < L)
    {
      const uint64_t w = lj92_ld((const uint64_t *)(d + k));
      for(int j = 0; j < 4; j++) dd[j] = (uint32_t)(w >> (16 * j)) & 0xffff;
    }
    else
      for(int j = 0; j < 4; j++) dd[j] = k + j < L ? lj92_ld(d + k + j) : 0u;
    for(int j = 0; j < 4; j++)
      if(k + j < C) dd[j] = (uint32_t)(c0 >> (16 * (k + j))) & 0xffff;
    uint64_t loc = 0;
    for(int j = 0; j < 4; j++) loc = lj92_add4(loc, (uint64_t)dd[j] << (16 * (j % C)));
    uint64_t inc = loc;
    for(int o = 1; o < 64; o <<= 1)
    {
      const uint64_t y = lj92_shfl_up(inc, o);
      if(lane >= o) inc = lj92_add4(inc, y);
    }
    uint64_t run = lj92_shfl_up(inc, 1);
    run = lj92_add4(carry, lane ? run : 0ull);
    const uint64_t last = __shfl(inc, 63);
    carry = lj92_add4(carry, last);
    uint32_t v4[4];
    for(int j = 0; j < 4; j++)
    {
      const int c = j % C;
      run = lj92_add4(run, (uint64_t)dd[j] << (16 * c));
      v4[j] = ((uint32_t)(run >> (16 * c)) & 0xffff) << p.pt;
    }
    uint16_t *dst = p.out + (long)(p.y0 + ty0) * p.pitch + p.x0 + k;
    if(direct && k + 3 < L && p.y0 + ty0 >= 0 && p.y0 + ty0 < p.img_ht && ((uintptr_t)dst & 7) == 0)
    {
      lj92_st((uint64_t *)dst, (uint64_t)(v4[0] & 0xffff) | ((uint64_t)(v4[1] & 0xffff) << 16) |
                                ((uint64_t)(v4[2] & 0xffff) << 32) | ((uint64_t)(v4[3] & 0xffff) << 48));
      continue;
    }
    for(int j = 0; j < 4; j++)
    {
      const int kk = k + j;
      if(kk >= L) break;
      const uint32_t v = v4[j];
      int tx, ty;
      if(direct) { tx = kk; ty = ty0; }
      else { const long sm = s0 + kk; tx = (int)(sm % p.tile_wd); ty = (int)(sm / p.tile_wd); }
      if(p.y0 + ty >= 0 && p.y0 + ty < p.img_ht && p.x0 + tx >= 0 && p.x0 + tx < p.img_wd)
        lj92_st(p.out + (long)(p.y0 + ty) * p.pitch + p.x0 + tx, (uint16_t)v);
    }
  }
}

__global__ void k_lj92_zero(int *p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0;
}

// ---- host side ---------------------------------------------------------------

extern "C" int vkdt_lj92_info(const uint8_t *data, size_t len, uint32_t *wd, uint32_t *ht, uint32_t *comp,
                              uint32_t *bits)
{
  lj92_hdr h;
  if(int e = lj92_parse(data, len, &h)) return e;
  if(wd) *wd = h.wd;
  if(ht) *ht = h.ht;
  if(comp) *comp = h.comp;
  if(bits) *bits = h.bits;
  return VKDT_HIP_OK;
}

extern "C" int vkdt_lj92_decode_host(const uint8_t *data, const vkdt_lj92_tile_t *t, uint16_t *out, uint32_t wd,
                                     uint32_t ht, uint32_t pitch)
{
  lj92_hdr h;
  if(int e = lj92_parse(data + t->offset, t->length, &h)) return e;
  if((long)h.wd * h.comp * h.ht < (long)t->tile_wd * t->tile_ht || t->tile_wd == 0)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: %dx%dx%d samples for a %ux%u tile", h.wd, h.ht, h.comp,
                          t->tile_wd, t->tile_ht);
  std::vector<lj92_tab> tabs(h.comp);
  std::vector<uint16_t> row((size_t)h.wd * h.comp, 0);
  lj92_job j = {};
  for(int c = 0; c < h.comp; c++)
  {
    if(int e = lj92_table(h.huff_bits[h.tbl[c]], h.huff_val[h.tbl[c]], &tabs[c])) return e;
    j.tab[c] = &tabs[c];
    if(h.tbl[c] == h.tbl[0]) j.same |= 1 << c;
  }
  // the segment in whole aligned words, as the reader takes it
  const size_t len = t->length - h.scan;
  std::vector<uint32_t> seg((len + 3) / 4 + 1, 0);
  memcpy(seg.data(), data + t->offset + h.scan, len);
  j.data = (const uint8_t *)seg.data();
  j.len = (long)len;
  j.bits = h.bits; j.fwd = h.wd; j.fht = h.ht; j.comp = h.comp; j.pred = h.pred; j.pt = h.pt;
  j.out = out; j.img_wd = wd; j.img_ht = ht; j.pitch = pitch;
  j.x0 = t->x0; j.y0 = t->y0; j.tile_wd = t->tile_wd;
  j.row = row.data();
  const int err = j.same == (1 << h.comp) - 1 ? lj92_run<1>(j, &tabs[0], row.data()) : lj92_run<0>(j, nullptr, row.data());
  if(err) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: corrupt entropy-coded data");
  return VKDT_HIP_OK;
}

extern "C" int vkdt_hip_lj92_decode(const uint8_t *data, size_t data_len, const vkdt_lj92_tile_t *tiles, int n,
                                    const vkdt_lj92_image_t *img, int nimg, void *stream)
{
  if(n <= 0) return VKDT_HIP_OK;
  std::vector<lj92_hdr> hdr(n);
  std::map<std::vector<uint8_t>, int> table_id;   // identical tables are uploaded once
  std::vector<std::vector<uint8_t>> tables;
  std::vector<int> tid((size_t)n * 4, 0);
  size_t rows = 0;
  // streams with one table for all components and predictor 1 go to the
  // many-lanes decode (kind 4); the others, one lane each, to four kernels
  // (line in LDS or not x one shared table or not)
  std::vector<int> kind(n);
  int count[5] = { 0, 0, 0, 0, 0 };
  for(int i = 0; i < n; i++)
  {
    const vkdt_lj92_tile_t &t = tiles[i];
    if(t.image >= (uint32_t)nimg || t.offset + t.length > data_len || t.tile_wd == 0)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92 tile %d: bad image, range or extent", i);
    if(int e = lj92_parse(data + t.offset, t.length, &hdr[i])) return e;
    const lj92_hdr &h = hdr[i];
    if((long)h.wd * h.comp * h.ht < (long)t.tile_wd * t.tile_ht)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92 tile %d: stream smaller than its tile", i);
    for(int c = 0; c < h.comp; c++)
    {
      std::vector<uint8_t> key(h.huff_bits[h.tbl[c]], h.huff_bits[h.tbl[c]] + 16);
      key.insert(key.end(), h.huff_val[h.tbl[c]], h.huff_val[h.tbl[c]] + 17);
      auto it = table_id.find(key);
      if(it == table_id.end()) { it = table_id.emplace(key, (int)tables.size()).first; tables.push_back(key); }
      tid[4 * i + c] = it->second;
    }
    int same = 1;
    for(int c = 1; c < h.comp; c++) same &= tid[4 * i + c] == tid[4 * i];
    const int lds = h.wd * h.comp <= LJ92_ROW;
    if(same && h.pred == 1 && t.length - h.scan < (1u << 28)) kind[i] = 4;
    else
    {
      if(!lds) rows += (size_t)h.wd * h.comp;
      kind[i] = 2 * lds + same;
    }
    count[kind[i]]++;
  }
  const int nser = n - count[4], npar = count[4];
  // the many-lanes streams: subsequences, scan lines, clean words, samples
  std::vector<lj92_pst> pst(npar);
  size_t nsub = 0, nrows = 0, cwords = 0, nsamp = 0, maxk = 0;
  bool all124 = true;                             // scan lines by waves (k_lj92_rows4)
  for(int i = 0, q = 0; i < n; i++)
  {
    if(kind[i] != 4) continue;
    const lj92_hdr &h = hdr[i];
    lj92_pst &p = pst[q++];
    memset(&p, 0, sizeof(p));
    p.len = (uint32_t)(tiles[i].length - h.scan);
    p.K = (uint32_t)(((size_t)p.len * 8 + LJ92_SUB - 1) / LJ92_SUB);
    p.cwords = (uint32_t)(((size_t)p.len + 16 + 15) / 16 * 4);
    p.S = (uint32_t)((size_t)h.wd * h.ht * h.comp);
    p.sbase = (uint32_t)nsub; nsub += p.K;
    p.rbase = (uint32_t)nrows; nrows += h.ht;
    maxk = p.K > maxk ? p.K : maxk;
    all124 &= h.comp != 3;
    cwords += p.cwords;
    nsamp += (p.S + 7) & ~7u;
  }
  if(nsub >= (1ull << 31) || nrows >= (1ull << 31))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: too many streams in one call");
  // device: [ data | tables | row caches | jobs, status, sync flags ] and for
  // the many-lanes streams [ descriptors | bit lengths | clean words |
  // differences | first columns | start, end x2, count, offset per subsequence ]
  const int maxit = (int)maxk + 2;
  const size_t tab_bytes = tables.size() * sizeof(lj92_tab);
  const size_t off_tab = (data_len + 255) & ~(size_t)255;
  const size_t off_row = off_tab + tab_bytes;
  const size_t off_job = (off_row + rows * 2 + 255) & ~(size_t)255;
  const size_t off_st = off_job + sizeof(lj92_job) * nser;
  const size_t off_flag = off_st + 256;
  const size_t off_pst = (off_flag + sizeof(int) * maxit + 255) & ~(size_t)255;
  const size_t off_nb = (off_pst + sizeof(lj92_pst) * npar + 255) & ~(size_t)255;
  const size_t off_cl = (off_nb + sizeof(uint32_t) * npar + 255) & ~(size_t)255;
  const size_t off_df = off_cl + cwords * 4;
  const size_t off_c0 = (off_df + nsamp * 2 + 255) & ~(size_t)255;
  const size_t off_sub = off_c0 + nrows * 8;
  const size_t total = off_sub + nsub * 4 * 5 + 256;
  std::vector<lj92_tab> tabs(tables.size());
  for(size_t k = 0; k < tables.size(); k++)
    if(int e = lj92_table(tables[k].data(), tables[k].data() + 16, &tabs[k])) return e;
  uint8_t *dev = nullptr;
  if(int e = hip_fail(hipMalloc((void **)&dev, total), "lj92 scratch")) return e;
  // the jobs grouped by kind, then a zero status word; the row caches need
  // no clearing: the first line reads none
  std::vector<uint8_t> jobbuf(off_st - off_job + sizeof(int), 0);
  lj92_job *jobs = (lj92_job *)jobbuf.data();
  int first[4], next[4];
  for(int k = 0, f = 0; k < 4; f += count[k++]) first[k] = next[k] = f;
  size_t r = 0, cw = 0, sp = 0, rw = 0;
  for(int i = 0, q = 0; i < n; i++)
  {
    const vkdt_lj92_tile_t &t = tiles[i];
    const lj92_hdr &h = hdr[i];
    const vkdt_lj92_image_t &im = img[t.image];
    const lj92_tab *dtab = (const lj92_tab *)(dev + off_tab);
    if(kind[i] == 4)
    {
      lj92_pst &p = pst[q++];
      p.src = dev + t.offset + h.scan;
      p.clean = (uint32_t *)(dev + off_cl) + cw; cw += p.cwords;
      p.tab = dtab + tid[4 * i];
      p.diff = (uint16_t *)(dev + off_df) + sp; sp += (p.S + 7) & ~7u;
      p.col0 = (uint64_t *)(dev + off_c0) + rw; rw += h.ht;
      p.fwd = h.wd; p.fht = h.ht; p.comp = h.comp; p.bits = h.bits; p.pt = h.pt;
      p.out = (uint16_t *)im.out; p.img_wd = im.wd; p.img_ht = im.ht; p.pitch = im.pitch;
      p.x0 = (int)t.x0 - (int)im.x0; p.y0 = (int)t.y0 - (int)im.y0; p.tile_wd = t.tile_wd;
      continue;
    }
    lj92_job &j = jobs[next[kind[i]]++];
    j.data = dev + t.offset + h.scan;
    j.len = (long)(t.length - h.scan);
    for(int c = 0; c < 4; c++)
    {
      const int id = tid[4 * i + (c < h.comp ? c : 0)];
      j.tab[c] = dtab + id;
      if(id == tid[4 * i]) j.same |= 1 << c;
    }
    j.bits = h.bits; j.fwd = h.wd; j.fht = h.ht; j.comp = h.comp; j.pred = h.pred; j.pt = h.pt;
    j.out = (uint16_t *)im.out; j.img_wd = im.wd; j.img_ht = im.ht; j.pitch = im.pitch;
    j.x0 = (int)t.x0 - (int)im.x0; j.y0 = (int)t.y0 - (int)im.y0; j.tile_wd = t.tile_wd;
    j.row = nullptr;
    if(!(kind[i] & 2))
    {
      j.row = (uint16_t *)(dev + off_row) + r;
      r += (size_t)h.wd * h.comp;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  int err = hip_fail(hipMemcpyAsync(dev, data, data_len, hipMemcpyHostToDevice, s), "lj92 upload");
  if(!err) err = hip_fail(hipMemcpyAsync(dev + off_tab, tabs.data(), tab_bytes, hipMemcpyHostToDevice, s), "lj92 tables");
  if(!err) err = hip_fail(hipMemcpyAsync(dev + off_job, jobbuf.data(), jobbuf.size(), hipMemcpyHostToDevice, s),
                          "lj92 jobs");
  if(!err && npar)
    err = hip_fail(hipMemcpyAsync(dev + off_pst, pst.data(), sizeof(lj92_pst) * npar, hipMemcpyHostToDevice, s),
                   "lj92 streams");
  const lj92_job *dj = (const lj92_job *)(dev + off_job);
  int *st = (int *)(dev + off_st);
  for(int k = 0; k < 4 && !err; k++)
  {
    if(!count[k]) continue;
    const dim3 grid((count[k] + 63) / 64), block(64);
    if(k == 3) hipLaunchKernelGGL((k_lj92<1, 1>), grid, block, 0, s, dj + first[k], count[k], st);
    else if(k == 2) hipLaunchKernelGGL((k_lj92<1, 0>), grid, block, 0, s, dj + first[k], count[k], st);
    else if(k == 1) hipLaunchKernelGGL((k_lj92<0, 1>), grid, block, 0, s, dj + first[k], count[k], st);
    else hipLaunchKernelGGL((k_lj92<0, 0>), grid, block, 0, s, dj + first[k], count[k], st);
    err = launch_post("lj92");
  }
  if(!err && npar)
  {
    const lj92_pst *dp = (const lj92_pst *)(dev + off_pst);
    uint32_t *nb = (uint32_t *)(dev + off_nb);
    uint32_t *sub = (uint32_t *)(dev + off_sub);
    uint32_t *start = sub, *endA = sub + nsub, *endB = sub + 2 * nsub, *cnt = sub + 3 * nsub, *off = sub + 4 * nsub;
    int *flags = (int *)(dev + off_flag);
    const uint32_t ns = (uint32_t)nsub, nr = (uint32_t)nrows;
    const dim3 gsub((ns + 255) / 256), grow((nr + 255) / 256), b256(256);
    hipLaunchKernelGGL(k_lj92_zero, dim3((maxit + 255) / 256), b256, 0, s, flags, maxit);
    hipLaunchKernelGGL(k_lj92_unstuff, dim3(npar), b256, 0, s, dp, nb);
    hipLaunchKernelGGL(k_lj92_first, gsub, b256, 0, s, dp, npar, ns, (const uint32_t *)nb, start, endA, cnt);
    err = launch_post("lj92 first");
    // resynchronise, two rounds per look at the flags, until a round changes nothing
    for(int it = 0; it < maxit && !err; it += 2)
    {
      int ch[2] = { 1, 1 };
      for(int u = 0; u < 2 && it + u < maxit; u++)
      {
        hipLaunchKernelGGL(k_lj92_sync, gsub, b256, 0, s, dp, npar, ns, (const uint32_t *)nb, start,
                           (const uint32_t *)endA, endB, cnt, flags + it + u);
        uint32_t *tmp = endA; endA = endB; endB = tmp;
      }
      err = launch_post("lj92 sync");
      if(!err) err = hip_fail(hipMemcpyAsync(ch, flags + it, sizeof(int) * (it + 1 < maxit ? 2 : 1),
                                             hipMemcpyDeviceToHost, s), "lj92 sync flags");
      if(!err) err = hip_fail(hipStreamSynchronize(s), "lj92 sync wait");
      if(!err && (ch[0] == 0 || (it + 1 < maxit && ch[1] == 0))) break;
      if(!err && it + 2 >= maxit) err = vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: subsequences did not synchronise");
    }
    if(!err)
    {
      hipLaunchKernelGGL(k_lj92_scan, dim3(npar), b256, 0, s, dp, (const uint32_t *)cnt, off);
      hipLaunchKernelGGL(k_lj92_write, gsub, b256, 0, s, dp, npar, ns, (const uint32_t *)nb, (const uint32_t *)start,
                         (const uint32_t *)off, st);
      hipLaunchKernelGGL(k_lj92_col0, dim3(npar), b256, 0, s, dp);
      if(all124) hipLaunchKernelGGL(k_lj92_rows4, dim3((nr + 3) / 4), b256, 0, s, dp, npar, nr);
      else hipLaunchKernelGGL(k_lj92_rows, grow, b256, 0, s, dp, npar, nr);
      err = launch_post("lj92 write");
    }
  }
  int bad = 0;
  if(!err) err = hip_fail(hipMemcpyAsync(&bad, dev + off_st, sizeof(int), hipMemcpyDeviceToHost, s), "lj92 status");
  if(!err) err = hip_fail(hipStreamSynchronize(s), "lj92 sync");
  (void)hipFree(dev);
  if(err) return err;
  if(bad) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: %d tiles with corrupt entropy-coded data", bad);
  return VKDT_HIP_OK;
}
