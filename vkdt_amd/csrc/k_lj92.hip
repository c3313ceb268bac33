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
      return VKDT_HIP_OK;
    }
    else if(m == 0xd9) break;
    else if(m >= 0xc0 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc)
      return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: not a lossless frame (SOF%d)", m - 0xc0);
    i += 2 + L;
  }
  return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: no scan");
}

// canonical codes (T.81 C.2) -> 16-bit lookup: entry = category << 8 | length,
// 0 for prefixes no code matches; fails on tables whose codes overflow
int lj92_lut(const uint8_t bits[16], const uint8_t *val, uint16_t *lut)
{
  memset(lut, 0, sizeof(uint16_t) << 16);
  int code = 0, k = 0;
  for(int len = 1; len <= 16; len++)
  {
    for(int j = 0; j < bits[len - 1]; j++, k++, code++)
    {
      if(code >= (1 << len) || val[k] > 16) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: bad Huffman table");
      const int lo = code << (16 - len), cnt = 1 << (16 - len);
      for(int e = 0; e < cnt; e++) lut[lo + e] = (uint16_t)((val[k] << 8) | len);
    }
    code <<= 1;
  }
  return VKDT_HIP_OK;
}

} // namespace

// ---- the decoder, host and device alike ------------------------------------

struct lj92_job
{
  const uint8_t *data;     // entropy-coded segment
  long     len;
  const uint16_t *lut[4];  // per component
  int      bits, fwd, fht, comp, pred, pt;
  uint16_t *out;           // the image
  int      img_wd, img_ht, pitch;
  int      x0, y0, tile_wd;
  uint16_t *row;           // fwd * comp samples: the previous line
};

// bit reader over a segment with 0xFF00 stuffing; a marker ends the data
// (zeros are fed past it)
struct lj92_bits
{
  const uint8_t *p;
  long n, i;
  uint64_t acc;
  int cnt;
};

__host__ __device__ static inline void lj92_fill(lj92_bits &b)
{
  while(b.cnt <= 56)
  {
    uint32_t byte = 0;
    if(b.i < b.n)
    {
      byte = b.p[b.i];
      if(byte == 0xff)
      {
        if(b.i + 1 < b.n && b.p[b.i + 1] == 0x00) b.i += 2;
        else { byte = 0; b.i = b.n; }   // a marker: the segment ends
      }
      else b.i++;
    }
    b.acc |= (uint64_t)byte << (56 - b.cnt);
    b.cnt += 8;
  }
}

__host__ __device__ static inline int lj92_get(lj92_bits &b, int k)
{
  const int v = (int)(b.acc >> (64 - k));
  b.acc <<= k;
  b.cnt -= k;
  return v;
}

// returns 0, or -1 on a prefix no code matches
__host__ __device__ static int lj92_run(const lj92_job &jg)
{
  const lj92_job j = jg;   // in registers: the stores below cannot alias it
  lj92_bits b = { j.data, j.len, 0, 0, 0 };
  const int C = j.comp, W = j.fwd;
  const int base = 1 << (j.bits - j.pt - 1);
  int tx = 0, ty = 0;      // tile site of the next sample
  for(int r = 0; r < j.fht; r++)
  {
    int left[4] = { 0, 0, 0, 0 }, ul[4] = { 0, 0, 0, 0 };
    for(int x = 0; x < W; x++)
      for(int c = 0; c < C; c++)
      {
        if(b.cnt < 32) lj92_fill(b);
        const uint16_t e = j.lut[c][(int)(b.acc >> 48)];
        if(!e) return -1;
        lj92_get(b, e & 0xff);
        const int t = e >> 8;
        int diff;
        if(t == 0) diff = 0;
        else if(t == 16) diff = 32768;
        else
        {
          if(b.cnt < 32) lj92_fill(b);
          diff = lj92_get(b, t);
          if(diff < (1 << (t - 1))) diff -= (1 << t) - 1;
        }
        const int k = x * C + c;
        const int up = j.row[k];
        int p;
        if(r == 0) p = x == 0 ? base : left[c];
        else if(x == 0) p = up;
        else
        {
          const int a = left[c], bb = up, cc = ul[c];
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
        ul[c] = up;
        left[c] = v;
        j.row[k] = (uint16_t)v;
        // the scan fills the tile row-major
        const int ox = j.x0 + tx, oy = j.y0 + ty;
        if(ox < j.img_wd && oy < j.img_ht) j.out[(long)oy * j.pitch + ox] = (uint16_t)(v << j.pt);
        if(++tx == j.tile_wd) { tx = 0; ty++; }
      }
  }
  return 0;
}

__global__ void __launch_bounds__(64) k_lj92(const lj92_job *jobs, int n, int *status)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  if(lj92_run(jobs[i])) atomicAdd(status, 1);
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
  std::vector<uint16_t> luts((size_t)h.comp << 16), row((size_t)h.wd * h.comp, 0);
  lj92_job j = {};
  for(int c = 0; c < h.comp; c++)
  {
    if(int e = lj92_lut(h.huff_bits[h.tbl[c]], h.huff_val[h.tbl[c]], luts.data() + ((size_t)c << 16))) return e;
    j.lut[c] = luts.data() + ((size_t)c << 16);
  }
  j.data = data + t->offset + h.scan;
  j.len = (long)(t->length - h.scan);
  j.bits = h.bits; j.fwd = h.wd; j.fht = h.ht; j.comp = h.comp; j.pred = h.pred; j.pt = h.pt;
  j.out = out; j.img_wd = wd; j.img_ht = ht; j.pitch = pitch;
  j.x0 = t->x0; j.y0 = t->y0; j.tile_wd = t->tile_wd;
  j.row = row.data();
  if(lj92_run(j)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: corrupt entropy-coded data");
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
    rows += (size_t)h.wd * h.comp;
  }
  // device: [ data | luts | row caches | jobs, status ]
  const size_t lut_bytes = tables.size() * (sizeof(uint16_t) << 16);
  const size_t off_lut = (data_len + 255) & ~(size_t)255;
  const size_t off_row = off_lut + lut_bytes;
  const size_t off_job = (off_row + rows * 2 + 255) & ~(size_t)255;
  const size_t off_st = off_job + sizeof(lj92_job) * n;
  const size_t total = off_st + 256;
  std::vector<uint16_t> luts(tables.size() << 16);
  for(size_t k = 0; k < tables.size(); k++)
    if(int e = lj92_lut(tables[k].data(), tables[k].data() + 16, luts.data() + (k << 16))) return e;
  uint8_t *dev = nullptr;
  if(int e = hip_fail(hipMalloc((void **)&dev, total), "lj92 scratch")) return e;
  // the jobs, then a zero status word (the row caches need no clearing: the
  // first line reads none)
  std::vector<uint8_t> jobbuf(off_st - off_job + sizeof(int), 0);
  lj92_job *jobs = (lj92_job *)jobbuf.data();
  size_t r = 0;
  for(int i = 0; i < n; i++)
  {
    const vkdt_lj92_tile_t &t = tiles[i];
    const lj92_hdr &h = hdr[i];
    const vkdt_lj92_image_t &im = img[t.image];
    lj92_job &j = jobs[i];
    j.data = dev + t.offset + h.scan;
    j.len = (long)(t.length - h.scan);
    for(int c = 0; c < 4; c++) j.lut[c] = (const uint16_t *)(dev + off_lut) + ((size_t)tid[4 * i + (c < h.comp ? c : 0)] << 16);
    j.bits = h.bits; j.fwd = h.wd; j.fht = h.ht; j.comp = h.comp; j.pred = h.pred; j.pt = h.pt;
    j.out = (uint16_t *)im.out; j.img_wd = im.wd; j.img_ht = im.ht; j.pitch = im.pitch;
    j.x0 = t.x0; j.y0 = t.y0; j.tile_wd = t.tile_wd;
    j.row = (uint16_t *)(dev + off_row) + r;
    r += (size_t)h.wd * h.comp;
  }
  hipStream_t s = (hipStream_t)stream;
  int err = hip_fail(hipMemcpyAsync(dev, data, data_len, hipMemcpyHostToDevice, s), "lj92 upload");
  if(!err) err = hip_fail(hipMemcpyAsync(dev + off_lut, luts.data(), lut_bytes, hipMemcpyHostToDevice, s), "lj92 luts");
  if(!err) err = hip_fail(hipMemcpyAsync(dev + off_job, jobbuf.data(), jobbuf.size(), hipMemcpyHostToDevice, s),
                          "lj92 jobs");
  if(!err)
  {
    hipLaunchKernelGGL(k_lj92, dim3((n + 63) / 64), dim3(64), 0, s, (const lj92_job *)(dev + off_job), n,
                       (int *)(dev + off_st));
    err = launch_post("lj92");
  }
  int bad = 0;
  if(!err) err = hip_fail(hipMemcpyAsync(&bad, dev + off_st, sizeof(int), hipMemcpyDeviceToHost, s), "lj92 status");
  if(!err) err = hip_fail(hipStreamSynchronize(s), "lj92 sync");
  hipFree(dev);
  if(err) return err;
  if(bad) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "lj92: %d tiles with corrupt entropy-coded data", bad);
  return VKDT_HIP_OK;
}
