/*
 * dng.c -- uncompressed CFA DNG reader (TIFF 6 / DNG 1.4 subset).
 *
 * Reads IFD0 and its SubIFDs, picks the full-resolution raw IFD
 * (NewSubFileType 0, PhotometricInterpretation CFA), and extracts the tags
 * i-raw needs: extent, CFA pattern, black/white levels, colour matrices
 * with their calibration illuminants, AsShotNeutral and DefaultCrop.
 * Raw data: 16-bit uncompressed strips, or lossless-JPEG compressed
 * (Compression 7) tiles or strips, kept compressed with their stream list
 * for a GPU decode (vkdt_hip_lj92_decode, from i-raw's read_source_hip) or a
 * host decode on demand (dng_decode_host); anything else is refused with a
 * message.
 */
#include "dng.h"
#include "pipe/dng_opcode.h"
#include "vkdt_hip.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct
{
  const uint8_t *d;
  size_t n;
  int be;   // big endian
}
tiff_t;

static uint16_t rd16(const tiff_t *t, size_t o)
{
  if(o + 2 > t->n) return 0;
  return t->be ? (uint16_t)(t->d[o] << 8 | t->d[o + 1]) : (uint16_t)(t->d[o] | t->d[o + 1] << 8);
}

static uint32_t rd32(const tiff_t *t, size_t o)
{
  if(o + 4 > t->n) return 0;
  return t->be ? ((uint32_t)t->d[o] << 24 | (uint32_t)t->d[o + 1] << 16 | (uint32_t)t->d[o + 2] << 8 | t->d[o + 3])
               : ((uint32_t)t->d[o] | (uint32_t)t->d[o + 1] << 8 | (uint32_t)t->d[o + 2] << 16 | (uint32_t)t->d[o + 3] << 24);
}

static int type_size(uint16_t type)
{
  switch(type)
  {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: return 4;
    case 5: case 10: case 12: return 8;
    default: return 0;
  }
}

typedef struct { uint16_t tag, type; uint32_t cnt; size_t off; } entry_t;

// value i of an entry as double
static double ent_val(const tiff_t *t, const entry_t *e, uint32_t i)
{
  if(i >= e->cnt) return 0.0;
  const int ts = type_size(e->type);
  const size_t o = e->off + (size_t)i * ts;
  switch(e->type)
  {
    case 1: case 7: return o < t->n ? t->d[o] : 0;
    case 6: return o < t->n ? (int8_t)t->d[o] : 0;
    case 3: return rd16(t, o);
    case 8: return (int16_t)rd16(t, o);
    case 4: return rd32(t, o);
    case 9: return (int32_t)rd32(t, o);
    case 5: { const uint32_t a = rd32(t, o), b = rd32(t, o + 4); return b ? (double)a / b : 0.0; }
    case 10: { const int32_t a = (int32_t)rd32(t, o), b = (int32_t)rd32(t, o + 4); return b ? (double)a / b : 0.0; }
    case 11: { uint32_t u = rd32(t, o); float f; memcpy(&f, &u, 4); return f; }
    case 12:
    {
      uint64_t u = t->be ? ((uint64_t)rd32(t, o) << 32 | rd32(t, o + 4)) : ((uint64_t)rd32(t, o + 4) << 32 | rd32(t, o));
      double f;
      memcpy(&f, &u, 8);
      return f;
    }
    default: return 0.0;
  }
}

#define MAX_ENTRIES 256
typedef struct { entry_t e[MAX_ENTRIES]; int n; } ifd_t;

static int read_ifd(const tiff_t *t, size_t off, ifd_t *ifd, size_t *next)
{
  ifd->n = 0;
  if(off + 2 > t->n) return 1;
  const uint16_t cnt = rd16(t, off);
  if(off + 2 + 12 * (size_t)cnt + 4 > t->n) return 1;
  for(int i = 0; i < cnt && ifd->n < MAX_ENTRIES; i++)
  {
    const size_t o = off + 2 + 12 * (size_t)i;
    entry_t *e = ifd->e + ifd->n;
    e->tag = rd16(t, o);
    e->type = rd16(t, o + 2);
    e->cnt = rd32(t, o + 4);
    const int ts = type_size(e->type);
    if(!ts) continue;
    const uint64_t bytes = (uint64_t)ts * e->cnt;
    e->off = bytes <= 4 ? o + 8 : rd32(t, o + 8);
    if(e->off + bytes > t->n) continue;
    ifd->n++;
  }
  if(next) *next = rd32(t, off + 2 + 12 * (size_t)cnt);
  return 0;
}

static const entry_t *find(const ifd_t *ifd, uint16_t tag)
{
  for(int i = 0; i < ifd->n; i++)
    if(ifd->e[i].tag == tag) return ifd->e + i;
  return 0;
}

static double tagv(const tiff_t *t, const ifd_t *a, const ifd_t *b, uint16_t tag, uint32_t i, double def)
{ // look in the raw IFD first, then IFD0
  const entry_t *e = find(a, tag);
  if(!e && b) e = find(b, tag);
  return e ? ent_val(t, e, i) : def;
}

static void copy_str(const tiff_t *t, const entry_t *e, char *dst, size_t n)
{
  dst[0] = 0;
  if(!e) return;
  size_t k = 0;
  for(; k < e->cnt && k + 1 < n; k++)
  {
    const char c = (char)t->d[e->off + k];
    if(!c) break;
    dst[k] = c;
  }
  dst[k] = 0;
}

// DNG LightSource codes -> i-raw's illuminant index (A D50 D55 D60 D65 D75 D93)
static int illuminant_index(int ls)
{
  switch(ls)
  {
    case 17: return 0;   // standard light A
    case 23: return 1;   // D50
    case 20: return 2;   // D55
    case 21: return 4;   // D65
    case 22: return 5;   // D75
    default: return -1;
  }
}

int dng_load(const char *filename, dng_image_t *img, char *err, size_t errlen)
{
  memset(img, 0, sizeof(*img));
  FILE *f = fopen(filename, "rb");
  if(!f) { snprintf(err, errlen, "cannot open %s", filename); return 1; }
  fseek(f, 0, SEEK_END);
  const long len = ftell(f);
  fseek(f, 0, SEEK_SET);
  if(len < 16) { fclose(f); snprintf(err, errlen, "%s: truncated", filename); return 1; }
  uint8_t *buf = malloc(len);
  if(!buf || fread(buf, 1, len, f) != (size_t)len) { free(buf); fclose(f); snprintf(err, errlen, "%s: read error", filename); return 1; }
  fclose(f);
  tiff_t t = { buf, (size_t)len, 0 };
  int ret = 1;
  static ifd_t ifd0, sub, raw;
  if(buf[0] == 'I' && buf[1] == 'I') t.be = 0;
  else if(buf[0] == 'M' && buf[1] == 'M') t.be = 1;
  else { snprintf(err, errlen, "%s: not a TIFF/DNG", filename); goto done; }
  if(rd16(&t, 2) != 42) { snprintf(err, errlen, "%s: bad TIFF magic", filename); goto done; }
  if(read_ifd(&t, rd32(&t, 4), &ifd0, 0)) { snprintf(err, errlen, "%s: bad IFD0", filename); goto done; }
  if(!find(&ifd0, 50706)) { snprintf(err, errlen, "%s: no DNGVersion tag", filename); goto done; }

  // raw IFD: IFD0 if it holds the CFA, else the first full-resolution CFA SubIFD
  const ifd_t *ri = 0;
  if((int)tagv(&t, &ifd0, 0, 254, 0, 0) == 0 && (int)tagv(&t, &ifd0, 0, 262, 0, 0) == 32803) ri = &ifd0;
  const entry_t *subs = find(&ifd0, 330);
  for(uint32_t s = 0; !ri && subs && s < subs->cnt; s++)
  {
    if(read_ifd(&t, (size_t)ent_val(&t, subs, s), &sub, 0)) continue;
    if((int)tagv(&t, &sub, 0, 254, 0, 0) == 0 && (int)tagv(&t, &sub, 0, 262, 0, 0) == 32803) { raw = sub; ri = &raw; }
  }
  if(!ri) { snprintf(err, errlen, "%s: no CFA raw IFD (linear/compressed DNGs are not supported)", filename); goto done; }

  const uint32_t W = (uint32_t)tagv(&t, ri, 0, 256, 0, 0), H = (uint32_t)tagv(&t, ri, 0, 257, 0, 0);
  const int bps = (int)tagv(&t, ri, 0, 258, 0, 0);
  const int comp = (int)tagv(&t, ri, 0, 259, 0, 1);
  const int spp = (int)tagv(&t, ri, 0, 277, 0, 1);
  if(W == 0 || H == 0 || W > 65535 || H > 65535) { snprintf(err, errlen, "%s: bad extent %ux%u", filename, W, H); goto done; }
  if(comp != 1 && comp != 7) { snprintf(err, errlen, "%s: compression %d not supported (none or lossless JPEG)", filename, comp); goto done; }
  if(spp != 1 || (comp == 1 && bps != 16) || bps > 16)
  { snprintf(err, errlen, "%s: %d bits x %d samples not supported (16 bit CFA, or lossless JPEG up to 16 bit)", filename, bps, spp); goto done; }
  if(comp == 7)
  { // lossless JPEG: one stream per tile (TileOffsets) or per strip
    const entry_t *to = find(ri, 324), *tcnt = find(ri, 325);
    const int tiled = to && tcnt;
    if(!tiled) { to = find(ri, 273); tcnt = find(ri, 279); }
    if(!to || !tcnt || to->cnt != tcnt->cnt) { snprintf(err, errlen, "%s: no tiles or strips", filename); goto done; }
    const uint32_t tw = tiled ? (uint32_t)tagv(&t, ri, 0, 322, 0, 0) : W;
    const uint32_t th = tiled ? (uint32_t)tagv(&t, ri, 0, 323, 0, 0) : (uint32_t)tagv(&t, ri, 0, 278, 0, H);
    if(tw == 0 || th == 0) { snprintf(err, errlen, "%s: bad tile extent %ux%u", filename, tw, th); goto done; }
    const uint32_t ntx = (W + tw - 1) / tw, nty = (H + th - 1) / th;
    if(to->cnt < ntx * nty) { snprintf(err, errlen, "%s: %u tiles for %ux%u", filename, to->cnt, ntx, nty); goto done; }
    img->tiles = calloc((size_t)ntx * nty, sizeof(vkdt_lj92_tile_t));
    if(!img->tiles) { snprintf(err, errlen, "out of memory"); goto done; }
    for(uint32_t k = 0; k < ntx * nty; k++)
    {
      const vkdt_lj92_tile_t tile = { (uint64_t)ent_val(&t, to, k), (uint64_t)ent_val(&t, tcnt, k),
                                      (k % ntx) * tw, (k / ntx) * th, tw, th, 0 };
      if(tile.offset + tile.length > t.n) { snprintf(err, errlen, "%s: truncated tile %u", filename, k); goto done; }
      uint32_t fw, fh, fc;   // the stream's header, checked now: decoding comes later
      if(vkdt_lj92_info(buf + tile.offset, tile.length, &fw, &fh, &fc, 0))
      { snprintf(err, errlen, "%s: tile %u: %s", filename, k, vkdt_hip_last_error()); goto done; }
      if((uint64_t)fw * fc * fh != (uint64_t)tw * th)
      { // a larger frame would write into the next tile's rows
        snprintf(err, errlen, "%s: tile %u: lj92 frame %ux%ux%u does not fill its %ux%u tile exactly", filename, k,
                 fw, fh, fc, tw, th);
        goto done;
      }
      img->tiles[k] = tile;
    }
    img->ntiles = ntx * nty;
    img->packed = buf;     // kept for the decode
    img->packed_len = t.n;
    buf = 0;
    goto decoded;
  }
  const entry_t *so = find(ri, 273), *sc = find(ri, 279);
  if(!so || !sc || so->cnt != sc->cnt) { snprintf(err, errlen, "%s: no strips (tiled DNGs are not supported)", filename); goto done; }
  const uint32_t rps = (uint32_t)tagv(&t, ri, 0, 278, 0, H);
  img->data = malloc((size_t)W * H * 2);
  if(!img->data) { snprintf(err, errlen, "out of memory"); goto done; }
  {
    uint32_t row = 0;
    for(uint32_t s = 0; s < so->cnt && row < H; s++)
    {
      const size_t off = (size_t)ent_val(&t, so, s);
      const uint32_t rows = (row + rps > H) ? H - row : rps;
      const size_t need = (size_t)rows * W * 2;
      if(off + need > t.n || (size_t)ent_val(&t, sc, s) < need)
      { snprintf(err, errlen, "%s: truncated strip %u", filename, s); goto done; }
      if(!t.be) memcpy(img->data + (size_t)row * W, buf + off, need);
      else for(size_t k = 0; k < (size_t)rows * W; k++) img->data[(size_t)row * W + k] = rd16(&t, off + 2 * k);
      row += rows;
    }
    if(row < H) { snprintf(err, errlen, "%s: strips cover %u of %u rows", filename, row, H); goto done; }
  }
decoded:
  img->width = W;
  img->height = H;
  img->stride = W;
  img->cpp = 1;

  // CFA pattern (colour codes 0 red, 1 green, 2 blue)
  img->cfa_w = (uint32_t)tagv(&t, ri, &ifd0, 33421, 1, 2);
  img->cfa_h = (uint32_t)tagv(&t, ri, &ifd0, 33421, 0, 2);
  if(!(img->cfa_w == 2 && img->cfa_h == 2) && !(img->cfa_w == 6 && img->cfa_h == 6))
  { snprintf(err, errlen, "%s: CFA %ux%u not supported (2x2 bayer, 6x6 x-trans)", filename, img->cfa_w, img->cfa_h); goto done; }
  const entry_t *pat = find(ri, 33422);
  if(!pat) pat = find(&ifd0, 33422);
  if(!pat || pat->cnt < img->cfa_w * img->cfa_h) { snprintf(err, errlen, "%s: no CFAPattern", filename); goto done; }
  for(uint32_t k = 0; k < img->cfa_w * img->cfa_h; k++)
  {
    const int v = (int)ent_val(&t, pat, k);
    if(v < 0 || v > 2) { snprintf(err, errlen, "%s: CFA colour %d not supported", filename, v); goto done; }
    img->cfa[k] = (uint8_t)v;
  }

  // levels: scalar or per CFA site of a 2x2 repeat -> 4 values
  {
    const entry_t *bl = find(ri, 50714);
    if(!bl) bl = find(&ifd0, 50714);
    for(int k = 0; k < 4; k++)
      img->blacklevels[k] = bl ? (float)ent_val(&t, bl, bl->cnt >= 4 ? (uint32_t)k : 0) : 0.0f;
    const entry_t *wl = find(ri, 50717);
    if(!wl) wl = find(&ifd0, 50717);
    for(int k = 0; k < 4; k++)
      img->whitelevels[k] = wl ? (float)ent_val(&t, wl, wl->cnt >= 4 ? (uint32_t)k : 0) : 65535.0f;
  }

  // white balance from AsShotNeutral
  {
    const entry_t *n = find(&ifd0, 50728);
    if(!n) n = find(ri, 50728);
    for(int k = 0; k < 3; k++)
    {
      const double v = n ? ent_val(&t, n, k) : 1.0;
      img->wb_coeffs[k] = v > 0 ? (float)(1.0 / v) : 1.0f;
    }
    img->wb_coeffs[3] = NAN;
  }

  // colour matrix: illuminant priority D65 > D55 > D50 > D75 > A
  {
    const entry_t *cm[2] = { find(&ifd0, 50721), find(&ifd0, 50722) };
    const int ill[2] = { (int)tagv(&t, &ifd0, 0, 50778, 0, 0), (int)tagv(&t, &ifd0, 0, 50779, 0, 0) };
    const int prio[5] = { 4, 2, 1, 5, 0 };
    int pick = -1;
    for(int p = 0; p < 5 && pick < 0; p++)
      for(int i = 0; i < 2; i++)
        if(cm[i] && cm[i]->cnt >= 9 && illuminant_index(ill[i]) == prio[p]) { pick = i; break; }
    if(pick < 0) { snprintf(err, errlen, "%s: no ColorMatrix with a supported illuminant", filename); goto done; }
    for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) img->xyz_to_cam[j][i] = (float)ent_val(&t, cm[pick], 3 * j + i);
    img->illuminant = (uint32_t)illuminant_index(ill[pick]);
  }

  // crop area: DefaultCropOrigin/Size, else the full buffer
  {
    const entry_t *co = find(ri, 50719), *cs = find(ri, 50720);
    if(co && cs)
    {
      img->crop_aabb[0] = (uint64_t)ent_val(&t, co, 0);
      img->crop_aabb[1] = (uint64_t)ent_val(&t, co, 1);
      img->crop_aabb[2] = img->crop_aabb[0] + (uint64_t)ent_val(&t, cs, 0);
      img->crop_aabb[3] = img->crop_aabb[1] + (uint64_t)ent_val(&t, cs, 1);
    }
    else
    {
      img->crop_aabb[0] = img->crop_aabb[1] = 0;
      img->crop_aabb[2] = W;
      img->crop_aabb[3] = H;
    }
    if(img->crop_aabb[2] > W || img->crop_aabb[3] > H ||
       img->crop_aabb[0] >= img->crop_aabb[2] || img->crop_aabb[1] >= img->crop_aabb[3])
    {
      snprintf(err, errlen, "%s: DefaultCrop %llu,%llu..%llu,%llu outside the %ux%u raw", filename,
               (unsigned long long)img->crop_aabb[0], (unsigned long long)img->crop_aabb[1],
               (unsigned long long)img->crop_aabb[2], (unsigned long long)img->crop_aabb[3], W, H);
      goto done;
    }
  }
  { // OpcodeList2 of the raw IFD (the list applied to the raw data after linearisation)
    const entry_t *ol = find(ri, 51009);
    if(ol && ol->cnt > 0 && type_size(ol->type) == 1)
    {
      img->oplist2 = malloc(ol->cnt);
      if(!img->oplist2) { snprintf(err, errlen, "out of memory"); goto done; }
      memcpy(img->oplist2, buf + ol->off, ol->cnt);
      img->oplist2_len = ol->cnt;
    }
  }
  copy_str(&t, find(&ifd0, 271), img->maker, sizeof(img->maker));
  copy_str(&t, find(&ifd0, 272), img->model, sizeof(img->model));
  img->iso = 0.0f;
  ret = 0;
done:
  if(ret) dng_free(img);
  free(buf);
  return ret;
}

void dng_free(dng_image_t *img)
{
  free(img->data);
  free(img->packed);
  free(img->tiles);
  free(img->oplist2);
  img->data = 0;
  img->packed = 0;
  img->tiles = 0;
  img->oplist2 = 0;
  img->oplist2_len = 0;
  img->packed_len = 0;
  img->ntiles = 0;
}

int dng_decode_host(dng_image_t *img, char *err, size_t errlen)
{
  if(img->data) return 0;
  if(!img->packed) { snprintf(err, errlen, "no raw data"); return 1; }
  const uint32_t W = (uint32_t)img->width, H = (uint32_t)img->height;
  img->data = calloc((size_t)W * H, 2);
  if(!img->data) { snprintf(err, errlen, "out of memory"); return 1; }
  for(uint32_t k = 0; k < img->ntiles; k++)
    if(vkdt_lj92_decode_host(img->packed, img->tiles + k, img->data, W, H, W))
    {
      snprintf(err, errlen, "tile %u: %s", k, vkdt_hip_last_error());
      free(img->data);
      img->data = 0;
      return 1;
    }
  return 0;
}

// ---------------------------------------------------------------------------
// DNG opcode lists (DNG specification 1.4, chapter 7): a big-endian count,
// then per opcode its id, the DNG version it needs, flags (bit 0 optional,
// bit 1 may be skipped for previews), the parameter size in bytes and the
// parameters.  The acceptance rules follow the raw input of the reference
// (i-raw/dng_opcode_decode.c:285-345): the records must tile the list
// exactly, and each known opcode must have the parameter size its counts
// imply (the formulas below), or the whole list is dropped.
// ---------------------------------------------------------------------------

static uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
static float be_f32(const uint8_t *p) { const uint32_t u = be32(p); float f; memcpy(&f, &u, 4); return f; }
static double be_f64(const uint8_t *p)
{
  const uint64_t u = (uint64_t)be32(p) << 32 | be32(p + 4);
  double f;
  memcpy(&f, &u, 8);
  return f;
}

// the parameter size an opcode of this id must have, given its first
// parameters (p, n bytes available); -1 for ids without a rule
static int64_t opcode_size(uint32_t id, const uint8_t *p, uint32_t n)
{
  if(id == 1 || id == 2 || id == 5 || id == 14) { if(n < 4) return -2; }
  switch(id)
  {
    case 1: return 20 + 48 * (int64_t)be32(p);                 // WarpRectilinear
    case 2: return 20 + 32 * (int64_t)be32(p);                 // WarpFisheye
    case 3: return 56;                                         // FixVignetteRadial
    case 4: return 8;                                          // FixBadPixelsConstant
    case 5: return n < 12 ? -2 : 12 + 8 * (int64_t)be32(p + 4) + 16 * (int64_t)be32(p + 8);   // FixBadPixelsList
    case 6: return 16;                                         // TrimBounds
    case 7: return n < 36 ? -2 : 36 + 2 * (int64_t)be32(p + 32);          // MapTable
    case 8: return n < 36 ? -2 : 36 + 8 * (int64_t)be32(p + 32);          // MapPolynomial
    case 9:                                                    // GainMap
    {
      if(n < 76) return -2;
      const uint64_t vh = (uint64_t)be32(p + 32) * be32(p + 36);
      if(vh > (1ull << 32)) return -2;                         // cannot match a 32-bit size
      const uint64_t cnt = vh * be32(p + 72);
      if(cnt > (1ull << 30)) return -2;
      return 76 + 4 * (int64_t)cnt;
    }
    case 14: return 24 + 76 * (int64_t)be32(p);                // WarpRectilinear2
    default: return -1;
  }
}

struct dt_dng_opcode_list_t *dng_opcode_list_decode(const uint8_t *data, size_t len)
{
  if(!data || len < 4) return 0;
  const uint32_t count = be32(data);
  if(count == 0) return 0;
  size_t o = 4;
  for(uint32_t i = 0; i < count; i++)
  { // the records must tile the list exactly
    if(o + 16 > len) return 0;
    const uint32_t sz = be32(data + o + 12);
    if(sz > len - o - 16) return 0;
    o += 16 + (size_t)sz;
  }
  if(o != len) return 0;
  dt_dng_opcode_list_t *list = calloc(1, sizeof(*list) + (size_t)count * sizeof(dt_dng_opcode_t));
  if(!list) return 0;
  list->count = (int)count;
  o = 4;
  for(uint32_t i = 0; i < count; i++)
  {
    const uint32_t id = be32(data + o), flags = be32(data + o + 8), sz = be32(data + o + 12);
    const uint8_t *p = data + o + 16;
    dt_dng_opcode_t *op = list->ops + i;
    op->id = (dt_dng_opcode_id_t)id;
    op->optional = (flags & 1) != 0;
    op->preview_skip = (flags & 2) != 0;
    const int64_t want = opcode_size(id, p, sz);
    if(want != -1 && want != (int64_t)sz) { dng_opcode_list_free(list); return 0; }
    if(id == s_dngop_gain_map)
    {
      const uint64_t cnt = (uint64_t)be32(p + 32) * be32(p + 36) * be32(p + 72);
      dt_dng_gain_map_t *gm = calloc(1, sizeof(*gm) + cnt * sizeof(float));
      if(!gm) { dng_opcode_list_free(list); return 0; }
      dt_dng_region_t *r = &gm->region;
      r->top = be32(p); r->left = be32(p + 4); r->bottom = be32(p + 8); r->right = be32(p + 12);
      r->plane = be32(p + 16); r->planes = be32(p + 20); r->row_pitch = be32(p + 24); r->col_pitch = be32(p + 28);
      gm->map_points_v = be32(p + 32);
      gm->map_points_h = be32(p + 36);
      gm->map_spacing_v = be_f64(p + 40);
      gm->map_spacing_h = be_f64(p + 48);
      gm->map_origin_v = be_f64(p + 56);
      gm->map_origin_h = be_f64(p + 64);
      gm->map_planes = be32(p + 72);
      for(uint64_t k = 0; k < cnt; k++) gm->map_gain[k] = be_f32(p + 76 + 4 * k);
      op->data = gm;
    }
    o += 16 + (size_t)sz;
  }
  return list;
}

void dng_opcode_list_free(struct dt_dng_opcode_list_t *list)
{
  if(!list) return;
  for(int i = 0; i < list->count; i++) free(list->ops[i].data);
  free(list);
}
