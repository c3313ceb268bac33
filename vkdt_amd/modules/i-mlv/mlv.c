/*
 * mlv.c -- Magic Lantern raw video clips: chunk files, block walk, frame
 * index, frame data and the host unpack of bit-packed frames.
 * Follows the reference's reader (i-mlv/video_mlv.c:139-196 chunks,
 * 200-283 frames, 306-634 the block walk).
 */
#include "mlv.h"

#include <stdlib.h>
#include <string.h>
#include <strings.h>

static const char *const mlv_known[] = {
  "MLVI", "VIDF", "AUDF", "RAWI", "RAWC", "WAVI", "EXPO", "LENS", "ELNS", "WBAL", "STYL", "RTCI", "IDNT",
  "INFO", "DISO", "MARK", "ELVL", "DEBG", "VERS", "DARK", "NULL", "BKUP" };

static int known(const uint8_t *t)
{
  for(size_t k = 0; k < sizeof(mlv_known) / sizeof(mlv_known[0]); k++)
    if(!memcmp(t, mlv_known[k], 4)) return 1;
  return 0;
}

// chunk files: name.MLV then name.M00 .. name.M98 while they exist (a name
// not ending in .mlv opens alone)
static int open_chunks(mlv_clip_t *c, const char *filename)
{
  const size_t n = strlen(filename);
  char *name = malloc(n + 1);
  if(!name) return 1;
  memcpy(name, filename, n + 1);
  c->file = calloc(100, sizeof(FILE *));
  if(!c->file) { free(name); return 1; }
  c->file[0] = fopen(name, "rb");
  if(!c->file[0]) { free(name); return 1; }
  c->files = 1;
  const char *dot = strrchr(name, '.');
  if(dot && !strcasecmp(dot + 1, "mlv") && n >= 2)
    for(int s = 0; s < 99; s++)
    {
      snprintf(name + n - 2, 3, "%02d", s);
      FILE *f = fopen(name, "rb");
      if(!f) break;
      c->file[c->files++] = f;
    }
  free(name);
  return 0;
}

// the next known block type at or after the current position (within 128 MiB)
static int seek_known(FILE *f)
{
  const size_t ahead = 128u << 20;
  uint8_t *buf = malloc(ahead);
  if(!buf) return 0;
  const long at = ftell(f);
  const size_t got = fread(buf, 1, ahead, f);
  int found = 0;
  for(size_t i = 0; i + 4 <= got && !found; i++)
    if(!memcmp(buf + i, "VIDF", 4) || !memcmp(buf + i, "AUDF", 4) || !memcmp(buf + i, "NULL", 4) ||
       !memcmp(buf + i, "RTCI", 4))
    {
      fseek(f, at + (long)i, SEEK_SET);
      found = 1;
    }
  free(buf);
  return found;
}

static int read_at(FILE *f, uint64_t off, void *dst, size_t bytes)
{
  return fseek(f, (long)off, SEEK_SET) || fread(dst, 1, bytes, f) != bytes;
}

static void sort_by_time(mlv_frame_t *fr, uint32_t cnt)
{
  for(uint32_t n = cnt; n > 1;)
  {
    uint32_t last = 1;
    for(uint32_t k = 0; k + 1 < n; k++)
      if(fr[k].timestamp > fr[k + 1].timestamp)
      {
        const mlv_frame_t t = fr[k];
        fr[k] = fr[k + 1];
        fr[k + 1] = t;
        last = k + 1;
      }
    n = last;
  }
}

// the audio blocks as one buffer synced to the first video frame
// (video_mlv.c:18-87): audio recorded before the first frame is cut, audio
// that starts later is preceded by silence, in whole samples of all
// channels.  The buffer is rounded up to blocks of 1024 samples, and what
// plays with the clip is at most the duration of its frames at the file's
// frame rate, rounded up the same way.  No audio when the header has no
// audio class, no WAVI channels or no AUDF block, when a block cannot be read,
// or when the cut would take all of it.
static void read_audio(mlv_clip_t *c)
{
  if(!c->hdr.audio_class || !c->audfs || !c->wavi.channels || !c->frames) return;
  uint64_t total = 0;
  for(uint32_t i = 0; i < c->audfs; i++) total += c->audf[i].size;
  uint8_t *raw = malloc(total ? total : 1);
  if(!raw) return;
  uint64_t at = 0;
  for(uint32_t i = 0; i < c->audfs; i++)
  {
    if(read_at(c->file[c->audf[i].chunk], c->audf[i].offset, raw + at, c->audf[i].size)) { free(raw); return; }
    at += c->audf[i].size;
  }
  const uint64_t sample = (uint64_t)c->wavi.channels * c->wavi.bits_per_sample / 8;   // bytes of all channels
  const uint16_t align = (uint16_t)(sample * 1024);
  if(!sample || !align) { free(raw); return; }
  const int64_t sync = (int64_t)(((double)c->frame[0].timestamp - (double)c->audf[0].timestamp)
                                 * (double)(c->wavi.sampling_rate * sample / 1000000.0));
  uint64_t cut = 0, pad = 0;
  if(sync >= 0) cut = (uint64_t)sync - (uint64_t)sync % sample;
  else pad = (uint64_t)(-sync) - (uint64_t)(-sync) % sample;
  if(cut >= total) { free(raw); return; }
  const uint64_t synced = total - cut + pad;
  const uint64_t synced_al = synced - synced % align + align;
  c->audio = calloc(synced_al, 1);
  if(!c->audio) { free(raw); return; }
  memcpy(c->audio + pad, raw + cut, total - cut);
  free(raw);
  // a zero denominator makes the rate infinite and the duration 0, as in the
  // reference's double arithmetic
  const double fps = (double)c->hdr.fps_nom / (double)c->hdr.fps_denom;
  const double bytes = (double)(c->wavi.sampling_rate * sample * c->frames) / fps;
  const uint64_t t = bytes > 0.0 ? (uint64_t)bytes : 0;
  const uint64_t t_al = t - t % align + align;
  c->audio_buffer_size = synced_al;
  c->audio_size = t_al < synced_al ? t_al : synced_al;
}

int mlv_open(mlv_clip_t *c, const char *filename, char *err, size_t errlen)
{
  memset(c, 0, sizeof(*c));
  if(open_chunks(c, filename)) { snprintf(err, errlen, "cannot open %s", filename); mlv_close(c); return 1; }
  uint32_t cap = 0;
  int lens = 0, rtci = 0;
  for(int i = 0; i < c->files; i++)
  {
    FILE *f = c->file[i];
    fseek(f, 0, SEEK_END);
    const uint64_t fsize = (uint64_t)ftell(f);
    mlv_block_t b;
    if(!fsize || read_at(f, 0, &b, sizeof(b)) || memcmp(b.type, "MLVI", 4))
    { snprintf(err, errlen, "%s: chunk %d is no MLV file", filename, i); mlv_close(c); return 1; }
    if(i == 0 && read_at(f, 0, &c->hdr, sizeof(c->hdr)))
    { snprintf(err, errlen, "%s: truncated header", filename); mlv_close(c); return 1; }
    uint64_t pos = 0;
    while(pos < fsize)
    {
      if(read_at(f, pos, &b, sizeof(b))) break;   // a partial block at the end
      if(b.size < sizeof(mlv_block_t)) { snprintf(err, errlen, "%s: block of %u bytes", filename, b.size); mlv_close(c); return 1; }
      if(!known(b.type))
      { // resynchronise on the next block type a writer emits often
        fseek(f, (long)pos, SEEK_SET);
        if(!seek_known(f)) { snprintf(err, errlen, "%s: unknown block at %llu", filename, (unsigned long long)pos); mlv_close(c); return 1; }
        pos = (uint64_t)ftell(f);
        continue;
      }
      int e = 0;
      if(!memcmp(b.type, "VIDF", 4))
      {
        mlv_vidf_t v;
        e = read_at(f, pos, &v, sizeof(v));
        if(!e && (uint64_t)sizeof(v) + v.frame_space <= b.size)
        {
          if(c->frames == cap)
          {
            cap = cap ? 2 * cap : 128;
            mlv_frame_t *nf = realloc(c->frame, cap * sizeof(mlv_frame_t));
            if(!nf) { snprintf(err, errlen, "out of memory"); mlv_close(c); return 1; }
            c->frame = nf;
          }
          c->frame[c->frames++] = (mlv_frame_t){ i, pos + sizeof(v) + v.frame_space, b.size - sizeof(v) - v.frame_space,
                                                 pos, b.timestamp };
        }
      }
      else if(!memcmp(b.type, "AUDF", 4))
      {
        mlv_audf_t a;
        e = read_at(f, pos, &a, sizeof(a));
        if(!e && (uint64_t)sizeof(a) + a.frame_space <= b.size)
        {
          if(c->audfs % 128 == 0)
          {
            mlv_frame_t *na = realloc(c->audf, (c->audfs + 128) * sizeof(mlv_frame_t));
            if(!na) { snprintf(err, errlen, "out of memory"); mlv_close(c); return 1; }
            c->audf = na;
          }
          c->audf[c->audfs++] = (mlv_frame_t){ i, pos + sizeof(a) + a.frame_space, b.size - sizeof(a) - a.frame_space,
                                               pos, b.timestamp };
        }
      }
      else if(!memcmp(b.type, "RAWI", 4)) e = read_at(f, pos, &c->rawi, sizeof(c->rawi));
      else if(!memcmp(b.type, "EXPO", 4)) e = read_at(f, pos, &c->expo, sizeof(c->expo));
      else if(!memcmp(b.type, "IDNT", 4)) e = read_at(f, pos, &c->idnt, sizeof(c->idnt));
      else if(!memcmp(b.type, "WAVI", 4)) e = read_at(f, pos, &c->wavi, sizeof(c->wavi));
      else if(!memcmp(b.type, "LENS", 4) && !lens++)
      {
        e = read_at(f, pos, &c->lens, sizeof(c->lens));
        c->lens.name[31] = 0;
      }
      else if(!memcmp(b.type, "RTCI", 4) && !rtci++) e = read_at(f, pos, &c->rtci, sizeof(c->rtci));
      if(e) { snprintf(err, errlen, "%s: truncated %.4s block", filename, (const char *)b.type); mlv_close(c); return 1; }
      pos += b.size;
    }
  }
  if(!c->frames) { snprintf(err, errlen, "%s: no video frames", filename); mlv_close(c); return 1; }
  if(c->rawi.xres == 0 || c->rawi.yres == 0 || c->rawi.raw.bits_per_pixel < 1 || c->rawi.raw.bits_per_pixel > 16)
  { snprintf(err, errlen, "%s: no usable RAWI block", filename); mlv_close(c); return 1; }
  // frames and audio blocks in timestamp order (stable: equal stamps keep
  // file order)
  sort_by_time(c->frame, c->frames);
  sort_by_time(c->audf, c->audfs);
  c->frame_rate = c->hdr.fps_denom ? (double)c->hdr.fps_nom / c->hdr.fps_denom : 0.0;
  read_audio(c);
  return 0;
}

void mlv_close(mlv_clip_t *c)
{
  if(c->file)
    for(int i = 0; i < 100; i++)
      if(c->file[i]) fclose(c->file[i]);
  free(c->file);
  free(c->frame);
  free(c->audf);
  free(c->audio);
  memset(c, 0, sizeof(*c));
}

uint8_t *mlv_frame_data(mlv_clip_t *c, uint32_t frame, uint64_t *size)
{
  if(frame >= c->frames) return 0;
  const mlv_frame_t *fr = c->frame + frame;
  uint8_t *d = malloc(fr->size + 8);   // room for the unpack's two-word reads
  if(!d) return 0;
  memset(d + fr->size, 0, 8);
  if(read_at(c->file[fr->chunk], fr->offset, d, fr->size)) { free(d); return 0; }
  *size = fr->size;
  return d;
}

// sample i takes `bits` bits from bit i*bits of the stream of 16-bit little
// endian words, each word read from its most significant bit
void mlv_unpack(const uint8_t *packed, uint64_t size, uint16_t *out, uint64_t n, int bits)
{
  const uint32_t mask = (1u << bits) - 1;
  for(uint64_t i = 0; i < n; i++)
  {
    const uint64_t bit = i * (uint64_t)bits, w = bit >> 4;
    const int s = (int)(bit & 15);
    uint32_t v = 0;
    if(2 * w + 1 < size) v = (uint32_t)(packed[2 * w] | packed[2 * w + 1] << 8) << 16;
    if(2 * w + 3 < size) v |= (uint32_t)(packed[2 * w + 2] | packed[2 * w + 3] << 8);
    out[i] = (uint16_t)((v >> (32 - s - bits)) & mask);
  }
}
