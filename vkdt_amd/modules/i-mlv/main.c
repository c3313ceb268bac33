/*
 * i-mlv -- Magic Lantern raw video input (reference i-mlv/main.c).
 *
 * Opens the clip named by the filename parameter, publishes its extent, raw
 * levels, frame count and rate and the capture metadata, and delivers frame
 * min(graph frame, frame count - 1) as a 16-bit rggb mosaic:
 *   read_source      on the host: the frame's bytes unpacked (bit-packed
 *                    frames) or decoded (LJ92) into the staging buffer, as
 *                    the reference does (video_mlv.c:200-283);
 *   read_source_hip  on the GPU: the stored bytes go up as they are and the
 *                    unpack or the lossless-JPEG decode runs on the device,
 *                    straight into the output buffer.
 * The camera's colour matrix comes from the Canon rows of the reference's
 * table of Adobe coefficients by model name (mlv_coeff.h), with the
 * reference's identity fallback for a model the table does not know
 * (main.c:172-181).  The clip's audio, synced to its first frame, is served
 * by the audio callback (main.c:214-230) and described by the snd_* image
 * parameters.
 */
#include "pipe/api.h"
#include "mlv.h"
#include "mlv_coeff.h"
#include "vkdt_hip.h"

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct
{
  char       filename[PATH_MAX];   // the open clip
  mlv_clip_t clip;
}
mlv_mod_t;

static int mat3inv_f(float *dst, const float *s);

static int open_clip(dt_module_t *mod)
{
  mlv_mod_t *d = mod->data;
  const char *fname = dt_module_param_string(mod, 0);
  char filename[PATH_MAX];
  if(dt_graph_get_resource_filename(mod, fname, 0, filename, sizeof(filename)))
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "i-mlv: cannot find %s", fname);
    return 1;
  }
  if(d->filename[0] && !strcmp(d->filename, filename)) return 0;
  mlv_close(&d->clip);
  d->filename[0] = 0;
  char err[512];
  if(mlv_open(&d->clip, filename, err, sizeof(err)))
  {
    fprintf(stderr, "[i-mlv] %s\n", err);
    snprintf(mod->graph->err, sizeof(mod->graph->err), "%s", err);
    return 1;
  }
  snprintf(d->filename, sizeof(d->filename), "%s", filename);
  return 0;
}

// main.c:88-91: the graph's frame, held at the last one the header announces
static uint32_t frame_index(const dt_module_t *mod)
{
  const mlv_clip_t *c = &((const mlv_mod_t *)mod->data)->clip;
  uint32_t f = mod->graph->frame < 0 ? 0 : (uint32_t)mod->graph->frame;
  if(c->hdr.video_frames && f > c->hdr.video_frames - 1) f = c->hdr.video_frames - 1;
  if(f > c->frames - 1) f = c->frames - 1;
  return f;
}

int init(dt_module_t *mod)
{
  mod->data = calloc(1, sizeof(mlv_mod_t));
  if(!mod->data) return 1;
  mod->flags = s_module_request_read_source;
  return 0;
}

void cleanup(dt_module_t *mod)
{
  if(!mod->data) return;
  mlv_close(&((mlv_mod_t *)mod->data)->clip);
  free(mod->data);
  mod->data = 0;
}

void modify_roi_out(dt_graph_t *graph, dt_module_t *mod)
{
  if(open_clip(mod)) return;
  const mlv_clip_t *c = &((mlv_mod_t *)mod->data)->clip;
  const uint32_t W = c->rawi.xres, H = c->rawi.yres;
  mod->connector[0].roi.full_wd = W;
  mod->connector[0].roi.full_ht = H;
  dt_image_params_t *ip = &mod->img_param;
  const float b = (float)c->rawi.raw.black_level, w = (float)c->rawi.raw.white_level;
  memset(ip, 0, sizeof(*ip));
  for(int k = 0; k < 4; k++) { ip->black[k] = b; ip->white[k] = w; ip->whitebalance[k] = 1.0f; }
  ip->filters = 0x5d5d5d5du;            // any value but 0 and 9: bayer from red (main.c:116)
  ip->crop_aabb[2] = W;
  ip->crop_aabb[3] = H;
  ip->exposure = (float)(c->expo.shutter_us * 1e-6);
  ip->aperture = c->lens.aperture * 0.01f;
  ip->iso = (float)c->expo.iso;
  ip->focal_length = c->lens.focal_length;
  ip->colour_primaries = s_colour_primaries_custom;
  ip->colour_trc = s_colour_trc_linear;
  ip->snd_samplerate = c->wavi.sampling_rate;   // main.c:136-138: S16_LE whatever WAVI says
  ip->snd_format = 2;
  ip->snd_channels = c->wavi.channels;
  ip->noise_a = 1.0f;                   // main.c:128-129, unless a noise profile says otherwise
  ip->noise_b = 1.0f;
  snprintf(ip->datetime, sizeof(ip->datetime), "%4d%2d%2d %2d:%2d%2d", c->rtci.year, c->rtci.mon, c->rtci.mday,
           c->rtci.hour, c->rtci.min, c->rtci.sec);
  char name[33];
  memcpy(name, c->idnt.camera_name, 32);
  name[32] = 0;
  snprintf(ip->model, sizeof(ip->model), "%s", name);
  snprintf(ip->maker, sizeof(ip->maker), "%s", name);
  char *sp = strchr(ip->maker, ' ');   // the maker is the name's first word
  if(sp) *sp = 0;
  graph->frame_cnt = c->hdr.video_frames ? (int)c->hdr.video_frames : (int)c->frames;
  graph->frame_rate = (float)c->frame_rate;

  char pname[512];   // main.c:145-160
  snprintf(pname, sizeof(pname), "nprof/%s-%s-%d.nprof", ip->maker, ip->model, (int)ip->iso);
  FILE *f = dt_graph_open_resource(graph, 0, pname, "rb");
  if(f)
  {
    float na = 0.0f, nb = 0.0f;
    if(fscanf(f, "%g %g", &na, &nb) == 2) { ip->noise_a = na; ip->noise_b = nb; }
    fclose(f);
  }
  // main.c:172-201: camera rgb -> xyz is the inverse of the body's xyz ->
  // camera matrix (the identity for an unknown body); the row sums of
  // cam -> xyz are the white balance baked into it, and cam -> rec2020 is
  // xyz -> rec2020 times cam -> xyz
  float xyz_to_cam[9], mat[9] = { 0 };
  if(mlv_coeff_lookup(ip->model, xyz_to_cam) || mat3inv_f(mat, xyz_to_cam))
  {
    fprintf(stderr, "[i-mlv] no colour matrix for camera '%s': camera rgb taken as xyz\n", ip->model);
    memset(mat, 0, sizeof(mat));
    mat[0] = mat[4] = mat[8] = 1.0f;
  }
  const double cam_to_xyz[9] = { mat[0], mat[1], mat[2], mat[3], mat[4], mat[5], mat[6], mat[7], mat[8] };
  ip->whitebalance[0] = (float)(cam_to_xyz[0] + cam_to_xyz[1] + cam_to_xyz[2]);
  ip->whitebalance[1] = (float)(cam_to_xyz[3] + cam_to_xyz[4] + cam_to_xyz[5]);
  ip->whitebalance[2] = (float)(cam_to_xyz[6] + cam_to_xyz[7] + cam_to_xyz[8]);
  ip->whitebalance[0] /= ip->whitebalance[1];
  ip->whitebalance[2] /= ip->whitebalance[1];
  ip->whitebalance[1] = 1.0f;
  static const float xyz_to_rec2020[9] = {
     1.7166511880f, -0.3556707838f, -0.2533662814f,
    -0.6666843518f,  1.6164812366f,  0.0157685458f,
     0.0176398574f, -0.0427706133f,  0.9421031212f };
  float c2r[9] = { 0 };
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) for(int k = 0; k < 3; k++)
    c2r[3 * j + i] += xyz_to_rec2020[3 * j + k] * cam_to_xyz[3 * k + i];
  for(int k = 0; k < 9; k++) ip->cam_to_rec2020[k] = c2r[k];
}

// core/mat3.h mat3inv: the inverse of a row-major 3x3 in float; 1 if singular
static int mat3inv_f(float *dst, const float *s)
{
  const float det = s[0] * (s[8] * s[4] - s[7] * s[5]) - s[3] * (s[8] * s[1] - s[7] * s[2])
                  + s[6] * (s[5] * s[1] - s[4] * s[2]);
  if(det > -1e-7f && det < 1e-7f) return 1;
  const float id = 1.0f / det;
  dst[0] =  id * (s[8] * s[4] - s[7] * s[5]);
  dst[1] = -id * (s[8] * s[1] - s[7] * s[2]);
  dst[2] =  id * (s[5] * s[1] - s[4] * s[2]);
  dst[3] = -id * (s[8] * s[3] - s[6] * s[5]);
  dst[4] =  id * (s[8] * s[0] - s[6] * s[2]);
  dst[5] = -id * (s[5] * s[0] - s[3] * s[2]);
  dst[6] =  id * (s[7] * s[3] - s[6] * s[4]);
  dst[7] = -id * (s[7] * s[0] - s[6] * s[1]);
  dst[8] =  id * (s[4] * s[0] - s[3] * s[1]);
  return 0;
}

static int frame_bytes(dt_module_t *mod, uint8_t **data, uint64_t *size, uint32_t *n, int *bits)
{
  mlv_clip_t *c = &((mlv_mod_t *)mod->data)->clip;
  *n = (uint32_t)c->rawi.xres * c->rawi.yres;
  *bits = c->rawi.raw.bits_per_pixel;
  *data = mlv_frame_data(c, frame_index(mod), size);
  if(!*data)
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "i-mlv: cannot read frame %u", frame_index(mod));
    return 1;
  }
  if(!(c->hdr.video_class & MLV_CLASS_FLAG_LJ92) && *size < ((uint64_t)*n * *bits + 7) / 8)
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "i-mlv: frame %u is short", frame_index(mod));
    free(*data);
    return 1;
  }
  return 0;
}

int read_source(dt_module_t *mod, void *mapped, dt_read_source_params_t *p)
{
  (void)p;
  if(open_clip(mod)) return 1;
  mlv_clip_t *c = &((mlv_mod_t *)mod->data)->clip;
  uint8_t *data;
  uint64_t size;
  uint32_t n;
  int bits;
  if(frame_bytes(mod, &data, &size, &n, &bits)) return 1;
  int err = 0;
  if(c->hdr.video_class & MLV_CLASS_FLAG_LJ92)
  { // one stream for the frame, its samples row-major over xres x yres
    const vkdt_lj92_tile_t t = { 0, size, 0, 0, c->rawi.xres, c->rawi.yres, 0 };
    err = vkdt_lj92_decode_host(data, &t, mapped, c->rawi.xres, c->rawi.yres, c->rawi.xres);
    if(err) snprintf(mod->graph->err, sizeof(mod->graph->err), "i-mlv: %s", vkdt_hip_last_error());
  }
  else mlv_unpack(data, size, mapped, n, bits);
  free(data);
  return err ? 1 : 0;
}

// No reference counterpart: the frame's stored bytes are uploaded and unpacked
// or decoded on the device (0 done, -1 error).
int read_source_hip(dt_module_t *mod, void *dev, void *stream)
{
  if(open_clip(mod)) return -1;
  mlv_clip_t *c = &((mlv_mod_t *)mod->data)->clip;
  uint8_t *data;
  uint64_t size;
  uint32_t n;
  int bits, err;
  if(frame_bytes(mod, &data, &size, &n, &bits)) return -1;
  if(c->hdr.video_class & MLV_CLASS_FLAG_LJ92)
  {
    const vkdt_lj92_tile_t t = { 0, size, 0, 0, c->rawi.xres, c->rawi.yres, 0 };
    const vkdt_lj92_image_t win = { dev, c->rawi.xres, c->rawi.yres, c->rawi.xres, 0, 0 };
    err = vkdt_hip_lj92_decode(data, size, &t, 1, &win, 1, stream);
  }
  else err = vkdt_hip_unpack_msb16(data, size, dev, n, bits, stream);
  free(data);
  if(err)
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "i-mlv: %s", vkdt_hip_last_error());
    return -1;
  }
  return 0;
}

// main.c:214-230: the clip's audio from sample `beg` on; returns how many
// samples are there, the smaller of `cnt` and the bytes left, in samples
// (the reference compares a sample count with a byte count)
int audio(dt_module_t *mod, uint64_t beg, uint32_t cnt, uint8_t **samples)
{
  mlv_mod_t *d = mod->data;
  if(!d || !d->filename[0] || !d->clip.audio || !d->clip.audio_size || !d->clip.wavi.sampling_rate) return 0;
  const uint64_t bps = d->clip.wavi.bytes_per_second / d->clip.wavi.sampling_rate;
  if(!bps) return 0;
  if(beg >= d->clip.audio_size / bps + 1) return 0;
  *samples = d->clip.audio + beg * bps;
  if(*samples >= d->clip.audio + d->clip.audio_size) return 0;
  const uint64_t left = (uint64_t)(d->clip.audio + d->clip.audio_size - *samples);
  return (int)((cnt < left ? cnt : left) / bps);
}
