/*
 * o-jpg -- baseline JPEG sink (reference o-jpg/main.c).
 *
 * The reference hands the 8-bit rgb scanlines to libjpeg on the host with
 * the quality parameter, luma sampling 2x2 up to quality 90, 2x1 up to 92,
 * 1x1 above (main.c:139-141), optimised Huffman codes and 300 dpi.  Here
 * write_sink_hip encodes the sink's device buffer on the GPU
 * (vkdt_hip_jpg_encode, the same published steps: the tables and the
 * entropy-coded data libjpeg writes with its islow DCT) and writes
 * <filename>.jpg; write_sink does the same from a host buffer (uploaded).
 * Not carried over: libjpeg's float DCT above quality 95, its ifast DCT
 * below 50 and smoothing below 80 (main.c:142-146) -- this encoder always
 * uses islow without smoothing -- and the ICC profile and exiftool copy of
 * the input's EXIF (main.c:152-156,169-210).
 */
#include "pipe/api.h"
#include "vkdt_hip.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

static void mkdir_p(const char *path)
{ // the directories of <filename> (fs_mkdir_p in the reference)
  char tmp[512];
  snprintf(tmp, sizeof(tmp), "%s", path);
  char *slash = strrchr(tmp, '/');
  if(!slash) return;
  *slash = 0;
  for(char *p = tmp + 1; *p; p++)
    if(*p == '/') { *p = 0; mkdir(tmp, 0755); *p = '/'; }
  mkdir(tmp, 0755);
}

static int encode(dt_module_t *module, const void *dev, void *stream)
{
  const char *basename = dt_module_param_string(module, 0);
  const float quality = dt_module_param_float(module, 1)[0];
  const int wd = module->connector[0].roi.wd, ht = module->connector[0].roi.ht;
  const int hs = quality > 92.0f ? 1 : 2, vs = quality > 90.0f ? 1 : 2;
  // the reference's libjpeg setup (o-jpg/main.c:142-146) takes the float DCT
  // above quality 95, the fast integer DCT below 50 and input smoothing
  // below 80; this encoder has the accurate integer DCT without smoothing
  // only, so files outside [80, 95] differ from the reference's
  if(quality > 95.0f || quality < 80.0f)
    fprintf(stderr, "[o-jpg] quality %g: encoded with the integer DCT and no smoothing; the reference's files "
            "differ outside quality 80..95\n", quality);
  uint8_t *jpg = 0;
  size_t len = 0;
  if(vkdt_hip_jpg_encode(dev, wd, ht, wd, (int)quality, hs, vs, &jpg, &len, stream))
  {
    fprintf(stderr, "[o-jpg] %s\n", vkdt_hip_last_error());
    return -1;
  }
  char filename[512];
  snprintf(filename, sizeof(filename), "%s.jpg", basename);
  mkdir_p(filename);
  FILE *f = fopen(filename, "wb");
  int ret = 0;
  if(!f || fwrite(jpg, 1, len, f) != len) { fprintf(stderr, "[o-jpg] cannot write '%s'\n", filename); ret = -1; }
  if(f) fclose(f);
  vkdt_jpg_free(jpg);
  return ret;
}

// No reference counterpart: encode straight from the device buffer.
int write_sink_hip(dt_module_t *module, const void *dev, void *stream)
{
  fprintf(stderr, "[o-jpg] writing '%s'\n", dt_module_param_string(module, 0));
  return encode(module, dev, stream);
}

void write_sink(dt_module_t *module, void *buf, dt_write_sink_params_t *p)
{
  (void)p;
  const size_t bytes = (size_t)module->connector[0].roi.wd * module->connector[0].roi.ht * 4;
  void *dev = 0;
  if(vkdt_hip_alloc(&dev, bytes) || vkdt_hip_upload(dev, buf, bytes, 0) || vkdt_hip_stream_sync(0))
  {
    fprintf(stderr, "[o-jpg] %s\n", vkdt_hip_last_error());
    if(dev) vkdt_hip_free(dev);
    return;
  }
  encode(module, dev, 0);
  vkdt_hip_free(dev);
}
