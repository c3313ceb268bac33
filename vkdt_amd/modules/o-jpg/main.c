/*
 * o-jpg -- baseline JPEG sink (reference o-jpg/main.c).
 *
 * The reference hands the 8-bit rgb scanlines to libjpeg on the host
 * (main.c:120-166): the quality parameter, luma sampling 2x2 up to quality
 * 90, 2x1 up to 92, 1x1 above; the float DCT above quality 95, the fast
 * integer DCT below 50, the accurate one between; input smoothing 20 below
 * quality 80, 40 below 60, 60 below 40; optimised Huffman codes; 300 dpi;
 * the Adobe-compatible ICC profile for Adobe primaries with a gamma curve and
 * the compact Rec.2020 one for Rec.2020 primaries with the BT.709 curve.
 * Here write_sink_hip encodes the sink's device buffer on the GPU with the
 * same settings (vkdt_hip_jpg_encode_ex) and writes <filename>.jpg;
 * write_sink does the same from a host buffer (uploaded).  With the exif
 * parameter set, exiftool (next to the modules) then copies the metadata of
 * the main i-raw or i-jpg input into the file, as the reference does
 * (main.c:169-210).
 */
#include "pipe/api.h"
#include "vkdt_hip.h"
#include "icc.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

const char *dt_pipe_basedir(void);

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

// the source file of the graph's main raw or jpg input for `frame`, or ""
static void main_input_file(dt_module_t *module, char *out, size_t size)
{
  out[0] = 0;
  dt_graph_t *g = module->graph;
  if(!g) return;
  for(int m = 0; m < g->num_modules; m++)
  {
    dt_module_t *in = g->module + m;
    if(in->inst != dt_token("main") || (in->name != dt_token("i-raw") && in->name != dt_token("i-jpg"))) continue;
    const int pid = dt_module_get_param(in->so, dt_token("startid"));
    const int pfn = dt_module_get_param(in->so, dt_token("filename"));
    if(pfn < 0) continue;
    const int id = pid >= 0 ? dt_module_param_int(in, pid)[0] : 0;
    if(dt_graph_get_resource_filename(in, dt_module_param_string(in, pfn), g->frame + id, out, size)) out[0] = 0;
    return;
  }
}

static void copy_exif(dt_module_t *module, const char *filename)
{
  const int pex = dt_module_get_param(module->so, dt_token("exif"));
  if(pex < 0 || !dt_module_param_int(module, pex)[0]) return;
  char src[1024];
  main_input_file(module, src, sizeof(src));
  if(!src[0]) return;
  char cmd[3200];
  if(snprintf(cmd, sizeof(cmd),
              "\"%s/exiftool\" -TagsFromFile \"%s\" \"-all:all>all:all\" -Software=\"vkdt\" -ModifyDate=\"now\" "
              "-*image*= -*orientation*= -orientation=normal -overwrite_original \"%s\"",
              dt_pipe_basedir(), src, filename) >= (int)sizeof(cmd)) return;
  FILE *f = popen(cmd, "r");
  int ret = -1;
  if(f)
  {
    while(!feof(f) && !ferror(f) && fgetc(f) != EOF);
    ret = pclose(f);
  }
  if(ret && module->graph)
    snprintf(module->graph->err, sizeof(module->graph->err),
             "o-jpg: unable to run exiftool to copy metadata! maybe you need to install it?");
}

static int encode(dt_module_t *module, const void *dev, void *stream)
{
  const char *basename = dt_module_param_string(module, 0);
  const float quality = dt_module_param_float(module, 1)[0];
  const int wd = module->connector[0].roi.wd, ht = module->connector[0].roi.ht;
  const int hs = quality > 92.0f ? 1 : 2, vs = quality > 90.0f ? 1 : 2;
  const int dct = quality > 95.0f ? VKDT_JPG_DCT_FLOAT : (quality < 50.0f ? VKDT_JPG_DCT_IFAST : VKDT_JPG_DCT_ISLOW);
  const int smooth = quality < 40.0f ? 60 : (quality < 60.0f ? 40 : (quality < 80.0f ? 20 : 0));
  const uint8_t *icc = 0;
  size_t icc_len = 0;
  const dt_image_params_t *ip = &module->img_param;
  if(ip->colour_primaries == s_colour_primaries_adobe && ip->colour_trc == s_colour_trc_gamma)
  { icc = ojpg_icc_adobe; icc_len = sizeof(ojpg_icc_adobe); }
  else if(ip->colour_primaries == s_colour_primaries_2020 && ip->colour_trc == s_colour_trc_709)
  { icc = ojpg_icc_rec2020; icc_len = sizeof(ojpg_icc_rec2020); }
  uint8_t *jpg = 0;
  size_t len = 0;
  // libjpeg takes the quality as an int (jpeg_set_quality)
  if(vkdt_hip_jpg_encode_ex(dev, wd, ht, wd, (int)quality, hs, vs, dct, smooth, icc, icc_len, &jpg, &len, stream))
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
  if(!ret) copy_exif(module, filename);
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
