/*
 * i-raw: raw source of the development graph (uncompressed CFA DNG).
 *
 * Callbacks follow src/pipe/modules/i-raw/main.c: modify_roi_out loads the
 * file and publishes extent and image metadata (main.c:138-259),
 * read_source copies the CFA rows into the mapped staging buffer (lossless-JPEG
 * raws are decoded on the GPU by read_source_hip instead, when the runtime has one)
 * (main.c:261-299).  The CFA offset normalisation and block rounding of the
 * raw loader (rawloader-c/lib.rs:200-283) are applied here, so downstream
 * kernels always see RGGB or the canonical X-Trans layout.
 */
#include "pipe/api.h"
#include "pipe/dng_opcode.h"
#include "dng.h"

#include <limits.h>
#include <stdlib.h>

typedef struct raw_t
{
  dng_image_t img;
  char     filename[PATH_MAX];
  int      frame;
  int      ox, oy;        // CFA offset into the decoded buffer
  uint64_t width, height; // after offset and block rounding
  uint32_t filters;
  uint64_t crop_aabb[4];
  dt_dng_opcode_list_t *ops[3];     // decoded OpcodeList1..3 (only 2 is read), or 0
  dt_image_metadata_dngop_t dngop;  // attached to img_param.meta while ops[] has a list
}
raw_t;

static void free_ops(raw_t *r)
{
  for(int i = 0; i < 3; i++)
  {
    dng_opcode_list_free(r->ops[i]);
    r->ops[i] = 0;
  }
}

// matrices.h: CAT16 adaptations of XYZ from the given white to D65
static const float cat16_a_to_d65[9]   = { 9.50674182e-01f, -1.87430902e-01f, 2.62831155e-01f, -2.56724729e-02f, 1.03231456e+00f, -1.15608371e-02f, -2.74089665e-03f, 9.09809774e-02f, 2.81290019e+00f };
static const float cat16_d50_to_d65[9] = { 9.89482020e-01f, -3.99715078e-02f, 4.40103520e-02f, -5.39737108e-03f, 1.00665157e+00f, -1.75423920e-03f, -4.03659609e-04f, 1.50625995e-02f, 1.30181385e+00f };
static const float cat16_d55_to_d65[9] = { 9.93809109e-01f, -2.36805157e-02f, 2.51833773e-02f, -3.19186083e-03f, 1.00396313e+00f, -9.86707812e-04f, -2.25775304e-04f, 8.60286147e-03f, 1.17257778e+00f };
static const float cat16_d60_to_d65[9] = { 9.97210818e-01f, -1.07310837e-02f, 1.10446315e-02f, -1.44406069e-03f, 1.00180486e+00f, -4.25428414e-04f, -9.67926098e-05f, 3.76602349e-03f, 1.07563353e+00f };
static const float cat16_d75_to_d65[9] = { 1.00410545e+00f, 1.61206883e-02f, -1.57697661e-02f, 2.16403273e-03f, 9.97220529e-01f, 5.90558927e-04f, 1.33059688e-04f, -5.36122439e-03f, 8.92084722e-01f };
static const float cat16_d93_to_d65[9] = { 1.00874225e+00f, 3.48643084e-02f, -3.24261661e-02f, 4.66935441e-03f, 9.93885810e-01f, 1.17801519e-03f, 2.62546309e-04f, -1.09895081e-02f, 7.78301142e-01f };
static const float xyz_to_rec2020[9]   = { 1.71665119f, -0.35567078f, -0.25336628f, -0.66668435f, 1.61648124f, 0.01576855f, 0.01763986f, -0.04277061f, 0.94210312f };

// 3x3 row-major product and inverse in float (core/mat3.h semantics)
static void mat3mul(float *dst, const float *a, const float *b)
{
  for(int i = 0; i < 9; i++) dst[i] = 0.0f;
  for(int k = 0; k < 3; k++)
    for(int i = 0; i < 3; i++)
      for(int j = 0; j < 3; j++) dst[3 * k + i] += a[3 * k + j] * b[3 * j + i];
}

static int mat3inv(float *dst, const float *s)
{ // adjugate / determinant
  const float det = s[0] * (s[8] * s[4] - s[7] * s[5])
                  - s[3] * (s[8] * s[1] - s[7] * s[2])
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

static int cfa_at(const dng_image_t *img, int row, int col)
{
  return img->cfa[(row % img->cfa_h) * img->cfa_w + (col % img->cfa_w)];
}

// rawloader-c/lib.rs:200-279: offset to RGGB / canonical x-trans, block
// rounding.  Returns non-zero (message in err) when the raw or its crop is
// too small to hold one CFA block past the offset.
static int normalise_cfa(raw_t *r, char *err, size_t errlen)
{
  const dng_image_t *img = &r->img;
  int ox = 0, oy = 0;
  if(img->cfa_w == 6)
  {
    r->filters = 9;
    for(int i = 0; i < 6; i++) if(cfa_at(img, 0, i) == 1) { ox = i; break; }
    if(cfa_at(img, 0, ox + 1) != 1 && cfa_at(img, 0, ox + 2) != 1) { oy = 2; ox = (ox + 2) % 3; }
    if(cfa_at(img, oy + 1, ox) == 1) oy = oy + 1;
    if(cfa_at(img, oy, ox + 1) == 1) { if(ox >= 2) ox -= 2; else ox += 1; }
    if(cfa_at(img, oy, ox + 1) == 2) { if(ox < oy) ox += 3; else oy += 3; }
  }
  else
  {
    r->filters = 0x49494949u;
    if(cfa_at(img, 0, 0) == 1)
    {
      if(cfa_at(img, 0, 1) == 0) ox = 1;
      if(cfa_at(img, 0, 1) == 2) oy = 1;
    }
    else if(cfa_at(img, 0, 0) == 2) { ox = 1; oy = 1; }
  }
  const uint64_t block = img->cfa_w == 6 ? 3 : 2, big = img->cfa_w;
  if((uint64_t)img->width < (uint64_t)ox + img->cfa_w || (uint64_t)img->height < (uint64_t)oy + img->cfa_h)
  {
    snprintf(err, errlen, "raw %llux%llu smaller than its %ux%u CFA at offset %d,%d",
             (unsigned long long)img->width, (unsigned long long)img->height, img->cfa_w, img->cfa_h, ox, oy);
    return 1;
  }
  if(img->crop_aabb[2] < (uint64_t)ox + block || img->crop_aabb[3] < (uint64_t)oy + block)
  {
    snprintf(err, errlen, "crop %llux%llu smaller than one CFA block at offset %d,%d",
             (unsigned long long)img->crop_aabb[2], (unsigned long long)img->crop_aabb[3], ox, oy);
    return 1;
  }
  uint64_t *b = r->crop_aabb;
  for(int k = 0; k < 4; k++) b[k] = img->crop_aabb[k];
  b[2] -= ox;
  b[3] -= oy;
  b[0] = ((b[0] + big - 1) / big) * big;
  b[1] = ((b[1] + big - 1) / big) * big;
  b[2] = (b[2] / block) * block;
  b[3] = (b[3] / block) * block;
  r->width = ((img->width - ox) / block) * block;
  r->height = ((img->height - oy) / block) * block;
  r->ox = ox;
  r->oy = oy;
  return 0;
}

static int load_raw(dt_module_t *mod, int frame, const char *filename)
{
  raw_t *r = mod->data;
  if(!strcmp(r->filename, filename) && r->frame == frame && (r->img.data || r->img.packed)) return 0;
  dng_free(&r->img);
  free_ops(r);
  r->filename[0] = 0;
  char err[512];
  if(dng_load(filename, &r->img, err, sizeof(err)))
  {
    fprintf(stderr, "[i-raw] %s\n", err);
    snprintf(mod->graph->err, sizeof(mod->graph->err), "%s", err);
    r->frame = -1;
    return 1;
  }
  if(normalise_cfa(r, err, sizeof(err)))
  {
    fprintf(stderr, "[i-raw] %s: %s\n", filename, err);
    snprintf(mod->graph->err, sizeof(mod->graph->err), "%s: %s", filename, err);
    dng_free(&r->img);
    r->frame = -1;
    return 1;
  }
  // OpcodeList2 of the raw IFD, decoded for the modules that apply it
  // (denoise's gain maps; i-raw/main.c:60-92)
  r->ops[1] = dng_opcode_list_decode(r->img.oplist2, r->img.oplist2_len);
  r->dngop = (dt_image_metadata_dngop_t){ .type = s_image_metadata_dngop, .next = 0,
                                          .ox = (uint32_t)r->ox, .oy = (uint32_t)r->oy,
                                          .op_list = { r->ops[0], r->ops[1], r->ops[2] } };
  snprintf(r->filename, sizeof(r->filename), "%s", filename);
  r->frame = frame;
  return 0;
}

int init(dt_module_t *mod)
{
  mod->data = calloc(1, sizeof(raw_t));
  return mod->data ? 0 : 1;
}

void cleanup(dt_module_t *mod)
{
  if(!mod->data) return;
  dng_free(&((raw_t *)mod->data)->img);
  free_ops(mod->data);
  free(mod->data);
  mod->data = 0;
}

dt_graph_run_t check_params(dt_module_t *mod, uint32_t parid, uint32_t num, void *oldval)
{
  (void)num; (void)oldval;
  if(parid == 1 || parid == 2)
  { // noise model: re-run modify_roi_out
    mod->img_param.noise_a = dt_module_param_float(mod, 1)[0];
    mod->img_param.noise_b = dt_module_param_float(mod, 2)[0];
    return s_graph_run_all;
  }
  return s_graph_run_record_cmd_buf;
}

void modify_roi_out(dt_graph_t *graph, dt_module_t *mod)
{
  const int id = dt_module_param_int(mod, 3)[0];
  const char *fname = dt_module_param_string(mod, 0);
  char filename[PATH_MAX];
  if(dt_graph_get_resource_filename(mod, fname, id + graph->frame, filename, sizeof(filename)))
  {
    snprintf(graph->err, sizeof(graph->err), "i-raw: cannot find %s", fname);
    return;
  }
  if(load_raw(mod, id + graph->frame, filename)) return;
  raw_t *r = mod->data;
  mod->connector[0].roi.full_wd = (uint32_t)r->width;
  mod->connector[0].roi.full_ht = (uint32_t)r->height;
  mod->img_param.colour_primaries = s_colour_primaries_custom;
  mod->img_param.colour_trc = s_colour_trc_linear;
  snprintf(mod->img_param.maker, sizeof(mod->img_param.maker), "%s", r->img.maker);
  snprintf(mod->img_param.model, sizeof(mod->img_param.model), "%s", r->img.model);
  mod->img_param.iso = r->img.iso;

  float *noise_a = (float *)dt_module_param_float(mod, 1);
  float *noise_b = (float *)dt_module_param_float(mod, 2);
  if(noise_a[0] == 0.0f && noise_b[0] == 0.0f)
  { // noise profile by camera and iso (i-raw/main.c:172-193)
    char pname[512];
    snprintf(pname, sizeof(pname), "nprof/%s-%s-%d.nprof", mod->img_param.maker, mod->img_param.model,
             (int)mod->img_param.iso);
    FILE *f = dt_graph_open_resource(graph, id, pname, "rb");
    if(f)
    {
      float a = 0.0f, b = 0.0f;
      if(fscanf(f, "%g %g", &a, &b) == 2)
      {
        noise_a[0] = mod->img_param.noise_a = a;
        noise_b[0] = mod->img_param.noise_b = b;
      }
      fclose(f);
    }
  }
  else
  {
    mod->img_param.noise_a = noise_a[0];
    mod->img_param.noise_b = noise_b[0];
  }
  for(int k = 0; k < 4; k++)
  {
    mod->img_param.black[k] = r->img.blacklevels[k];
    mod->img_param.white[k] = r->img.whitelevels[k];
    mod->img_param.whitebalance[k] = r->img.wb_coeffs[k];
    mod->img_param.crop_aabb[k] = (uint32_t)r->crop_aabb[k];
  }
  mod->img_param.whitebalance[0] /= mod->img_param.whitebalance[1];
  mod->img_param.whitebalance[2] /= mod->img_param.whitebalance[1];
  mod->img_param.whitebalance[3] /= mod->img_param.whitebalance[1];
  mod->img_param.whitebalance[1] = 1.0f;
  mod->img_param.filters = r->filters;
  mod->connector[0].chan = dt_token("rggb");
  mod->img_param.meta = (r->ops[0] || r->ops[1] || r->ops[2]) ? (dt_image_metadata_t *)&r->dngop : 0;

  // camera rgb -> rec2020: xyz_to_rec2020 * inverse(xyz_to_cam), with a
  // CAT16 adaptation to D65 if the matrix is for another illuminant
  float xyz_to_cam[9], mat[9];
  for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) xyz_to_cam[3 * j + i] = r->img.xyz_to_cam[j][i];
  const float *adj = 0;
  switch(r->img.illuminant)
  {
    case 0: adj = cat16_a_to_d65; break;
    case 1: adj = cat16_d50_to_d65; break;
    case 2: adj = cat16_d55_to_d65; break;
    case 3: adj = cat16_d60_to_d65; break;
    case 5: adj = cat16_d75_to_d65; break;
    case 6: adj = cat16_d93_to_d65; break;
    default: break;
  }
  if(adj)
  {
    memcpy(mat, xyz_to_cam, sizeof(mat));
    mat3mul(xyz_to_cam, mat, adj);
  }
  float cam_to_xyz[9] = {0}, cam_to_rec2020[9];
  if(mat3inv(cam_to_xyz, xyz_to_cam)) fprintf(stderr, "[i-raw] singular colour matrix\n");
  mat3mul(cam_to_rec2020, xyz_to_rec2020, cam_to_xyz);
  for(int k = 0; k < 9; k++) mod->img_param.cam_to_rec2020[k] = cam_to_rec2020[k];
}

int read_source(dt_module_t *mod, void *mapped, dt_read_source_params_t *p)
{
  (void)p;
  const int id = dt_module_param_int(mod, 3)[0];
  const char *fname = dt_module_param_string(mod, 0);
  char filename[PATH_MAX];
  if(dt_graph_get_resource_filename(mod, fname, id + mod->graph->frame, filename, sizeof(filename))) return 1;
  if(load_raw(mod, id + mod->graph->frame, filename)) return 1;
  raw_t *r = mod->data;
  char err[512];
  if(dng_decode_host(&r->img, err, sizeof(err)))
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "%s: %s", filename, err);
    return 1;
  }
  uint16_t *buf = (uint16_t *)mapped;
  const uint64_t wd = r->width, ht = r->height, stride = r->img.stride;
  for(uint64_t j = 0; j < ht; j++)
    memcpy(buf + j * wd, r->img.data + (j + r->oy) * stride + r->ox, sizeof(uint16_t) * wd);
  return 0;
}

// The raw straight into the output buffer on the device: lossless-JPEG raws
// are uploaded compressed and decoded on the GPU, cropped to the CFA-aligned
// window (ox, oy, width x height) by the decode itself.  Returns 1 for raws
// that take read_source's path (uncompressed: a plain copy and upload), 0
// when done, -1 on error.  No reference counterpart: the reference decodes
// every raw on the host inside read_source (main.c:240-259).
int read_source_hip(dt_module_t *mod, void *dev, void *stream)
{
  const int id = dt_module_param_int(mod, 3)[0];
  const char *fname = dt_module_param_string(mod, 0);
  char filename[PATH_MAX];
  if(dt_graph_get_resource_filename(mod, fname, id + mod->graph->frame, filename, sizeof(filename))) return -1;
  if(load_raw(mod, id + mod->graph->frame, filename)) return -1;
  const raw_t *r = mod->data;
  if(!r->img.packed || r->img.data) return 1;
  const vkdt_lj92_image_t win = { dev, (uint32_t)r->width, (uint32_t)r->height, (uint32_t)r->width,
                                  (uint32_t)r->ox, (uint32_t)r->oy };
  if(vkdt_hip_lj92_decode(r->img.packed, r->img.packed_len, r->img.tiles, (int)r->img.ntiles, &win, 1, stream))
  {
    snprintf(mod->graph->err, sizeof(mod->graph->err), "%s: %s", filename, vkdt_hip_last_error());
    return -1;
  }
  return 0;
}
