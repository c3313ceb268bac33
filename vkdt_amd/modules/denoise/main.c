/*
 * denoise: noise-profile driven decimated wavelet denoising on the CFA.
 *
 * Node graph and push constants as in src/pipe/modules/denoise/main.c:
 * strength <= 0 -> one `noop` node (crop + black/white scaling, :202-225),
 * whose output is the one-channel mosaic the demosaic and hilite kernels
 * read (the reference declares it rgba and its readers sample .r);
 * otherwise half -> downcov -> down x3 -> assemble on the block grid and
 * doub back at full resolution (:227-326).  A Bayer DNG whose OpcodeList2
 * starts with four GainMap opcodes, one per CFA site (:12-43), adds a
 * `gainmap` source node (the four maps interleaved as rgba f32, uploaded by
 * read_source, :62-78) that noop and doub sample; otherwise their gain-map
 * binding is a dummy and the push constants say so.
 */
#include "pipe/api.h"
#include "pipe/dng_opcode.h"

#include <stdlib.h>

typedef struct dn_data_t
{
  const dt_dng_gain_map_t *gm[4];   // by CFA site of the emitted buffer: (0,0) (1,0) (0,1) (1,1)
}
dn_data_t;

int init(dt_module_t *mod)
{
  mod->data = calloc(1, sizeof(dn_data_t));
  return mod->data ? 0 : 1;
}

void cleanup(dt_module_t *mod)
{
  free(mod->data);
  mod->data = 0;
}

// The four GainMap opcodes of a Bayer raw, sorted by the CFA site of the
// emitted (offset-normalised) buffer they apply to.  As the reference
// checks it (denoise/main.c:12-43): the list's first four opcodes must all
// be gain maps over one plane with row and column pitch 2 and one map
// plane, at least one axis with two or more points, one per site, on the
// same grid (points, spacing, origin) over the same 2x2 block region.
// Returns 1 and fills d->gm, else 0.
static int bayer_gain_maps(const dt_dng_opcode_list_t *ops, uint32_t ox, uint32_t oy, dn_data_t *d)
{
  for(int k = 0; k < 4; k++) d->gm[k] = 0;
  if(!ops || ops->count < 4) return 0;
  for(int i = 0; i < 4; i++) if(ops->ops[i].id != s_dngop_gain_map || !ops->ops[i].data) return 0;
  for(int i = 0; i < 4; i++)
  {
    const dt_dng_gain_map_t *g = ops->ops[i].data;
    const dt_dng_region_t *r = &g->region;
    if(r->plane != 0 || r->planes != 1 || g->map_planes != 1 || r->row_pitch != 2 || r->col_pitch != 2) return 0;
    if(g->map_points_h < 2 && g->map_points_v < 2) return 0;
    d->gm[(((r->top + oy) & 1) << 1) + ((r->left + ox) & 1)] = g;
  }
  for(int k = 0; k < 4; k++) if(!d->gm[k]) return 0;
  const dt_dng_gain_map_t *a = d->gm[0];
  for(int k = 1; k < 4; k++)
  {
    const dt_dng_gain_map_t *b = d->gm[k];
    if(a->map_points_h != b->map_points_h || a->map_points_v != b->map_points_v ||
       a->map_spacing_h != b->map_spacing_h || a->map_spacing_v != b->map_spacing_v ||
       a->map_origin_h != b->map_origin_h || a->map_origin_v != b->map_origin_v ||
       a->region.top / 2 != b->region.top / 2 || a->region.left / 2 != b->region.left / 2 ||
       a->region.bottom / 2 != b->region.bottom / 2 || a->region.right / 2 != b->region.right / 2)
      return 0;
  }
  return 1;
}

// the gain-map source: gains of the four sites interleaved as rgba
int read_source(dt_module_t *mod, void *mapped, dt_read_source_params_t *p)
{
  const dn_data_t *d = mod->data;
  if(p->node->kernel != dt_token("gainmap") || !d || !d->gm[0]) return 0;
  const uint32_t wd = d->gm[0]->map_points_h, ht = d->gm[0]->map_points_v;
  float *out = mapped;
  for(uint32_t j = 0; j < ht; j++)
    for(uint32_t i = 0; i < wd; i++)
      for(int c = 0; c < 4; c++) out[4 * ((size_t)j * wd + i) + c] = d->gm[c]->map_gain[(size_t)j * wd + i];
  return 0;
}

void modify_roi_in(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  if(ip->filters)
  { // the full uncropped mosaic: borders feed the noise statistics
    module->connector[0].roi.wd = module->connector[0].roi.full_wd;
    module->connector[0].roi.ht = module->connector[0].roi.full_ht;
    module->connector[0].roi.marker = s_roi_mark_hard_bck;
  }
  else module->connector[0].roi = module->connector[1].roi;
}

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  const uint32_t *b = ip->crop_aabb;
  module->connector[1].roi = module->connector[0].roi;
  if(ip->filters)
  {
    module->connector[1].roi.full_wd = b[2] - b[0];
    module->connector[1].roi.full_ht = b[3] - b[1];
  }
}

dt_graph_run_t check_params(dt_module_t *module, uint32_t parid, uint32_t num, void *oldval)
{
  (void)num;
  if(parid == 0)
  { // strength crossing zero changes the topology
    const float o = *(float *)oldval, n = dt_module_param_float(module, 0)[0];
    if((o <= 0.0f && n > 0.0f) || (o > 0.0f && n <= 0.0f)) return s_graph_run_all;
  }
  return s_graph_run_record_cmd_buf;
}

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  for(int k = 0; k < 4; k++)
  { // black/white/crop are consumed here
    module->img_param.black[k] = 0.0f;
    module->img_param.white[k] = 1.0f;
  }
  module->img_param.crop_aabb[0] = 0;
  module->img_param.crop_aabb[1] = 0;
  module->img_param.crop_aabb[2] = module->connector[1].roi.full_wd;
  module->img_param.crop_aabb[3] = module->connector[1].roi.full_ht;

  const float nowb[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
  const float *wb = ip->filters ? ip->whitebalance : nowb;
  float black[4], white[4];
  const float cs = module->connector[0].roi.wd / (float)module->connector[0].roi.full_wd;
  const uint32_t crop[4] = {
    (uint32_t)(ip->crop_aabb[0] * cs), (uint32_t)(ip->crop_aabb[1] * cs),
    (uint32_t)(ip->crop_aabb[2] * cs), (uint32_t)(ip->crop_aabb[3] * cs) };
  for(int k = 0; k < 4; k++) black[k] = ip->black[k] / 65535.0f;
  for(int k = 0; k < 4; k++) white[k] = ip->white[k] / 65535.0f;
  const uint32_t na = fbits(ip->noise_a), nb = fbits(ip->noise_b);

  // DNG gain maps (denoise/main.c:172-200): map origin and inverse extent
  dn_data_t *dat = module->data;
  const dt_image_metadata_dngop_t *dngop = dt_metadata_find(ip->meta, s_image_metadata_dngop);
  uint32_t gainmap = dngop && dat ? (uint32_t)bayer_gain_maps(dngop->op_list[1], dngop->ox, dngop->oy, dat) : 0;
  uint32_t gm_ox = 0, gm_oy = 0, gm_sx = 0, gm_sy = 0;
  int id_gm = -1;
  if(gainmap)
  {
    const dt_dng_gain_map_t *g = dat->gm[0];
    dt_roi_t groi = { .wd = g->map_points_h, .ht = g->map_points_v };
    id_gm = dt_node_add(graph, module, "denoise", "gainmap", g->map_points_h, g->map_points_v, 1, 0, 0, 1,
                        "source", "source", "rgba", "f32", &groi);
    gm_ox = fbits((float)g->map_origin_h);
    gm_oy = fbits((float)g->map_origin_v);
    gm_sx = fbits((float)(1.0 / (g->map_spacing_h * (g->map_points_h - 1))));
    gm_sy = fbits((float)(1.0 / (g->map_spacing_v * (g->map_points_v - 1))));
  }
  else if(dat) for(int k = 0; k < 4; k++) dat->gm[k] = 0;

  const float strength = dt_module_param_float(module, dt_module_get_param(module->so, dt_token("strength")))[0];
  if(strength <= 0.0f)
  {
    if(ip->filters == 0) { dt_connector_bypass(graph, module, 0, 1); return; }
    const int32_t pc[] = {
      (int32_t)crop[0], (int32_t)crop[1], (int32_t)crop[2], (int32_t)crop[3],
      (int32_t)fbits(black[0]), (int32_t)fbits(black[1]), (int32_t)fbits(black[2]), (int32_t)fbits(black[3]),
      (int32_t)fbits(white[0]), (int32_t)fbits(white[1]), (int32_t)fbits(white[2]), (int32_t)fbits(white[3]),
      (int32_t)gm_ox, (int32_t)gm_oy, (int32_t)gm_sx, (int32_t)gm_sy, (int32_t)ip->filters, (int32_t)gainmap };
    const int id = dt_node_add(graph, module, "denoise", "noop", module->connector[1].roi.wd,
                               module->connector[1].roi.ht, 1, sizeof(pc), pc, 3,
                               "input", "read", "rgba", "f16", dt_no_roi,
                               "output", "write", "rggb", "f16", &module->connector[1].roi,
                               "gainmap", "read", "rgba", "*", dt_no_roi);
    dt_connector_copy(graph, module, 0, id, 0);
    if(gainmap) CONN(dt_node_connect(graph, id_gm, 0, id, 2));
    else dt_connector_copy(graph, module, 0, id, 2);
    dt_connector_copy(graph, module, 1, id, 1);
    return;
  }

  const int block = ip->filters == 0 ? 1 : (ip->filters == 9u ? 3 : 2);
  dt_roi_t rh = module->connector[1].roi;
  rh.full_wd /= block;
  rh.full_ht /= block;
  rh.wd /= block;
  rh.ht /= block;
  const int wd = rh.wd, ht = rh.ht;

  // the shared head of every push constant block: wb, black, white
  int32_t head[12];
  for(int k = 0; k < 4; k++)
  {
    head[k] = (int32_t)fbits(wb[k]);
    head[4 + k] = (int32_t)fbits(black[k]);
    head[8 + k] = (int32_t)fbits(white[k]);
  }

  int id_down[4];
  for(int i = 0; i < 4; i++)
  {
    int32_t pc[20];
    memcpy(pc, head, sizeof(head));
    for(int k = 0; k < 4; k++) pc[12 + k] = (i == 0 && block == 1) ? (int32_t)crop[k] : 0;
    pc[16] = (int32_t)na;
    pc[17] = (int32_t)nb;
    pc[18] = i;
    pc[19] = block;
    const int cov = ip->filters && i == 0;
    id_down[i] = dt_node_add(graph, module, "denoise", cov ? "downcov" : "down", wd, ht, 1, sizeof(pc), pc, cov ? 3 : 2,
                             "input", "read", "rgba", "f16", dt_no_roi,
                             "output", "write", "rgba", "f16", &rh,
                             "cov", "write", "rgba", "f16", &rh);
  }
  for(int i = 1; i < 4; i++) CONN(dt_node_connect(graph, id_down[i - 1], 1, id_down[i], 0));

  int32_t pca[19];
  memcpy(pca, head, sizeof(head));
  for(int k = 0; k < 4; k++) pca[12 + k] = block == 1 ? (int32_t)crop[k] : 0;
  pca[16] = (int32_t)na;
  pca[17] = (int32_t)nb;
  pca[18] = (int32_t)ip->filters;
  const int id_asm = dt_node_add(graph, module, "denoise", "assemble", wd, ht, 1, sizeof(pca), pca, 6,
                                 "s0", "read", "rgba", "f16", dt_no_roi,
                                 "s1", "read", "rgba", "f16", dt_no_roi,
                                 "s2", "read", "rgba", "f16", dt_no_roi,
                                 "s3", "read", "rgba", "f16", dt_no_roi,
                                 "s4", "read", "rgba", "f16", dt_no_roi,
                                 "output", "write", "rgba", "f16", &rh);
  for(int i = 0; i < 4; i++) CONN(dt_node_connect(graph, id_down[i], 1, id_asm, i + 1));

  if(ip->filters)
  { // mosaic: wrap the wavelet into the block-averaging half and the full-res doub
    int32_t pch[17];
    memcpy(pch, head, sizeof(head));
    for(int k = 0; k < 4; k++) pch[12 + k] = (int32_t)crop[k];
    pch[16] = (int32_t)ip->filters;
    const int id_half = dt_node_add(graph, module, "denoise", "half", rh.full_wd, rh.full_ht, 1, sizeof(pch), pch, 2,
                                    "input", "read", "rggb", "ui16", dt_no_roi,
                                    "output", "write", "rgba", "f16", &rh);
    int32_t pc[24];
    memcpy(pc, head, sizeof(head));
    for(int k = 0; k < 4; k++) pc[12 + k] = (int32_t)crop[k];
    pc[16] = (int32_t)ip->filters;
    pc[17] = (int32_t)na;
    pc[18] = (int32_t)nb;
    pc[19] = (int32_t)gainmap;
    pc[20] = (int32_t)gm_ox;
    pc[21] = (int32_t)gm_oy;
    pc[22] = (int32_t)gm_sx;
    pc[23] = (int32_t)gm_sy;
    const int id_doub = dt_node_add(graph, module, "denoise", "doub", module->connector[1].roi.wd,
                                    module->connector[1].roi.ht, 1, sizeof(pc), pc, 5,
                                    "orig", "read", "rggb", "f16", dt_no_roi,
                                    "crs0", "read", "rgba", "f16", dt_no_roi,
                                    "crs1", "read", "rgba", "f16", dt_no_roi,
                                    "output", "write", "rggb", "f16", &module->connector[1].roi,
                                    "gainmap", "read", "rgba", "*", dt_no_roi);
    CONN(dt_node_connect(graph, id_asm, 5, id_doub, 1));
    CONN(dt_node_connect(graph, id_half, 1, id_doub, 2));
    if(gainmap) CONN(dt_node_connect(graph, id_gm, 0, id_doub, 4));
    else CONN(dt_node_connect(graph, id_half, 1, id_doub, 4));   // dummy gain map binding
    dt_connector_copy(graph, module, 0, id_doub, 0);
    dt_connector_copy(graph, module, 1, id_doub, 3);
    CONN(dt_node_connect(graph, id_half, 1, id_down[0], 0));
    CONN(dt_node_connect(graph, id_half, 1, id_asm, 0));
    dt_connector_copy(graph, module, 0, id_half, 0);
  }
  else
  {
    dt_connector_copy(graph, module, 0, id_down[0], 0);
    dt_connector_copy(graph, module, 0, id_asm, 0);
    dt_connector_copy(graph, module, 1, id_asm, 5);
  }
}
