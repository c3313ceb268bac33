/*
 * denoise: noise-profile driven decimated wavelet denoising on the CFA.
 *
 * Node graph and push constants as in src/pipe/modules/denoise/main.c:
 * strength <= 0 -> one `noop` node (crop + black/white scaling, :202-225);
 * otherwise half -> downcov -> down x3 -> assemble on the block grid and
 * doub back at full resolution (:227-326).  DNG gain maps are not carried
 * by this path's raw input, so the gain-map source node is never built.
 */
#include "pipe/api.h"

#include <stdlib.h>

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
  const uint32_t gainmap = 0;

  const float strength = dt_module_param_float(module, dt_module_get_param(module->so, dt_token("strength")))[0];
  if(strength <= 0.0f)
  {
    if(ip->filters == 0) { dt_connector_bypass(graph, module, 0, 1); return; }
    const int32_t pc[] = {
      (int32_t)crop[0], (int32_t)crop[1], (int32_t)crop[2], (int32_t)crop[3],
      (int32_t)fbits(black[0]), (int32_t)fbits(black[1]), (int32_t)fbits(black[2]), (int32_t)fbits(black[3]),
      (int32_t)fbits(white[0]), (int32_t)fbits(white[1]), (int32_t)fbits(white[2]), (int32_t)fbits(white[3]),
      0, 0, 0, 0, (int32_t)ip->filters, (int32_t)gainmap };
    const int id = dt_node_add(graph, module, "denoise", "noop", module->connector[1].roi.wd,
                               module->connector[1].roi.ht, 1, sizeof(pc), pc, 3,
                               "input", "read", "rgba", "f16", dt_no_roi,
                               "output", "write", "rgba", "f16", &module->connector[1].roi,
                               "gainmap", "read", "rgba", "*", dt_no_roi);
    dt_connector_copy(graph, module, 0, id, 0);
    dt_connector_copy(graph, module, 0, id, 2);
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
    pc[20] = pc[21] = pc[22] = pc[23] = 0;
    const int id_doub = dt_node_add(graph, module, "denoise", "doub", module->connector[1].roi.wd,
                                    module->connector[1].roi.ht, 1, sizeof(pc), pc, 5,
                                    "orig", "read", "rggb", "f16", dt_no_roi,
                                    "crs0", "read", "rgba", "f16", dt_no_roi,
                                    "crs1", "read", "rgba", "f16", dt_no_roi,
                                    "output", "write", "rggb", "f16", &module->connector[1].roi,
                                    "gainmap", "read", "rgba", "*", dt_no_roi);
    CONN(dt_node_connect(graph, id_asm, 5, id_doub, 1));
    CONN(dt_node_connect(graph, id_half, 1, id_doub, 2));
    CONN(dt_node_connect(graph, id_half, 1, id_doub, 4));   // dummy gain map binding
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
