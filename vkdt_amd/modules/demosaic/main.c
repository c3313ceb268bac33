/*
 * demosaic: gaussian-splat demosaicing of Bayer and X-Trans mosaics.
 *
 * ROI callbacks and node graph as in src/pipe/modules/demosaic/main.c:
 * method 0 (default) builds down -> gauss -> splat -> fix (:158-202);
 * half-size (method 2, or an output scaled down by 1.5 blocks or more)
 * builds halfsize [-> shared/resample] (:92-110); a scaled output of the
 * full-resolution paths ends in shared/resample.  RCD (method 1, Bayer)
 * builds rcd_conv -> rcd_fill with the reference's 64x32 fill tiles
 * (k_rcd.hip).
 */
#include "pipe/api.h"

#define RCD_LOCAL 32
#define RCD_TILE_X 64
#define RCD_TILE_Y 32
#define RCD_BORDER 3

void modify_roi_in(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  dt_roi_t *ri = &module->connector[0].roi;
  ri->wd = ri->full_wd;
  ri->ht = ri->full_ht;
  ri->marker = s_roi_mark_hard_bck;
}

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  dt_roi_t *ri = &module->connector[0].roi;
  dt_roi_t *ro = &module->connector[1].roi;
  const int method = dt_module_param_int(module, 1)[0];
  const int block = module->img_param.filters == 9u ? 3 : 2;
  const float scale = ro->full_wd > 0 ? (float)ri->full_wd / (float)ro->full_wd : 1.0f;
  const int halfsize = (method == 2) || (scale >= 1.5 * block);
  ro->marker = ri->marker;
  if(halfsize)
  {
    ro->full_wd = (ri->full_wd + 1) / 2;
    ro->full_ht = (ri->full_ht + 1) / 2;
    // a request for block size or less: the consumer's roi wins the
    // negotiation, and the half-size image is resampled to it
    if(scale >= block) ro->marker = s_roi_mark_soft_fwd;
  }
  else
  {
    ro->full_wd = ri->full_wd;
    ro->full_ht = ri->full_ht;
  }
  module->img_param.filters = 0u;
}

dt_graph_run_t check_params(dt_module_t *module, uint32_t parid, uint32_t num, void *oldval)
{
  (void)num;
  if(parid == 1 && *(int *)oldval != dt_module_param_int(module, 1)[0]) return s_graph_run_all;
  return s_graph_run_record_cmd_buf;
}

// a shared/resample node to the module's output extent, wired to the
// module output (demosaic/main.c:98-107,144-152)
static int add_resample(dt_graph_t *graph, dt_module_t *module)
{
  const int id = dt_node_add(graph, module, "shared", "resample", module->connector[1].roi.wd,
                             module->connector[1].roi.ht, 1, 0, 0, 2,
                             "input", "read", "rgba", "f16", dt_no_roi,
                             "output", "write", "rgba", "f16", &module->connector[1].roi);
  dt_connector_copy(graph, module, 1, id, 1);
  return id;
}

// the full-resolution paths end in a resample when the output extent differs
// from the input's (a scaled export).  The reference keys this on a hard roi
// marker on the output (demosaic/main.c:144,195), which its roi negotiation
// sets on this chain; at equal extents that resample is the identity
// (catmull-rom weights 0 1 0 0) and is left out here.
static int output_scaled(const dt_module_t *module)
{
  return module->connector[0].roi.wd != module->connector[1].roi.wd ||
         module->connector[0].roi.ht != module->connector[1].roi.ht;
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  if(!ip->filters)
  { // nothing to demosaic
    if(module->connector[0].roi.wd == module->connector[1].roi.wd && module->connector[0].roi.ht == module->connector[1].roi.ht)
      dt_connector_bypass(graph, module, 0, 1);
    return;
  }
  const int block = ip->filters == 9u ? 3 : 2;
  module->img_param.filters = 0u;
  const int wd = module->connector[0].roi.wd, ht = module->connector[0].roi.ht;
  dt_roi_t rfull = module->connector[0].roi, rhalf = module->connector[0].roi;
  rhalf.full_wd /= block;
  rhalf.full_ht /= block;
  rhalf.wd /= block;
  rhalf.ht /= block;
  int32_t pc[5];
  memcpy(pc, ip->whitebalance, 4 * sizeof(float));
  pc[4] = (int32_t)ip->filters;
  const int method = dt_module_param_int(module, 1)[0];
  const float scale = (float)module->connector[0].roi.wd / (float)module->connector[1].roi.wd;
  const int halfsize = (scale >= 1.5 * block) || (method == 2);
  if(halfsize)
  { // block means, then resampled to the output unless that is the block grid
    const int id = dt_node_add(graph, module, "demosaic", "halfsize", rhalf.wd, rhalf.ht, 1, sizeof(pc), pc, 2,
                               "input", "read", "rggb", "*", dt_no_roi,
                               "output", "write", "rgba", "f16", &rhalf);
    dt_connector_copy(graph, module, 0, id, 0);
    if(block != scale)
    {
      const int id_res = add_resample(graph, module);
      CONN(dt_node_connect(graph, id, 1, id_res, 0));
    }
    else dt_connector_copy(graph, module, 1, id, 1);
    return;
  }
  if(method == 1)
  {
    dt_roi_t hr = module->connector[0].roi;
    hr.wd /= 2;
    const int id_conv = dt_node_add(graph, module, "demosaic", "rcd_conv", wd, ht, 1, 0, 0, 4,
                                    "cfa", "read", "*", "*", dt_no_roi,
                                    "vh", "write", "r", "f16", &module->connector[0].roi,
                                    "pq", "write", "r", "f16", &hr,
                                    "lp", "write", "r", "f16", &hr);
    const uint32_t tx = RCD_TILE_X - 2 * RCD_BORDER, ty = RCD_TILE_Y - 2 * RCD_BORDER;
    const int id_fill = dt_node_add(graph, module, "demosaic", "rcd_fill", ((wd + tx - 1) / tx) * DT_LOCAL_SIZE_X,
                                    ((ht + ty - 1) / ty) * DT_LOCAL_SIZE_Y, 1, sizeof(pc), pc, 5,
                                    "cfa", "read", "*", "*", dt_no_roi,
                                    "vh", "read", "*", "*", dt_no_roi,
                                    "pq", "read", "*", "*", dt_no_roi,
                                    "lp", "read", "*", "*", dt_no_roi,
                                    "output", "write", "rgba", "f16", &module->connector[0].roi);
    CONN(dt_node_connect_named(graph, id_conv, "vh", id_fill, "vh"));
    CONN(dt_node_connect_named(graph, id_conv, "pq", id_fill, "pq"));
    CONN(dt_node_connect_named(graph, id_conv, "lp", id_fill, "lp"));
    dt_connector_copy(graph, module, 0, id_conv, 0);
    dt_connector_copy(graph, module, 0, id_fill, 0);
    if(output_scaled(module))
    {
      const int id_res = add_resample(graph, module);
      CONN(dt_node_connect(graph, id_fill, 4, id_res, 0));
    }
    else dt_connector_copy(graph, module, 1, id_fill, 4);
    return;
  }
  // gaussian splats
  const int id_down = dt_node_add(graph, module, "demosaic", "down", wd / block, ht / block, 1, sizeof(pc), pc, 2,
                                  "input", "read", "rggb", "*", dt_no_roi,
                                  "output", "write", "y", "f16", &rhalf);
  const int id_gauss = dt_node_add(graph, module, "demosaic", "gauss", wd / block, ht / block, 1, sizeof(pc), pc, 3,
                                   "input", "read", "y", "f16", dt_no_roi,
                                   "orig", "read", "rggb", "*", dt_no_roi,
                                   "output", "write", "rgba", "f16", &rhalf);
  CONN(dt_node_connect(graph, id_down, 1, id_gauss, 0));
  dt_connector_copy(graph, module, 0, id_gauss, 1);
  const int id_splat = dt_node_add(graph, module, "demosaic", "splat", wd, ht, 1, sizeof(pc), pc, 3,
                                   "input", "read", "rggb", "*", dt_no_roi,
                                   "gauss", "read", "rgba", "f16", dt_no_roi,
                                   "output", "write", "g", "f16", &rfull);
  dt_connector_copy(graph, module, 0, id_splat, 0);
  CONN(dt_node_connect(graph, id_gauss, 2, id_splat, 1));
  dt_connector_copy(graph, module, 0, id_down, 0);
  const int id_fix = dt_node_add(graph, module, "demosaic", "fix", wd, ht, 1, sizeof(pc), pc, 4,
                                 "input", "read", "rggb", "*", dt_no_roi,
                                 "green", "read", "g", "*", dt_no_roi,
                                 "cov", "read", "rgba", "f16", dt_no_roi,
                                 "output", "write", "rgba", "f16", &rfull);
  dt_connector_copy(graph, module, 0, id_fix, 0);
  CONN(dt_node_connect(graph, id_splat, 2, id_fix, 1));
  CONN(dt_node_connect(graph, id_gauss, 2, id_fix, 2));
  if(output_scaled(module))
  {
    const int id_res = add_resample(graph, module);
    CONN(dt_node_connect(graph, id_fix, 3, id_res, 0));
  }
  else dt_connector_copy(graph, module, 1, id_fix, 3);
}
