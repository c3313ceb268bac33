/*
 * rawhist: noise statistics of a raw mosaic for noise profiling.
 *
 * As src/pipe/modules/rawhist/main.c: `collect` bins every CFA block's green
 * sites by brightness (count, sum of means, sum of second moments) into a
 * 1000 x 5 accumulator connector `data`, cleared before it runs; `map` plots
 * it (1000 x 600).  The accumulators are FP64 on this path (the "atom"
 * format), which nprof reads.
 */
#include "pipe/api.h"

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  // fixed plot and accumulator extents (rawhist/main.c:3-15)
  module->connector[1].roi.marker = s_roi_mark_soft_fwd;
  module->connector[2].roi.marker = s_roi_mark_hard_fwd;
  module->connector[1].roi.full_wd = 1000;
  module->connector[1].roi.full_ht = 600;
  module->connector[2].roi.full_wd = 1000;
  module->connector[2].roi.full_ht = 5;
}

void commit_params(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  float *f = (float *)module->committed_param;
  uint32_t *u = (uint32_t *)(f + 2);
  f[0] = module->img_param.noise_a;
  f[1] = module->img_param.noise_b;
  u[0] = module->img_param.filters;
  const float black = module->img_param.black[1];
  u[1] = black > 1.0f ? (uint32_t)black : 1u;
  u[2] = (uint32_t)module->img_param.white[1];
}

int init(dt_module_t *mod)
{
  mod->committed_param_size = 2 * sizeof(float) + 3 * sizeof(uint32_t);
  return 0;
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  const int block = ip->filters == 9u ? 3 : 2;
  dt_roi_t racc = module->connector[2].roi;
  const int id_collect = dt_node_add(graph, module, "rawhist", "collect",
                                     module->connector[0].roi.wd / block, module->connector[0].roi.ht / block, 1,
                                     0, 0, 2,
                                     "input", "read", "rggb", "ui16", &module->connector[0].roi,
                                     "output", "write", "r", "atom", &racc);
  const int id_map = dt_node_add(graph, module, "rawhist", "map",
                                 module->connector[1].roi.wd, module->connector[1].roi.ht, 1, 0, 0, 2,
                                 "input", "read", "r", "atom", &racc,
                                 "output", "write", "rgba", "f16", &module->connector[1].roi);
  dt_connector_copy(graph, module, 0, id_collect, 0);
  CONN(dt_node_connect(graph, id_collect, 1, id_map, 0));
  dt_connector_copy(graph, module, 1, id_map, 1);
  dt_connector_copy(graph, module, 2, id_collect, 1);
  graph->node[id_collect].connector[1].flags = s_conn_clear;
}
