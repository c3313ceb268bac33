/*
 * colenc: output colour encoding, linear rec2020 -> the export's primaries
 * and transfer function (src/pipe/modules/colenc).  The default node
 * colenc/main does the work; modify_roi_out publishes the encoding.
 */
#include "pipe/api.h"

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  const int32_t prim = dt_module_param_int(module, dt_module_get_param(module->so, dt_token("prim")))[0];
  const int32_t trc = dt_module_param_int(module, dt_module_get_param(module->so, dt_token("trc")))[0];
  module->img_param.colour_primaries = prim;
  module->img_param.colour_trc = trc;
  module->connector[1].roi.full_wd = module->connector[0].roi.full_wd;
  module->connector[1].roi.full_ht = module->connector[0].roi.full_ht;
}
