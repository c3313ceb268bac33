/*
 * hist -- waveform histogram (reference hist/main.c): per column of a
 * 1000x600 output the counts of r, g and b per row, then a display map.
 */
#include "pipe/api.h"

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{ // main.c:3-12: a constant histogram size
  (void)graph;
  module->connector[1].roi = module->connector[0].roi;
  module->connector[1].roi.full_wd = 1000;
  module->connector[1].roi.full_ht = 600;
}

void modify_roi_in(dt_graph_t *graph, dt_module_t *module)
{ // main.c:14-21: the whole input
  (void)graph;
  module->connector[0].roi.wd = module->connector[0].roi.full_wd;
  module->connector[0].roi.ht = module->connector[0].roi.full_ht;
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{ // main.c:23-43
  const int id_collect = dt_node_add(graph, module, "hist", "collect",
    module->connector[0].roi.wd, module->connector[0].roi.ht, 1, 0, 0, 2,
    "input",  "read",  "rgba", "f16",  dt_no_roi,
    "output", "write", "r",    "ui32", &module->connector[1].roi);
  graph->node[id_collect].connector[1].flags = s_conn_clear;
  const int id_map = dt_node_add(graph, module, "hist", "map",
    module->connector[1].roi.wd, module->connector[1].roi.ht, 1, 0, 0, 2,
    "input",  "read",  "r",    "ui32", dt_no_roi,
    "output", "write", "rgba", "f16",  &module->connector[1].roi);
  dt_connector_copy(graph, module, 0, id_collect, 0);
  dt_node_connect(graph, id_collect, 1, id_map, 0);
  dt_connector_copy(graph, module, 1, id_map, 1);
}
