/*
 * autoexp -- automatic exposure by a moving average of the mean log
 * luminance (reference autoexp/main.c): hist counts a radially thinned
 * pixel set into 256 log2 luminance bins, lum turns the mean bin into a
 * luminance smoothed over frames (kept in a protected buffer), exp scales
 * the image to put that luminance at 0.18.
 */
#include "pipe/api.h"

#include <string.h>

int init(dt_module_t *mod)
{ // the shaders' uniforms: params (vup, vdown, timediff, reset) and the frame
  mod->committed_param_size = 3 * sizeof(float) + 2 * sizeof(uint32_t);
  return 0;
}

void commit_params(dt_graph_t *graph, dt_module_t *module)
{
  memcpy(module->committed_param, module->param, 3 * sizeof(float) + sizeof(int32_t));
  const uint32_t frame = graph->frame < 0 ? 0u : (uint32_t)graph->frame;
  memcpy(module->committed_param + 3 * sizeof(float) + sizeof(int32_t), &frame, sizeof(frame));
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{ // main.c:3-34
  const int wd = module->connector[0].roi.wd;
  const int ht = module->connector[0].roi.ht;
  dt_roi_t hroi = (dt_roi_t){ .wd = 16 + 1, .ht = 16 };
  dt_roi_t lroi = (dt_roi_t){ .wd = 8, .ht = 8 };
  const int id_hist = dt_node_add(graph, module, "autoexp", "hist", (wd + 15) / 16 * 8, (ht + 15) / 16 * 8, 1, 0, 0, 2,
      "input", "read",  "*",    "*",   dt_no_roi,
      "hist",  "write", "ssbo", "u32", &hroi);
  const int id_lum = dt_node_add(graph, module, "autoexp", "lum", 1, 1, 1, 0, 0, 3,
      "input", "read",  "*",    "*",   dt_no_roi,
      "hist",  "read",  "ssbo", "u32", dt_no_roi,
      "lum",   "write", "ssbo", "f32", &lroi);
  const int id_exp = dt_node_add(graph, module, "autoexp", "exp", wd, ht, 1, 0, 0, 4,
      "input",  "read",  "*",    "*",   dt_no_roi,
      "hist",   "read",  "ssbo", "u32", dt_no_roi,
      "lum",    "read",  "ssbo", "f32", dt_no_roi,
      "output", "write", "rgba", "*",   &module->connector[0].roi);
  dt_connector_copy(graph, module, 0, id_hist, 0);
  dt_connector_copy(graph, module, 0, id_lum,  0);
  dt_connector_copy(graph, module, 0, id_exp,  0);
  dt_connector_copy(graph, module, 1, id_exp,  3);
  dt_node_connect(graph, id_hist, 1, id_lum, 1);
  dt_node_connect(graph, id_hist, 1, id_exp, 1);
  dt_node_connect(graph, id_lum,  2, id_exp, 2);
  graph->node[id_hist].connector[1].flags |= s_conn_clear;
  graph->node[id_lum].connector[2].flags |= s_conn_protected;
}
