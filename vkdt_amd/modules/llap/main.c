/*
 * llap: local laplacian pyramid local contrast on luminance.
 *
 * Node graph as src/pipe/modules/llap/main.c:14-114: curve (num_gamma
 * remapped planes + luminance), reduce on every coarser level, assemble
 * from coarse to fine, then colour.  The reference fixes the number of
 * curves (config.h NUM_GAMMA 10) and the pyramid depth (12); here they are
 * the params `numgamma` and `levels`, whose defaults are those constants.
 */
#include "pipe/api.h"

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const int wd = module->connector[0].roi.wd, ht = module->connector[0].roi.ht;
  int num_gamma = dt_module_param_int(module, dt_module_get_param(module->so, dt_token("numgamma")))[0];
  int max_nl = dt_module_param_int(module, dt_module_get_param(module->so, dt_token("levels")))[0];
  num_gamma = CLAMP(num_gamma, 2, 32);
  max_nl = CLAMP(max_nl, 2, 12);

  int pcc[] = { num_gamma };
  const int id_curve = dt_node_add(graph, module, "llap", "curve", wd, ht, 1, sizeof(pcc), pcc, 2,
                                   "input", "read", "rgba", "f16", dt_no_roi,
                                   "output", "write", "y", "f16", &module->connector[0].roi);
  graph->node[id_curve].connector[1].array_length = num_gamma + 1;

  dt_roi_t rf = module->connector[0].roi, rc = module->connector[0].roi;
  rc.wd = (rc.wd - 1) / 2 + 1;
  rc.ht = (rc.ht - 1) / 2 + 1;
  rc.full_wd = (rc.full_wd - 1) / 2 + 1;
  rc.full_ht = (rc.full_ht - 1) / 2 + 1;
  int nl = max_nl;
  int id_reduce[12] = { -1 }, id_assemble[12] = { -1 };
  id_reduce[0] = id_curve;
  for(int l = 1; l < nl; l++)
  {
    id_reduce[l] = dt_node_add(graph, module, "llap", "reduce", rc.wd, rc.ht, num_gamma + 1, 0, 0, 2,
                               "inhi", "read", "y", "f16", dt_no_roi,
                               "outlo", "write", "y", "f16", &rc);
    graph->node[id_reduce[l]].connector[1].array_length = num_gamma + 1;
    CONN(dt_node_connect(graph, id_reduce[l - 1], 1, id_reduce[l], 0));
    int pca[] = { num_gamma, 0 };
    id_assemble[l] = dt_node_add(graph, module, "llap", "assemble", rf.wd, rf.ht, 1, sizeof(pca), pca, 4,
                                 "coarse", "read", "y", "f16", dt_no_roi,
                                 "currlo", "read", "y", "f16", dt_no_roi,
                                 "currhi", "read", "y", "f16", dt_no_roi,
                                 "fine", "write", "y", "f16", &rf);
    CONN(dt_node_connect(graph, id_reduce[l - 1], 1, id_assemble[l], 1));
    CONN(dt_node_connect(graph, id_reduce[l], 1, id_assemble[l], 2));
    if(l > 1) CONN(dt_node_connect(graph, id_assemble[l], 3, id_assemble[l - 1], 0));
    rf = rc;
    rc.wd = (rc.wd - 1) / 2 + 1;
    rc.ht = (rc.ht - 1) / 2 + 1;
    rc.full_wd = (rc.full_wd - 1) / 2 + 1;
    rc.full_ht = (rc.full_ht - 1) / 2 + 1;
    if(rc.wd <= 1 || rc.ht <= 1) { nl = l + 1; break; }
  }
  // the coarsest assemble expands the luminance plane of its own level
  graph->node[id_assemble[nl - 1]].push_constant[1] = 1;
  CONN(dt_node_connect(graph, id_curve, 1, id_assemble[nl - 1], 0));
  const int id_col = dt_node_add(graph, module, "llap", "colour", wd, ht, 1, 0, 0, 3,
                                 "lum", "read", "y", "f16", dt_no_roi,
                                 "input", "read", "rgba", "f16", dt_no_roi,
                                 "output", "write", "rgba", "f16", &module->connector[0].roi);
  CONN(dt_node_connect(graph, id_assemble[1], 3, id_col, 0));
  dt_connector_copy(graph, module, 0, id_curve, 0);
  dt_connector_copy(graph, module, 0, id_col, 1);
  dt_connector_copy(graph, module, 1, id_col, 2);
}
