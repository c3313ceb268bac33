/*
 * hilite: highlight inpainting on the mosaic (wavelet pyramid of the
 * half-size rgb image, clipping-aware blur, colour-ratio rescaling).
 *
 * Node graph as in src/pipe/modules/hilite/main.c:6-86: half (mosaic ->
 * rgba on the block grid), then per coarseness level a reduce and an
 * assemble node (at most 14 levels, until an extent reaches 1), and doub
 * back onto the mosaic.  Non-mosaic inputs are passed through.  Default ROI
 * callbacks: the output has the input's extent.
 */
#include "pipe/api.h"

#include <string.h>

#define HILITE_MAX_LEVELS 15

static dt_roi_t half_extent(dt_roi_t r)
{
  r.wd = (r.wd - 1) / 2 + 1;
  r.ht = (r.ht - 1) / 2 + 1;
  r.full_wd = (r.full_wd - 1) / 2 + 1;
  r.full_ht = (r.full_ht - 1) / 2 + 1;
  return r;
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  if(!ip->filters)
  {
    dt_connector_bypass(graph, module, 0, 1);
    return;
  }
  const int block = ip->filters == 9u ? 3 : 2;
  const int wd = module->connector[0].roi.wd, ht = module->connector[0].roi.ht;
  // push constants: the white balance coefficients as raw float bits + filters
  int32_t pc[5];
  memcpy(pc, module->img_param.whitebalance, 4 * sizeof(float));
  pc[4] = (int32_t)ip->filters;

  dt_roi_t rgrid = module->connector[0].roi;
  rgrid.wd /= block;
  rgrid.ht /= block;
  const int id_half = dt_node_add(graph, module, "hilite", "half", wd / block, ht / block, 1, sizeof(pc), pc, 2,
                                  "input", "read", "rggb", "ui16", dt_no_roi,
                                  "output", "write", "rgba", "f16", &rgrid);
  const int id_doub = dt_node_add(graph, module, "hilite", "doub", wd / block, ht / block, 1, sizeof(pc), pc, 3,
                                  "input", "read", "rggb", "ui16", dt_no_roi,
                                  "coarse", "read", "rgba", "f16", dt_no_roi,
                                  "output", "write", "rggb", "ui16", &module->connector[0].roi);
  // the module's connectors carry the negotiated formats onto these nodes
  dt_connector_copy(graph, module, 0, id_half, 0);
  dt_connector_copy(graph, module, 0, id_doub, 0);
  dt_connector_copy(graph, module, 1, id_doub, 2);

  // fine level feeds reduce and assemble; each assemble feeds the one above
  int fine_node = id_half, fine_conn = 1;
  int up_node = id_doub, up_conn = 1;
  dt_roi_t rfine = rgrid, rcoarse = half_extent(rgrid);
  for(int l = 1; l < HILITE_MAX_LEVELS; l++)
  {
    const int id_reduce = dt_node_add(graph, module, "hilite", "reduce", rcoarse.wd, rcoarse.ht, 1, sizeof(pc), pc, 2,
                                      "input", "read", "rgba", "f16", dt_no_roi,
                                      "output", "write", "rgba", "f16", &rcoarse);
    const int id_assemble = dt_node_add(graph, module, "hilite", "assemble", rfine.wd, rfine.ht, 1, sizeof(pc), pc, 3,
                                        "fine", "read", "rgba", "f16", dt_no_roi,
                                        "coarse", "read", "rgba", "f16", dt_no_roi,
                                        "output", "write", "rgba", "f16", &rfine);
    CONN(dt_node_connect(graph, fine_node, fine_conn, id_reduce, 0));
    CONN(dt_node_connect(graph, fine_node, fine_conn, id_assemble, 0));
    CONN(dt_node_connect(graph, id_assemble, 2, up_node, up_conn));
    fine_node = id_reduce;
    fine_conn = 1;
    up_node = id_assemble;
    up_conn = 1;
    rfine = rcoarse;
    rcoarse = half_extent(rcoarse);
    if(rcoarse.wd <= 1 || rcoarse.ht <= 1 || l + 1 == HILITE_MAX_LEVELS)
    { // coarsest level: the last reduce is the last assemble's coarse input
      CONN(dt_node_connect(graph, id_reduce, 1, id_assemble, 1));
      break;
    }
  }
}
