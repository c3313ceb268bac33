/*
 * colour: camera rgb -> linear rec2020, white balance by CAT16, exposure.
 *
 * commit_params writes the same uniform block as
 * src/pipe/modules/colour/main.c:219-365 (242 floats, std140: mul, mat3
 * columns, rbf, temp/mode/sat/pick/gamut/primaries/trc/clip, auto wb).
 * The data-driven rbf (mode 1) coefficients need a linear solve and the
 * clut/spectral paths need their LUT inputs; neither is part of the
 * raw-development graph, so this module writes N = 0 and the HIP kernel
 * refuses blocks that ask for them.
 */
#include "pipe/api.h"

#include <math.h>

void modify_roi_in(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  module->connector[0].roi = module->connector[1].roi;
  for(int i = 2; i < module->num_connectors; i++) module->connector[i].roi.marker = s_roi_mark_uninited;
}

void modify_roi_out(dt_graph_t *graph, dt_module_t *module)
{
  module->connector[1].roi = module->connector[0].roi;
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  for(int k = 0; k < 9; k++) module->img_param.cam_to_rec2020[k] = (k % 4 == 0) ? 1.0f : 0.0f;
  module->img_param.whitebalance[0] = module->img_param.whitebalance[1] = module->img_param.whitebalance[2] = 1.0f;
  module->img_param.colour_primaries = s_colour_primaries_2020;
  module->img_param.colour_trc = s_colour_trc_linear;
}

int init(dt_module_t *mod)
{
  mod->committed_param_size = sizeof(float) * (4 + 12 + 4 + 12 + 4 * 24 + 4 * 24 + 5 + 8 + 5);
  return 0;
}

void commit_params(dt_graph_t *graph, dt_module_t *module)
{
  const dt_image_params_t *ip = dt_module_get_input_img_param(graph, module, dt_token("input"));
  if(!ip) return;
  float *f = (float *)module->committed_param;
  uint32_t *u = (uint32_t *)module->committed_param;
  memset(f, 0, module->committed_param_size);
  const dt_module_so_t *so = module->so;
  float *p_wb = (float *)dt_module_param_float(module, dt_module_get_param(so, dt_token("white")));
  const float p_tmp = dt_module_param_float(module, dt_module_get_param(so, dt_token("temp")))[0];
  const int p_mat = dt_module_param_int(module, dt_module_get_param(so, dt_token("matrix")))[0];
  const float *p_mtx = dt_module_param_float(module, dt_module_get_param(so, dt_token("mat")));
  const int p_gam = dt_module_param_int(module, dt_module_get_param(so, dt_token("gamut")))[0];
  const int p_mod = dt_module_param_int(module, dt_module_get_param(so, dt_token("mode")))[0];
  const float p_sat = dt_module_param_float(module, dt_module_get_param(so, dt_token("sat")))[0];
  const int p_pck = dt_module_param_int(module, dt_module_get_param(so, dt_token("picked")))[0];
  const int p_clp = dt_module_param_int(module, dt_module_get_param(so, dt_token("clip")))[0];
  const float p_clm = dt_module_param_float(module, dt_module_get_param(so, dt_token("clipmax")))[0];

  if(p_wb[0] == 0.0f && p_wb[1] == 0.0f && p_wb[2] == 0.0f)
  { // camera coefficients, expressed through the camera matrix
    float w0[3] = { 0 }, w[3] = { ip->whitebalance[0], ip->whitebalance[1], ip->whitebalance[2] };
    for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) w0[j] += ip->cam_to_rec2020[3 * j + i] / w[i];
    w0[0] /= w0[1];
    w0[2] /= w0[1];
    w0[1] = 1.0f;
    p_wb[0] = 1 / w0[0];
    p_wb[1] = 1;
    p_wb[2] = 1 / w0[2];
  }
  if(!(p_wb[0] == p_wb[0]) || p_wb[0] == 0.0f || p_wb[1] == 0.0f || p_wb[2] == 0.0f)
    p_wb[0] = p_wb[1] = p_wb[2] = 1.0f;
  f[0] = p_wb[0] / p_wb[1];
  f[1] = 1.0f;
  f[2] = p_wb[2] / p_wb[1];
  f[3] = powf(2.0f, ((const float *)module->param)[0]);
  const int off = 4 + 12 + 4 + 12 + 4 * 24 + 4 * 24;
  if(p_tmp <= 0.0f) f[off + 0] = -1.0f;
  else f[off + 0] = 1.0f - CLAMP(tanf(asinhf(46.3407f + p_tmp)) + (-0.0287128f * cosf(0.000798585f * (714.855f - p_tmp))) + 0.942275f, 0.0f, 1.0f);
  u[off + 1] = p_mat == 4 ? 1 : 0;
  f[off + 2] = p_sat;
  u[off + 3] = p_pck;
  u[off + 4] = p_gam;
  u[off + 5] = ip->colour_primaries;
  u[off + 6] = ip->colour_trc;
  f[off + 7] = p_clp ? p_clm : 0.0f;
  float awb[3] = { ip->whitebalance[0], ip->whitebalance[1], ip->whitebalance[2] };
  if(!(awb[0] > 0.0f) || !(awb[1] > 0.0f) || !(awb[2] > 0.0f)) awb[0] = awb[1] = awb[2] = 1.0f;
  f[off + 8] = awb[0] / awb[1];
  f[off + 9] = 1.0f;
  f[off + 10] = awb[2] / awb[1];
  f[off + 11] = 1.0f;
  if(p_mat == 1)
  { // the matrix the source node attached to the image
    for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) f[4 + 4 * i + j] = ip->cam_to_rec2020[3 * j + i];
  }
  else if(p_mat == 2) { u[off + 5] = s_colour_primaries_XYZ; u[off + 6] = s_colour_trc_linear; }
  else if(p_mat == 3) { u[off + 5] = s_colour_primaries_srgb; u[off + 6] = s_colour_trc_linear; }
  else if(p_mat == 5)
  {
    u[off + 5] = s_colour_primaries_custom;
    u[off + 6] = s_colour_trc_linear;
    for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) f[4 + 4 * i + j] = p_mtx[3 * j + i];
  }
  else
  {
    u[off + 5] = s_colour_primaries_2020;
    u[off + 6] = s_colour_trc_linear;
    for(int j = 0; j < 3; j++) for(int i = 0; i < 3; i++) f[4 + 4 * j + i] = i == j ? 1.0f : 0.0f;
  }
  // rbf disabled (parametric mode); data-driven mode is not on this path
  u[16] = u[17] = u[18] = u[19] = 0;
  if(p_mod == 1) fprintf(stderr, "[colour] data-driven mode is not supported on this path, using the matrix\n");
}

dt_graph_run_t check_params(dt_module_t *module, uint32_t parid, uint32_t num, void *oldval)
{
  (void)module; (void)parid; (void)num; (void)oldval;
  return s_graph_run_record_cmd_buf;
}

void create_nodes(dt_graph_t *graph, dt_module_t *module)
{
  const int have_clut = dt_connected(module->connector + 2);
  const int have_pick = dt_connected(module->connector + 3);
  const int have_abney = dt_connected(module->connector + 4) && dt_connected(module->connector + 5);
  const int pc[] = { have_clut, have_pick, have_abney };
  const int id = dt_node_add(graph, module, "colour", "main", module->connector[0].roi.wd, module->connector[0].roi.ht, 1,
                             sizeof(pc), pc, 7,
                             "input", "read", "rgba", "f16", dt_no_roi,
                             "output", "write", "rgba", "f16", &module->connector[0].roi,
                             "clut", "read", "rgba", "f16", dt_no_roi,
                             "picked", "read", "r", "f16", dt_no_roi,
                             "abney", "read", "rg", "f16", dt_no_roi,
                             "spectra", "read", "rgba", "f16", dt_no_roi,
                             "autotemp", "read", "y", "f32", dt_no_roi);
  dt_connector_copy(graph, module, 0, id, 0);
  dt_connector_copy(graph, module, 1, id, 1);
  // unconnected optional inputs bind the main input as a dummy
  for(int i = 2; i < 6; i++) dt_connector_copy(graph, module, dt_connected(module->connector + i) ? i : 0, id, i);
  dt_connector_copy(graph, module, 0, id, 6);
}
