// develop.c -- one image of the batched develop path, read off a planned
// graph.  The batch (vkdt_hip_batch_*) runs the raw-development chain
// i-raw -> denoise -> [hilite ->] demosaic -> colour -> filmcurv -> llap with
// the values this graph's nodes would receive, so that its result equals
// the graph's.
#include "api.h"

static const dt_node_t *find_node(const dt_graph_t *g, const char *name, const char *kernel, int *count)
{
  const dt_node_t *r = 0;
  int c = 0;
  for(int n = 0; n < g->num_nodes; n++)
    if(g->node[n].name == dt_token(name) && g->node[n].kernel == dt_token(kernel))
    {
      if(!r) r = g->node + n;
      c++;
    }
  if(count) *count = c;
  return r;
}

static const void *node_params(const dt_node_t *n, int *size)
{
  const dt_module_t *m = n->module;
  if(m->committed_param_size) { *size = m->committed_param_size; return m->committed_param; }
  *size = m->param_size;
  return m->param;
}

// the module instance whose output feeds module m's connector `input`
static const dt_module_t *upstream(const dt_graph_t *g, const dt_module_t *m)
{
  const int c = dt_module_get_connector(m, dt_token("input"));
  if(c < 0 || dt_cid_unset(m->connector[c].connected)) return 0;
  return g->module + m->connector[c].connected.i;
}

static int fail(dt_graph_t *g, const char *msg)
{
  snprintf(g->err, sizeof(g->err), "develop image: %s", msg);
  return 1;
}

int dt_graph_develop_image(dt_graph_t *g, vkdt_develop_image_t *d)
{
  memset(d, 0, sizeof(*d));
  dt_graph_commit(g);
  const dt_node_t *doub = find_node(g, "denoise", "doub", 0);
  const dt_node_t *half = find_node(g, "denoise", "half", 0);
  const dt_node_t *gauss = find_node(g, "demosaic", "gauss", 0);
  const dt_node_t *fix = find_node(g, "demosaic", "fix", 0);
  const dt_node_t *col = find_node(g, "colour", "main", 0);
  const dt_node_t *fc = find_node(g, "filmcurv", "main", 0);
  const dt_node_t *curve = find_node(g, "llap", "curve", 0);
  const dt_node_t *llc = find_node(g, "llap", "colour", 0);
  int nreduce = 0;
  find_node(g, "llap", "reduce", &nreduce);
  if(!doub || !half) return fail(g, "denoise is not the mosaic wavelet path (strength > 0 on a raw)");
  if(!gauss || !fix) return fail(g, "demosaic is not the gaussian splat path");
  if(!col || !fc || !curve || !llc) return fail(g, "colour, filmcurv or llap missing");
  // the chain must be exactly i-raw -> denoise -> [hilite ->] demosaic ->
  // colour -> filmcurv -> llap
  static const char *chain_msg =
      "the modules are not the chain i-raw > denoise > [hilite >] demosaic > colour > filmcurv > llap";
  const dt_module_t *chain[4] = { llc->module, fc->module, col->module, fix->module };
  for(int k = 0; k < 3; k++)
    if(upstream(g, chain[k]) != chain[k + 1]) return fail(g, chain_msg);
  const dt_module_t *hl = 0, *up = upstream(g, fix->module);
  if(up && up->name == dt_token("hilite"))
  {
    hl = up;
    up = upstream(g, hl);
  }
  if(up != doub->module) return fail(g, chain_msg);
  const dt_module_t *raw = upstream(g, doub->module);
  if(!raw || raw->name != dt_token("i-raw")) return fail(g, "denoise does not read i-raw");
  for(int n = 0; n < g->num_nodes; n++)
  {
    const dt_node_t *nd = g->node + n;
    if(nd->culled) continue;
    const dt_token_t mn = nd->name;
    if(mn != dt_token("denoise") && mn != dt_token("demosaic") && mn != dt_token("colour") &&
       mn != dt_token("filmcurv") && mn != dt_token("llap") && !(hl && nd->module == hl) && nd->module != raw &&
       nd->module->connector[0].type != dt_token("sink"))
      return fail(g, "the graph runs nodes outside the raw-development chain");
  }

  const uint32_t *pc = doub->push_constant;
  d->cfa_wd = raw->connector[0].roi.wd;
  d->cfa_ht = raw->connector[0].roi.ht;
  d->cfa_pitch = d->cfa_wd;
  d->wd = doub->wd;
  d->ht = doub->ht;
  for(int k = 0; k < 4; k++)
  {
    memcpy(d->wb + k, pc + k, 4);
    memcpy(d->black + k, pc + 4 + k, 4);
    memcpy(d->white + k, pc + 8 + k, 4);
    d->crop[k] = (int32_t)pc[12 + k];
  }
  d->filters = pc[16];
  memcpy(&d->noise_a, pc + 17, 4);
  memcpy(&d->noise_b, pc + 18, 4);

  int sz = 0;
  const void *p = node_params(doub, &sz);
  if(sz < 32) return fail(g, "denoise params");
  // a Bayer gain map (push constant and module parameter both 1) is applied
  // by the per-node doub only
  if(pc[19] == 1 && d->filters != 9 && (sz < 36 || ((const int32_t *)p)[8] == 1))
    return fail(g, "the batched path does not apply dng gain maps (the per-node graph does)");
  memcpy(d->dn_params, p, 32);
  p = node_params(fix, &sz);
  d->dm_fixup = sz >= 4 ? ((const int32_t *)p)[0] : 0;
  p = node_params(col, &sz);
  if(sz < (int)sizeof(d->co_params)) return fail(g, "colour params");
  memcpy(d->co_params, p, sizeof(d->co_params));
  p = node_params(fc, &sz);
  if(sz < 16) return fail(g, "filmcurv params");
  // the colour modes 0, 2, 4 and 5 run on the per-node filmcurv only
  const int32_t fc_mode = ((const int32_t *)p)[3];
  if(fc_mode != 1 && fc_mode != 3)
    return fail(g, "the batched path runs filmcurv colour modes 1 and 3 (the per-node graph runs all)");
  memcpy(d->fc_params, p, 16);
  p = node_params(curve, &sz);
  if(sz < 16) return fail(g, "llap params");
  memcpy(d->ll_params, p, 16);
  d->ll_numgamma = (int32_t)curve->push_constant[0];
  d->ll_levels = nreduce + 1;
  if(llc->wd != d->wd || llc->ht != d->ht || fix->wd != d->wd || fix->ht != d->ht)
    return fail(g, "demosaic or llap extent differs from the denoise output");
  if(hl)
  { // hilite/main.c: half carries wb[4] + filters, a reduce per coarse level
    const dt_node_t *hh = 0;
    int nhr = 0;
    for(int n = 0; n < g->num_nodes; n++)
    {
      const dt_node_t *nd = g->node + n;
      if(nd->module != hl) continue;
      if(nd->kernel == dt_token("half")) hh = nd;
      if(nd->kernel == dt_token("reduce")) nhr++;
    }
    if(!hh) return fail(g, "hilite passes its input through (not a mosaic)");
    const uint32_t block = d->filters == 9u ? 3 : 2;
    if(hh->wd != d->wd / block || hh->ht != d->ht / block || hh->push_constant[4] != d->filters)
      return fail(g, "hilite extent or filters differ from the denoise output");
    memcpy(d->hl_wb, hh->push_constant, 16);
    p = node_params(hh, &sz);
    if(sz < 12) return fail(g, "hilite params");
    memcpy(d->hl_params, p, 12);
    d->hl_levels = nhr;
    d->hl_on = 1;
  }
  return 0;
}
