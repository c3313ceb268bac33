/*
 * graph.c -- host graph runtime of the raw-development path: modules,
 * connections, cfg files, region-of-interest passes, node creation,
 * buffer allocation and execution through the HIP C ABI.
 *
 * Follows the reference's runtime (src/pipe/graph.c, graph-run-modules.h,
 * graph-run-nodes-*.h, connector.inc, graph-io.c) for the subset the path
 * needs: full-resolution regions (no LOD resampling), compute nodes only.
 * Every node becomes one vkdt_hip_node_launch on the graph's stream; the
 * node buffers live in one device arena whose offsets are reused once the
 * last reader of a buffer has run (graph-run-nodes-allocate.h:604-616).
 */
#include "api.h"
#include "ascii.h"
#include "vkdt_hip.h"

#include <stdarg.h>
#include <stdlib.h>

const char *dt_pipe_basedir(void);

static int gerr(dt_graph_t *g, int code, const char *fmt, ...)
{
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if(g)
  { // a module's own message of this pass (graph->err, gui_msg) comes first
    const size_t n = strnlen(g->err, sizeof(g->err));
    if(n && n + 3 < sizeof(g->err)) snprintf(g->err + n, sizeof(g->err) - n, "; %s", buf);
    else if(!n) snprintf(g->err, sizeof(g->err), "%s", buf);
  }
  fprintf(stderr, "[pipe] %s\n", buf);
  return code;
}

const char *dt_graph_last_error(const dt_graph_t *g) { return g->err; }

/* ------------------------------------------------------------------------
 * connectors
 * --------------------------------------------------------------------- */

int dt_connector_channels(const dt_connector_t *c)
{ // connector.h:229-238: rggb/rgbx mosaics are one channel, 3 channels pad to 4
  if(c->chan == dt_token("ssbo")) return 1;
  if(c->chan == dt_token("rggb") || c->chan == dt_token("rgbx")) return 1;
  return c->chan <= 0xff ? 1 : (c->chan <= 0xffff ? 2 : 4);
}

size_t dt_connector_bytes_per_channel(const dt_connector_t *c)
{
  if(c->format == dt_token("atom")) return 8;   // FP64 accumulators (vkdt_hip.h VKDT_FMT_ATOM)
  if(c->format == dt_token("f32") || c->format == dt_token("ui32") || c->format == dt_token("u32")) return 4;
  if(c->format == dt_token("f16") || c->format == dt_token("ui16") || c->format == dt_token("u16")) return 2;
  if(c->format == dt_token("ui8") || c->format == dt_token("u8")) return 1;
  return 4;
}

size_t dt_connector_bufsize(const dt_connector_t *c, uint32_t wd, uint32_t ht)
{
  const size_t n = (size_t)dt_connector_channels(c) * dt_connector_bytes_per_channel(c) * wd * ht;
  return n > 0 ? n : 16;
}

uint32_t dt_connector_hip_format(const dt_connector_t *c)
{
  if(c->format == dt_token("ui16") || c->format == dt_token("u16")) return VKDT_FMT_UI16;
  if(c->format == dt_token("f16")) return VKDT_FMT_F16;
  if(c->format == dt_token("f32")) return VKDT_FMT_F32;
  if(c->format == dt_token("atom")) return VKDT_FMT_ATOM;
  if(c->format == dt_token("ui32") || c->format == dt_token("u32")) return VKDT_FMT_UI32;
  if(c->format == dt_token("ui8") || c->format == dt_token("u8")) return VKDT_FMT_UI8;
  return 0;
}

int dt_connected(const dt_connector_t *c) { return !dt_cid_unset(c->connected); }

/* ------------------------------------------------------------------------
 * graph lifetime
 * --------------------------------------------------------------------- */

int dt_graph_init(dt_graph_t *g)
{
  memset(g, 0, sizeof(*g));
  g->params_max = 16u << 20;
  g->params_pool = calloc(1, g->params_max);
  g->max_nodes = DT_GRAPH_MAX_NODES;
  g->frame_cnt = 1;
  g->skip_dead = 1;
  g->runflags = s_graph_run_all;
  return g->params_pool ? 0 : 1;
}

static void free_device(dt_graph_t *g)
{
  if(g->rec) vkdt_hip_graph_destroy(g->rec);
  g->rec = 0;
  if(g->arena) vkdt_hip_free(g->arena);
  if(g->staging) vkdt_hip_host_free(g->staging);
  g->arena = g->staging = 0;
  g->arena_size = g->staging_size = 0;
}

void dt_graph_cleanup(dt_graph_t *g)
{
  for(int m = 0; m < g->num_modules; m++)
    if(g->module[m].name && g->module[m].so && g->module[m].so->cleanup)
      g->module[m].so->cleanup(g->module + m);
  if(g->stream) vkdt_hip_stream_sync(g->stream);
  free_device(g);
  for(int i = 0; i < DT_GRAPH_MAX_NODES + 1; i++)
    if(g->ev[i]) { vkdt_hip_event_destroy(g->ev[i]); g->ev[i] = 0; }
  if(g->stream) vkdt_hip_stream_destroy(g->stream);
  g->stream = 0;
  free(g->params_pool);
  g->params_pool = 0;
}

/* ------------------------------------------------------------------------
 * modules
 * --------------------------------------------------------------------- */

int dt_module_get(const dt_graph_t *g, dt_token_t name, dt_token_t inst)
{
  for(int m = 0; m < g->num_modules; m++)
    if(g->module[m].name == name && g->module[m].inst == inst) return m;
  return -1;
}

int dt_module_get_connector(const dt_module_t *m, dt_token_t conn)
{
  for(int c = 0; c < m->num_connectors; c++)
    if(m->connector[c].name == conn) return c;
  return -1;
}

int dt_module_get_param(const dt_module_so_t *so, dt_token_t param)
{
  for(int i = 0; i < so->num_params; i++)
    if(so->param[i]->name == param) return i;
  return -1;
}

const float *dt_module_param_float(const dt_module_t *m, int parid)
{
  if(parid < 0 || parid >= m->so->num_params) return 0;
  return (const float *)(m->param + m->so->param[parid]->offset);
}

const int32_t *dt_module_param_int(const dt_module_t *m, int parid)
{
  if(parid < 0 || parid >= m->so->num_params) return 0;
  return (const int32_t *)(m->param + m->so->param[parid]->offset);
}

const char *dt_module_param_string(const dt_module_t *m, int parid)
{
  if(parid < 0 || parid >= m->so->num_params) return 0;
  return (const char *)(m->param + m->so->param[parid]->offset);
}

int dt_module_set_param_string(dt_module_t *m, dt_token_t param, const char *str)
{
  const int p = dt_module_get_param(m->so, param);
  if(p < 0 || m->so->param[p]->type != dt_token("string")) return 1;
  char *d = (char *)(m->param + m->so->param[p]->offset);
  snprintf(d, m->so->param[p]->cnt, "%s", str);
  return 0;
}

int dt_module_set_param_float(dt_module_t *m, dt_token_t param, float val)
{
  const int p = dt_module_get_param(m->so, param);
  if(p < 0 || m->so->param[p]->type != dt_token("float")) return 1;
  *(float *)(m->param + m->so->param[p]->offset) = val;
  return 0;
}

int dt_module_set_param_int(dt_module_t *m, dt_token_t param, int32_t val)
{
  const int p = dt_module_get_param(m->so, param);
  if(p < 0 || m->so->param[p]->type != dt_token("int")) return 1;
  *(int32_t *)(m->param + m->so->param[p]->offset) = val;
  return 0;
}

int dt_module_add(dt_graph_t *g, dt_token_t name, dt_token_t inst)
{
  const int existing = dt_module_get(g, name, inst);
  if(existing >= 0) return existing;
  dt_module_so_t *so = dt_pipe_get_so(name);
  if(!so) return gerr(g, -1, "no such module %" PRItkn, dt_token_str(name));
  if(g->num_modules >= DT_GRAPH_MAX_MODULES) return gerr(g, -1, "too many modules");
  const int id = g->num_modules++;
  dt_module_t *m = g->module + id;
  memset(m, 0, sizeof(*m));
  m->so = so;
  m->name = name;
  m->inst = inst;
  m->graph = g;
  m->num_connectors = so->num_connectors;
  for(int c = 0; c < so->num_connectors; c++)
  {
    m->connector[c] = so->connector[c];
    m->connector[c].connected = dt_connector_input(m->connector + c) ? s_cid_unset : dt_cid(0, -1);
    m->connector[c].associated = s_cid_unset;
    m->connector[c].bypass = s_cid_unset;
  }
  size_t psize = 0;
  for(int i = 0; i < so->num_params; i++)
    psize = so->param[i]->offset + (so->param[i]->type == dt_token("string") ? 1 : 4) * so->param[i]->cnt;
  if(g->params_end + psize + 16 > g->params_max) return gerr(g, -1, "params pool exhausted");
  m->param = g->params_pool + g->params_end;
  m->param_size = (int)psize;
  g->params_end += (psize + 15) & ~(size_t)15;
  for(int i = 0; i < so->num_params; i++)
  {
    const dt_ui_param_t *p = so->param[i];
    const size_t sz = (p->type == dt_token("string") ? 1 : 4) * p->cnt;
    memcpy(m->param + p->offset, p->val, sz);
  }
  if(so->init && so->init(m)) return gerr(g, -1, "init failed for %" PRItkn, dt_token_str(name));
  if(m->committed_param_size)
  {
    if(g->params_end + m->committed_param_size + 16 > g->params_max) return gerr(g, -1, "params pool exhausted");
    m->committed_param = g->params_pool + g->params_end;
    g->params_end += (m->committed_param_size + 15) & ~15;
  }
  return id;
}

int dt_module_remove(dt_graph_t *g, int mid)
{ // disconnect everything and mark the module removed (name 0)
  if(mid < 0 || mid >= g->num_modules) return 1;
  dt_module_t *m = g->module + mid;
  for(int c = 0; c < m->num_connectors; c++)
    if(dt_connector_input(m->connector + c) && !dt_cid_unset(m->connector[c].connected))
      dt_module_connect(g, -1, -1, mid, c);
  for(int o = 0; o < g->num_modules; o++)
    for(int c = 0; c < g->module[o].num_connectors; c++)
      if(dt_connector_input(g->module[o].connector + c) && g->module[o].connector[c].connected.i == mid)
        dt_module_connect(g, -1, -1, o, c);
  if(m->so && m->so->cleanup) m->so->cleanup(m);
  m->name = 0;
  m->inst = 0;
  return 0;
}

// connection with channel/format negotiation (connector.inc), shared by
// modules and nodes
static int connect_impl(dt_graph_t *g, dt_connector_t *cn0, dt_connector_t *cn1,
                        dt_connector_t *old0, int m0, int c0, int is_module, const dt_connector_t *so1,
                        const dt_connector_t *so_old)
{
  if(cn1->type != dt_token("read") && cn1->type != dt_token("sink") && cn1->type != dt_token("modify")) return 3;
  if(old0)
  {
    cn1->connected = s_cid_unset;
    if(is_module && so1) { cn1->format = so1->format; cn1->chan = so1->chan; }
    if(old0->connected.i > 0) old0->connected.i--;
    if(is_module && so_old && old0->connected.i == 0) { old0->format = so_old->format; old0->chan = so_old->chan; }
  }
  cn1->associated = s_cid_unset;
  if(!cn0) return 0;
  cn0->associated = s_cid_unset;
  if(cn0->type != dt_token("write") && cn0->type != dt_token("source") && cn0->type != dt_token("modify")) return 9;
  const int c0ref = dt_token_str(cn0->chan)[0] == '&';
  if(cn1->chan == dt_token("*") && !c0ref) cn1->chan = cn0->chan;
  if(cn0->chan == dt_token("*")) cn0->chan = cn1->chan;
  if(cn1->chan == dt_token("*")) cn1->chan = dt_token("rgba");
  if(cn0->chan == dt_token("*")) cn0->chan = dt_token("rgba");
  if(!c0ref && cn1->chan != cn0->chan) return s_conn_err_chan_mismatch;
  if(cn1->format == dt_token("*")) cn1->format = cn0->format;
  if(cn0->format == dt_token("*")) cn0->format = cn1->format;
  if(cn1->format == dt_token("*")) cn1->format = dt_token("f16");
  if(cn0->format == dt_token("*")) cn0->format = dt_token("f16");
  if(cn1->format != cn0->format) return s_conn_err_format_mismatch;
  cn1->connected = dt_cid(m0, c0);
  cn1->array_length = cn0->array_length;
  cn1->flags = cn0->flags & ~s_conn_feedback;
  cn1->roi = cn0->roi;
  if(cn0->type == dt_token("write") || cn0->type == dt_token("source")) cn0->connected.i++;
  (void)g;
  return 0;
}

int dt_module_connect(dt_graph_t *g, int m0, int c0, int m1, int c1)
{
  if(m1 < 0 || m1 >= g->num_modules) return s_conn_err_module_ids;
  if(c1 < 0 || c1 >= g->module[m1].num_connectors) return s_conn_err_connector_ids;
  dt_connector_t *cn1 = g->module[m1].connector + c1;
  if(cn1->connected.i == m0 && cn1->connected.c == c0 && m0 >= 0) return 0;
  dt_connector_t *old0 = 0;
  const dt_connector_t *so_old = 0;
  if(!dt_cid_unset(cn1->connected))
  {
    old0 = g->module[cn1->connected.i].connector + cn1->connected.c;
    so_old = g->module[cn1->connected.i].so ? g->module[cn1->connected.i].so->connector + cn1->connected.c : 0;
  }
  if(m0 >= g->num_modules) return s_conn_err_module_ids;
  if(m0 >= 0 && (c0 < 0 || c0 >= g->module[m0].num_connectors)) return s_conn_err_connector_ids;
  dt_connector_t *cn0 = m0 >= 0 ? g->module[m0].connector + c0 : 0;
  if(m0 >= 0 && m0 == m1) return s_conn_err_cycle;
  return connect_impl(g, cn0, cn1, old0, m0, c0, 1, g->module[m1].so->connector + c1, so_old);
}

/* ------------------------------------------------------------------------
 * node-level API for create_nodes (modules/api.h:87-230)
 * --------------------------------------------------------------------- */

int dt_node_add(dt_graph_t *g, dt_module_t *module, const char *name, const char *kernel,
                int wd, int ht, int dp, int pc_size, const int *pc, int nc, ...)
{
  if(g->num_nodes >= g->max_nodes || nc > DT_MAX_CONNECTORS || pc_size > (int)sizeof(((dt_node_t *)0)->push_constant))
    return -1;
  const int id = g->num_nodes++;
  dt_node_t *n = g->node + id;
  memset(n, 0, sizeof(*n));
  n->name = dt_token(name);
  n->kernel = dt_token(kernel);
  n->module = module;
  n->num_connectors = nc;
  n->flags = module->flags;
  n->wd = wd;
  n->ht = ht;
  n->dp = dp;
  n->push_constant_size = pc_size;
  if(pc && pc_size > 0) memcpy(n->push_constant, pc, pc_size);
  va_list args;
  va_start(args, nc);
  for(int c = 0; c < nc; c++)
  {
    const char *cname = va_arg(args, const char *), *ctype = va_arg(args, const char *);
    const char *cchan = va_arg(args, const char *), *cfmt = va_arg(args, const char *);
    dt_roi_t *roi = va_arg(args, dt_roi_t *);
    dt_connector_t *cc = n->connector + c;
    cc->name = dt_token(cname);
    cc->type = dt_token(ctype);
    cc->chan = dt_token(cchan);
    cc->format = dt_token(cfmt);
    cc->associated = s_cid_unset;
    cc->bypass = s_cid_unset;
    cc->connected = dt_connector_input(cc) ? s_cid_unset : dt_cid(0, -1);
    if(dt_connector_owner(cc) && roi != dt_no_roi && roi) cc->roi = *roi;
  }
  va_end(args);
  return id;
}

int dt_node_connect(dt_graph_t *g, int n0, int c0, int n1, int c1)
{
  if(n1 < 0 || n1 >= g->num_nodes) return s_conn_err_module_ids;
  if(c1 < 0 || c1 >= g->node[n1].num_connectors) return s_conn_err_connector_ids;
  if(n0 < 0 || n0 >= g->num_nodes) return s_conn_err_module_ids;
  if(c0 < 0 || c0 >= g->node[n0].num_connectors) return s_conn_err_connector_ids;
  dt_connector_t *cn1 = g->node[n1].connector + c1;
  if(cn1->connected.i == n0 && cn1->connected.c == c0) return 0;
  dt_connector_t *old0 = dt_cid_unset(cn1->connected) ? 0 : g->node[cn1->connected.i].connector + cn1->connected.c;
  return connect_impl(g, g->node[n0].connector + c0, cn1, old0, n0, c0, 0, 0, 0);
}

int dt_node_connect_named(dt_graph_t *g, int n0, const char *c0, int n1, const char *c1)
{
  int i0 = -1, i1 = -1;
  for(int c = 0; c < g->node[n0].num_connectors; c++) if(g->node[n0].connector[c].name == dt_token(c0)) { i0 = c; break; }
  for(int c = 0; c < g->node[n1].num_connectors; c++) if(g->node[n1].connector[c].name == dt_token(c1)) { i1 = c; break; }
  if(i0 < 0) return -100;
  if(i1 < 0) return -101;
  return dt_node_connect(g, n0, i0, n1, i1);
}

void dt_connector_copy(dt_graph_t *g, dt_module_t *module, int mc, int nid, int nc)
{
  module->connector[mc].associated = dt_cid(nid, nc);
  dt_connector_t *c0 = module->connector + mc;
  dt_connector_t *c1 = g->node[nid].connector + nc;
  if(c1->name == 0) *c1 = *c0;
  else
  {
    c1->frames = MAX(c1->frames, c0->frames);
    c1->flags = c0->flags;
    c1->format = c0->format;
    if(!(c1->type == dt_token("write") && c1->roi.marker != s_roi_mark_uninited && c0->roi.marker == s_roi_mark_uninited))
      c1->roi = c0->roi;
    c1->connected = c0->connected;
    c1->array_length = c0->array_length;
  }
  c1->associated = dt_cid((int)(module - g->module), mc);
}

void dt_connector_bypass(dt_graph_t *g, dt_module_t *module, int mc_in, int mc_out)
{
  const int mid = (int)(module - g->module);
  module->connector[mc_out].bypass = dt_cid(mid, mc_in);
  module->connector[mc_in].bypass = dt_cid(mid, mc_out);
}

const dt_image_params_t *dt_module_get_input_img_param(dt_graph_t *g, dt_module_t *module, dt_token_t conn)
{
  for(int c = 0; c < module->num_connectors; c++)
    if(module->connector[c].name == conn)
    {
      const int mid = module->connector[c].connected.i;
      if(mid < 0) return 0;
      return &g->module[mid].img_param;
    }
  return 0;
}

/* ------------------------------------------------------------------------
 * resources and cfg files (graph-io.c)
 * --------------------------------------------------------------------- */

FILE *dt_graph_open_resource(dt_graph_t *g, int frame, const char *fname, const char *mode)
{
  (void)frame;
  if(fname[0] == '/') return fopen(fname, mode);
  char path[2300];
  if(g && g->searchpath[0])
  {
    snprintf(path, sizeof(path), "%s/%s", g->searchpath, fname);
    FILE *f = fopen(path, mode);
    if(f) return f;
  }
  snprintf(path, sizeof(path), "%s/%s", dt_pipe_basedir(), fname);
  FILE *f = fopen(path, mode);
  if(f) return f;
  return fopen(fname, mode);
}

int dt_graph_get_resource_filename(dt_module_t *mod, const char *fname, int frame, char *out, size_t size)
{
  char name[1024];
  if(strchr(fname, '%')) snprintf(name, sizeof(name), fname, frame);
  else snprintf(name, sizeof(name), "%s", fname);
  if(name[0] == '/') { snprintf(out, size, "%s", name); return 0; }
  dt_graph_t *g = mod->graph;
  if(g && g->searchpath[0])
  {
    snprintf(out, size, "%s/%s", g->searchpath, name);
    FILE *f = fopen(out, "rb");
    if(f) { fclose(f); return 0; }
  }
  snprintf(out, size, "%s", name);
  FILE *f = fopen(out, "rb");
  if(f) { fclose(f); return 0; }
  return 1;
}

static int read_param_values(dt_graph_t *g, char *line, dt_token_t name, dt_token_t inst, dt_token_t parm,
                             int beg, int end, int mode)
{
  const int mid = dt_module_get(g, name, inst);
  if(mid < 0) return gerr(g, 1, "no such module/instance %" PRItkn "/%" PRItkn, dt_token_str(name), dt_token_str(inst));
  const int pid = dt_module_get_param(g->module[mid].so, parm);
  if(pid < 0) return 2;
  const dt_ui_param_t *p = g->module[mid].so->param[pid];
  if(beg < 0 || beg >= p->cnt || end < 0 || end > p->cnt) return 4;
  if(end == 0) end = p->cnt;
  uint8_t *data = g->module[mid].param + p->offset;
  if(p->type == dt_token("float"))
  {
    float *b = (float *)data + beg;
    for(int i = beg; i < end; i++, b++)
    {
      const float v = dt_read_float(line, &line);
      *b = mode == 0 ? v : (mode == 1 ? *b + v : *b - v);
    }
  }
  else if(p->type == dt_token("int"))
  {
    int32_t *b = (int32_t *)data + beg;
    for(int i = beg; i < end; i++, b++)
    {
      const int v = dt_read_int(line, &line);
      *b = mode == 0 ? v : (mode == 1 ? *b + v : *b - v);
    }
  }
  else if(p->type == dt_token("string"))
  {
    char *s = (char *)data;
    int i = beg;
    while(line[0] && i < end - 1) s[i++] = *(line++);
    s[i] = 0;
  }
  return 0;
}

int dt_graph_read_config_line(dt_graph_t *g, char *c)
{
  if(c[0] == '#' || c[0] == 0) return 0;
  const dt_token_t cmd = dt_read_token(c, &c);
  if(cmd == dt_token("module"))
  {
    const dt_token_t name = dt_read_token(c, &c);
    const dt_token_t inst = dt_read_token(c, &c);
    return dt_module_add(g, name, inst) < 0 ? -1 : 0;
  }
  if(cmd == dt_token("param") || cmd == dt_token("paramsub") || cmd == dt_token("paraminc") ||
     cmd == dt_token("paramdec"))
  {
    const dt_token_t name = dt_read_token(c, &c);
    const dt_token_t inst = dt_read_token(c, &c);
    const dt_token_t parm = dt_read_token(c, &c);
    int beg = 0, end = 0, mode = 0;
    if(cmd != dt_token("param"))
    {
      beg = dt_read_int(c, &c);
      end = dt_read_int(c, &c);
      mode = cmd == dt_token("paraminc") ? 1 : (cmd == dt_token("paramdec") ? 2 : 0);
    }
    return read_param_values(g, c, name, inst, parm, beg, end, mode);
  }
  if(cmd == dt_token("connect") || cmd == dt_token("feedback"))
  {
    const dt_token_t mod0 = dt_read_token(c, &c), inst0 = dt_read_token(c, &c), conn0 = dt_read_token(c, &c);
    const dt_token_t mod1 = dt_read_token(c, &c), inst1 = dt_read_token(c, &c), conn1 = dt_read_token(c, &c);
    if(cmd == dt_token("feedback")) return gerr(g, 1, "feedback connections are not supported");
    const int m0 = dt_module_get(g, mod0, inst0), m1 = dt_module_get(g, mod1, inst1);
    if(m0 < 0 || m1 < 0) return gerr(g, s_conn_err_module_ids, "connect: no such module");
    const int c0 = dt_module_get_connector(g->module + m0, conn0), c1 = dt_module_get_connector(g->module + m1, conn1);
    const int err = dt_module_connect(g, m0, c0, m1, c1);
    if(err) gerr(g, err, "connect %" PRItkn ":%" PRItkn ":%" PRItkn " -> %" PRItkn ":%" PRItkn ":%" PRItkn " failed: %d",
                 dt_token_str(mod0), dt_token_str(inst0), dt_token_str(conn0),
                 dt_token_str(mod1), dt_token_str(inst1), dt_token_str(conn1), err);
    return err;
  }
  if(cmd == dt_token("frames")) { g->frame_cnt = atoi(c); return 0; }
  if(cmd == dt_token("fps")) return 0;
  if(dt_token_str(cmd)[0] == 'k' || dt_token_str(cmd)[0] == 'K') return 1;   // keyframes: ignored
  return 1;
}

int dt_graph_read_config_ascii(dt_graph_t *g, const char *filename)
{
  FILE *f = dt_graph_open_resource(g, 0, filename, "rb");
  if(!f) return gerr(g, 1, "cannot open %s", filename);
  // search path: the cfg's directory
  snprintf(g->searchpath, sizeof(g->searchpath), "%s", filename);
  char *slash = strrchr(g->searchpath, '/');
  if(slash) *slash = 0;
  else snprintf(g->searchpath, sizeof(g->searchpath), ".");
  char *line = malloc(300000);
  unsigned lno = 0;
  int ret = 0;
  while(fgets(line, 300000, f))
  {
    lno++;
    size_t len = strlen(line);
    while(len && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
    if(dt_graph_read_config_line(g, line) < 0)
    {
      gerr(g, 1, "failed in line %u: '%s'", lno, line);
      ret = 1;
      break;
    }
  }
  free(line);
  fclose(f);
  return ret;
}

/* ------------------------------------------------------------------------
 * export helpers (graph-export.c:24-106)
 * --------------------------------------------------------------------- */

int dt_graph_replace_display(dt_graph_t *g, dt_token_t inst, dt_token_t mod)
{
  return dt_graph_replace_display_colour(g, inst, mod, s_colour_primaries_2020, s_colour_trc_linear);
}

int dt_graph_replace_display_colour(dt_graph_t *g, dt_token_t inst, dt_token_t mod, int prim, int trc)
{
  if(inst == 0) inst = dt_token("main");
  const int mid = dt_module_get(g, dt_token("display"), inst);
  if(mid < 0) return -1;
  const int cid = dt_module_get_connector(g->module + mid, dt_token("input"));
  const int m0 = g->module[mid].connector[cid].connected.i;
  const int o0 = g->module[mid].connector[cid].connected.c;
  if(m0 < 0) return -2;
  const int m2 = dt_module_add(g, mod, inst);
  if(m2 < 0) return -3;
  const int i2 = dt_module_get_connector(g->module + m2, dt_token("input"));
  if(g->module[m2].connector[i2].format == dt_token("ui8") || prim != s_colour_primaries_2020 ||
     trc != s_colour_trc_linear)
  { // encode on the way out: m0 -> colenc -> output module
    const int m1 = dt_module_add(g, dt_token("colenc"), inst);
    if(m1 < 0) return gerr(g, -3, "no colenc module");
    const int i1 = dt_module_get_connector(g->module + m1, dt_token("input"));
    const int o1 = dt_module_get_connector(g->module + m1, dt_token("output"));
    if(prim == s_colour_primaries_unknown) prim = g->module[m0].img_param.colour_primaries;
    if(trc == s_colour_trc_unknown) trc = g->module[m0].img_param.colour_trc;
    dt_module_t *enc = g->module + m1;
    *(int32_t *)dt_module_param_int(enc, dt_module_get_param(enc->so, dt_token("prim"))) = prim;
    *(int32_t *)dt_module_param_int(enc, dt_module_get_param(enc->so, dt_token("trc"))) = trc;
    g->module[m1].connector[o1].format = g->module[m2].connector[i2].format;
    g->module[m1].connector[o1].chan = g->module[m2].connector[i2].chan;
    int err = dt_module_connect(g, m0, o0, m1, i1);
    if(!err) err = dt_module_connect(g, m1, o1, m2, i2);
    if(err) return gerr(g, -4, "connecting colenc failed: %d", err);
    return m2;
  }
  // linear rec2020 output: no colour encoding module needed, the last
  // module writes the sink's format directly
  g->module[m0].connector[o0].format = g->module[m2].connector[i2].format;
  g->module[m0].connector[o0].chan = g->module[m2].connector[i2].chan;
  const int err = dt_module_connect(g, m0, o0, m2, i2);
  if(err) return gerr(g, -4, "connecting output module failed: %d", err);
  return m2;
}

void dt_graph_disconnect_display_modules(dt_graph_t *g)
{
  for(int m = 0; m < g->num_modules; m++)
    if(g->module[m].name == dt_token("display")) dt_module_remove(g, m);
}

/* ------------------------------------------------------------------------
 * module passes (graph-run-modules.h)
 * --------------------------------------------------------------------- */

static int visit_module(dt_graph_t *g, int m, int *mark, int *order, int *cnt)
{
  if(mark[m] == 2) return 0;
  if(mark[m] == 1) return gerr(g, -1, "cycle in module graph");
  mark[m] = 1;
  dt_module_t *mod = g->module + m;
  for(int c = 0; c < mod->num_connectors; c++)
    if(dt_connector_input(mod->connector + c) && !dt_cid_unset(mod->connector[c].connected))
      if(visit_module(g, mod->connector[c].connected.i, mark, order, cnt)) return -1;
  mark[m] = 2;
  order[(*cnt)++] = m;
  return 0;
}

static void roi_out(dt_graph_t *g, dt_module_t *m)
{
  const int input = dt_module_get_connector(m, dt_token("input"));
  if(!(m->inst == dt_token("main") && !strncmp(dt_token_str(m->name), "i-", 2)))
    memset(&m->img_param, 0, sizeof(m->img_param));
  if(input >= 0 && !dt_cid_unset(m->connector[input].connected))
    m->img_param = g->module[m->connector[input].connected.i].img_param;
  for(int i = 0; i < m->num_connectors; i++)
  { // &name channel references
    dt_connector_t *c = m->connector + i;
    if(dt_token_str(c->chan)[0] == '&')
    {
      const dt_token_t ref = c->chan >> 8;
      for(int j = 0; j < m->num_connectors; j++)
        if(m->connector[j].name == ref) { c->chan = m->connector[j].chan; break; }
    }
  }
  for(int i = 0; i < m->num_connectors; i++)
  {
    dt_connector_t *c = m->connector + i;
    if(dt_connector_input(c) && !dt_cid_unset(c->connected))
      c->roi = g->module[c->connected.i].connector[c->connected.c].roi;
  }
  for(int i = 0; i < m->num_connectors; i++)
    if(dt_connector_owner(m->connector + i)) m->connector[i].roi.marker = s_roi_mark_uninited;
  if(!m->disabled && m->so->modify_roi_out)
  {
    m->so->modify_roi_out(g, m);
    for(int i = 0; i < m->num_connectors; i++)
    {
      if(m->connector[i].type == dt_token("source") && m->connector[i].roi.marker == s_roi_mark_uninited)
        m->connector[i].roi.marker = s_roi_mark_hard_fwd;
      else if(m->connector[i].roi.marker == s_roi_mark_uninited)
        m->connector[i].roi.marker = s_roi_mark_soft_fwd;
    }
  }
  else
  {
    dt_roi_t roi = { .marker = s_roi_mark_soft_fwd };
    if(input < 0) { roi.full_wd = 1024; roi.full_ht = 1024; }
    else if(!dt_cid_unset(m->connector[input].connected))
      roi = g->module[m->connector[input].connected.i].connector[m->connector[input].connected.c].roi;
    for(int i = 0; i < m->num_connectors; i++)
      if(m->connector[i].type == dt_token("write"))
      {
        m->connector[i].roi.full_wd = roi.full_wd;
        m->connector[i].roi.full_ht = roi.full_ht;
        m->connector[i].roi.marker = s_roi_mark_soft_fwd;
      }
  }
  if(m->connector[0].type == dt_token("source")) m->img_param.input_name = m->inst;
  for(int i = 0; i < m->num_connectors; i++)
  {
    dt_connector_t *c = m->connector + i;
    if(dt_connector_owner(c)) { c->roi.wd = c->roi.full_wd; c->roi.ht = c->roi.full_ht; }
  }
}

static void roi_in(dt_graph_t *g, dt_module_t *m)
{
  if(!m->disabled && m->so->modify_roi_in) m->so->modify_roi_in(g, m);
  else
  {
    int output = dt_module_get_connector(m, dt_token("output"));
    if(output < 0 && m->connector[0].type == dt_token("sink"))
    { // sinks ask for the full image, scaled to fit a requested export size
      // (graph-run-modules.h:459-474)
      output = 0;
      dt_roi_t *r = &m->connector[0].roi;
      const float max_wd = m->connector[0].max_wd, max_ht = m->connector[0].max_ht;
      const float sx = max_wd > 0 ? r->full_wd / max_wd : 1.0f;
      const float sy = max_ht > 0 ? r->full_ht / max_ht : 1.0f;
      const float scale = sx > sy ? sx : sy;
      r->wd = (int)(r->full_wd / scale + 0.5f);
      r->ht = (int)(r->full_ht / scale + 0.5f);
      r->marker = m->inst == dt_token("main") ? s_roi_mark_hard_bck : s_roi_mark_soft_bck;
    }
    if(output >= 0)
      for(int i = 0; i < m->num_connectors; i++)
        if(dt_connector_input(m->connector + i)) m->connector[i].roi = m->connector[output].roi;
  }
  // hand the request to the connected outputs (graph-run-modules.h:491-541).
  // The request's extent always travels upstream (a scaled export scales
  // every module on its path); its marker replaces the output's where it
  // constrains more, so a module that keys nodes on a hard marker of its
  // output (demosaic's resample, demosaic/main.c:145,193) sees the request.
  for(int i = 0; i < m->num_connectors; i++)
  {
    dt_connector_t *c = m->connector + i;
    if(!dt_connector_input(c) || dt_cid_unset(c->connected)) continue;
    dt_connector_t *up = g->module[c->connected.i].connector + c->connected.c;
    if(up->type == dt_token("source"))
    { // sources give what they have
      up->roi.wd = up->roi.full_wd;
      up->roi.ht = up->roi.full_ht;
      if(up->roi.marker == s_roi_mark_uninited) up->roi.marker = s_roi_mark_hard_bck;
      c->roi = up->roi;
    }
    else if(c->roi.wd > 0)
    {
      up->roi.wd = c->roi.wd;
      up->roi.ht = c->roi.ht;
      if(dt_roi_stronger(&c->roi, &up->roi)) up->roi.marker = c->roi.marker;
    }
    c->array_length = up->array_length;
  }
}

static void create_nodes(dt_graph_t *g, dt_module_t *m)
{
  for(int i = 0; i < m->num_connectors; i++) m->connector[i].bypass = s_cid_unset;
  if(m->disabled)
  {
    const int ci = dt_module_get_connector(m, dt_token("input")), co = dt_module_get_connector(m, dt_token("output"));
    if(ci >= 0 && co >= 0) dt_connector_bypass(g, m, ci, co);
  }
  else if(m->so->create_nodes) m->so->create_nodes(g, m);
  else
  { // default: one node running <module>/main on the module's connectors
    if(g->num_nodes >= g->max_nodes) return;
    const int nid = g->num_nodes++;
    dt_node_t *n = g->node + nid;
    memset(n, 0, sizeof(*n));
    n->name = m->name;
    n->kernel = dt_token("main");
    n->num_connectors = m->num_connectors;
    n->module = m;
    n->flags = m->flags;
    const int output = dt_module_get_connector(m, dt_token("output"));
    if(output >= 0) { n->wd = m->connector[output].roi.wd; n->ht = m->connector[output].roi.ht; n->dp = 1; }
    for(int i = 0; i < m->num_connectors; i++) dt_connector_copy(g, m, i, nid, i);
  }
}

static void repoint_nodes(dt_graph_t *g)
{ // graph-run-modules.h:762-807: node input -> module input -> module output -> node output
  for(int ni = 0; ni < g->num_nodes; ni++)
  {
    dt_node_t *n = g->node + ni;
    for(int i = 0; i < n->num_connectors; i++)
    {
      if(dt_cid_unset(n->connector[i].associated)) continue;
      const dt_cid_t m0 = n->connector[i].associated;
      dt_cid_t id;
      if(dt_connector_input(n->connector + i))
      {
        dt_cid_t m1 = g->module[m0.i].connector[m0.c].connected;
        if(dt_cid_unset(m1)) continue;
        int guard = 0;
        while(!dt_cid_unset(g->module[m1.i].connector[m1.c].bypass) && guard++ < DT_GRAPH_MAX_MODULES)
        {
          const dt_cid_t m2 = g->module[m1.i].connector[m1.c].bypass;
          const dt_cid_t m3 = g->module[m2.i].connector[m2.c].connected;
          if(dt_cid_unset(m3)) break;
          m1 = m3;
        }
        id = g->module[m1.i].connector[m1.c].associated;
      }
      else id = g->module[m0.i].connector[m0.c].associated;
      n->connector[i].connected = id;
    }
  }
}

/* ------------------------------------------------------------------------
 * node scheduling and allocation
 * --------------------------------------------------------------------- */

static int node_is_source(const dt_node_t *n) { return n->connector[0].type == dt_token("source"); }
static int node_is_sink(const dt_node_t *n) { return n->connector[0].type == dt_token("sink"); }

// bindings a kernel declares but never samples (their producers are dead
// code): demosaic/gauss.comp binds "input" (binding 0) and reads only "orig"
static int binding_unused(const dt_node_t *n, int c)
{
  return n->name == dt_token("demosaic") && n->kernel == dt_token("gauss") && c == 0;
}

static int visit_node(dt_graph_t *g, int n, int *mark, int *order, int *cnt)
{
  if(mark[n] == 2) return 0;
  if(mark[n] == 1) return gerr(g, -1, "cycle in node graph");
  mark[n] = 1;
  dt_node_t *node = g->node + n;
  for(int c = 0; c < node->num_connectors; c++)
  {
    const dt_connector_t *cc = node->connector + c;
    if(!dt_connector_input(cc)) continue;
    if(dt_cid_unset(cc->connected) || cc->connected.i >= g->num_nodes)
      return gerr(g, -1, "node %" PRItkn "/%" PRItkn " connector %" PRItkn " is not connected",
                  dt_token_str(node->name), dt_token_str(node->kernel), dt_token_str(cc->name));
    if(visit_node(g, cc->connected.i, mark, order, cnt)) return -1;
  }
  mark[n] = 2;
  order[(*cnt)++] = n;
  return 0;
}

typedef struct buf_slot { int node, conn; uint64_t size, off; int beg, end; } buf_slot;

static int allocate(dt_graph_t *g)
{
  static buf_slot slot[DT_GRAPH_MAX_NODES * 4];
  int ns = 0;
  int pos[DT_GRAPH_MAX_NODES];
  for(int i = 0; i < g->num_nodes; i++) pos[i] = -1;
  for(int k = 0; k < g->num_exec; k++) pos[g->exec[k]] = k;
  const int INF = 1 << 30;
  // readers of every owner connector
  for(int k = 0; k < g->num_exec; k++)
  {
    const int ni = g->exec[k];
    dt_node_t *n = g->node + ni;
    for(int c = 0; c < n->num_connectors; c++)
    {
      dt_connector_t *cc = n->connector + c;
      cc->dev = 0;
      if(!dt_connector_owner(cc)) continue;
      int last = -1;
      for(int r = 0; r < g->num_exec; r++)
      {
        const dt_node_t *rn = g->node + g->exec[r];
        for(int rc = 0; rc < rn->num_connectors; rc++)
          if(dt_connector_input(rn->connector + rc) && rn->connector[rc].connected.i == ni &&
             rn->connector[rc].connected.c == c && !binding_unused(rn, rc))
            last = MAX(last, node_is_sink(rn) ? INF : r);
      }
      if(last < 0 && !node_is_source(n)) continue;   // dead store: left unbound
      if(ns >= (int)(sizeof(slot) / sizeof(slot[0]))) return gerr(g, -1, "too many buffers");
      const int planes = MAX(1, cc->array_length);
      slot[ns].node = ni;
      slot[ns].conn = c;
      slot[ns].size = (dt_connector_bufsize(cc, cc->roi.wd, cc->roi.ht) * planes + 255) & ~(uint64_t)255;
      // sources keep their upload for repeated runs, protected buffers their
      // contents from frame to frame: never aliased
      const int keep = node_is_source(n) || (cc->flags & s_conn_protected);
      slot[ns].beg = keep ? -1 : k;
      slot[ns].end = keep ? INF : last;
      ns++;
    }
  }
  // linear scan: lowest offset not overlapping a live buffer
  uint64_t total = 0;
  for(int s = 0; s < ns; s++)
  {
    uint64_t off = 0;
    int moved = 1;
    while(moved)
    {
      moved = 0;
      for(int o = 0; o < s; o++)
      {
        const int live = !(slot[o].end < slot[s].beg || slot[s].end < slot[o].beg);
        if(live && off < slot[o].off + slot[o].size && slot[o].off < off + slot[s].size)
        { off = slot[o].off + slot[o].size; moved = 1; }
      }
    }
    slot[s].off = off;
    total = MAX(total, off + slot[s].size);
  }
  if(total > g->arena_size || !g->arena)
  {
    if(g->arena) vkdt_hip_free(g->arena);
    g->arena = 0;
    g->arena_size = 0;
    if(vkdt_hip_alloc(&g->arena, total))
      return gerr(g, -2, "out of device memory (%.1f MB): %s", total / 1e6, vkdt_hip_last_error());
    g->arena_size = total;
  }
  uint64_t staging = 0;
  for(int s = 0; s < ns; s++)
  {
    dt_connector_t *cc = g->node[slot[s].node].connector + slot[s].conn;
    cc->offset = slot[s].off;
    cc->bytes = slot[s].size;
    cc->dev = (uint8_t *)g->arena + slot[s].off;
    if(node_is_source(g->node + slot[s].node) || slot[s].end == INF) staging += slot[s].size;
  }
  // inputs point at their producers' buffers
  for(int k = 0; k < g->num_exec; k++)
  {
    dt_node_t *n = g->node + g->exec[k];
    for(int c = 0; c < n->num_connectors; c++)
    {
      dt_connector_t *cc = n->connector + c;
      if(dt_connector_input(cc) && !dt_cid_unset(cc->connected))
        cc->dev = binding_unused(n, c) ? 0 : g->node[cc->connected.i].connector[cc->connected.c].dev;
    }
  }
  if(staging > g->staging_size)
  {
    if(g->staging) vkdt_hip_host_free(g->staging);
    g->staging = 0;
    g->staging_size = 0;
    if(vkdt_hip_host_alloc(&g->staging, staging)) return gerr(g, -2, "out of pinned host memory");
    g->staging_size = staging;
  }
  return 0;
}

static void fill_buf(const dt_graph_t *g, const dt_node_t *n, int c, vkdt_buf_t *b)
{
  const dt_connector_t *cc = n->connector + c;
  const dt_connector_t *own = cc;
  if(dt_connector_input(cc) && !dt_cid_unset(cc->connected))
    own = g->node[cc->connected.i].connector + cc->connected.c;
  memset(b, 0, sizeof(*b));
  b->ptr = cc->dev;
  b->wd = own->roi.wd;
  b->ht = own->roi.ht;
  b->pitch = own->roi.wd;
  b->chan = dt_connector_channels(own);
  b->format = dt_connector_hip_format(own);
  b->array_length = MAX(1, own->array_length);
  b->plane_stride = dt_connector_bufsize(own, own->roi.wd, own->roi.ht);
}

// FNV-1a over what a run's launches see: every executed node's kernel,
// extent, push constants, uniform blob and connector buffers
static uint64_t fnv(uint64_t h, const void *p, size_t n)
{
  const uint8_t *b = (const uint8_t *)p;
  for(size_t i = 0; i < n; i++) { h ^= b[i]; h *= 0x100000001b3ull; }
  return h;
}

static uint64_t record_hash(const dt_graph_t *g)
{
  uint64_t h = 0xcbf29ce484222325ull;
  for(int k = 0; k < g->num_exec; k++)
  {
    const dt_node_t *n = g->node + g->exec[k];
    if(node_is_source(n) || node_is_sink(n)) continue;
    const dt_module_t *m = n->module;
    h = fnv(h, &n->name, sizeof(n->name));
    h = fnv(h, &n->kernel, sizeof(n->kernel));
    h = fnv(h, &n->wd, sizeof(n->wd));
    h = fnv(h, &n->ht, sizeof(n->ht));
    h = fnv(h, &n->dp, sizeof(n->dp));
    h = fnv(h, n->push_constant, n->push_constant_size);
    if(m->committed_param_size) h = fnv(h, m->committed_param, m->committed_param_size);
    else if(m->param_size) h = fnv(h, m->param, m->param_size);
    for(int c = 0; c < n->num_connectors; c++)
    {
      vkdt_buf_t b;
      memset(&b, 0, sizeof(b));
      fill_buf(g, n, c, &b);
      h = fnv(h, &b, sizeof(b));
    }
  }
  return h;
}

/* ------------------------------------------------------------------------
 * run
 * --------------------------------------------------------------------- */

static int module_order(dt_graph_t *g, int *order, int *cnt)
{
  int mark[DT_GRAPH_MAX_MODULES] = {0};
  *cnt = 0;
  for(int m = 0; m < g->num_modules; m++)
    if(g->module[m].name && g->module[m].connector[0].type == dt_token("sink"))
      if(visit_module(g, m, mark, order, cnt)) return -1;
  if(!*cnt) return gerr(g, -1, "graph has no sink");
  return 0;
}

int dt_graph_plan(dt_graph_t *g, dt_graph_run_t run)
{
  int order[DT_GRAPH_MAX_MODULES], cnt = 0;
  g->err[0] = 0;
  if(module_order(g, order, &cnt)) return -1;

  if(run & s_graph_run_roi)
  {
    for(int i = 0; i < cnt; i++)
      for(int c = 0; c < g->module[order[i]].num_connectors; c++)
      {
        dt_roi_t zero = {0};
        g->module[order[i]].connector[c].roi = zero;
      }
    for(int i = 0; i < cnt; i++) roi_out(g, g->module + order[i]);
    for(int i = cnt - 1; i >= 0; i--)
      if(g->module[order[i]].connector[0].type == dt_token("sink") && g->module[order[i]].connector[0].roi.full_wd == 0)
        return gerr(g, -1, "module %" PRItkn " %" PRItkn " connector %" PRItkn " has uninited size!",
                    dt_token_str(g->module[order[i]].name), dt_token_str(g->module[order[i]].inst),
                    dt_token_str(g->module[order[i]].connector[0].name));
  }
  if(run & (s_graph_run_roi | s_graph_run_create_nodes))
  {
    g->num_nodes = 0;
    for(int i = 0; i < cnt; i++)
      for(int c = 0; c < g->module[order[i]].num_connectors; c++) g->module[order[i]].connector[c].associated = s_cid_unset;
    for(int i = cnt - 1; i >= 0; i--)
      if(g->module[order[i]].connector[0].roi.full_wd > 0) roi_in(g, g->module + order[i]);
    for(int i = 0; i < cnt; i++)
      if(g->module[order[i]].connector[0].roi.full_wd > 0) create_nodes(g, g->module + order[i]);
    repoint_nodes(g);
    // node schedule: depth first from the sinks, inputs before consumers
    int nmark[DT_GRAPH_MAX_NODES] = {0};
    g->num_exec = 0;
    for(int n = 0; n < g->num_nodes; n++)
      if(node_is_sink(g->node + n))
        if(visit_node(g, n, nmark, g->exec, &g->num_exec)) return -1;
    for(int n = 0; n < g->num_nodes; n++) g->node[n].culled = nmark[n] != 2;
    // drop nodes whose outputs no kernel samples (repeat until stable)
    if(g->skip_dead)
    {
      int changed = 1;
      while(changed)
      {
        changed = 0;
        for(int k = 0; k < g->num_exec; k++)
        {
          const int ni = g->exec[k];
          dt_node_t *n = g->node + ni;
          if(node_is_sink(n) || node_is_source(n)) continue;
          int used = 0;
          for(int r = 0; r < g->num_exec && !used; r++)
          {
            const dt_node_t *rn = g->node + g->exec[r];
            for(int rc = 0; rc < rn->num_connectors; rc++)
              if(dt_connector_input(rn->connector + rc) && rn->connector[rc].connected.i == ni && !binding_unused(rn, rc))
              { used = 1; break; }
          }
          if(!used)
          {
            n->culled = 1;
            memmove(g->exec + k, g->exec + k + 1, sizeof(int) * (g->num_exec - k - 1));
            g->num_exec--;
            changed = 1;
            break;
          }
        }
      }
    }
    for(int k = 0; k < g->num_exec; k++) g->node[g->exec[k]].order = k;
  }
  return 0;
}

int dt_graph_run(dt_graph_t *g, dt_graph_run_t run)
{
  int order[DT_GRAPH_MAX_MODULES], cnt = 0;
  g->err[0] = 0;
  if(run & (s_graph_run_roi | s_graph_run_create_nodes))
  {
    const int e = dt_graph_plan(g, run);
    if(e) return e;
    run |= s_graph_run_alloc;
  }
  if(module_order(g, order, &cnt)) return -1;
  if(!g->stream && vkdt_hip_stream_create(&g->stream)) return gerr(g, -3, "stream: %s", vkdt_hip_last_error());
  if(run & s_graph_run_alloc)
  {
    int e = allocate(g);
    if(e) return e;
  }
  // upload sources (graph-run-nodes-upload.h)
  uint64_t stoff = 0;
  if(run & s_graph_run_upload_source)
  {
    for(int k = 0; k < g->num_exec; k++)
    {
      dt_node_t *n = g->node + g->exec[k];
      if(!node_is_source(n) || !n->module->so->read_source) continue;
      dt_connector_t *cc = n->connector + 0;
      dt_read_source_params_t p = { .node = n, .a = 0 };
      uint8_t *mapped = (uint8_t *)g->staging + stoff;
      stoff += cc->bytes;
      if(n->module->so->read_source_hip)
      { // the module fills its device buffer itself (GPU unpack)
        const int rc = n->module->so->read_source_hip(n->module, cc->dev, g->stream);
        if(rc < 0) return gerr(g, -1, "read_source_hip failed for %" PRItkn, dt_token_str(n->module->name));
        if(rc == 0) continue;
      }
      if(n->module->so->read_source(n->module, mapped, &p))
        return gerr(g, -1, "read_source failed for %" PRItkn, dt_token_str(n->module->name));
      const size_t bytes = dt_connector_bufsize(cc, cc->roi.wd, cc->roi.ht) * MAX(1, cc->array_length);
      if(vkdt_hip_upload(cc->dev, mapped, bytes, g->stream)) return gerr(g, -3, "upload: %s", vkdt_hip_last_error());
    }
  }
  // uniforms
  for(int i = 0; i < cnt; i++)
  {
    dt_module_t *m = g->module + order[i];
    if(m->so->commit_params) m->so->commit_params(g, m);
  }
  // record: one launch per node.  With replay on (and no per-node timing),
  // a run whose launches see the same buffers, push constants and uniforms
  // as the recorded one replays that recording as one hipGraph instead.
  if(g->perf)
    for(int k = 0; k <= g->num_exec; k++)
      if(!g->ev[k] && vkdt_hip_event_create(&g->ev[k])) return gerr(g, -3, "event: %s", vkdt_hip_last_error());
  const int replay = g->replay && !g->perf;
  const uint64_t h = replay ? record_hash(g) : 0;
  if(replay && g->rec && g->rec_hash == h)
  {
    if(vkdt_hip_graph_launch(g->rec, g->stream)) return gerr(g, -3, "replay: %s", vkdt_hip_last_error());
  }
  else
  {
    if(g->rec) vkdt_hip_graph_destroy(g->rec);
    g->rec = 0;
    if(replay && vkdt_hip_capture_begin(g->stream)) return gerr(g, -3, "capture: %s", vkdt_hip_last_error());
    if(g->perf) vkdt_hip_event_record(g->ev[0], g->stream);
    int e = 0;
    for(int k = 0; k < g->num_exec && !e; k++)
    {
      dt_node_t *n = g->node + g->exec[k];
      if(!node_is_source(n) && !node_is_sink(n))
      {
        vkdt_buf_t conn[DT_MAX_CONNECTORS];
        for(int c = 0; c < n->num_connectors; c++) fill_buf(g, n, c, conn + c);
        for(int c = 0; c < n->num_connectors && !e; c++)
          if((n->connector[c].flags & s_conn_clear) && dt_connector_output(n->connector + c) && conn[c].ptr &&
             vkdt_hip_buf_clear(conn + c, g->stream))
            e = gerr(g, -3, "clear %" PRItkn ": %s", dt_token_str(n->name), vkdt_hip_last_error());
        const dt_module_t *m = n->module;
        const void *params = m->committed_param_size ? (const void *)m->committed_param : (const void *)m->param;
        const size_t psize = m->committed_param_size ? (size_t)m->committed_param_size : (size_t)m->param_size;
        if(!e && (e = vkdt_hip_node_launch(n->name, n->kernel, n->push_constant, n->push_constant_size,
                                           params, psize, conn, n->num_connectors, n->wd, n->ht, n->dp, g->stream)))
          e = gerr(g, e, "node %" PRItkn "/%" PRItkn ": %s", dt_token_str(n->name), dt_token_str(n->kernel),
                   vkdt_hip_last_error());
      }
      if(g->perf) vkdt_hip_event_record(g->ev[k + 1], g->stream);
    }
    if(replay)
    { // end the capture whatever happened, so the stream leaves capture mode
      void *rec = 0;
      const int ec = vkdt_hip_capture_end(g->stream, &rec);
      if(e) { vkdt_hip_graph_destroy(rec); return e; }
      if(ec) return gerr(g, -3, "capture: %s", vkdt_hip_last_error());
      g->rec = rec;
      g->rec_hash = h;
      if(vkdt_hip_graph_launch(g->rec, g->stream)) return gerr(g, -3, "replay: %s", vkdt_hip_last_error());
    }
    else if(e) return e;
  }
  // download sinks (graph-run-nodes-download.h)
  if(run & s_graph_run_download_sink)
  {
    for(int k = 0; k < g->num_exec; k++)
    {
      dt_node_t *n = g->node + g->exec[k];
      if(!node_is_sink(n) || !(n->module->so->write_sink || n->module->so->write_sink_hip)) continue;
      dt_connector_t *cc = n->connector + 0;
      if(n->module->so->write_sink_hip)
      { // the module consumes its device buffer itself (GPU encode)
        const int rc = n->module->so->write_sink_hip(n->module, cc->dev, g->stream);
        if(rc < 0) return gerr(g, -1, "write_sink_hip failed for %" PRItkn, dt_token_str(n->module->name));
        if(rc == 0 || !n->module->so->write_sink) continue;
      }
      const dt_connector_t *own = g->node[cc->connected.i].connector + cc->connected.c;
      const size_t bytes = dt_connector_bufsize(own, own->roi.wd, own->roi.ht) * MAX(1, own->array_length);
      uint8_t *mapped = (uint8_t *)g->staging + stoff;
      if(vkdt_hip_download(mapped, cc->dev, bytes, g->stream)) return gerr(g, -3, "download: %s", vkdt_hip_last_error());
      if(vkdt_hip_stream_sync(g->stream)) return gerr(g, -3, "sync: %s", vkdt_hip_last_error());
      dt_write_sink_params_t p = { .node = n, .a = 0 };
      n->module->so->write_sink(n->module, mapped, &p);
    }
  }
  if(run & (s_graph_run_wait_done | s_graph_run_download_sink | s_graph_run_upload_source) || g->perf)
    if(vkdt_hip_stream_sync(g->stream)) return gerr(g, -3, "sync: %s", vkdt_hip_last_error());
  if(g->perf)
    for(int k = 0; k < g->num_exec; k++)
      vkdt_hip_event_elapsed_ms(&g->node_ms[k], g->ev[k], g->ev[k + 1]);
  return 0;
}
