/*
 * access.c -- flat accessors over the graph for foreign callers (Python
 * ctypes in the tests and benchmark), so they need no knowledge of the
 * struct layouts in api.h.
 */
#include "api.h"
#include "vkdt_hip.h"

#include <stdlib.h>

dt_graph_t *dt_graph_new(void)
{
  dt_graph_t *g = malloc(sizeof(dt_graph_t));
  if(g && dt_graph_init(g)) { free(g); return 0; }
  return g;
}

void dt_graph_free(dt_graph_t *g)
{
  if(!g) return;
  dt_graph_cleanup(g);
  free(g);
}

void dt_graph_set_perf(dt_graph_t *g, int on) { g->perf = on; }
void dt_graph_set_skip_dead(dt_graph_t *g, int on) { g->skip_dead = on; }
// record repeated runs' node launches once and replay them as a hipGraph
// (no effect while per-node timing is on)
void dt_graph_set_replay(dt_graph_t *g, int on) { g->replay = on; }
// the frame a run delivers (graph-export.c:268-271 sets graph->frame per frame)
void dt_graph_set_frame(dt_graph_t *g, int frame) { g->frame = frame; }
int  dt_graph_frame_cnt(const dt_graph_t *g) { return g->frame_cnt; }
float dt_graph_frame_rate(const dt_graph_t *g) { return g->frame_rate; }
int  dt_graph_num_nodes(const dt_graph_t *g) { return g->num_nodes; }
int  dt_graph_num_exec(const dt_graph_t *g) { return g->num_exec; }
int  dt_graph_exec_node(const dt_graph_t *g, int k) { return (k >= 0 && k < g->num_exec) ? g->exec[k] : -1; }
float dt_graph_exec_ms(const dt_graph_t *g, int k) { return (k >= 0 && k < g->num_exec) ? g->node_ms[k] : 0.0f; }
void *dt_graph_stream(dt_graph_t *g) { return g->stream; }
uint64_t dt_graph_arena_bytes(const dt_graph_t *g) { return g->arena_size; }

// module, kernel, extent, push constants, connector count, culled flag
int dt_graph_node_info(const dt_graph_t *g, int n, uint64_t *name, uint64_t *kernel, uint32_t dim[3],
                       const void **pc, uint64_t *pc_size, int *nconn, int *culled)
{
  if(n < 0 || n >= g->num_nodes) return 1;
  const dt_node_t *node = g->node + n;
  *name = node->name;
  *kernel = node->kernel;
  dim[0] = node->wd;
  dim[1] = node->ht;
  dim[2] = node->dp;
  *pc = node->push_constant;
  *pc_size = node->push_constant_size;
  *nconn = node->num_connectors;
  *culled = node->culled;
  return 0;
}

// the uniform block the node's kernel sees (committed params or params)
const void *dt_graph_node_params(const dt_graph_t *g, int n, uint64_t *size)
{
  if(n < 0 || n >= g->num_nodes) return 0;
  const dt_module_t *m = g->node[n].module;
  if(m->committed_param_size) { *size = m->committed_param_size; return m->committed_param; }
  *size = m->param_size;
  return m->param;
}

// one connector: names, the node/connector it reads, and its buffer view
int dt_graph_node_conn(const dt_graph_t *g, int n, int c, uint64_t tok[4], int link[2], vkdt_buf_t *buf)
{
  if(n < 0 || n >= g->num_nodes || c < 0 || c >= g->node[n].num_connectors) return 1;
  const dt_connector_t *cc = g->node[n].connector + c;
  tok[0] = cc->name;
  tok[1] = cc->type;
  tok[2] = cc->chan;
  tok[3] = cc->format;
  link[0] = cc->connected.i;
  link[1] = cc->connected.c;
  const dt_connector_t *own = cc;
  if(dt_connector_input(cc) && !dt_cid_unset(cc->connected) && cc->connected.i < g->num_nodes)
    own = g->node[cc->connected.i].connector + cc->connected.c;
  memset(buf, 0, sizeof(*buf));
  buf->ptr = cc->dev;
  buf->wd = own->roi.wd;
  buf->ht = own->roi.ht;
  buf->pitch = own->roi.wd;
  buf->chan = dt_connector_channels(own);
  buf->format = dt_connector_hip_format(own);
  buf->array_length = own->array_length > 1 ? own->array_length : 1;
  buf->plane_stride = dt_connector_bufsize(own, own->roi.wd, own->roi.ht);
  return 0;
}

// export size limit on a sink module's input (graph-export.c:207-208):
// the image is scaled to fit max_wd x max_ht (0: no limit on that axis)
int dt_graph_set_export_size(dt_graph_t *g, int m, int max_wd, int max_ht)
{
  if(m < 0 || m >= g->num_modules || g->module[m].connector[0].type != dt_token("sink")) return -1;
  g->module[m].connector[0].max_wd = max_wd > 0 ? max_wd : 0;
  g->module[m].connector[0].max_ht = max_ht > 0 ? max_ht : 0;
  return 0;
}

// module lookup and parameter setting by name, for scripted graphs
int dt_graph_module(dt_graph_t *g, const char *name, const char *inst)
{
  return dt_module_get(g, dt_token(name), dt_token(inst));
}

int dt_graph_set_param(dt_graph_t *g, int m, const char *param, const char *value)
{ // value in cfg syntax (':' separated for arrays)
  if(m < 0 || m >= g->num_modules) return 1;
  char line[2048];
  snprintf(line, sizeof(line), "param:%.8s:%.8s:%s:%s", dt_token_str(g->module[m].name),
           dt_token_str(g->module[m].inst), param, value);
  return dt_graph_read_config_line(g, line);
}

const dt_image_params_t *dt_graph_module_img_param(const dt_graph_t *g, int m)
{
  if(m < 0 || m >= g->num_modules) return 0;
  return &g->module[m].img_param;
}

// the metadata a module publishes: levels, wb, matrix, CFA, crop, noise
int dt_graph_img_param_get(const dt_graph_t *g, int m, float black[4], float white[4], float wb[4],
                           float cam[9], uint32_t *filters, uint32_t crop[4], float noise[2], uint32_t roi[2])
{
  if(m < 0 || m >= g->num_modules) return 1;
  const dt_image_params_t *p = &g->module[m].img_param;
  memcpy(black, p->black, sizeof(p->black));
  memcpy(white, p->white, sizeof(p->white));
  memcpy(wb, p->whitebalance, sizeof(p->whitebalance));
  memcpy(cam, p->cam_to_rec2020, sizeof(p->cam_to_rec2020));
  memcpy(crop, p->crop_aabb, sizeof(p->crop_aabb));
  *filters = p->filters;
  noise[0] = p->noise_a;
  noise[1] = p->noise_b;
  roi[0] = g->module[m].connector[0].roi.full_wd;
  roi[1] = g->module[m].connector[0].roi.full_ht;
  return 0;
}

// capture metadata: exposure (s), aperture, iso, focal length; maker, model, datetime
int dt_graph_img_meta_get(const dt_graph_t *g, int m, float meta[4], char maker[32], char model[32], char datetime[20])
{
  if(m < 0 || m >= g->num_modules) return 1;
  const dt_image_params_t *p = &g->module[m].img_param;
  meta[0] = p->exposure;
  meta[1] = p->aperture;
  meta[2] = p->iso;
  meta[3] = p->focal_length;
  memcpy(maker, p->maker, 32);
  memcpy(model, p->model, 32);
  memcpy(datetime, p->datetime, 20);
  return 0;
}

// audio of a video source: snd_format, snd_channels, snd_samplerate
int dt_graph_img_snd_get(const dt_graph_t *g, int m, int snd[3])
{
  if(m < 0 || m >= g->num_modules) return 1;
  const dt_image_params_t *p = &g->module[m].img_param;
  snd[0] = p->snd_format;
  snd[1] = p->snd_channels;
  snd[2] = p->snd_samplerate;
  return 0;
}

// host part of a source module's upload: read_source into caller memory
// (the staging copy that precedes the device transfer)
int dt_graph_read_source(dt_graph_t *g, int m, void *buf)
{
  if(m < 0 || m >= g->num_modules || !g->module[m].so->read_source) return 1;
  dt_read_source_params_t p = { 0 };
  for(int n = 0; n < g->num_nodes && !p.node; n++)   // the module's first source node
    if(g->node[n].module == g->module + m && g->node[n].connector[0].type == dt_token("source")) p.node = g->node + n;
  return g->module[m].so->read_source(g->module + m, buf, &p);
}

// read_source of one source node (a module with several sources tells them
// apart by the node, as the runtime's upload passes it)
int dt_graph_node_read_source(dt_graph_t *g, int n, void *buf)
{
  if(n < 0 || n >= g->num_nodes || g->node[n].connector[0].type != dt_token("source")) return 1;
  dt_module_t *m = g->node[n].module;
  if(!m || !m->so->read_source) return 1;
  dt_read_source_params_t p = { .node = g->node + n, .a = 0 };
  return m->so->read_source(m, buf, &p);
}

int dt_graph_module_audio(dt_graph_t *g, int m, uint64_t beg, uint32_t cnt, uint8_t **samples)
{
  *samples = 0;
  if(m < 0 || m >= g->num_modules || !g->module[m].so || !g->module[m].so->audio) return 0;
  return g->module[m].so->audio(g->module + m, beg, cnt, samples);
}

// run every module's commit_params (the uniform upload) without the device
void dt_graph_commit(dt_graph_t *g)
{
  for(int m = 0; m < g->num_modules; m++)
    if(g->module[m].name && g->module[m].so && g->module[m].so->commit_params)
      g->module[m].so->commit_params(g, g->module + m);
}

// device address of a node connector's buffer after allocation
void *dt_graph_node_buffer(const dt_graph_t *g, int n, int c)
{
  if(n < 0 || n >= g->num_nodes || c < 0 || c >= g->node[n].num_connectors) return 0;
  return g->node[n].connector[c].dev;
}
