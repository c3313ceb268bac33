/*
 * global.c -- module class registry: reads every module's `params` and
 * `connectors` files and resolves its DSO callbacks.
 *
 * Same file formats and lookup as the reference (src/pipe/global.c:36-150,
 * 331-360): <basedir>/modules/<name>/{params,connectors,lib<name>.so};
 * params lines are name:type:cnt:defaults..., packed back to back in the
 * params blob; connector lines are name:type:chan:format.
 */
#include "api.h"
#include "ascii.h"

#include <dirent.h>
#include <dlfcn.h>
#include <stdlib.h>
#include <sys/stat.h>

#define DT_MAX_MODULE_SO 64

static struct
{
  char basedir[1024];
  dt_module_so_t so[DT_MAX_MODULE_SO];
  int num;
}
g_pipe;

static size_t param_type_size(dt_token_t type)
{
  if(type == dt_token("string")) return 1;
  return 4;   // float, int
}

static dt_ui_param_t *read_param_line(char *line)
{
  dt_token_t name = dt_read_token(line, &line);
  dt_token_t type = dt_read_token(line, &line);
  int cnt = dt_read_int(line, &line);
  if(cnt < 1) cnt = 1;
  dt_ui_param_t *p = calloc(1, sizeof(*p) + param_type_size(type) * cnt + 8);
  p->name = name;
  p->type = type;
  p->cnt = cnt;
  if(type == dt_token("float"))
    for(int i = 0; i < cnt; i++) p->val[i] = dt_read_float(line, &line);
  else if(type == dt_token("int"))
    for(int i = 0; i < cnt; i++) p->vali[i] = dt_read_int(line, &line);
  else if(type == dt_token("string"))
  {
    int i = 0;
    while(line[0] && i < cnt - 1) p->str[i++] = *(line++);
    p->str[i] = 0;
  }
  else
  {
    fprintf(stderr, "[pipe] unknown parameter type %" PRItkn "\n", dt_token_str(type));
  }
  return p;
}

static int read_lines(const char *path, int (*fn)(void *, char *), void *user)
{
  FILE *f = fopen(path, "rb");
  if(!f) return -1;
  char line[8192];
  int n = 0;
  while(fgets(line, sizeof(line), f))
  {
    size_t len = strlen(line);
    while(len && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
    if(!len) continue;
    if(fn(user, line)) break;
    n++;
  }
  fclose(f);
  return n;
}

static int add_param(void *user, char *line)
{
  dt_module_so_t *so = user;
  if(so->num_params >= DT_MAX_PARAMS) return 1;
  so->param[so->num_params++] = read_param_line(line);
  return 0;
}

static int add_connector(void *user, char *line)
{
  dt_module_so_t *so = user;
  if(so->num_connectors >= DT_MAX_CONNECTORS) return 1;
  dt_connector_t *c = so->connector + so->num_connectors++;
  memset(c, 0, sizeof(*c));
  c->name = dt_read_token(line, &line);
  c->type = dt_read_token(line, &line);
  c->chan = dt_read_token(line, &line);
  c->format = dt_read_token(line, &line);
  c->connected = s_cid_unset;
  c->associated = s_cid_unset;
  c->bypass = s_cid_unset;
  return 0;
}

static int load_so(dt_module_so_t *so, const char *dirname)
{
  memset(so, 0, sizeof(*so));
  so->name = dt_token(dirname);
  char path[2048];
  snprintf(path, sizeof(path), "%s/modules/%s/connectors", g_pipe.basedir, dirname);
  if(read_lines(path, add_connector, so) <= 0) return 1;   // a module needs connectors
  snprintf(path, sizeof(path), "%s/modules/%s/params", g_pipe.basedir, dirname);
  read_lines(path, add_param, so);
  int off = 0;
  for(int i = 0; i < so->num_params; i++)
  {
    so->param[i]->offset = off;
    off += (int)param_type_size(so->param[i]->type) * so->param[i]->cnt;
  }
  snprintf(path, sizeof(path), "%s/modules/%s/lib%s.so", g_pipe.basedir, dirname, dirname);
  struct stat st;
  if(!stat(path, &st))
  {
    so->dlhandle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if(!so->dlhandle)
    {
      fprintf(stderr, "[pipe] %s\n", dlerror());
      return 1;
    }
    so->create_nodes   = dlsym(so->dlhandle, "create_nodes");
    so->modify_roi_out = dlsym(so->dlhandle, "modify_roi_out");
    so->modify_roi_in  = dlsym(so->dlhandle, "modify_roi_in");
    so->init           = dlsym(so->dlhandle, "init");
    so->cleanup        = dlsym(so->dlhandle, "cleanup");
    so->commit_params  = dlsym(so->dlhandle, "commit_params");
    so->check_params   = dlsym(so->dlhandle, "check_params");
    so->read_source    = dlsym(so->dlhandle, "read_source");
    so->read_source_hip = dlsym(so->dlhandle, "read_source_hip");
    so->write_sink_hip = dlsym(so->dlhandle, "write_sink_hip");
    so->write_sink     = dlsym(so->dlhandle, "write_sink");
    so->audio          = dlsym(so->dlhandle, "audio");
  }
  return 0;
}

const dt_qvk_caps_t qvk = { .float_atomics_supported = 1 };

int dt_pipe_global_init(const char *basedir)
{
  if(g_pipe.num) dt_pipe_global_cleanup();
  snprintf(g_pipe.basedir, sizeof(g_pipe.basedir), "%s", basedir);
  char path[2048];
  snprintf(path, sizeof(path), "%s/modules", basedir);
  DIR *dp = opendir(path);
  if(!dp) return 1;
  struct dirent *ep;
  char names[DT_MAX_MODULE_SO][256];
  int n = 0;
  while((ep = readdir(dp)) && n < DT_MAX_MODULE_SO)
  {
    if(ep->d_name[0] == '.' || !strcmp(ep->d_name, "shared")) continue;
    snprintf(names[n++], 256, "%s", ep->d_name);
  }
  closedir(dp);
  for(int i = 0; i < n; i++)
    if(!load_so(g_pipe.so + g_pipe.num, names[i])) g_pipe.num++;
  return g_pipe.num > 0 ? 0 : 1;
}

void dt_pipe_global_cleanup(void)
{
  for(int i = 0; i < g_pipe.num; i++)
  {
    for(int p = 0; p < g_pipe.so[i].num_params; p++) free(g_pipe.so[i].param[p]);
    if(g_pipe.so[i].dlhandle) dlclose(g_pipe.so[i].dlhandle);
  }
  memset(&g_pipe, 0, sizeof(g_pipe));
}

dt_module_so_t *dt_pipe_get_so(dt_token_t name)
{
  for(int i = 0; i < g_pipe.num; i++)
    if(g_pipe.so[i].name == name) return g_pipe.so + i;
  return 0;
}

const char *dt_pipe_basedir(void) { return g_pipe.basedir; }
