/*
 * vkdt-cli for the raw-development path: read a .cfg graph, replace its
 * display:main by an output module, run it on the GPU and write the file.
 *
 * Arguments as src/cli/main.c:37-159 for what the path supports:
 *   -g <graph.cfg>            graph to process
 *   --format o-pfm            output module (o-pfm only)
 *   --filename <base>         output file name without extension
 *   --colour-prim <p>         sRGB bt2020 adobergb P3 XYZ (default sRGB)
 *   --colour-trc <t>          linear 709 sRGB PQ DCI gamma2.2; the
 *                             reference's default is its parse of "sRGB" as
 *                             primaries, i.e. trc 1 (709) -- kept as is
 *   --width/--height <px>     scale the export to fit
 *   -d perf                   per-node and total GPU time
 *   --device <n>              HIP device
 */
#include "api.h"
#include "vkdt_hip.h"

#include <libgen.h>
#include <strings.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

// cli/main.c:14-35
static int parse_prim(const char *s)
{
  if(!strcasecmp(s, "sRGB")) return s_colour_primaries_srgb;
  if(!strcasecmp(s, "bt2020")) return s_colour_primaries_2020;
  if(!strcasecmp(s, "adobergb")) return s_colour_primaries_adobe;
  if(!strcasecmp(s, "P3")) return s_colour_primaries_P3;
  if(!strcasecmp(s, "XYZ")) return s_colour_primaries_XYZ;
  return s_colour_primaries_unknown;
}

static int parse_trc(const char *s)
{
  if(!strcasecmp(s, "linear")) return s_colour_trc_linear;
  if(!strcasecmp(s, "709")) return s_colour_trc_709;
  if(!strcasecmp(s, "sRGB")) return s_colour_trc_srgb;
  if(!strcasecmp(s, "PQ")) return s_colour_trc_PQ;
  if(!strcasecmp(s, "DCI")) return s_colour_trc_DCI;
  if(!strcasecmp(s, "gamma2.2")) return s_colour_trc_gamma;
  return s_colour_trc_unknown;
}

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int main(int argc, char *argv[])
{
  const char *cfg = 0, *fmt = "o-pfm", *fname = "output";
  int prim = parse_prim("sRGB"), trc = parse_prim("sRGB");   // cli/main.c:58-59
  int perf = 0, dev = 0, max_wd = 0, max_ht = 0, last_only = 0;
  float quality = 0.0f;   // > 0: the output's quality parameter
  for(int i = 1; i < argc; i++)
  {
    if(!strcmp(argv[i], "-g") && i + 1 < argc) cfg = argv[++i];
    else if(!strcmp(argv[i], "--format") && i + 1 < argc) fmt = argv[++i];
    else if(!strcmp(argv[i], "--filename") && i + 1 < argc) fname = argv[++i];
    else if(!strcmp(argv[i], "--colour-prim") && i + 1 < argc) prim = parse_prim(argv[++i]);
    else if(!strcmp(argv[i], "--colour-trc") && i + 1 < argc) trc = parse_trc(argv[++i]);
    else if(!strcmp(argv[i], "-d") && i + 1 < argc) { if(!strcmp(argv[++i], "perf") || !strcmp(argv[i], "all")) perf = 1; }
    else if(!strcmp(argv[i], "--device") && i + 1 < argc) dev = atoi(argv[++i]);
    else if(!strcmp(argv[i], "--width") && i + 1 < argc) max_wd = (int)atof(argv[++i]);
    else if(!strcmp(argv[i], "--height") && i + 1 < argc) max_ht = (int)atof(argv[++i]);
    else if(!strcmp(argv[i], "--last-frame-only")) last_only = 1;
    else if(!strcmp(argv[i], "--quality") && i + 1 < argc) quality = (float)atof(argv[++i]);
    else if(!strcmp(argv[i], "--version")) { printf("vkdt-cli (raw-development path, HIP gfx950)\n"); return 0; }
  }
  if(!cfg)
  {
    fprintf(stderr, "usage: %s -g <graph.cfg> [--format o-pfm] [--filename <base>] [--width <px>] [--height <px>]"
                    " [--quality <q>] [--last-frame-only] [-d perf]\n", argv[0]);
    return 1;
  }
  if(strcmp(fmt, "o-pfm") && strcmp(fmt, "o-jpg"))
  { fprintf(stderr, "[cli] output format %s not supported (o-pfm, o-jpg)\n", fmt); return 1; }

  char exe[4096] = {0}, base[4096];
  if(readlink("/proc/self/exe", exe, sizeof(exe) - 1) <= 0) snprintf(exe, sizeof(exe), "%s", argv[0]);
  snprintf(base, sizeof(base), "%s/..", dirname(exe));   // <pkg>/bin/.. holds modules/
  if(dt_pipe_global_init(base)) { fprintf(stderr, "[cli] no modules under %s\n", base); return 1; }
  if(vkdt_hip_init(dev)) { fprintf(stderr, "[cli] %s\n", vkdt_hip_last_error()); return 1; }

  dt_graph_t *g = malloc(sizeof(dt_graph_t));
  if(!g || dt_graph_init(g)) return 1;
  g->perf = perf;
  int ret = 1;
  if(dt_graph_read_config_ascii(g, cfg)) goto out;
  const int m = dt_graph_replace_display_colour(g, dt_token("main"), dt_token(fmt), prim, trc);
  if(m < 0) { fprintf(stderr, "[cli] graph has no connected display:main: %s\n", dt_graph_last_error(g)); goto out; }
  dt_graph_disconnect_display_modules(g);
  dt_module_set_param_string(g->module + m, dt_token("filename"), fname);
  if(quality > 0.0f)
  { // graph-export.c:240-241
    const int qp = dt_module_get_param(g->module[m].so, dt_token("quality"));
    if(qp >= 0) *(float *)dt_module_param_float(g->module + m, qp) = quality;
  }
  dt_graph_set_export_size(g, m, max_wd, max_ht);
  // a video source publishes its frame count in the roi pass: plan first, then
  // name the outputs <base>_%04d per frame unless only the last is wanted
  // (graph-export.c:213-230,251-316)
  if(dt_graph_plan(g, s_graph_run_all)) { fprintf(stderr, "[cli] plan failed: %s\n", dt_graph_last_error(g)); goto out; }
  const int frames = g->frame_cnt > 1 ? g->frame_cnt : 1;
  char name[2048];
  if(frames > 1 && !last_only)
  {
    snprintf(name, sizeof(name), "%s_%04d", fname, 0);
    dt_module_set_param_string(g->module + m, dt_token("filename"), name);
  }
  const double t0 = now_ms();
  g->frame = 0;
  if(dt_graph_run(g, s_graph_run_all & ~(frames > 1 && last_only ? s_graph_run_download_sink : 0)))
  { fprintf(stderr, "[cli] run failed: %s\n", dt_graph_last_error(g)); goto out; }
  for(int f = 1; f < frames; f++)
  {
    g->frame = f;
    if(!last_only) snprintf(name, sizeof(name), "%s_%04d", fname, f);
    else snprintf(name, sizeof(name), "%s", fname);
    dt_module_set_param_string(g->module + m, dt_token("filename"), name);
    const int dl = !last_only || f == frames - 1;
    if(dt_graph_run(g, s_graph_run_record_cmd_buf | s_graph_run_upload_source | s_graph_run_wait_done |
                         (dl ? s_graph_run_download_sink : 0)))
    { fprintf(stderr, "[cli] frame %d failed: %s\n", f, dt_graph_last_error(g)); goto out; }
    fprintf(stderr, "\r[cli] processing frame %d/%d", f, frames - 1);
  }
  if(frames > 1) fprintf(stderr, "\n");
  const double t1 = now_ms();
  if(perf)
  {
    float total = 0.0f;
    for(int k = 0; k < g->num_exec; k++)
    {
      const dt_node_t *n = g->node + g->exec[k];
      total += g->node_ms[k];
      fprintf(stderr, "[perf] %" PRItkn "\t%" PRItkn "\t: %8.3f ms\n", dt_token_str(n->name), dt_token_str(n->kernel),
              g->node_ms[k]);
    }
    fprintf(stderr, "[perf] total time:\t%8.3f ms (end to end %.1f ms)\n", total, t1 - t0);
  }
  ret = 0;
out:
  dt_graph_cleanup(g);
  free(g);
  dt_pipe_global_cleanup();
  return ret;
}
