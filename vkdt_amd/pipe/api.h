/*
 * api.h -- the host-side module API of the raw-development path.
 *
 * Module sources (vkdt_amd/modules/<name>/main.c) compile against this
 * header.  It keeps the names, argument meaning and error behaviour of
 * vkdt's module API (src/pipe/modules/api.h, module.h, node.h,
 * connector.h, graph.h) for the subset the path uses, so a module's
 * create_nodes / modify_roi_in / modify_roi_out / commit_params /
 * check_params / read_source / write_sink read like the reference's.
 * There is no Vulkan: nodes run through the HIP C ABI (include/vkdt_hip.h).
 */
#pragma once
#include <assert.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include "vkdt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tokens: up to 8 chars packed little endian (pipe/token.h) ---------- */
typedef uint64_t dt_token_t;

static inline dt_token_t dt_token(const char *s)
{
  dt_token_t t = 0;
  for(int i = 0; i < 8 && s[i]; i++) t |= (dt_token_t)(uint8_t)s[i] << (8 * i);
  return t;
}
#define dt_token_str(A) ((const char *)&(A))
#define PRItkn ".8s"

#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif
#ifndef CLAMP
#define CLAMP(x, a, b) MIN(MAX((x), (a)), (b))
#endif

#define DT_MAX_CONNECTORS   30
#define DT_MAX_PARAMS       64
#define DT_LOCAL_SIZE_X      8
#define DT_LOCAL_SIZE_Y      8
#define dt_no_roi ((dt_roi_t *)(uintptr_t)UINT64_C(-1))

typedef struct dt_graph_t dt_graph_t;
typedef struct dt_module_t dt_module_t;
typedef struct dt_node_t dt_node_t;

/* ---- regions of interest (connector.h:11-23) --------------------------- */
typedef enum dt_roi_mark_t
{
  s_roi_mark_uninited  = 0,
  s_roi_mark_soft      = 1,
  s_roi_mark_hard      = 2,
  s_roi_mark_bck       = 4,
  s_roi_mark_fwd       = 8,
  s_roi_mark_soft_bck  = 1 | 4,
  s_roi_mark_hard_bck  = 2 | 4,
  s_roi_mark_soft_fwd  = 1 | 8,
  s_roi_mark_hard_fwd  = 2 | 8,
  s_roi_mark_dontcare  = 0xff,
}
dt_roi_mark_t;

typedef struct dt_roi_t
{
  uint32_t full_wd, full_ht;  // full input size
  uint32_t wd, ht;            // dimensions of the processed region
  uint32_t marker;            // dt_roi_mark_t
}
dt_roi_t;

// non-zero if the marker of a constrains more than the marker of b: hard
// over soft over uninited, forward over backward at equal strength, and a
// don't-care marker never (connector.h:32-39)
static inline int dt_roi_stronger(const dt_roi_t *a, const dt_roi_t *b)
{
  if(a->marker == s_roi_mark_dontcare) return 0;
  if(b->marker == s_roi_mark_dontcare) return 1;
  const int strength = (int)(a->marker & 3u) - (int)(b->marker & 3u);
  if(strength) return strength > 0;
  return (a->marker & 12u) > (b->marker & 12u);
}

/* ---- connectors (connector.h) ------------------------------------------ */
typedef struct dt_cid_t { int i, c; } dt_cid_t;
static const dt_cid_t s_cid_unset = { -1, -1 };
static inline dt_cid_t dt_cid(int i, int c) { dt_cid_t r = { i, c }; return r; }
static inline int dt_cid_unset(dt_cid_t c) { return c.i < 0 || c.c < 0; }

// connector flags (connector.h:60-71)
enum
{
  s_conn_none = 0, s_conn_smooth = 1, s_conn_clear = 2, s_conn_feedback = 4, s_conn_dynamic_array = 8,
  s_conn_protected = 16, s_conn_double_buffer = 32, s_conn_mipmap = 64, s_conn_clear_once = 128,
};

typedef struct dt_connector_t
{
  dt_token_t name;      // connector name
  dt_token_t type;      // read write source sink modify
  dt_token_t chan;      // rgba rggb y g ... or &name
  dt_token_t format;    // ui16 f16 f32 ...
  dt_roi_t   roi;
  dt_cid_t   connected;  // inputs: the output we read; outputs: .i counts readers
  dt_cid_t   associated; // module <-> node link
  dt_cid_t   bypass;     // module connector bypass
  int        array_length;
  int        flags;
  int        frames;
  uint32_t   max_wd, max_ht;   // export size limit on sinks
  // runtime state (set by the graph during allocation)
  void      *dev;        // device address of the buffer this connector reads/writes
  uint64_t   bytes;      // buffer size (owners)
  uint64_t   offset;     // arena offset (owners)
}
dt_connector_t;

// connect error codes (connector.h:179-199)
typedef enum dt_connector_error_t
{
  s_conn_ok                  = 0,
  s_conn_err_module_ids      = 1,
  s_conn_err_connector_ids   = 2,
  s_conn_err_not_output      = 3,
  s_conn_err_not_input       = 4,
  s_conn_err_already_in      = 5,
  s_conn_err_cycle           = 8,
  s_conn_err_chan_mismatch   = 10,
  s_conn_err_format_mismatch = 11,
  s_conn_err_unknown         = 12,
}
dt_connector_error_t;

int dt_connector_channels(const dt_connector_t *c);
size_t dt_connector_bytes_per_channel(const dt_connector_t *c);
size_t dt_connector_bufsize(const dt_connector_t *c, uint32_t wd, uint32_t ht);
uint32_t dt_connector_hip_format(const dt_connector_t *c);   // VKDT_FMT_* of a connector
static inline int dt_connector_owner(const dt_connector_t *c)
{ return c->type == dt_token("write") || c->type == dt_token("source"); }
static inline int dt_connector_input(const dt_connector_t *c)
{ return c->type == dt_token("read") || c->type == dt_token("sink") || c->type == dt_token("modify"); }
static inline int dt_connector_output(const dt_connector_t *c)
{ return c->type == dt_token("write") || c->type == dt_token("source") || c->type == dt_token("modify"); }

/* ---- image metadata travelling with the graph (module.h:55-100) ------- */
typedef enum dt_colour_primaries_t   // module.h:21-42
{
  s_colour_primaries_custom       = 0,
  s_colour_primaries_srgb         = 1,
  s_colour_primaries_2020         = 2,
  s_colour_primaries_adobe        = 3,
  s_colour_primaries_P3           = 4,
  s_colour_primaries_XYZ          = 5,
  s_colour_primaries_aces2065_1   = 6,
  s_colour_primaries_acescg       = 7,
  s_colour_primaries_arri3        = 8,
  s_colour_primaries_arri4        = 9,
  s_colour_primaries_red          = 10,
  s_colour_primaries_unknown      = 0xffff,
}
dt_colour_primaries_t;

typedef enum dt_colour_trc_t          // module.h:44-60
{
  s_colour_trc_linear  = 0,
  s_colour_trc_709     = 1,
  s_colour_trc_srgb    = 2,
  s_colour_trc_PQ      = 3,
  s_colour_trc_DCI     = 4,
  s_colour_trc_HLG     = 5,
  s_colour_trc_gamma   = 6,
  s_colour_trc_unknown = 0xffff,
}
dt_colour_trc_t;

// extra metadata an input module attaches to its image params, as a singly
// linked list of typed records it owns (metadata.h)
typedef struct dt_image_metadata_t
{
  uint32_t type;
  void    *next;
}
dt_image_metadata_t;

#define s_image_metadata_text 1
typedef struct dt_image_metadata_text_t
{
  uint32_t    type;           // s_image_metadata_text
  void       *next;
  const char *text;
}
dt_image_metadata_text_t;

#define s_image_metadata_dngop 2
typedef struct dt_image_metadata_dngop_t
{
  uint32_t type;              // s_image_metadata_dngop
  void    *next;
  uint32_t ox, oy;            // CFA offset of the emitted buffer in the raw
  void    *op_list[3];        // dt_dng_opcode_list_t of OpcodeList1..3, or 0
}
dt_image_metadata_dngop_t;

static inline dt_image_metadata_t *dt_metadata_append(dt_image_metadata_t *meta, dt_image_metadata_t *add)
{
  add->next = meta;
  return add;
}

static inline void *dt_metadata_find(dt_image_metadata_t *meta, uint32_t type)
{
  for(; meta; meta = (dt_image_metadata_t *)meta->next)
    if(meta->type == type) return meta;
  return 0;
}

static inline void *dt_metadata_remove(dt_image_metadata_t *meta, dt_image_metadata_t *rem)
{
  if(!meta) return 0;
  if(meta == rem) return rem->next;
  meta->next = dt_metadata_remove((dt_image_metadata_t *)meta->next, rem);
  return meta;
}

typedef struct dt_image_params_t
{
  float    black[4];          // black point (raw units)
  float    white[4];          // clipping threshold (raw units)
  float    whitebalance[4];   // camera white balance coefficients
  uint32_t filters;           // 0 full colour, 9 x-trans, else bayer (rggb after normalisation)
  uint32_t crop_aabb[4];      // x y X Y of the usable area
  float    cam_to_rec2020[9]; // row major
  uint32_t orientation;
  char     datetime[20];
  char     maker[32];
  char     model[32];
  float    exposure, aperture, iso, focal_length;
  int      colour_primaries;
  int      colour_trc;
  float    noise_a;           // raw noise, gaussian part
  float    noise_b;           // raw noise, poissonian part
  dt_token_t input_name;
  dt_image_metadata_t *meta;  // owned by the input module that attached it
  // audio of a video source (module.h:96-99)
  int      snd_format;        // alsa's snd_pcm_format_t (2: S16_LE)
  int      snd_channels;      // 0 none, 1, 2 ...
  int      snd_samplerate;    // Hz
}
dt_image_params_t;

/* ---- module classes and instances (global.h, module.h) ---------------- */
typedef enum dt_graph_run_t
{
  s_graph_run_none           = 0,
  s_graph_run_roi            = 1 << 0,
  s_graph_run_create_nodes   = 1 << 1,
  s_graph_run_alloc          = 1 << 2,
  s_graph_run_record_cmd_buf = 1 << 3,
  s_graph_run_upload_source  = 1 << 4,
  s_graph_run_download_sink  = 1 << 5,
  s_graph_run_wait_done      = 1 << 6,
  s_graph_run_all            = -1,
}
dt_graph_run_t;

typedef enum dt_module_flags_t
{
  s_module_request_none        = 0,
  s_module_request_read_source = 1,
  s_module_request_write_sink  = 2,
  s_module_request_all         = 4,
}
dt_module_flags_t;

typedef struct dt_ui_param_t
{
  dt_token_t name;
  dt_token_t type;   // float int string
  int        cnt;
  int        offset; // byte offset in the params blob
  union { float val[1]; int32_t vali[1]; char str[1]; };
}
dt_ui_param_t;

typedef struct dt_read_source_params_t { dt_node_t *node; int a; } dt_read_source_params_t;
typedef struct dt_write_sink_params_t  { dt_node_t *node; int a; } dt_write_sink_params_t;

typedef struct dt_module_so_t
{
  dt_token_t name;
  void *dlhandle;
  void (*create_nodes)(dt_graph_t *, dt_module_t *);
  void (*modify_roi_out)(dt_graph_t *, dt_module_t *);
  void (*modify_roi_in)(dt_graph_t *, dt_module_t *);
  int  (*init)(dt_module_t *);
  void (*cleanup)(dt_module_t *);
  void (*commit_params)(dt_graph_t *, dt_module_t *);
  dt_graph_run_t (*check_params)(dt_module_t *, uint32_t, uint32_t, void *);
  int  (*read_source)(dt_module_t *, void *, dt_read_source_params_t *);
  // optional, no reference counterpart: fill the source's device buffer on
  // `stream` directly (GPU unpack); 0 done, 1 use read_source, -1 error
  int  (*read_source_hip)(dt_module_t *, void *dev, void *stream);
  // optional, no reference counterpart: consume the sink's device buffer on
  // `stream` directly (GPU encode); 0 done, 1 use download + write_sink, -1 error
  int  (*write_sink_hip)(dt_module_t *, const void *dev, void *stream);
  void (*write_sink)(dt_module_t *, void *, dt_write_sink_params_t *);
  // optional: *samples points at the clip's audio from sample `beg` on;
  // returns how many of `cnt` samples are there (module.h audio callback)
  int  (*audio)(dt_module_t *, uint64_t beg, uint32_t cnt, uint8_t **samples);
  dt_ui_param_t *param[DT_MAX_PARAMS];
  int num_params;
  dt_connector_t connector[DT_MAX_CONNECTORS];
  int num_connectors;
}
dt_module_so_t;

struct dt_module_t
{
  dt_module_so_t *so;
  dt_token_t name, inst;
  dt_graph_t *graph;
  dt_connector_t connector[DT_MAX_CONNECTORS];
  int num_connectors;
  dt_image_params_t img_param;
  int disabled;
  uint8_t *param;
  int param_size;
  uint8_t *committed_param;
  int committed_param_size;
  int flags;
  void *data;
};

/* ---- nodes (node.h) ---------------------------------------------------- */
struct dt_node_t
{
  dt_token_t name;     // module directory of the kernel
  dt_token_t kernel;   // kernel name
  dt_module_t *module;
  dt_connector_t connector[DT_MAX_CONNECTORS];
  int num_connectors;
  int flags;
  uint32_t wd, ht, dp;
  uint32_t push_constant[64];
  size_t   push_constant_size;
  int      culled;      // not reachable from a sink
  int      order;       // execution position
};

/* ---- the graph (graph.h) ----------------------------------------------- */
#define DT_GRAPH_MAX_MODULES 64
#define DT_GRAPH_MAX_NODES   512

struct dt_graph_t
{
  dt_module_t module[DT_GRAPH_MAX_MODULES];
  int num_modules;
  dt_node_t node[DT_GRAPH_MAX_NODES];
  int num_nodes;
  int max_nodes;

  uint8_t *params_pool;
  size_t   params_end, params_max;

  char searchpath[1024];
  char basedir[1024];

  int frame, frame_cnt;
  float frame_rate;         // frames per second of a video source (0: stills)
  dt_graph_run_t runflags;

  // device state
  void    *stream;
  void    *arena;           // one device allocation for all node buffers
  uint64_t arena_size;
  void    *staging;         // pinned host staging (sources and sinks)
  uint64_t staging_size;
  int      exec[DT_GRAPH_MAX_NODES];   // node ids in execution order
  int      num_exec;
  int      perf;            // record per-node timings
  void    *ev[DT_GRAPH_MAX_NODES + 1];
  float    node_ms[DT_GRAPH_MAX_NODES];
  int      skip_dead;       // skip nodes whose outputs no kernel samples
  int      replay;          // record a run's launches as one hipGraph and replay it
  void    *rec;             // the recorded launches (a hipGraph executable), or 0
  uint64_t rec_hash;        // hash of what the recorded launches saw
  char     err[256];        // last error text ("gui_msg")
};

/* ---- graph API (graph.h, graph-io.h, module.h) ------------------------- */
// load module classes from <basedir>/modules/*/ (global.c:86-150, 331-360)
int  dt_pipe_global_init(const char *basedir);
void dt_pipe_global_cleanup(void);
dt_module_so_t *dt_pipe_get_so(dt_token_t name);

int  dt_graph_init(dt_graph_t *graph);
void dt_graph_cleanup(dt_graph_t *graph);
int  dt_graph_read_config_ascii(dt_graph_t *graph, const char *filename);
int  dt_graph_read_config_line(dt_graph_t *graph, char *line);
int  dt_graph_run(dt_graph_t *graph, dt_graph_run_t run);
// module passes and node schedule only (no device work): roi, create_nodes
int  dt_graph_plan(dt_graph_t *graph, dt_graph_run_t run);
// every module's commit_params (the uniform blocks) without the device
void dt_graph_commit(dt_graph_t *graph);
// after dt_graph_plan: the batched-develop description of this graph's image
// (commits the params first)
// every field but the device pointers; the chain must be exactly
// i-raw > denoise > demosaic > colour > filmcurv > llap.  0 on success.
int  dt_graph_develop_image(dt_graph_t *graph, vkdt_develop_image_t *img);
int  dt_module_add(dt_graph_t *graph, dt_token_t name, dt_token_t inst);
int  dt_module_get(const dt_graph_t *graph, dt_token_t name, dt_token_t inst);
int  dt_module_connect(dt_graph_t *graph, int m0, int c0, int m1, int c1);
int  dt_module_remove(dt_graph_t *graph, int m);
int  dt_module_get_connector(const dt_module_t *m, dt_token_t conn);
int  dt_module_set_param_string(dt_module_t *m, dt_token_t param, const char *str);
int  dt_module_set_param_float(dt_module_t *m, dt_token_t param, float val);
int  dt_module_set_param_int(dt_module_t *m, dt_token_t param, int32_t val);
// replace display:<inst> by an output module (graph-export.c:24-96)
int  dt_graph_replace_display(dt_graph_t *graph, dt_token_t inst, dt_token_t mod);
// the same with an output colour encoding: a colenc module goes in between
// unless the encoding is linear rec2020 (graph-export.c:60-90)
int  dt_graph_replace_display_colour(dt_graph_t *graph, dt_token_t inst, dt_token_t mod, int prim, int trc);
void dt_graph_disconnect_display_modules(dt_graph_t *graph);
int  dt_graph_set_export_size(dt_graph_t *graph, int module, int max_wd, int max_ht);
// resolve a resource relative to the graph's search path, then basedir
FILE *dt_graph_open_resource(dt_graph_t *graph, int frame, const char *fname, const char *mode);
int  dt_graph_get_resource_filename(dt_module_t *mod, const char *fname, int frame, char *out, size_t size);
// the snd_format, snd_channels and snd_samplerate a module publishes
int  dt_graph_img_snd_get(const dt_graph_t *graph, int m, int snd[3]);
// a module's audio callback: samples from `beg` (count returned, 0 none)
int  dt_graph_module_audio(dt_graph_t *graph, int m, uint64_t beg, uint32_t cnt, uint8_t **samples);
const char *dt_graph_last_error(const dt_graph_t *graph);

/* ---- module-side helpers (modules/api.h) ------------------------------ */
int  dt_node_add(dt_graph_t *graph, dt_module_t *module, const char *name, const char *kernel,
                 int wd, int ht, int dp, int pc_size, const int *pc, int nc, ...);
int  dt_node_connect(dt_graph_t *graph, int n0, int c0, int n1, int c1);
int  dt_node_connect_named(dt_graph_t *graph, int n0, const char *c0, int n1, const char *c1);
void dt_connector_copy(dt_graph_t *graph, dt_module_t *module, int mc, int nid, int nc);
void dt_connector_bypass(dt_graph_t *graph, dt_module_t *module, int mc_in, int mc_out);
const dt_image_params_t *dt_module_get_input_img_param(dt_graph_t *graph, dt_module_t *module, dt_token_t conn);
int  dt_module_get_param(const dt_module_so_t *so, dt_token_t param);
const float   *dt_module_param_float(const dt_module_t *m, int parid);
const int32_t *dt_module_param_int(const dt_module_t *m, int parid);
const char    *dt_module_param_string(const dt_module_t *m, int parid);
int  dt_connected(const dt_connector_t *c);

/* ---- device capabilities modules consult (qvk/qvk.h) ------------------ */
// Modules pick kernels by `qvk.float_atomics_supported` (colour, rawhist):
// gfx950 adds floats atomically in global memory, so it is always 1 here.
typedef struct dt_qvk_caps_t
{
  int float_atomics_supported;
}
dt_qvk_caps_t;
extern const dt_qvk_caps_t qvk;

#define CONN(A) do { int err = (A); if(err) \
  fprintf(stderr, "[create_nodes] %s:%d connection failed: %d\n", __FILE__, __LINE__, err); } while(0)

#ifdef __cplusplus
}
#endif
