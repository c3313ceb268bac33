// launch.h -- host-side plumbing shared by the node launchers: the decoded
// call, buffer checks and the thread-local error text.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/vkdt_hip.h"
#include "common.h"

struct node_call
{
  const void       *pc;
  size_t            pc_size;
  const void       *params;
  size_t            params_size;
  const vkdt_buf_t *conn;
  int               nconn;
  uint32_t          wd, ht, dp;
  hipStream_t       stream;
};

typedef int (*launch_fn)(const node_call &);

// records the error text for vkdt_hip_last_error() and returns `code`
int vkdt_set_error(int code, const char *fmt, ...);

// bytes per channel of a buffer format
static inline size_t vkdt_fmt_bytes(uint32_t fmt)
{
  return fmt == VKDT_FMT_ATOM ? 8 : ((fmt == VKDT_FMT_F32 || fmt == VKDT_FMT_UI32) ? 4 : (fmt == VKDT_FMT_UI8 ? 1 : 2));
}

// element-pitch device view of a connector
static inline dimg to_dimg(const vkdt_buf_t &b)
{
  dimg d;
  d.p = b.ptr;
  d.wd = (int)b.wd;
  d.ht = (int)b.ht;
  d.pitch = (int)b.pitch;
  d.chan = (int)b.chan;
  const size_t esz = vkdt_fmt_bytes(b.format) * (size_t)b.chan;
  d.pstride = (long)(b.plane_stride / (esz / b.chan));  // in scalar elements
  d.fmt = (int)b.format;
  return d;
}

// host-side validation of one connector before any launch: the kernels
// index within [0,wd) x [0,ht) x [0,array_length) and rely on these.
int check_conn(const node_call &c, int idx, uint32_t fmt, uint32_t chan, uint32_t min_planes, const char *what);

// the same for a connector that may hold either of two formats
int check_conn2(const node_call &c, int idx, uint32_t fmt_a, uint32_t fmt_b, uint32_t chan, uint32_t min_planes,
                const char *what);

// pc/params blob size checks
int check_blob(const node_call &c, size_t pc_min, size_t params_min, const char *what);

static inline dim3 grid2d(uint32_t wd, uint32_t ht, uint32_t dp)
{
  return dim3((wd + BX - 1) / BX, (ht + BY - 1) / BY, dp ? dp : 1);
}

int launch_post(const char *what);

// maps a HIP error to VKDT_HIP_OOM / VKDT_HIP_DEVICE with its text (OK on success)
int hip_fail(hipError_t e, const char *what);

// launchers, one per reference kernel
int launch_denoise_noop(const node_call &);
int launch_denoise_half(const node_call &);
int launch_denoise_downcov(const node_call &);
int launch_denoise_down(const node_call &);
int launch_denoise_assemble(const node_call &);
int launch_denoise_doub(const node_call &);
int launch_demosaic_down(const node_call &);
int launch_demosaic_gauss(const node_call &);
int launch_demosaic_splat(const node_call &);
int launch_demosaic_fix(const node_call &);
int launch_colour_main(const node_call &);
int launch_filmcurv_main(const node_call &);
int launch_llap_curve(const node_call &);
int launch_llap_reduce(const node_call &);
int launch_llap_assemble(const node_call &);
int launch_llap_colour(const node_call &);
int launch_hilite_half(const node_call &);
int launch_hilite_reduce(const node_call &);
int launch_hilite_assemble(const node_call &);
int launch_hilite_doub(const node_call &);
int launch_rawhist_collect(const node_call &);
int launch_hist_collect(const node_call &);
int launch_hist_map(const node_call &);
int launch_autoexp_hist(const node_call &);
int launch_autoexp_lum(const node_call &);
int launch_autoexp_exp(const node_call &);
int launch_rawhist_map(const node_call &);
int launch_demosaic_halfsize(const node_call &);
int launch_shared_resample(const node_call &);
int launch_colenc_main(const node_call &);
int launch_demosaic_rcd_conv(const node_call &);
int launch_demosaic_rcd_fill(const node_call &);
int launch_mcraw_prefix(const node_call &);
int launch_mcraw_decode(const node_call &);
