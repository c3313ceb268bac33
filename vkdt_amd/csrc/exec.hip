// exec.hip -- the C-ABI entry points of libvkdt_hip.so (include/vkdt_hip.h):
// node registry keyed by (module token, kernel token), device/pinned memory,
// copies, streams, events and error text.
#include "launch.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static thread_local char g_err[512] = "";

int vkdt_set_error(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *what)
{
  if(e == hipSuccess) return VKDT_HIP_OK;
  return vkdt_set_error(e == hipErrorOutOfMemory ? VKDT_HIP_OOM : VKDT_HIP_DEVICE,
                        "%s: %s", what, hipGetErrorString(e));
}

int launch_post(const char *what) { return hip_fail(hipGetLastError(), what); }

int check_blob(const node_call &c, size_t pc_min, size_t params_min, const char *what)
{
  if(pc_min && (!c.pc || c.pc_size < pc_min))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: push constants %zu < %zu bytes", what, c.pc_size, pc_min);
  if(params_min && (!c.params || c.params_size < params_min))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: params %zu < %zu bytes", what, c.params_size, params_min);
  return VKDT_HIP_OK;
}

int check_conn2(const node_call &c, int idx, uint32_t fmt_a, uint32_t fmt_b, uint32_t chan, uint32_t min_planes,
                const char *what)
{
  if(idx < c.nconn && c.conn[idx].format == fmt_b) return check_conn(c, idx, fmt_b, chan, min_planes, what);
  return check_conn(c, idx, fmt_a, chan, min_planes, what);
}

int check_conn(const node_call &c, int idx, uint32_t fmt, uint32_t chan, uint32_t min_planes, const char *what)
{
  if(idx >= c.nconn) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d missing", what, idx);
  const vkdt_buf_t &b = c.conn[idx];
  if(!b.ptr) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d unbound", what, idx);
  if(b.format != fmt || b.chan != chan)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d is fmt %u chan %u, want fmt %u chan %u",
                          what, idx, b.format, b.chan, fmt, chan);
  if(b.wd == 0 || b.ht == 0 || b.pitch < b.wd)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d extent %ux%u pitch %u", what, idx, b.wd, b.ht, b.pitch);
  if(b.array_length < (min_planes ? min_planes : 1))
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d has %u planes, need %u", what, idx, b.array_length, min_planes);
  const size_t esz = vkdt_fmt_bytes(fmt) * (size_t)chan;
  if(b.array_length > 1 && b.plane_stride < (size_t)b.pitch * b.ht * esz)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d plane stride too small", what, idx);
  if(b.array_length > 1 && b.plane_stride % esz)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d plane stride not a texel multiple", what, idx);
  if(((uintptr_t)b.ptr) % esz)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "%s: connector %d misaligned", what, idx);
  return VKDT_HIP_OK;
}

static constexpr uint64_t tkn(const char *s)
{
  uint64_t t = 0;
  for(int i = 0; i < 8 && s[i]; i++) t |= (uint64_t)(uint8_t)s[i] << (8 * i);
  return t;
}

struct reg_entry { uint64_t module, kernel; launch_fn fn; };

static const reg_entry g_registry[] = {
  { tkn("denoise"),  tkn("noop"),     launch_denoise_noop },
  { tkn("denoise"),  tkn("half"),     launch_denoise_half },
  { tkn("denoise"),  tkn("downcov"),  launch_denoise_downcov },
  { tkn("denoise"),  tkn("down"),     launch_denoise_down },
  { tkn("denoise"),  tkn("assemble"), launch_denoise_assemble },
  { tkn("denoise"),  tkn("doub"),     launch_denoise_doub },
  { tkn("demosaic"), tkn("down"),     launch_demosaic_down },
  { tkn("demosaic"), tkn("gauss"),    launch_demosaic_gauss },
  { tkn("demosaic"), tkn("splat"),    launch_demosaic_splat },
  { tkn("demosaic"), tkn("fix"),      launch_demosaic_fix },
  { tkn("colour"),   tkn("main"),     launch_colour_main },
  { tkn("filmcurv"), tkn("main"),     launch_filmcurv_main },
  { tkn("llap"),     tkn("curve"),    launch_llap_curve },
  { tkn("llap"),     tkn("reduce"),   launch_llap_reduce },
  { tkn("llap"),     tkn("assemble"), launch_llap_assemble },
  { tkn("llap"),     tkn("colour"),   launch_llap_colour },
  { tkn("hilite"),   tkn("half"),     launch_hilite_half },
  { tkn("hilite"),   tkn("reduce"),   launch_hilite_reduce },
  { tkn("hilite"),   tkn("assemble"), launch_hilite_assemble },
  { tkn("hilite"),   tkn("doub"),     launch_hilite_doub },
  { tkn("rawhist"),  tkn("collect"),  launch_rawhist_collect },
  { tkn("rawhist"),  tkn("map"),      launch_rawhist_map },
  { tkn("hist"),     tkn("collect"),  launch_hist_collect },
  { tkn("hist"),     tkn("map"),      launch_hist_map },
  { tkn("autoexp"),  tkn("hist"),     launch_autoexp_hist },
  { tkn("autoexp"),  tkn("lum"),      launch_autoexp_lum },
  { tkn("autoexp"),  tkn("exp"),      launch_autoexp_exp },
  { tkn("demosaic"), tkn("halfsize"), launch_demosaic_halfsize },
  { tkn("shared"),   tkn("resample"), launch_shared_resample },
  { tkn("colenc"),   tkn("main"),     launch_colenc_main },
  { tkn("demosaic"), tkn("rcd_conv"), launch_demosaic_rcd_conv },
  { tkn("demosaic"), tkn("rcd_fill"), launch_demosaic_rcd_fill },
  { tkn("i-mcraw"),  tkn("prefix"),   launch_mcraw_prefix },
  { tkn("i-mcraw"),  tkn("decode"),   launch_mcraw_decode },
};

static launch_fn find(uint64_t module, uint64_t kernel)
{
  for(const reg_entry &e : g_registry)
    if(e.module == module && e.kernel == kernel) return e.fn;
  return nullptr;
}

extern "C" {

int vkdt_hip_abi_version(void) { return VKDT_HIP_ABI_VERSION; }

int vkdt_hip_init(int dev)
{
  int n = 0;
  if(int e = hip_fail(hipGetDeviceCount(&n), "hipGetDeviceCount")) return e;
  if(dev < 0 || dev >= n) return vkdt_set_error(VKDT_HIP_DEVICE, "device %d of %d", dev, n);
  return hip_fail(hipSetDevice(dev), "hipSetDevice");
}

int vkdt_hip_node_supported(uint64_t module, uint64_t kernel) { return find(module, kernel) != nullptr; }

int vkdt_hip_node_launch(uint64_t module, uint64_t kernel, const void *pc, size_t pc_size,
                         const void *params, size_t params_size, const vkdt_buf_t *conn, int nconn,
                         uint32_t wd, uint32_t ht, uint32_t dp, void *stream)
{
  launch_fn fn = find(module, kernel);
  if(!fn)
  {
    char m[9] = {0}, k[9] = {0};
    memcpy(m, &module, 8);
    memcpy(k, &kernel, 8);
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "no HIP kernel for %s/%s", m, k);
  }
  if(nconn < 0 || (nconn && !conn)) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "bad connector list");
  if(wd == 0 || ht == 0) return VKDT_HIP_OK;   // empty dispatch
  // node extents: images up to 2^20 texels a side; one-row nodes up to 2^28
  // (i-mcraw/decode's is blocks x 8, a million for a 4096 x 2160 frame)
  if(wd > (ht == 1 ? (1u << 28) : (1u << 20)) || ht > (1u << 20) || dp > 4096)
    return vkdt_set_error(VKDT_HIP_INCOMPLETE, "dispatch %ux%ux%u too large", wd, ht, dp);
  node_call c = { pc, pc_size, params, params_size, conn, nconn, wd, ht, dp ? dp : 1, (hipStream_t)stream };
  return fn(c);
}

int vkdt_hip_alloc(void **ptr, size_t bytes) { return hip_fail(hipMalloc(ptr, bytes ? bytes : 16), "hipMalloc"); }
int vkdt_hip_free(void *ptr) { return hip_fail(hipFree(ptr), "hipFree"); }
int vkdt_hip_host_alloc(void **ptr, size_t bytes) { return hip_fail(hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault), "hipHostMalloc"); }
int vkdt_hip_host_free(void *ptr) { return hip_fail(hipHostFree(ptr), "hipHostFree"); }

int vkdt_hip_upload(void *dst_dev, const void *src_host, size_t bytes, void *stream)
{
  return hip_fail(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "upload");
}

int vkdt_hip_download(void *dst_host, const void *src_dev, size_t bytes, void *stream)
{
  return hip_fail(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "download");
}

// zero fill: 16-byte stores over the aligned body, 2-byte stores for the rest
__global__ void __launch_bounds__(256) k_zero(uint4 *body, size_t n16, uint16_t *tail, size_t n2)
{
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) body[i] = make_uint4(0, 0, 0, 0);
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) tail[i] = 0;
}

int vkdt_hip_buf_clear(const vkdt_buf_t *buf, void *stream)
{
  if(!buf || !buf->ptr) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "clear: no buffer");
  const size_t bytes = buf->plane_stride * (buf->array_length ? buf->array_length : 1);
  if(bytes % 2) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "clear: odd byte count %zu", bytes);
  const int aligned = ((uintptr_t)buf->ptr % 16) == 0;
  const size_t n16 = aligned ? bytes / 16 : 0;
  uint16_t *tail = (uint16_t *)((uint8_t *)buf->ptr + n16 * 16);
  const size_t n2 = (bytes - n16 * 16) / 2;
  const size_t n = n16 > n2 ? n16 : n2;
  const unsigned blocks = (unsigned)(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (uint4 *)buf->ptr, n16, tail, n2);
  return launch_post("clear");
}

int vkdt_hip_stream_create(void **stream)
{
  hipStream_t s;
  int e = hip_fail(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = e ? nullptr : (void *)s;
  return e;
}
int vkdt_hip_stream_destroy(void *stream) { return hip_fail(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy"); }
int vkdt_hip_stream_sync(void *stream) { return hip_fail(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize"); }

int vkdt_hip_event_create(void **event)
{
  hipEvent_t e;
  int r = hip_fail(hipEventCreate(&e), "hipEventCreate");
  *event = r ? nullptr : (void *)e;
  return r;
}
int vkdt_hip_event_destroy(void *event) { return hip_fail(hipEventDestroy((hipEvent_t)event), "hipEventDestroy"); }
int vkdt_hip_event_record(void *event, void *stream) { return hip_fail(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord"); }
int vkdt_hip_event_elapsed_ms(float *ms, void *b, void *e) { return hip_fail(hipEventElapsedTime(ms, (hipEvent_t)b, (hipEvent_t)e), "hipEventElapsedTime"); }
int vkdt_hip_event_query(void *event)
{
  const hipError_t e = hipEventQuery((hipEvent_t)event);
  if(e == hipSuccess) return 1;
  if(e == hipErrorNotReady) return 0;
  return hip_fail(e, "hipEventQuery");
}

// stream capture of a run's launches into a replayable graph
int vkdt_hip_capture_begin(void *stream)
{
  return hip_fail(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

int vkdt_hip_capture_end(void *stream, void **graph_exec)
{
  hipGraph_t gr = nullptr;
  *graph_exec = nullptr;
  if(int e = hip_fail(hipStreamEndCapture((hipStream_t)stream, &gr), "hipStreamEndCapture")) return e;
  hipGraphExec_t ex = nullptr;
  const int e = hip_fail(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(gr);
  if(!e) *graph_exec = (void *)ex;
  return e;
}

int vkdt_hip_graph_launch(void *graph_exec, void *stream)
{
  return hip_fail(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
}

int vkdt_hip_graph_destroy(void *graph_exec)
{
  return graph_exec ? hip_fail(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy") : 0;
}

const char *vkdt_hip_last_error(void) { return g_err; }

} // extern "C"
