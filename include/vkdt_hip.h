/*
 * vkdt_hip.h -- the C ABI between vkdt's host-side graph runtime and the
 * hand-written HIP kernels of the raw-development path on MI355X (gfx950).
 *
 * Plain C, plain pointers and sizes.  The host side (libvkdt_pipe.so and the
 * module DSOs) never sees a HIP or torch type: streams travel as void*.
 *
 * In the reference this boundary is the Vulkan back end of the graph
 * runtime; each entry point below names the reference interface it takes
 * over.  Errors map onto the three outcomes the reference distinguishes
 * (modules/api.h VkResult use, graph.c:776-793):
 *   VKDT_HIP_OK          success
 *   VKDT_HIP_INCOMPLETE  bad node/graph configuration (unknown kernel,
 *                        wrong connector count, format or extent)
 *   VKDT_HIP_OOM         device allocation failed
 *   VKDT_HIP_DEVICE      any other HIP runtime error
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VKDT_HIP_ABI_VERSION 1

enum
{
  VKDT_HIP_OK         =  0,
  VKDT_HIP_INCOMPLETE = -1,
  VKDT_HIP_OOM        = -2,
  VKDT_HIP_DEVICE     = -3,
};

// buffer formats, as the connector "format" tokens of the reference
// (graph-run-nodes-allocate.h:39-96): ui16 = R16_UNORM, f16 = R16F/RGBA16F,
// f32 = R32F/RGBA32F.  atom: the accumulators that kernels add into with
// atomics (the reference's R32F float atomics, :53), held here as FP64 --
// 8 bytes per channel -- so per-frame sums do not lose precision.  ui32:
// R32_UINT counters (histograms the reference adds into with imageAtomicAdd).
// ui8: R8G8B8A8_UNORM, the 8-bit export images.
enum
{
  VKDT_FMT_UI16 = 1,
  VKDT_FMT_F16  = 2,
  VKDT_FMT_F32  = 3,
  VKDT_FMT_ATOM = 4,
  VKDT_FMT_UI32 = 5,
  VKDT_FMT_UI8  = 6,
};

// one connector's device buffer.  Replaces the descriptor-set image
// binding of a node connector (graph-run-nodes-allocate.h:628-940).
// Texels are stored row major, channels interleaved, rows `pitch` texels
// apart, planes of an array connector `plane_stride` bytes apart.
typedef struct vkdt_buf_t
{
  void     *ptr;           // device address (NULL: connector unused/dead store)
  uint32_t  wd, ht;        // extent in texels (imageSize/textureSize)
  uint32_t  pitch;         // texels per row (>= wd)
  uint32_t  chan;          // 1 or 4
  uint32_t  format;        // VKDT_FMT_*
  uint32_t  array_length;  // number of planes (1 for plain images)
  uint64_t  plane_stride;  // bytes between planes
}
vkdt_buf_t;

// ABI version of the loaded library (for loaders that dlsym by name).
int vkdt_hip_abi_version(void);

// Select device `dev` and initialise the kernel registry.  Takes over
// qvk_init (qvk/qvk.c:138) for the compute path: no instance, queues or
// samplers, only the device.
int vkdt_hip_init(int dev);

// Run one node.  Takes over one vkCmdDispatch of
// graph-run-nodes-record-cmd.h:372-380 (bind pipeline, push constants,
// dispatch ceil(wd/8) x ceil(ht/8) x dp): the (module, kernel) token pair
// picks the HIP kernel that replaces src/pipe/modules/<module>/<kernel>.comp,
// `pc` is the push-constant blob dt_node_add stored (modules/api.h:160-210),
// `params` the module's uniform block (committed_param if the module has
// one, else its params blob: graph-run-modules.h:6-31), and `conn` the
// node's connectors in declaration order.  Asynchronous on `stream`
// (a hipStream_t, NULL = default stream).
int vkdt_hip_node_launch(
    uint64_t          module,
    uint64_t          kernel,
    const void       *pc,
    size_t            pc_size,
    const void       *params,
    size_t            params_size,
    const vkdt_buf_t *conn,
    int               nconn,
    uint32_t          wd,
    uint32_t          ht,
    uint32_t          dp,
    void             *stream);

// 1 if (module, kernel) has a HIP implementation, else 0.
int vkdt_hip_node_supported(uint64_t module, uint64_t kernel);

// Device memory for the graph's arena.  Takes over the vkAllocateMemory
// of the image pool (graph-run-nodes-allocate.h:17-37, alloc.c).
int vkdt_hip_alloc(void **ptr, size_t bytes);
int vkdt_hip_free(void *ptr);

// Pinned host staging.  Takes over the mapped staging buffers that
// read_source fills and write_sink reads (graph-run-nodes-upload.h:23-80,
// graph-run-nodes-download.h:3-32).
int vkdt_hip_host_alloc(void **ptr, size_t bytes);
int vkdt_hip_host_free(void *ptr);

// Zero every plane of `buf` on `stream` (a fill kernel, so the call may be
// captured into a hipGraph).  Takes over the clear of connectors flagged
// s_conn_clear before their node runs (graph-run-nodes-record-cmd.h:100-175).
int vkdt_hip_buf_clear(const vkdt_buf_t *buf, void *stream);

// Asynchronous copies on `stream`.  Take over the staging<->image copies
// recorded in graph-run-nodes-record-cmd.h:255-320.
int vkdt_hip_upload(void *dst_dev, const void *src_host, size_t bytes, void *stream);
int vkdt_hip_download(void *dst_host, const void *src_dev, size_t bytes, void *stream);

// Streams and timing.  Take over the command buffer submit + timeline
// semaphore wait (graph.c:842-874) and the timestamp queries of -d perf
// (graph.c:96-112, 881-932).
int vkdt_hip_stream_create(void **stream);
int vkdt_hip_stream_destroy(void *stream);
int vkdt_hip_stream_sync(void *stream);
int vkdt_hip_event_create(void **event);
int vkdt_hip_event_destroy(void *event);
int vkdt_hip_event_record(void *event, void *stream);
int vkdt_hip_event_elapsed_ms(float *ms, void *ev_begin, void *ev_end);
// Non-blocking completion test: 1 when the work recorded before `event`
// has finished, 0 while it runs, < 0 on error.  No reference counterpart
// (the reference waits on its timeline semaphore, graph.c:842-874); the
// serving loop polls it to retire batches while it admits new requests.
int vkdt_hip_event_query(void *event);

// Stream capture: the launches recorded on `stream` between begin and end
// become one replayable graph (hipGraph), launched again with
// vkdt_hip_graph_launch.  Takes over the reuse of a recorded Vulkan command
// buffer when nothing changed between runs (graph.c:719-800, the
// s_graph_run_record_cmd_buf flag): the host runtime records a run's node
// launches once and replays them while the buffers and the push-constant
// and uniform blobs stay the same.  Capture records kernels only (buffer
// clears are fill kernels).
int vkdt_hip_capture_begin(void *stream);
int vkdt_hip_capture_end(void *stream, void **graph_exec);
int vkdt_hip_graph_launch(void *graph_exec, void *stream);
int vkdt_hip_graph_destroy(void *graph_exec);

// Human readable text of the last error on this thread.
const char *vkdt_hip_last_error(void);

// ---- lossless JPEG raw data (compressed DNG, Compression = 7) -------------
// One stream per DNG tile or strip (ITU-T T.81 process 14).  The reference
// decodes these on the host inside rawler (i-raw/rawloader-c, a dependency
// outside its tree) and liblj92 (i-mlv/liblj92/lj92.c).  The scan's samples
// fill the tile row-major; sites outside the image are dropped.
typedef struct vkdt_lj92_tile_t
{
  uint64_t offset, length;    // the stream's bytes in the data buffer
  uint32_t x0, y0;            // the tile's first site in the image
  uint32_t tile_wd, tile_ht;  // its extent (TileWidth/TileLength, or width x RowsPerStrip)
  uint32_t image;             // index into the image list (GPU decode)
}
vkdt_lj92_tile_t;

typedef struct vkdt_lj92_image_t
{
  void    *out;               // device: uint16 mosaic, the window's first site
  uint32_t wd, ht, pitch;     // window extent and texels per row of out
  uint32_t x0, y0;            // the window's first site in the tiles' image (a crop; 0, 0 for all of it)
}
vkdt_lj92_image_t;

// Frame of one stream: samples per line per component, lines, components, bits.
int vkdt_lj92_info(const uint8_t *data, size_t length, uint32_t *wd, uint32_t *ht, uint32_t *comp, uint32_t *bits);

// Decode one tile on the host into `out` (wd x ht, `pitch` texels per row):
// i-raw's path, taking over the decode inside rawler::decode_file (rawloader-c/lib.rs:102-124).
int vkdt_lj92_decode_host(const uint8_t *data, const vkdt_lj92_tile_t *tile, uint16_t *out,
                          uint32_t wd, uint32_t ht, uint32_t pitch);

// Decode n tiles of nimg images on the GPU into their windows: the host
// parses the headers and builds the code tables, uploads the compressed
// bytes (`data`, data_len) and waits for the decode on `stream`.  Streams
// with one table and predictor 1 are decoded by many lanes each, the others
// by one lane each.  No reference counterpart (the reference unpacks on the
// host); i-raw calls it from read_source_hip.
int vkdt_hip_lj92_decode(const uint8_t *data, size_t data_len, const vkdt_lj92_tile_t *tiles, int n,
                         const vkdt_lj92_image_t *img, int nimg, void *stream);

// Unpack n samples of `bits` bits (1..16) from a bit-packed frame on the
// host (`packed`, bytes) into a uint16 device buffer: sample i is bits
// [i*bits, (i+1)*bits) of the stream of 16-bit little-endian words, each
// word read from its most significant bit (Magic Lantern raw video frames).
// Replaces the host loop of i-mlv/video_mlv.c:262-276 (mlv_get_frame); the
// bytes are uploaded as stored and unpacked on `stream`, which it waits for.
int vkdt_hip_unpack_msb16(const uint8_t *packed, size_t bytes, void *out, uint32_t n, int bits, void *stream);

// Baseline JPEG of an 8-bit rgba image on the device (`pitch` pixels a row):
// the o-jpg sink's encoder.  Replaces the libjpeg calls of
// o-jpg/main.c:101-160 (jpeg_set_quality, sampling, optimised Huffman
// codes, 300 dpi JFIF) with the same published steps on the GPU: the file
// holds the tables and entropy-coded data libjpeg writes for the same image
// and settings.  hs x vs luma sampling: 1x1 (4:4:4), 2x1 (4:2:2) or 2x2
// (4:2:0).  *out (malloc'd, free with vkdt_jpg_free) gets the whole file.
int vkdt_hip_jpg_encode(const void *rgba, uint32_t wd, uint32_t ht, uint32_t pitch, int quality, int hs, int vs,
                        uint8_t **out, size_t *out_len, void *stream);
// vkdt_hip_jpg_encode with the rest of o-jpg/main.c:142-156: libjpeg's
// dct_method (VKDT_JPG_DCT_*, JDCT_ISLOW / JDCT_IFAST / JDCT_FLOAT),
// smoothing_factor (0..100; full-size components and 2x2 chroma) and an ICC
// profile written in APP2 markers after the JFIF header (write_icc_profile,
// main.c:36-93; none when icc_len is 0)
#define VKDT_JPG_DCT_ISLOW 0
#define VKDT_JPG_DCT_IFAST 1
#define VKDT_JPG_DCT_FLOAT 2
int vkdt_hip_jpg_encode_ex(const void *rgba, uint32_t wd, uint32_t ht, uint32_t pitch, int quality, int hs, int vs,
                           int dct, int smoothing, const uint8_t *icc, size_t icc_len, uint8_t **out, size_t *out_len,
                           void *stream);
void vkdt_jpg_free(uint8_t *p);
// IJG quality scaling of the T.81 Annex K tables (natural order), as
// jpeg_set_quality(quality, force_baseline) computes them
void vkdt_jpg_quant_tables(int quality, uint8_t luma[64], uint8_t chroma[64]);

// ---- batched develop ------------------------------------------------------
// The raw-development graph i-raw -> denoise -> [hilite ->] demosaic ->
// colour -> filmcurv -> llap of many stills or frames, of any sizes and CFAs, in one
// launch per stage.  No reference counterpart: in the reference each image
// is its own graph run (graph.c:719 dt_graph_run per export), which is what
// vkdt_hip_node_launch replaces one dispatch at a time.  The batch runs the
// same arithmetic as that per-node graph (stage outputs rounded to f16
// where the graph stores f16), so every image's result equals its plain
// graph run.
//
// One image of a batch: the values the per-node graph's nodes receive.
// dt_graph_develop_image() (vkdt_amd/pipe/api.h) fills every field but the
// two device pointers from a planned graph (its push constants, params and
// node extents).
typedef struct vkdt_develop_image_t
{
  const void *cfa;            // device: the ui16 mosaic i-raw uploads
  uint32_t cfa_wd, cfa_ht;    // its extent
  uint32_t cfa_pitch;         // texels per row
  void    *out;               // device: rgba f32, wd x ht, tightly packed (llap/colour output)
  uint32_t wd, ht;            // developed extent (denoise output: the crop box)
  uint32_t filters;           // 9: X-Trans; other non-zero: Bayer RGGB
  float    wb[4], black[4], white[4];   // denoise push constants (black/white in UNORM)
  int32_t  crop[4];
  float    noise_a, noise_b;
  float    dn_params[8];      // denoise/params blob: strength, luma, detail, (pad), edges[4]
  int32_t  dm_fixup;          // demosaic/params word 0
  float    co_params[242];    // colour's committed uniform block
  float    fc_params[4];      // filmcurv/params: light, contrast, bias, colour
  float    ll_params[4];      // llap/params: sigma, shadows, hilights, clarity
  int32_t  ll_numgamma;       // curve planes G-1 (llap/curve push constant)
  int32_t  ll_levels;         // pyramid levels nl (reduce nodes + 1)
  int32_t  hl_on;             // 1: hilite runs between denoise and demosaic
  float    hl_wb[4];          // hilite push constants: white balance
  float    hl_params[3];      // hilite/params: white, desat, soft
  int32_t  hl_levels;         // hilite reduce nodes (coarse levels of the block grid)
}
vkdt_develop_image_t;

typedef struct vkdt_batch_t vkdt_batch_t;

// Device scratch a batch of these images needs: their intermediates and the
// batch's own tables (the caller owns the mosaics and the outputs).
int vkdt_hip_batch_scratch_bytes(const vkdt_develop_image_t *img, int n, size_t *bytes);

// Plan a batch: validate the images, lay out their intermediates in
// `scratch` (at least vkdt_hip_batch_scratch_bytes, 256-byte aligned) and
// build the per-stage tile lists.  `tune` selects the launch shape per
// stage group: NULL takes the built-in choice, else VKDT_BATCH_STAGES
// entries as written by vkdt_hip_batch_autotune, each
//   persistent work groups per CU (0: one tile per work group)
//   | rows per thread of the per-pixel stages (1..8) << 8.
#define VKDT_BATCH_STAGES 10
int vkdt_hip_batch_create(vkdt_batch_t **batch, const vkdt_develop_image_t *img, int n,
                          void *scratch, size_t scratch_bytes, const int32_t *tune);

// Enqueue the whole batch on `stream`.  The first run (and the first after
// a retune) also uploads the batch's tables into its scratch, in stream
// order.  The first run with `graph` != 0 captures the launch sequence into
// a hipGraph; later runs replay it.
int vkdt_hip_batch_run(vkdt_batch_t *batch, void *stream, int graph);

// Time each stage's launch shapes on this batch and keep the fastest
// (written to `tune`, VKDT_BATCH_STAGES entries, when not NULL).
int vkdt_hip_batch_autotune(vkdt_batch_t *batch, void *stream, int32_t *tune);

// Per-stage GPU time of the last vkdt_hip_batch_run (ms, VKDT_BATCH_STAGES
// entries, when timing was on) and the stage names.
int vkdt_hip_batch_set_timing(vkdt_batch_t *batch, int on);
int vkdt_hip_batch_stage_ms(vkdt_batch_t *batch, float *ms);
const char *vkdt_hip_batch_stage_name(int stage);

// Develop a batch without allocating: no counterpart in the reference.  For
// callers that cannot allocate device memory (code under hipGraph capture
// elsewhere in the process, serving loops and harnesses that forbid
// hipMalloc).  The intermediates live in the library's static device
// scratch of VKDT_BATCH_STATIC_BYTES per GPU (held until the process ends)
// instead of a caller workspace; images whose scratch
// (vkdt_hip_batch_scratch_bytes) exceeds it are refused with an error.  The
// caller still passes the mosaics and outputs.  The batch runs eagerly with
// the built-in launch shape and the call returns when `stream` has drained.
// Calls that overlap on one GPU share its scratch and race.
#define VKDT_BATCH_STATIC_BYTES (1ull << 30)
int vkdt_hip_batch_develop_static(const vkdt_develop_image_t *img, int n, void *stream);

// Release the host side of a batch (after its work has completed; the
// scratch stays the caller's).
void vkdt_hip_batch_destroy(vkdt_batch_t *batch);

#ifdef __cplusplus
}
#endif
