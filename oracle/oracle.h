/*
 * oracle.h -- CPU restatement of vkdt's raw-development node kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (vkdt_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * One source, three builds:
 *   liboracle64.so   arithmetic in double ("FP64 oracle")
 *   liboracle32.so   arithmetic in float  (same operation order as the GLSL)
 *   liboracle32p.so  the float build with every transcendental and square
 *                    root moved by one ulp by a hash of its argument: a
 *                    second float implementation for the error-bound tests
 * Buffers are always double in host memory; with storage mode on (the
 * default) every f16 image store is rounded to the nearest binary16 value
 * (round-to-nearest-even, overflow to inf), which is what the reference's
 * R16F/RGBA16F storage images do.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_UI16 = 1, OR_F16 = 2, OR_F32 = 3 };

typedef struct or_buf_t
{
  void    *data;          // uint16_t* for OR_UI16, double* otherwise
  int32_t  wd, ht;        // texel extent
  int32_t  chan;          // 1 or 4 (interleaved)
  int32_t  format;        // OR_UI16, OR_F16 or OR_F32
  int32_t  array_length;  // planes, >= 1, plane i at offset i*wd*ht*chan
}
or_buf_t;

// storage rounding of f16 images (1 = binary16 round trip on store, 0 = keep full precision)
void or_set_storage_f16(int on);
// 8 for the float build, 16 for the double build
int or_real_bytes(void);
// host threads of the per-pixel nodes' row loops (OpenMP)
void or_set_threads(int n);
int or_threads(void);

// near-tie flags of the next or_node calls: a node with a data-dependent
// branch sets mask[y*wd+x] = 1 for each output texel whose decision margin
// is at most eps times the magnitude of the decision's operands (mask 0: off)
void or_set_ties(uint8_t *mask, int wd, int ht, double eps);
// a near tie counts where its other branch moves the node's result by more
// than `effect` (absolute), or by more than one binary16 ulp with effect 0
void or_set_tie_effect(double effect);
// the same for or_develop, gathered over its nodes into one full-resolution
// mask (W x H): denoise flags dilated by r_dn half-resolution texels,
// demosaic flags by r_dm pixels (mask 0: off).  The effect threshold of
// or_set_tie_effect applies.
void or_set_develop_ties(uint8_t *mask, double eps, int r_dn, int r_dm);
// liboracle32p.so only: the seed of the hash that moves its transcendentals
void or_set_perturb_seed(uint32_t seed);

// run one node with the push-constant and params blobs exactly as the
// module's create_nodes/commit_params produced them (std140 layouts of the
// GLSL kernels).  returns 0 on success, <0 on unknown node or bad buffers.
int or_node(
    const char *module, const char *kernel,
    const void *pc, int pc_size,
    const void *params, int params_size,
    const or_buf_t *conn, int nconn,
    int wd, int ht, int dp);

// half precision rounding helper (binary16, RNE), exposed for tests
double or_f16_round(double x);
// mirrored-repeat texel index as the sampler computes it (exposed for tests)
int or_mirror(int i, int n);
// float -> R16_UNORM store conversion
uint16_t or_unorm16(double v);
// bilinear sample of a single-channel double image at texture-space (u, v)
double or_bilinear(const double *img, int wd, int ht, double u, double v);
// textureGather component 0 of a ui16 image: (i0j1, i1j1, i1j0, i0j0)
void or_gather(const uint16_t *img, int wd, int ht, double u, double v, double out[4]);
// 2x2 symmetric eigen decomposition as the reference computes it
void or_evd2x2(const double M[4], double eval[2], double evec0[2], double evec1[2]);

// nprof's noise-model fit (nprof/main.c:84-140) on rawhist accumulators h
// (wd bins x 3 rows: count, sum of means, sum of second moments); returns
// the number of bins that entered the line, a = intercept, b = slope
int or_nprof_fit(const double *h, int wd, float *a, float *b);

// whole-graph restatement of i-raw -> denoise -> demosaic -> colour ->
// filmcurv -> llap for one uncompressed CFA image (already block rounded,
// CFA offset normalised).  parameters follow the modules' params files.
typedef struct or_develop_t
{
  // image parameters as i-raw sets them
  int32_t filters;            // 9 = x-trans, anything else non-zero = bayer rggb
  float   black[4], white[4]; // raw units (0..65535)
  float   wb[4];              // camera white balance, g-normalised
  float   cam_to_rec2020[9];  // row major
  float   noise_a, noise_b;   // vkdt units (gaussian, poissonian)
  // denoise params
  float   dn_strength, dn_luma, dn_detail, dn_edges[4];
  // demosaic params
  int32_t dm_colour;
  // colour params
  float   co_exposure;
  float   co_white[3];        // 0,0,0 = use camera coefficients
  // filmcurv params
  float   fc_light, fc_contrast, fc_bias;
  int32_t fc_colour;
  // llap params
  float   ll_sigma, ll_shadows, ll_hilights, ll_clarity;
  int32_t ll_num_gamma;       // curve samples (10 in the reference)
  int32_t ll_max_levels;      // pyramid depth cap (12 in the reference)
  // hilite between denoise and demosaic (bin/default-darkroom.i-raw:17-18)
  int32_t hl_on;              // 0: denoise feeds demosaic directly
  float   hl_white, hl_desat, hl_soft;
}
or_develop_t;

// i-mcraw's GPU codec (i-mcraw/prefix.comp, decode.comp).  bits: 4-bit
// counts, two per byte, low nibble first; or_mcraw_prefix writes the
// inclusive prefix sum of the n blocks' data sizes; or_mcraw_block decodes
// one block at byte `off` of dat (len bytes, past them zeros) into its 64
// offsets from the reference, lane order (value 2s+h of register lane s);
// or_mcraw_decode writes the ui16 frame (wd x ht, the R16 UNORM store of
// (offset + ref)/65536) of an owd x oht coded frame shifted by (ox, oy).
void or_mcraw_prefix(const uint8_t *bits, uint32_t n, uint32_t *off);
void or_mcraw_block(const uint8_t *dat, size_t len, uint32_t off, uint32_t cnt, uint32_t out[64]);
// n blocks at once: block i at byte off[i] with count cnt[i] into out[64 i ..]
void or_mcraw_blocks(const uint8_t *dat, size_t len, const uint32_t *off, const uint32_t *cnt, uint32_t n,
                     uint32_t *out);
void or_mcraw_decode(const uint8_t *bits, const uint32_t *off, const uint16_t *ref, const uint8_t *dat, size_t len,
                     int ox, int oy, int owd, int oht, uint16_t *out, int wd, int ht);

// filmcurv's colour spaces, n triples each way (pins only): which = 0
// rec2020 -> oklab, 1 oklab -> rec2020, 2 xyY -> dt UCS JCh, 3 JCh -> xyY,
// 4 xy -> Munsell (hue, chroma, 0), 5 Munsell (hue, chroma) -> xy,
// 6 rec2020 -> xyY.  Returns -1 for another `which`.
int or_fc_space(int which, const double *in, double *out, int n);

// out_rgba: W*H*4 doubles (rgba f32 output of llap/colour).  stage buffers
// are optional (may be 0): denoise W*H*1, demosaic/colour/filmcurv W*H*4,
// llap W*H*1 (the assembled luminance that llap/colour transfers).
int or_develop(
    const uint16_t *cfa, int W, int H,
    const or_develop_t *p,
    double *out_rgba,
    double *st_denoise, double *st_demosaic,
    double *st_colour, double *st_filmcurv, double *st_llap);

#ifdef __cplusplus
}
#endif
