/*
 * dng.h -- reader for uncompressed CFA DNGs: the raw input of the
 * raw-development path.  Produces what i-raw's raw loader hands to the
 * module (i-raw/rawloader-c/rawloader.h rawimage_t) for this file type.
 */
#pragma once
#include <stdint.h>
#include <stddef.h>

#include "vkdt_hip.h"

typedef struct dng_image_t
{
  char     maker[32], model[32];
  uint64_t width, height;     // raw buffer extent (before CFA offset/rounding)
  uint32_t stride;            // texels per row of data
  uint32_t cpp;               // components per pixel (1 for CFA)
  float    wb_coeffs[4];      // 1/AsShotNeutral
  float    whitelevels[4];
  float    blacklevels[4];
  float    xyz_to_cam[4][3];  // colour matrix of the chosen illuminant
  uint32_t illuminant;        // 0 A, 1 D50, 2 D55, 3 D60, 4 D65, 5 D75, 6 D93, 666 unknown
  uint32_t cfa_w, cfa_h;      // CFA repeat dims
  uint8_t  cfa[36];           // 0 red, 1 green, 2 blue
  uint64_t crop_aabb[4];      // x y X Y
  float    iso;
  uint16_t *data;             // height x stride samples, owned (0 for a lossless-JPEG raw until decoded)
  // lossless-JPEG raws: the file's bytes and one stream per tile or strip,
  // decoded on the GPU into the module's buffer (read_source_hip) or on the
  // host on demand (dng_decode_host)
  uint8_t  *packed;
  size_t    packed_len;
  vkdt_lj92_tile_t *tiles;
  uint32_t  ntiles;
  // OpcodeList2 of the raw IFD as stored (big-endian records), or 0
  uint8_t  *oplist2;
  uint32_t  oplist2_len;
}
dng_image_t;

// returns 0 on success; on failure `err` holds a message
int  dng_load(const char *filename, dng_image_t *img, char *err, size_t errlen);
void dng_free(dng_image_t *img);
// fills data from the lossless-JPEG streams (no-op when data is there)
int  dng_decode_host(dng_image_t *img, char *err, size_t errlen);

// decodes a stored opcode list into pipe/dng_opcode.h structs (GainMap
// parameters decoded, other opcodes' data left 0).  Returns 0 for an empty
// list, or when the list's length or any opcode's parameter size is
// inconsistent (the whole list is then ignored).
struct dt_dng_opcode_list_t;
struct dt_dng_opcode_list_t *dng_opcode_list_decode(const uint8_t *data, size_t len);
void dng_opcode_list_free(struct dt_dng_opcode_list_t *list);
