/*
 * dng_opcode.h -- DNG opcode lists as the raw input hands them to modules.
 *
 * A DNG's OpcodeList1/2/3 tags hold big-endian opcode records (DNG
 * specification 1.4, chapter 7): id, DNG version, flags, parameter size,
 * parameters.  i-raw decodes OpcodeList2 of the raw IFD into these structs
 * and attaches them to the image metadata (dt_image_metadata_dngop_t);
 * denoise applies the four per-CFA-site GainMap opcodes of Bayer raws.
 * Type and field names are those module sources use
 * (src/pipe/dng_opcode.h); the opcode ids are the specification's.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

typedef enum dt_dng_opcode_id_t
{
  s_dngop_warp_rectilinear        = 1,
  s_dngop_warp_fisheye            = 2,
  s_dngop_fix_vignette_radial     = 3,
  s_dngop_fix_bad_pixels_constant = 4,
  s_dngop_fix_bad_pixels_list     = 5,
  s_dngop_trim_bounds_t           = 6,
  s_dngop_map_table_t             = 7,
  s_dngop_map_polynomial_t        = 8,
  s_dngop_gain_map                = 9,
  s_dngop_warp_rectilinear_2      = 14,
}
dt_dng_opcode_id_t;

typedef struct dt_dng_opcode_t
{
  dt_dng_opcode_id_t id;
  uint32_t optional;       // flags bit 0
  uint32_t preview_skip;   // flags bit 1
  void    *data;           // dt_dng_gain_map_t for s_dngop_gain_map, else 0 (not decoded)
}
dt_dng_opcode_t;

typedef struct dt_dng_opcode_list_t
{
  int             count;
  dt_dng_opcode_t ops[];
}
dt_dng_opcode_list_t;

// the area an opcode applies to, in raw buffer coordinates
typedef struct dt_dng_region_t
{
  uint32_t top, left, bottom, right;
  uint32_t plane, planes;
  uint32_t row_pitch, col_pitch;
}
dt_dng_region_t;

// GainMap (opcode 9): a map_points_v x map_points_h grid of gains per plane
// over the region, at map_origin + i * map_spacing in relative image
// coordinates
typedef struct dt_dng_gain_map_t
{
  dt_dng_region_t region;
  uint32_t map_points_v, map_points_h;
  double   map_spacing_v, map_spacing_h;
  double   map_origin_v, map_origin_h;
  uint32_t map_planes;
  float    map_gain[];     // [v][h][plane]
}
dt_dng_gain_map_t;
