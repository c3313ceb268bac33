/*
 * mlv.h -- reader for Magic Lantern raw video (MLV v2.0): the raw-video
 * input of the raw-development path (reference i-mlv/video_mlv.c, mlv.h).
 *
 * An MLV clip is a chain of blocks, each starting with a four-letter type, its
 * size and a timestamp; the first block of every chunk file is the file
 * header (MLVI).  The on-disk layouts below are the format's (little endian,
 * packed), restated for the blocks this reader uses.  Frames (VIDF) carry one
 * raw image each: bit-packed samples (16-bit little-endian words filled from
 * the most significant bit) or, with the LJ92 class flag, one lossless-JPEG
 * stream.
 */
#pragma once
#include <stdint.h>
#include <stdio.h>

#include "vkdt_hip.h"

#define MLV_CLASS_FLAG_LJ92 0x20

#pragma pack(push, 1)
typedef struct { uint8_t type[4]; uint32_t size; uint64_t timestamp; } mlv_block_t;

typedef struct
{
  uint8_t  magic[4];          // "MLVI"
  uint32_t size;
  uint8_t  version[8];
  uint64_t guid;
  uint16_t file_num, file_count;
  uint32_t file_flags;
  uint16_t video_class, audio_class;
  uint32_t video_frames, audio_frames;
  uint32_t fps_nom, fps_denom;
}
mlv_file_t;

typedef struct
{
  mlv_block_t b;              // "VIDF"
  uint32_t frame_number;
  uint16_t crop_x, crop_y, pan_x, pan_y;
  uint32_t frame_space;       // padding before the frame's data
}
mlv_vidf_t;

typedef struct
{
  mlv_block_t b;              // "AUDF"
  uint32_t frame_number;
  uint32_t frame_space;       // padding before the audio data
}
mlv_audf_t;

typedef struct
{
  uint32_t api_version, buffer;
  int32_t  height, width, pitch, frame_size, bits_per_pixel, black_level, white_level;
  int32_t  crop[4];
  int32_t  active_area[4];
  int32_t  exposure_bias[2];
  int32_t  cfa_pattern, calibration_illuminant1;
  int32_t  color_matrix1[18];
  int32_t  dynamic_range;
}
mlv_raw_info_t;

typedef struct { mlv_block_t b; uint16_t xres, yres; mlv_raw_info_t raw; } mlv_rawi_t;   // "RAWI"
typedef struct { mlv_block_t b; uint32_t iso_mode, iso, iso_analog, digital_gain; uint64_t shutter_us; } mlv_expo_t;
typedef struct
{
  mlv_block_t b;              // "LENS"
  uint16_t focal_length, focal_dist, aperture;
  uint8_t  stabilizer, autofocus;
  uint32_t flags, lens_id;
  uint8_t  name[32], serial[32];
}
mlv_lens_t;
typedef struct { mlv_block_t b; uint16_t sec, min, hour, mday, mon, year, wday, yday, isdst, gmtoff; uint8_t zone[8]; } mlv_rtci_t;
typedef struct { mlv_block_t b; uint8_t camera_name[32]; uint32_t camera_model; uint8_t camera_serial[32]; } mlv_idnt_t;
typedef struct
{
  mlv_block_t b;              // "WAVI"
  uint16_t format, channels;
  uint32_t sampling_rate, bytes_per_second;
  uint16_t block_align, bits_per_sample;
}
mlv_wavi_t;
#pragma pack(pop)

typedef struct
{
  int      chunk;             // file of the chunk list
  uint64_t offset, size;      // the frame's data
  uint64_t block;             // its VIDF block
  uint64_t timestamp;
}
mlv_frame_t;

typedef struct
{
  FILE      **file;           // chunks: name.MLV, name.M00, name.M01, ...
  int         files;
  mlv_file_t  hdr;            // of the first chunk
  mlv_rawi_t  rawi;
  mlv_expo_t  expo;
  mlv_lens_t  lens;
  mlv_rtci_t  rtci;
  mlv_idnt_t  idnt;
  mlv_wavi_t  wavi;
  mlv_frame_t *frame;         // in timestamp order
  uint32_t    frames;
  double      frame_rate;
  mlv_frame_t *audf;          // audio blocks, in timestamp order
  uint32_t    audfs;
  uint8_t    *audio;          // the clip's audio synced to its first frame (or 0)
  uint64_t    audio_size;     // bytes that play with the clip's frames
  uint64_t    audio_buffer_size;
}
mlv_clip_t;

// 0 on success, else a message in err
int  mlv_open(mlv_clip_t *clip, const char *filename, char *err, size_t errlen);
void mlv_close(mlv_clip_t *clip);
// the frame's bytes as stored (packed or one LJ92 stream); caller frees
uint8_t *mlv_frame_data(mlv_clip_t *clip, uint32_t frame, uint64_t *size);
// host unpack of a bit-packed frame (xres x yres samples of bits_per_pixel)
void mlv_unpack(const uint8_t *packed, uint64_t size, uint16_t *out, uint64_t n, int bits);
