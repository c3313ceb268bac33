/*
 * o-pfm: sink writing the rgb channels of an rgba f32 buffer as a
 * little-endian PFM.  Byte layout as src/pipe/modules/o-pfm/main.c:8-42:
 * "PF\n<w> <h>\n-1.0", '0' padding so the header plus its final newline
 * end on a 16-byte boundary, then rows top to bottom.
 */
#include "pipe/api.h"

#include <stdlib.h>

void write_sink(dt_module_t *module, void *buf, dt_write_sink_params_t *p)
{
  (void)p;
  const char *basename = dt_module_param_string(module, 0);
  const int width = module->connector[0].roi.wd, height = module->connector[0].roi.ht;
  char filename[512];
  snprintf(filename, sizeof(filename), "%s.pfm", basename);
  FILE *f = fopen(filename, "wb");
  if(!f)
  {
    fprintf(stderr, "[o-pfm] cannot write '%s'\n", filename);
    return;
  }
  char header[128];
  int len = snprintf(header, sizeof(header), "PF\n%d %d\n-1.0", width, height);
  while((len + 1) & 0xf) header[len++] = '0';
  header[len++] = '\n';
  fwrite(header, 1, len, f);
  const float *pf = buf;
  const size_t n = (size_t)width * height;
  float *row = malloc(sizeof(float) * 3 * (size_t)width);
  for(size_t k0 = 0; row && k0 < n; k0 += width)
  {
    for(int i = 0; i < width; i++)
      for(int c = 0; c < 3; c++) row[3 * i + c] = pf[4 * (k0 + i) + c];
    fwrite(row, sizeof(float), 3 * (size_t)width, f);
  }
  free(row);
  fclose(f);
}
