/*
 * nprof: fits the raw noise model  variance = a + b * brightness  to the
 * statistics rawhist collected and writes it as "<maker>-<model>-<iso>.nprof"
 * ("a b"), the file i-raw reads when its noise parameters are zero.
 *
 * The fit is the one of src/pipe/modules/nprof/main.c:84-140: bins are
 * screened by pairs (a bin is kept once enough later bins give a line with
 * positive slope and intercept through it, with the demands lowered up to
 * seven times), then the kept bins enter a count-weighted least-squares
 * line.  The statistics arrive here as FP64 sums (the "atom" connector).
 * GUI callbacks (test/install/remove buttons) are not part of this path.
 */
#include "pipe/api.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void modify_roi_in(dt_graph_t *graph, dt_module_t *module)
{
  (void)graph;
  dt_roi_t *r = &module->connector[0].roi;
  r->wd = r->full_wd;
  r->ht = r->full_ht;
  r->marker = s_roi_mark_hard_bck;
}

// brightness x and variance y of bin i: first moment and central second
// moment of the bin's block values
static void bin_stats(const double *h, int wd, int i, double *c, double *x, double *y)
{
  *c = h[i];
  const double m1 = h[wd + i], m2 = h[2 * wd + i];
  *x = m1 / *c;
  *y = m2 / *c - (m1 / *c) * (m1 / *c);
}

// the fit on accumulators h (wd bins x 3 rows); returns the number of bins kept
int nprof_fit(const double *h, int wd, float *a_out, float *b_out)
{
  int *keep = malloc(sizeof(int) * (wd > 0 ? wd : 1));
  int nkeep = 0;
  float a = 0.0f, b = 0.0f;
  for(int f = 0; f < 7 && !nkeep; f++)
  { // demands: at least fk*64 samples per bin, more than fk*32 agreeing partners
    const double fk = 1.0 / pow(2, f);
    for(int i = 0; i < wd; i++)
    {
      int score = 0;
      for(int j = i + 1; j < wd * (0.8 - fk); j++)
      {
        double c1, x1, y1, c2, x2, y2;
        bin_stats(h, wd, i, &c1, &x1, &y1);
        if(c1 < fk * 64) continue;
        bin_stats(h, wd, j, &c2, &x2, &y2);
        if(c2 < fk * 64) continue;
        if(y1 <= 0 || y2 <= 0) continue;
        const double eb = (y2 - y1) / (x2 - x1);
        const double ea = y1 - x1 * eb;
        if(!(eb > 0.0) || !(ea > 0.0)) continue;
        if(++score > fk * 32)
        {
          a = ea;
          b = eb;
          keep[nkeep++] = i;
          break;
        }
      }
    }
    if(!nkeep) fprintf(stderr, "[nprof] lowering the demands %dx: too few valid bins\n", f + 1);
  }
  double sc = 0.0, sx = 0.0, sx2 = 0.0, sy = 0.0, sxy = 0.0;
  for(int k = 0; k < nkeep; k++)
  { // (the variance here is evaluated as m2/c - m1/c*m1/c, left to right)
    const int i = keep[k];
    const double c = h[i], m1 = h[wd + i], m2 = h[2 * wd + i];
    const double x = m1 / c;
    double y = m2 / c - m1 / c * m1 / c;
    y = 0.0 > y ? 0.0 : y;                   // MAX(0, y): NaN passes
    sc += c;
    sx += x * c;
    sx2 += x * x * c;
    sy += y * c;
    sxy += x * y * c;
  }
  const float denom = sc * sx2 - sx * sx;   // single precision, as the reference
  if(fabsf(denom) > 1e-10f)
  {
    const double bb = (sc * sxy - sx * sy) / denom, aa = (sy * sx2 - sx * sxy) / denom;
    b = 0.0 > bb ? 0.0f : (float)bb;   // NaN passes, as MAX(0, NaN)
    a = 0.0 > aa ? 0.0f : (float)aa;
  }
  if(nkeep <= 0)
  { // no usable bin: the reference's placeholder model
    fprintf(stderr, "[nprof] no valid bin\n");
    a = 100.0f;
    b = 1.0f;
  }
  free(keep);
  *a_out = a;
  *b_out = b;
  return nkeep;
}

void write_sink(dt_module_t *module, void *buf, dt_write_sink_params_t *p)
{
  (void)p;
  const int wd = module->connector[0].roi.wd;
  float a, b;
  nprof_fit((const double *)buf, wd, &a, &b);
  char filename[512];
  if(module->img_param.maker[0])
    snprintf(filename, sizeof(filename), "%s-%s-%d.nprof", module->img_param.maker, module->img_param.model,
             (int)module->img_param.iso);
  else
    snprintf(filename, sizeof(filename), "unknown-%s-%d.nprof", module->img_param.model, (int)module->img_param.iso);
  fprintf(stdout, "[nprof] writing '%s'\n", filename);
  FILE *f = fopen(filename, "wb");
  if(f)
  {
    fprintf(f, "%g %g\n", a, b);
    fclose(f);
  }
}
