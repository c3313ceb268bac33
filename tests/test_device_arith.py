"""Arithmetic identities the HIP kernels rely on in place of a slower written
form, checked exhaustively on the CPU with the same IEEE binary32 operations
(round to nearest even; fmaf is the single-rounding fused multiply-add the
GPU's v_fma_f32 performs).

- R16_UNORM loads (common.h ld_u16): the reference reads UNORM16 textures as
  v/65535.  The kernels form q = v*r with r the rounded 1/65535 and correct it
  once, fmaf(fmaf(-q, 65535, v), r, q); that equals the correctly rounded
  quotient for every one of the 65536 codes, while the bare product differs
  on some (so the correction step is needed)."""
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
int main(void)
{
  const float r = 1.0f / 65535.0f;
  int bad = 0, bare = 0;
  for(int v = 0; v < 65536; v++)
  {
    const float x = (float)v, q = x / 65535.0f;
    const float q0 = x * r;
    const float q1 = fmaf(fmaf(-q0, 65535.0f, x), r, q0);
    bad += q1 != q;
    bare += q0 != q;
  }
  printf("%d %d\n", bad, bare);
  return 0;
}
"""


def test_unorm16_corrected_product_equals_division(tmp_path):
    c = tmp_path / "u16.c"
    c.write_text(SRC)
    exe = tmp_path / "u16"
    try:
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True,
                       capture_output=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"no C compiler: {e}")
    bad, bare = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert bad == 0
    assert bare > 0
