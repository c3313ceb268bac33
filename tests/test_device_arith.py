"""Arithmetic identities the HIP kernels rely on in place of a slower written
form, checked exhaustively on the CPU with the same IEEE binary32 operations
(round to nearest even; fmaf is the single-rounding fused multiply-add the
GPU's v_fma_f32 performs).

- R16_UNORM loads (common.h ld_u16): the reference reads UNORM16 textures as
  v/65535.  The kernels form q = v*r with r the rounded 1/65535 and correct it
  once, fmaf(fmaf(-q, 65535, v), r, q); that equals the correctly rounded
  quotient for every one of the 65536 codes, while the bare product differs
  on some (so the correction step is needed).
- llap's expand filter (ops.h ll_expand5): the reference's 5x5 weighted sum
  (weights 1/2 1/2 1 1/2 1/2 and 0 1 1/2 1/2 1, summed from zero) equals a
  quarter of the kernels' sum of exact doublings, bit for bit (zero signs
  aside), on random f16-valued taps for every texel parity."""
import subprocess

import numpy as np
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


# llap expand (ops.h ll_expand5): the reference's weighted 5x5 sum against
# the kernels' form, bit for bit in binary32 on random f16-valued taps

WE = np.array([0.5, 0.5, 1.0, 0.5, 0.5], np.float32)
WO = np.array([0.0, 1.0, 0.5, 0.5, 1.0], np.float32)


def _weighted(w, t):
    """sum from zero in tap order: ((0 + w0 t0) + w1 t1) + ..."""
    acc = np.zeros(t.shape[1:], np.float32)
    for i in range(5):
        acc = (acc + (w[i] * t[i]).astype(np.float32)).astype(np.float32)
    return acc


def _expand5(odd, t):
    """(((t0 z0 + t1 m1) + t2 m2) + t3) + t4 m4, each product by 0, 1 or 2
    exact and each sum rounded once (what the FMAs compute)"""
    z0, m1, m2, m4 = (0.0, 2.0, 1.0, 2.0) if odd else (1.0, 1.0, 2.0, 1.0)
    f = np.float32
    a = (t[0] * f(z0)).astype(f)
    a = (a + (t[1] * f(m1)).astype(f)).astype(f)
    a = (a + (t[2] * f(m2)).astype(f)).astype(f)
    a = (a + t[3]).astype(f)
    return (a + (t[4] * f(m4)).astype(f)).astype(f)


def test_llap_expand_sum_form():
    rng = np.random.default_rng(7)
    n = 200000
    mag = 10.0 ** rng.uniform(-4, 1, size=(5, 5, n))
    t = (rng.choice([-1.0, 1.0], size=(5, 5, n), p=[0.2, 0.8]) * mag).astype(np.float16).astype(np.float32)
    t[:, :, :1000] = 0.0                                    # flat zero patches
    t[:, :, 1000:2000] = t[0, 0, 1000:2000]                 # flat patches
    for ox in (0, 1):
        for oy in (0, 1):
            wx, wy = (WO if ox else WE), (WO if oy else WE)
            ref = _weighted(wy, np.stack([_weighted(wx, t[j]) for j in range(5)]))
            rows = np.stack([_expand5(ox, t[j]) for j in range(5)])
            got = (np.float32(0.25) * _expand5(oy, rows)).astype(np.float32)
            assert np.array_equal(got, ref), (ox, oy, int((got != ref).sum()))
