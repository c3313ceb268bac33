"""The error bar of the GPU path against the float oracle.

Two float implementations of the reference's graph (here: the GPU and the
oracle's float build) differ in the last place of their transcendentals
and divisions.  That difference is amplified in two ways only:

* near ties: a data-dependent branch whose decision margin is at the
  rounding level takes the other side.  The oracle flags those texels
  (or_set_ties / or_set_develop_ties); a flip inside the denoise wavelet
  spreads over its polyphase levels, so the whole-graph mask dilates the
  flags.
* f16 storage: where one binary16 quantum of llap/colour's inputs (the
  filmcurv rgba and the assembled luminance) moves the output by more than
  a third of 1e-3 -- values above 0.5, where the quantum is 4.9e-4, and
  near-black pixels, whose chroma is the ratio of subnormal f16 values --
  the storage the reference prescribes cannot resolve 1e-3, and the bound
  there is three such quanta.

Everywhere else the bound is the north star's max abs 1e-3.
"""
import numpy as np

LUM = np.array([2.62700212e-01, 6.77998072e-01, 5.93017165e-02])


def f16_quantum(x):
    """one binary16 ulp at |x| (the subnormal spacing 2^-24 below 2^-14)"""
    a = np.abs(np.asarray(x, np.float64))
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -14)))
    return np.exp2(e - 10)


def llap_colour(rgb, y):
    """llap/colour.comp:18-37: black rolloff and luminance transfer"""
    yo = np.maximum(rgb @ LUM, 1e-8)
    y = np.where(y < yo, yo * np.exp(np.minimum(y - yo, 0.0)), y)
    return np.maximum(0.0, rgb * (y / yo)[..., None])


def storage_sensitivity(fc, yv):
    """largest change of llap/colour's rgb when one of its f16 inputs
    (a channel of `fc`, H x W x 3, or the luminance `yv`, H x W) moves by one
    binary16 quantum either way"""
    fc = np.asarray(fc, np.float64)[..., :3]
    yv = np.asarray(yv, np.float64)
    base = llap_colour(fc, yv)
    s = np.zeros(yv.shape)
    for k in range(3):
        q = f16_quantum(fc[..., k])
        for sgn in (1.0, -1.0):
            f2 = fc.copy()
            f2[..., k] = np.maximum(0.0, f2[..., k] + sgn * q)
            s = np.maximum(s, np.abs(llap_colour(f2, yv) - base).max(-1))
    q = f16_quantum(yv)
    for sgn in (1.0, -1.0):
        s = np.maximum(s, np.abs(llap_colour(fc, yv + sgn * q) - base).max(-1))
    return s


def tolerance(fc, yv, abs_tol=1e-3, quanta=3.0):
    """per-pixel bound: 1e-3, or three storage quanta where f16 cannot resolve 1e-3"""
    return np.maximum(abs_tol, quanta * storage_sensitivity(fc, yv))
