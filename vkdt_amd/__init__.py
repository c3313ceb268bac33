"""vkdt_amd -- MI355X-native (gfx950, HIP) raw-development path of vkdt.

The path is i-raw -> denoise -> demosaic -> colour -> filmcurv -> llap ->
o-pfm behind vkdt's own module API (per-module callbacks, ``connectors`` /
``params`` files, ``.cfg`` graphs).  Host C calls hand-written HIP kernels
through the C ABI declared in ``include/vkdt_hip.h``.

Python here is plumbing for tests and the benchmark: ``vkdt_amd.hip`` binds
the C ABI with ctypes, ``vkdt_amd.pipe`` binds the host runtime.  Nothing
falls back to a CPU implementation: if the native libraries are missing the
import of those modules fails.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
REPO_DIR = os.path.dirname(PKG_DIR)
INCLUDE_DIR = os.path.join(REPO_DIR, "include")


def lib_path(name):
    """Absolute path of a built shared library inside the package."""
    return os.path.join(LIB_DIR, name)
