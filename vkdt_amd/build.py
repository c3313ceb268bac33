"""Build the native parts in-tree.

  vkdt_amd/lib/libvkdt_hip.so    HIP kernels + C ABI (hipcc, gfx950)
  vkdt_amd/lib/libvkdt_calib.so  instruction-rate calibration kernels (bench.py)
  vkdt_amd/lib/libvkdt_pipe.so   host graph runtime (C, gcc)
  vkdt_amd/modules/<m>/lib<m>.so module DSOs (C, gcc)
  vkdt_amd/bin/vkdt-cli          command line front end
  oracle/_build/liboracle{64,32}.so  the test oracle (make)

Every call compiles; nothing is skipped because an output already exists.
"""
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
PIPE = os.path.join(PKG, "pipe")
MODS = os.path.join(PKG, "modules")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VKDT_HIP_ARCH", "gfx950")
# device code: no packed-FP32 instruction selection, and no FMA contraction so
# the float arithmetic follows the written operation order (as the float
# oracle build does with -ffp-contract=off)
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           "-Wno-unused-but-set-variable"]

HIP_SOURCES = ["exec.hip", "k_denoise.hip", "k_demosaic.hip", "k_colour.hip", "k_llap.hip"]


def run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    extra = [os.path.join(CSRC, s) for s in sorted(os.listdir(CSRC))
             if s.endswith(".hip") and s not in HIP_SOURCES]
    out = os.path.join(LIB, "libvkdt_hip.so")
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs + extra)
    os.replace(out + ".tmp", out)
    return out


def build_calib():
    """the instruction-rate calibration kernels bench.py --calibrate-valu times"""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libvkdt_calib.so")
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", out + ".tmp", os.path.join(CSRC, "calib", "calib.hip")])
    os.replace(out + ".tmp", out)
    return out


def build_pipe():
    """host runtime + module DSOs + cli (only when their sources exist)"""
    if not os.path.isdir(PIPE) or not glob.glob(os.path.join(PIPE, "*.c")):
        return None
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PIPE, "*.c")))
    srcs = [s for s in srcs if not s.endswith("cli.c")]
    out = os.path.join(LIB, "libvkdt_pipe.so")
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + PKG]
    run(["gcc"] + C_FLAGS + inc + ["-shared", "-o", out + ".tmp"] + srcs +
        ["-L" + LIB, "-lvkdt_hip", "-Wl,-rpath,$ORIGIN", "-ldl", "-lm", "-lpthread"])
    os.replace(out + ".tmp", out)
    for mdir in sorted(glob.glob(os.path.join(MODS, "*"))):
        msrc = sorted(glob.glob(os.path.join(mdir, "*.c")))
        if not msrc:
            continue
        name = os.path.basename(mdir)
        mo = os.path.join(mdir, f"lib{name}.so")
        run(["gcc"] + C_FLAGS + inc + ["-shared", "-o", mo + ".tmp"] + msrc +
            ["-L" + LIB, "-lvkdt_pipe", "-Wl,-rpath,$ORIGIN/../../lib", "-lm"])
        os.replace(mo + ".tmp", mo)
    cli = os.path.join(PIPE, "cli.c")
    if os.path.exists(cli):
        os.makedirs(BIN, exist_ok=True)
        exe = os.path.join(BIN, "vkdt-cli")
        run(["gcc"] + C_FLAGS + inc + ["-o", exe + ".tmp", cli, "-L" + LIB, "-lvkdt_pipe", "-lvkdt_hip",
             "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
        os.replace(exe + ".tmp", exe)
    return out


def build_oracle():
    run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])


def build_all(verbose=True):
    steps = [("hip", build_hip), ("calib", build_calib), ("pipe", build_pipe), ("oracle", build_oracle)]
    for name, fn in steps:
        fn()
        if verbose:
            print(f"[build] {name} ok", flush=True)


if __name__ == "__main__":
    build_all()
