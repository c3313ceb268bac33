"""The C-ABI libraries load on a machine without a GPU and export every
entry point include/*.h declares; the module DSOs export the callbacks the
runtime resolves with dlsym (pipe/global.c).  No compute call is made."""
import ctypes as C
import glob
import os
import re

import pytest

from vkdt_amd import INCLUDE_DIR, PKG_DIR, lib_path


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_][a-z0-9_ ]*\**\s*\**(vkdt_[a-z0-9_]+)\s*\(", text, flags=re.M)))


HEADERS = sorted(glob.glob(os.path.join(INCLUDE_DIR, "*.h")))


@pytest.mark.parametrize("header", HEADERS, ids=[os.path.basename(h) for h in HEADERS])
def test_header_symbols_exported(header):
    names = declared(header)
    assert names, header
    lib = C.CDLL(lib_path("libvkdt_hip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_registry():
    lib = C.CDLL(lib_path("libvkdt_hip.so"))
    assert lib.vkdt_hip_abi_version() == 1
    tok = lambda s: int.from_bytes(s.encode().ljust(8, b"\0"), "little")
    lib.vkdt_hip_node_supported.argtypes = [C.c_uint64, C.c_uint64]
    for m, k in [("denoise", "half"), ("denoise", "downcov"), ("denoise", "down"), ("denoise", "assemble"),
                 ("denoise", "doub"), ("denoise", "noop"), ("demosaic", "gauss"), ("demosaic", "splat"),
                 ("demosaic", "fix"), ("demosaic", "down"), ("colour", "main"), ("filmcurv", "main"),
                 ("llap", "curve"), ("llap", "reduce"), ("llap", "assemble"), ("llap", "colour"),
                 ("hilite", "half"), ("hilite", "reduce"), ("hilite", "assemble"), ("hilite", "doub"),
                 ("rawhist", "collect"), ("rawhist", "map"), ("demosaic", "halfsize"), ("shared", "resample"),
                 ("colenc", "main"), ("demosaic", "rcd_conv"), ("demosaic", "rcd_fill"), ("hist", "collect"),
                 ("hist", "map"), ("autoexp", "hist"), ("autoexp", "lum"), ("autoexp", "exp"),
                 ("i-mcraw", "prefix"), ("i-mcraw", "decode")]:
        assert lib.vkdt_hip_node_supported(tok(m), tok(k)) == 1, (m, k)
    # kernels of modules outside this path are refused by name
    assert lib.vkdt_hip_node_supported(tok("demosaic"), tok("coldumb")) == 0
    assert lib.vkdt_hip_node_supported(tok("i-vid"), tok("decode")) == 0


def test_launch_rejects_bad_nodes_without_touching_the_device():
    """unknown kernels and malformed connector lists fail in host checks"""
    from vkdt_amd import hip
    L = hip.lib()
    tok = hip.token
    rc = L.vkdt_hip_node_launch(tok("nosuch"), tok("main"), None, 0, None, 0, None, 0, 8, 8, 1, None)
    assert rc == hip.INCOMPLETE and b"no HIP kernel" in L.vkdt_hip_last_error()
    conn = (hip.Buf * 2)()
    rc = L.vkdt_hip_node_launch(tok("denoise"), tok("half"), None, 0, None, 0, conn, 2, 8, 8, 1, None)
    assert rc == hip.INCOMPLETE and b"push constants" in L.vkdt_hip_last_error()
    pc = (C.c_int32 * 17)()
    rc = L.vkdt_hip_node_launch(tok("denoise"), tok("half"), pc, 68, None, 0, conn, 2, 8, 8, 1, None)
    assert rc == hip.INCOMPLETE and b"unbound" in L.vkdt_hip_last_error()


MODULE_CALLBACKS = {
    "i-raw": ["modify_roi_out", "read_source", "read_source_hip", "init", "cleanup", "check_params"],
    "i-mlv": ["modify_roi_out", "read_source", "read_source_hip", "init", "cleanup"],
    "hist": ["create_nodes", "modify_roi_in", "modify_roi_out"],
    "autoexp": ["create_nodes", "commit_params", "init"],
    "o-jpg": ["write_sink", "write_sink_hip"],
    "denoise": ["create_nodes", "modify_roi_in", "modify_roi_out", "check_params"],
    "demosaic": ["create_nodes", "modify_roi_in", "modify_roi_out", "check_params"],
    "colour": ["create_nodes", "modify_roi_in", "modify_roi_out", "commit_params", "init", "check_params"],
    "filmcurv": ["create_nodes", "modify_roi_out"],
    "llap": ["create_nodes"],
    "o-pfm": ["write_sink"],
}


@pytest.mark.parametrize("mod", sorted(MODULE_CALLBACKS))
def test_module_dso_callbacks(mod):
    from vkdt_amd import pipe
    pipe.lib()
    so = C.CDLL(os.path.join(PKG_DIR, "modules", mod, f"lib{mod}.so"))
    for cb in MODULE_CALLBACKS[mod]:
        assert hasattr(so, cb), (mod, cb)
    for f in ("connectors", "params") if mod not in ("display", "hist") else ("connectors",):   # hist has no params
        assert os.path.exists(os.path.join(PKG_DIR, "modules", mod, f)), (mod, f)


def test_pipe_exports_module_api():
    lib = C.CDLL(lib_path("libvkdt_pipe.so"))
    for n in ["dt_node_add", "dt_node_connect", "dt_node_connect_named", "dt_connector_copy", "dt_connector_bypass",
              "dt_module_get_input_img_param", "dt_module_param_float", "dt_module_param_int",
              "dt_module_param_string", "dt_module_get_param", "dt_graph_read_config_ascii", "dt_graph_run",
              "dt_graph_init", "dt_graph_cleanup", "dt_pipe_global_init", "dt_graph_replace_display"]:
        assert hasattr(lib, n), n
