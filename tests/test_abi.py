"""The C ABI library loads and exports every entry point include/p7_gpu.h declares."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "p7_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p7_gpu_\w+)\s*\(", text)))


def test_header_declares_the_drop_ins():
    names = declared_functions()
    for f in ("p7_gpu_MSVFilter", "p7_gpu_SSVFilter", "p7_gpu_ViterbiFilter", "p7_gpu_ForwardParser",
              "p7_gpu_BackwardParser", "p7_gpu_oprofile_Convert", "p7_gpu_oprofile_ReconfigLength",
              "p7_gpu_Pipeline", "p7_gpu_bg_FilterScore"):
        assert f in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "biobench2_amd", "libp7gpu.so"))
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_product_does_not_link_the_oracle():
    so = open(os.path.join(ROOT, "biobench2_amd", "libp7gpu.so"), "rb").read()
    assert b"orc_msv_filter" not in so and b"libp7oracle" not in so
    for dirpath, _, files in os.walk(os.path.join(ROOT, "biobench2_amd")):
        for fn in files:
            if fn.endswith((".py", ".c", ".h", ".hip")):
                src = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in src and "from oracle" not in src, fn
