"""The C ABI libraries load without a GPU and export every function their
public headers declare; the Python prototypes cover the same set."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

import examinimd_amd as emd
from examinimd_amd import abi


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emd_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader,protos", [("emd_hip.h", emd.hip_lib, abi.HIP_PROTOS),
                                                  ("emd_app.h", emd.app_lib, abi.APP_PROTOS)])
def test_exports(header, loader, protos):
    lib = loader()
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert sorted(protos) == names


def test_no_gpu_fails_loudly():
    """without a device emd_create reports an error instead of falling back"""
    import subprocess, sys
    code = ("import ctypes as C, examinimd_amd as e\n"
            "h=C.c_void_p(); st=e.hip_lib().emd_create(0, C.byref(h)); print(st)")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=ROOT, env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"})
    assert out.stdout.strip() not in ("", "0") or out.returncode != 0


def test_version_string():
    assert b"gfx950" in emd.hip_lib().emd_version()
