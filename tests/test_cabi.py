"""The C-ABI library: it loads, exports what include/cathy_gpu.h declares,
and the product path has no CPU fallback.  No compute calls (CPU only)."""
from __future__ import annotations

import ast
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cathy_gpu.h")
PKG = os.path.join(ROOT, "pycathy_wrapper_amd")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cathy_gpu_\w+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("cathy_gpu_symslv", "cathy_gpu_flow3d_create", "cathy_gpu_flow3d_advance",
                 "cathy_gpu_flow3d_get_state", "cathy_gpu_flow3d_destroy", "cathy_gpu_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pycathy_wrapper_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libcathy_gpu.so not built (run __graft_entry__.build())")
    L = C.CDLL(gpu.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    from pycathy_wrapper_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libcathy_gpu.so not built")
    blob = open(gpu.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_missing_library_raises(monkeypatch, tmp_path):
    """No silent fallback: without the HIP library every entry raises."""
    from pycathy_wrapper_amd import gpu
    monkeypatch.setattr(gpu, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(gpu, "_lib", None)
    with pytest.raises(gpu.CathyGpuError):
        gpu.device_info()
    with pytest.raises(gpu.CathyGpuError):
        gpu.symslv([1, 2], [1], [1.0], [1.0])


def test_product_never_imports_the_oracle():
    """The package (product path) imports neither oracle/ nor tests/."""
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if not f.endswith(".py"):
                continue
            tree = ast.parse(open(os.path.join(dirpath, f)).read())
            for node in ast.walk(tree):
                mods = []
                if isinstance(node, ast.Import):
                    mods = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom):
                    mods = [node.module or ""]
                for mname in mods:
                    assert "oracle" not in mname and "refrun" not in mname and "systems" not in mname, (f, mname)
    for dirpath, _, files in os.walk(os.path.join(PKG, "csrc")):
        for f in files:
            txt = open(os.path.join(dirpath, f), errors="replace").read()
            assert "cathy_oracle" not in txt and "orc_" not in txt, f


def test_route_network_checks_run_before_the_device():
    """cathy_gpu_route_create checks the drainage network on the host before
    any HIP call, so a network the reference could not route (a cell draining
    to one before it in QOI_SN, or off the grid) is refused with a message
    even without a GPU."""
    from pycathy_wrapper_amd import routing as RT
    net = RT.synthetic_network(6, 5, dx=10.0, seed=1)
    q = net.qoi.copy()
    q[[0, -2]] = q[[-2, 0]]
    with pytest.raises(RT.CathyGpuError, match="not after it in QOI_SN"):
        RT.RouteGPU(RT.Network(**{**net.__dict__, "qoi": q}))
    with pytest.raises(RT.CathyGpuError, match="drains off the grid"):
        RT.RouteGPU(RT.Network(**{**net.__dict__, "qoi": net.qoi[::-1].copy()}))
    dup = net.qoi.copy()
    dup[1] = dup[0]
    with pytest.raises(RT.CathyGpuError, match="twice"):
        RT.RouteGPU(RT.Network(**{**net.__dict__, "qoi": dup}))
