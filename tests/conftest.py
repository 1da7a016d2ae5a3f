import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    # Build what is missing (the in-tree builds are reused where they exist).
    lib = os.path.join(ROOT, "vbd_tetmesh_amd", "libvbd_hip.so")
    orc = os.path.join(ROOT, "oracle", "_oracle.so")
    if not os.path.exists(orc):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(lib):
        from vbd_tetmesh_amd.build import build_library  # noqa: E402 -- builds before the package import
        build_library()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
