"""Test configuration: markers, import path, and the CPU checker build."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
os.chdir(ROOT)  # setup files name paths relative to the repository root


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    lib = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    srcs = [os.path.join(ROOT, "oracle", f) for f in os.listdir(os.path.join(ROOT, "oracle"))
            if f.endswith((".c", ".h"))]
    srcs += [os.path.join(ROOT, "fargo3d_amd", "csrc", "host", f)
             for f in os.listdir(os.path.join(ROOT, "fargo3d_amd", "csrc", "host"))]
    if (not os.path.exists(lib)) or any(os.path.getmtime(s) > os.path.getmtime(lib) for s in srcs):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    yield
