"""Shared GPU fixture: the device, or a loud failure (the reference's YBUS_REQUIRE_CUDA=1,
ybus_cuda/tests/common/mod.rs:9-18): a GPU test never passes by skipping."""
import os

import pytest

import ybus_amd as Y


@pytest.fixture(scope="session")
def gpu():
    g = Y.Gpu.new()
    if g is None:
        if os.environ.get("YBUS_HIP_ALLOW_NO_DEVICE") == "1":
            pytest.skip("no HIP device")
        raise RuntimeError("no HIP device: the gpu tests need an MI355X")
    return g


@pytest.fixture(params=["columns", "wide"])
def path(request, monkeypatch):
    """Runs a test on both kernel paths: factors uploaded while YBUS_HIP_PATH=wide take the wide
    (global-memory tile) path for every block, those uploaded while YBUS_HIP_PATH=columns the LDS
    column path for every block wherever the factor fits."""
    monkeypatch.setenv("YBUS_HIP_PATH", request.param)
    return request.param
