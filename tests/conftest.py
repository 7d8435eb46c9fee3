import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

P = 0xFFFFFFFF00000001


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle, build

    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        build(with_ref=False)
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference

    if not Reference.available():
        pytest.skip("oracle/_ref/libtvm_ref.so not built (needs the reference sources)")
    r = Reference()
    r.set_threads(1)
    return r


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


def tvh_module():
    """triton-vm-prover_amd/tvh.py (the package directory name is not importable)."""
    import importlib.util

    if "tvh" in sys.modules:
        return sys.modules["tvh"]
    spec = importlib.util.spec_from_file_location(
        "tvh", os.path.join(ROOT, "triton-vm-prover_amd", "tvh.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tvh"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def tvhmod():
    return tvh_module()


@pytest.fixture(scope="session")
def tvh(tvhmod):
    """The HIP library through its C ABI (GPU tests only).  Fails loudly when the
    library or the device is missing: there is no fallback."""
    lib = tvhmod.load()
    rc = lib.tvh_init(0)
    if rc != 0:
        raise RuntimeError(f"tvh_init(0) failed: {lib.status_string(rc)}: {lib.last_error()}")
    return lib


@pytest.fixture(scope="session")
def T(tvh, tvhmod):
    hip_rt = tvhmod.hip_rt

    class Ctx:
        lib = tvh
        mod = tvhmod

        @staticmethod
        def dev(a):
            return tvhmod.to_device(a)

        @staticmethod
        def host(buf, n=None):
            return buf.download(n)

        @staticmethod
        def empty(n):
            return tvhmod.empty(n)

        @staticmethod
        def sync():
            hip_rt.synchronize()

    return Ctx


def u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)
