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


@pytest.fixture(scope="session")
def tvh():
    """The HIP library through its C ABI (GPU tests only)."""
    import tvh_loader

    lib = tvh_loader.load()
    rc = lib.tvh_init(0)
    if rc != 0:
        raise RuntimeError(f"tvh_init(0) failed: {lib.status_string(rc)}")
    return lib


def u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)
