import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def _ensure_built():
    from chiaro_amd import build
    from chiaro_amd import _abi
    if not _abi.SCENE_LIB.exists():
        build.build_scene()
    from oracle import binding
    if not binding.LIB_PATH.exists():
        binding.build()


_ensure_built()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import binding
    return binding.lib()


@pytest.fixture(scope="session")
def gpu_context():
    from chiaro_amd import _abi
    from chiaro_amd.depth import Context
    if not _abi.DEPTH_LIB.exists():
        pytest.fail("libchiaro_depth.so is not built (python -m chiaro_amd.build)")
    ctx = Context(0)
    yield ctx
    ctx.close()
