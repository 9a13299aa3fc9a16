import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "tissue-forge_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")


def load_package():
    """Imports the tissue-forge_amd directory as the module tissue_forge_amd."""
    if "tissue_forge_amd" in sys.modules:
        return sys.modules["tissue_forge_amd"]
    spec = importlib.util.spec_from_file_location("tissue_forge_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tissue_forge_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if ORACLE_DIR not in sys.path:
        sys.path.insert(0, ORACLE_DIR)
    import oracle  # noqa: E402
    return oracle


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def tf():
    return load_package()


@pytest.fixture(scope="session")
def W(tf):
    from tissue_forge_amd import workloads
    return workloads


@pytest.fixture(scope="session")
def O():
    return load_oracle()


@pytest.fixture(scope="session")
def gpu(tf):
    if tf.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return True
