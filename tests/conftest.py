import os
import sys
from pathlib import Path

import pytest

try:
    # torch's HIP runtime has to open the device before the first test loads
    # libgicops_hip.so, whichever test files are selected and in whatever order
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    """Load what is built; build only when a library is missing."""
    from gicops_amd import _build

    if not (_build.LIBDIR / "libgicops_host.so").exists():
        _build.build_host()
    if not (_build.ORACLE / "liboracle.so").exists():
        _build.build_oracle()


_ensure_built()


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device probe only)

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def tiny_index():
    from gicops_amd import slm

    return slm.SlmIndex(slm.DB_PRESETS["tiny"])


@pytest.fixture(scope="session")
def small_index():
    from gicops_amd import slm

    return slm.SlmIndex(slm.DB_PRESETS["small"])


@pytest.fixture(scope="session")
def small_batch(small_index):
    from gicops_amd import slm

    return slm.generate_spectra(small_index, slm.SPEC_PRESETS["small"])


@pytest.fixture(scope="session")
def open_index():
    """200k tryptic peptides + Met oxidation (<=2): an open-search index of
    ~0.56M entries that the CPU search finishes in seconds."""
    from gicops_amd import slm

    cfg = slm.DBConfig(n_peptides=200_000, seed=11, mods=[("M", 15.9949, 2)], max_mods_per_pep=2)
    return slm.SlmIndex(cfg)


@pytest.fixture(scope="session")
def searcher_small(small_index):
    from gicops_amd.search import Searcher

    s = Searcher(small_index)
    yield s
    s.close()


@pytest.fixture(scope="session")
def searcher_open(open_index):
    from gicops_amd.search import Searcher

    s = Searcher(open_index)
    yield s
    s.close()
