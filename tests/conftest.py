import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO / "corefl-cpc_amd", REPO / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

os.environ.setdefault("OMP_NUM_THREADS", str(min(8, os.cpu_count() or 1)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_modifyitems(config, items):
    # keep GPU tests in one process and in file order; nothing to reorder here
    return


@pytest.fixture(scope="session")
def h2_oracle_tables():
    from oracle import h2_tables
    return h2_tables()


@pytest.fixture(scope="session")
def h2_tables_dict():
    from chem_tables import build_tables
    root = REPO / "corefl-cpc_amd" / "data" / "h2_li2004"
    return build_tables(str(root / "mechanism.inp"), str(root / "therm.dat"), str(root / "tran.dat"))


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
