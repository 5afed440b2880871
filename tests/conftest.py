import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run with -m gpu on an MI355X)")


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / f"ref_mt_{name}.npz") as z:
        return {k: z[k] for k in z.files}


GOLDEN_CASES = sorted(p.stem[len("ref_mt_"):] for p in GOLDEN.glob("ref_mt_*.npz"))
MT_WALK_CASES = sorted(p.stem[len("ref_rwmt_"):] for p in GOLDEN.glob("ref_rwmt_*.npz"))


def load_mt_walk_golden(name: str) -> dict:
    with np.load(GOLDEN / f"ref_rwmt_{name}.npz") as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    """One context for the whole GPU session (tests run in one process)."""
    from tgtp_amd import Context

    ctx = Context(0)
    yield ctx
    ctx.close()
