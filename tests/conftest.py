import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "j-pet-mlem-old_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_present() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pystrip_mod():
    import pystrip
    pystrip.load_library()
    return pystrip


@pytest.fixture(scope="session")
def phantom_mod():
    from pystrip import phantom
    return phantom


def print_case(*parts):
    """Names the case on stdout (flushed) before a GPU step."""
    print("[case]", *parts, flush=True)


def expected_digest(name, n, seed=None):
    """SHA-256 of a seeded event set (tests/golden/event_digests.json,
    written by tests/golden/make_event_digests.py)."""
    import json
    path = os.path.join(ROOT, "tests", "golden", "event_digests.json")
    with open(path) as f:
        table = json.load(f)
    return table[f"{name}:{n}:{'default' if seed is None else seed}"]


def check_events(phantom_mod, events, name, seed=None):
    """Asserts the generated events are the pinned input set: a changed
    generator stream fails here, not as a parity error."""
    got = phantom_mod.events_digest(events)
    assert got == expected_digest(name, len(events), seed), (
        f"{name}: the seeded events differ from the pinned set ({got})")
