import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "misa-akmc-gpu_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libakmc on cuda:0)")


def make_cfg(n=16, va=6, T=573.0, sector_T=3e-7, physics=10.0, alloy=None, seeds=None, box=None):
    """A config in the reference's config.yaml terms (flattened)."""
    return dict(
        box=box if box is not None else (n, n, n),
        alloy=alloy if alloy is not None else [("Fe", 9666), ("Cu", 134), ("Mn", 130), ("Ni", 70)],
        va_count=va,
        seeds=seeds if seeds is not None else dict(create_types=100565, create_vacancy=16650,
                                                   event_selection=563264, time_inc=594689),
        temperature=T, attempt_freq=6e12, sector_T=sector_T, physics_time=physics)


@pytest.fixture
def cfg_factory():
    return make_cfg
