import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libeffm_hip.so)")


_WL = {}


def workload(name, **kw):
    """Cached synthetic workloads (effmass_amd.synth.build)."""
    from effmass_amd import synth
    key = (name, tuple(sorted(kw.items())))
    if key not in _WL:
        _WL[key] = synth.build(name, **kw)
    return _WL[key]
