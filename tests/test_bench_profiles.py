"""The bench line's roofline object reads the committed rocprofv3 summaries
under profiles/ (HBM traffic per push call, VALU/LDS utilisation): every
push profile the bench can name parses to numbers."""
import os

import pytest

import bench

PROFILES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles")


def _names(suffix):
    return sorted(f for f in os.listdir(PROFILES) if f.startswith("push_") and f.endswith(suffix))


@pytest.mark.parametrize("name", _names("_pmc.txt"))
def test_traffic_parses(name):
    t = bench.load_traffic(name)
    assert t is not None and t > 1e8, (name, t)


@pytest.mark.parametrize("name", _names("_util.txt"))
def test_util_parses(name):
    u = bench.load_util(name)
    assert u is not None
    for k in ("VALUBusy", "LdsUtil", "flop_per_push"):
        assert k in u and u[k] > 0, (name, k, u)
