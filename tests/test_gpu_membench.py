"""The measurement microbenchmarks behind the roofline denominators: the
random-gather rate and the random-line rate by load shape return finite
positive rates, and reject arguments outside their documented ranges."""
import math
import sys

import pytest

import zgraph as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("chains", [1, 2, 4])
def test_random_gather_rate_positive(chains):
    say(f"[membench] gather 256 MiB chains={chains}")
    gbs, ms = Z.random_gather_bandwidth(256 << 20, chains, 1 << 22)
    assert math.isfinite(gbs) and gbs > 0 and ms > 0


@pytest.mark.parametrize("shape", [1, 2, 5, 8, "coop"])
def test_random_line_rate_positive(shape):
    coop = shape == "coop"
    say(f"[membench] lines 256 MiB shape={shape}")
    gbs, ms = Z.random_line_bandwidth(256 << 20, 8 if coop else shape, coop, 1 << 22)
    assert math.isfinite(gbs) and gbs > 0 and ms > 0


def test_membench_rejects_bad_arguments():
    with pytest.raises(Z.ZGraphError):
        Z.random_gather_bandwidth(256 << 20, 3, 1 << 20)
    with pytest.raises(Z.ZGraphError):
        Z.random_line_bandwidth(256 << 20, 9, False, 1 << 20)
    with pytest.raises(Z.ZGraphError):
        Z.random_line_bandwidth(256 << 20, 0, False, 1 << 20)
    with pytest.raises(Z.ZGraphError):
        Z.random_line_bandwidth(1024, 1, False, 1 << 20)
