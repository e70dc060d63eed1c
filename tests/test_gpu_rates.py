"""The two reference-rate utilities the roofline figures are divided by:
a plain device copy (HBM) and XOR-popcount chains (integer VALU).  Sanity
bounds only: the rates are measurements of the device at hand."""
import pytest

pytestmark = pytest.mark.gpu


def test_copy_bandwidth(ctx):
    gbps = ctx.copy_bandwidth(1 << 28, 3)
    assert 100.0 < gbps < 20000.0


def test_popcount_rate(ctx):
    gops = ctx.popcount_rate(3)
    # 256 CUs x 64 lanes x ~2.4 GHz is ~39 Top/s; anything far outside is a
    # broken measurement
    assert 1000.0 < gops < 200000.0
