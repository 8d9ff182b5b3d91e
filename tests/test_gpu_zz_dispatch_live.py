"""The learned dispatcher judged live on call shapes it never saw, by exhaustive
timing on the device.  Timing-based, so it sits in a file that is collected after
every functional GPU test: a noisy run cannot hide their results."""
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")
from rsgpu import dispatch as D  # noqa: E402


def test_dispatcher_live_regret_against_exhaustive_autotune():
    """The shipped network, installed as variant tables on fresh codecs, judged
    live on call shapes it never saw: exhaustive timing of every valid variant is
    the judge; the default tiling is the method without the network."""
    announce("dispatcher live regret")
    disp = D.Dispatcher.load()
    specs = D.sample_specs(48, seed=20261203)
    r = D.evaluate_live(torch, disp, specs)
    for kind, v in r["by_kind"].items():
        print(f"  {kind}: {v}", flush=True)
    print(f"live: {r}", flush=True)
    assert r["regret_mean"] <= 0.10
    # per call, the network beats the default tiling (the GF(2^8) default, one dword
    # per lane, is itself near the best variant for most shapes, so the margin is small)
    assert r["gain_vs_default_geomean"] >= 1.05
    assert r["gain_vs_default_total"] >= 0.98    # and never costs total time
