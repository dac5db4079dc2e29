"""mhmhip_run_noalloc: the time loop without device allocation (per-call
buffers in the library's fixed device scratch) gives what mhmhip_run gives,
bit for bit, on two calls in a row (the second shorter), in a context that
never ran mhmhip_run (its device footprint stays where it was); a call whose
scratch would exceed the array is refused with a message."""
import sys

import numpy as np
import pytest

from mhm_amd import synth
from pairs import make_pair

pytestmark = pytest.mark.gpu

STATES = ("INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("route_mode,window", [(3, 16), (1, 0), (2, 0)])
def test_run_noalloc_matches_run(route_mode, window):
    _log(f"case: run_noalloc route_mode={route_mode} window={window}")
    dom = synth.make_domain(20, 24, seed=91)
    ref, _, steps, _ = make_pair([dom], n_sets=2, n_steps=60, seed=92)
    sut, _, _, _ = make_pair([dom], n_sets=2, n_steps=60, seed=92)
    kw = dict(mode=1, steps_per_launch=8, route_mode=route_mode, route_window=window)
    bytes0 = sut.device_bytes
    for lo, hi in ((0, 40), (40, 60)):  # the second call shorter
        g_ref = ref.run(steps[lo:hi], **kw)
        g = sut.run(steps[lo:hi], noalloc=True, **kw)
        assert np.array_equal(g, g_ref), (lo, hi)
        for s in range(2):
            for k in STATES:
                assert np.array_equal(sut.get(k, s), ref.get(k, s)), (lo, s, k)
            a, b = sut.routing_state(s), ref.routing_state(s)
            for k in ("qOUT", "qTIN", "qTR", "Qmod"):
                assert np.array_equal(a[k], b[k]), (lo, s, k)
    assert sut.device_bytes == bytes0  # nothing grown on the device
    ref.close()
    sut.close()


def test_run_noalloc_refuses_what_does_not_fit():
    _log("case: run_noalloc size limit")
    dom = synth.make_domain(256, 256, seed=93)
    model, _, steps, _ = make_pair([dom], n_sets=1, n_steps=200, seed=94)
    with pytest.raises(RuntimeError, match="scratch"):
        model.run(steps, mode=1, steps_per_launch=24, route_window=64, noalloc=True)
    with pytest.raises(RuntimeError, match="use_graph"):
        model.run(steps[:8], mode=1, use_graph=True, noalloc=True)
    model.close()
