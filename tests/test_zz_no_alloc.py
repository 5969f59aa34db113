"""hlb_reduce_noalloc, the allocation-free monitor, against hlb_reduce.

Both run the same kernel; the allocation-free one writes the library's
static device scratch instead of solver-owned buffers.  The bar is bitwise
equality of the whole hlb_monitor record, on two calls in a row with a
smaller domain the second time (the scratch keeps the previous call's
contents), for the pending (pulled) and the post-stream state, every layout
and every storage form.
"""
import ctypes as C

import pytest

from hemelb_amd import _native as N
from tests import cases as K

pytestmark = pytest.mark.gpu


def _both(lbm):
    a, b = N.hlb_monitor(), N.hlb_monitor()
    N.check(N.lib().hlb_reduce(lbm._h, C.byref(a)))
    N.check(N.lib().hlb_reduce_noalloc(lbm._h, C.byref(b)))
    return a, b


def _check(lbm, label):
    a, b = _both(lbm)
    print(f"[no-alloc] {label}: n={lbm.domain.n} mass={a.mass!r} flux0={a.iolet_flux[0]!r}",
          flush=True)
    assert a.stable == 1 and a.mass > 0 and a.n_iolets == 2
    assert bytes(a) == bytes(b), (label, a.mass, b.mass, a.velocity_max, b.velocity_max)


@pytest.mark.parametrize("layout", ["blocks", "pull_table", "compact_blocks"])
@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_noalloc_monitor_equals_monitor(layout, storage):
    big = K.cylinder_case(7.0, 28, 19, "LBGK", "BFL", "womersley", tau=0.7)
    small = K.cylinder_case(4.0, 12, 15, "LBGK", "SBB", "cosine")
    assert small.domain.n < big.domain.n
    for label, case, steps in (("big", big, 30), ("small", small, 20)):
        lbm = case.product(layout=layout, storage=storage)
        _check(lbm, f"{label} {layout} {storage} post-stream")
        lbm.step(steps)
        _check(lbm, f"{label} {layout} {storage} after {steps} steps")
        lbm.step(1)
        _check(lbm, f"{label} {layout} {storage} after {steps + 1} steps")
        assert any(v != 0.0 for v in _both(lbm)[1].iolet_flux[:2])
