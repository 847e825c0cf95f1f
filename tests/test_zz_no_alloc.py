"""uc_run_lle_noalloc (tables in the GPU's static scratch array, nothing
allocated) against uc_run_lle on the same requests: identical values and
flags, on two calls in a row with the second smaller, in both precisions and
both launch shapes (two cells per lane on the larger FP32C map); a module
that is not loaded and a grid beyond the scratch size are errors."""
import numpy as np
import pytest

import ucuda


def _pair(sysm, axes, **kw):
    ref = ucuda.run_lyap(sysm, axes, [1.0, 1.0, 1.0], sysm.values(), 0.01, 2.0, 8.0, 0.1, **kw)
    got = ucuda.run_lyap(sysm, axes, [1.0, 1.0, 1.0], sysm.values(), 0.01, 2.0, 8.0, 0.1, noalloc=True, **kw)
    return ref, got


def _same(ref, got):
    np.testing.assert_array_equal(got["flags"].cpu().numpy(), ref["flags"].cpu().numpy())
    a, b = got["values"].cpu().numpy(), ref["values"].cpu().numpy()
    assert np.array_equal(a, b, equal_nan=True), np.nanmax(np.abs(a - b))


@pytest.mark.gpu
def test_noalloc_lle_equals_uc_run_lle():
    print("case: uc_run_lle_noalloc vs uc_run_lle, Rossler 480x480 F32C then 40x30 F64", flush=True)
    eng = ucuda.Engine(0)
    ros = ucuda.preset("rossler", eng)
    axes_big = [ucuda.axis(1, 2, 480, 0.05, 2.0), ucuda.axis(1, 3, 480, 1.0, 20.0)]
    axes_small = [ucuda.axis(1, 2, 40, 0.05, 2.0), ucuda.axis(1, 3, 30, 1.0, 20.0)]
    # the module is not loaded on this new engine yet: an error, no launch
    with pytest.raises(ucuda.UcError, match="not loaded"):
        ucuda.run_lyap(ros, axes_small, [1.0, 1.0, 1.0], ros.values(), 0.01, 2.0, 8.0, 0.1, noalloc=True)
    ros.compile(strict=False)
    ref, got = _pair(ros, axes_big, precision=ucuda.PREC_F32C)
    _same(ref, got)
    assert (got["flags"].cpu().numpy() == 1).mean() > 0.5
    ref, got = _pair(ros, axes_small, precision=ucuda.PREC_F64)
    _same(ref, got)
    # strict modules are separate: load, then a third call
    ros.compile(strict=True)
    ref, got = _pair(ros, axes_small, precision=ucuda.PREC_F64, strict=True)
    _same(ref, got)
    # beyond the scratch array: rejected before any device work
    too_big = [ucuda.axis(1, 2, 65000, 0.05, 2.0), ucuda.axis(1, 3, 441, 1.0, 20.0)]
    with pytest.raises(ucuda.UcError, match="maximum is 65536"):
        ucuda.run_lyap(ros, too_big, [1.0, 1.0, 1.0], ros.values(), 0.01, 2.0, 8.0, 0.1, noalloc=True,
                       out=dict(values=got["values"], flags=got["flags"], final_state=None))
