"""The C++ drop-in facade (ParametricEngine with the reference's Request /
Result structs) returns exactly what the C ABI returns for the same request."""
import json
import os
import subprocess

import numpy as np
import pytest

import ucuda

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_matches_c_abi(gpu_engine):
    print("case: facade_demo (bif 16 rho, LLE2D 8x8, LS2D 4x4) vs C ABI", flush=True)
    out = subprocess.run([os.path.join(ROOT, "u-cuda_amd", "facade_demo")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    lor = ucuda.preset("lorenz", gpu_engine)
    g = ucuda.run_bif(lor, [ucuda.axis(ucuda.AXIS_PARAM, 2, 16, 0.0, 250.0)], [0.1, 0.1, 0.1], lor.values(), 0.01,
                      20.0, 20.0, max_peaks=64)
    flags = g["flags"].cpu().numpy()
    assert res["bif_flags"] == flags.tolist()
    first = np.where(flags > 0, g["peaks"].cpu().numpy()[0], 0.0)
    assert res["bif_first_peaks"] == first.tolist()
    ros = ucuda.preset("rossler", gpu_engine)
    axes = [ucuda.axis(ucuda.AXIS_PARAM, 2, 8, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 8, 1.0, 20.0)]
    l = ucuda.run_lyap(ros, axes, [1.0, 1.0, 1.0], ros.values(), 0.01, 5.0, 50.0, 0.1)["values"].cpu().numpy()[0]
    # separate processes compile the module separately: equal to the last bits
    np.testing.assert_allclose(res["lle2d"], l, rtol=1e-12, atol=0)
    axes4 = [ucuda.axis(ucuda.AXIS_PARAM, 2, 4, 0.05, 2.0), ucuda.axis(ucuda.AXIS_PARAM, 3, 4, 1.0, 20.0)]
    ls = ucuda.run_lyap(ros, axes4, [1.0, 1.0, 1.0], ros.values(), 0.01, 5.0, 50.0, 0.1, spectrum=True)
    s = ls["values"].cpu().numpy()
    np.testing.assert_allclose(res["ls2d_sum"], s[0] + s[1] + s[2], rtol=0, atol=1e-15)
