"""GPU basins of attraction (adaptive Dormand-Prince 5(4) + Poincare-section
signatures) vs the FP64 CPU oracle.

Bar: integer labels identical on every cell more than 2 grid spacings from a
basin boundary (boundary = a label change within Chebyshev distance 2 in the
oracle's map); the oracle's per-cell attractor is mapped to the GPU
catalogue by (period, canonical point within 3 quanta).
"""
import math

import numpy as np
import pytest

import oracle
import ucuda

pytestmark = pytest.mark.gpu


def _duffing(eng, n, gamma=0.3):
    s = ucuda.preset("duffing", eng)
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, n, -2.5, 2.5), ucuda.axis(ucuda.AXIS_VAR, 1, n, -2.5, 2.5)]
    return s, axes, s.values(gamma=gamma)


def _oracle_labels(ref, attractors, quantum):
    per, pt = ref["period"], ref["point"]
    out = np.zeros(len(per), dtype=np.int32)
    for c in range(len(per)):
        p = per[c]
        if p < 0:
            continue
        for a in attractors:
            if a["period"] != p:
                continue
            if p == 0 or (abs(pt[0, c] - a["point"][0]) <= 3 * quantum and abs(pt[1, c] - a["point"][1]) <= 3 * quantum):
                out[c] = a["label"]
                break
        else:
            out[c] = -99  # attractor missing from the GPU catalogue
    return out


def _interior(labels2d, r=2):
    n0, n1 = labels2d.shape
    inner = np.ones_like(labels2d, dtype=bool)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            sh = np.full_like(labels2d, -12345)
            ys = slice(max(0, dy), n0 + min(0, dy))
            yd = slice(max(0, -dy), n0 + min(0, -dy))
            xs = slice(max(0, dx), n1 + min(0, dx))
            xd = slice(max(0, -dx), n1 + min(0, -dx))
            sh[yd, xd] = labels2d[ys, xs]
            inner &= (sh == labels2d) | (sh == -12345)
    return inner


def test_basins_duffing_forced_f64(gpu_engine):
    print("case: forced Duffing basins 32x32, 200 periods, DP5(4) tol 1e-9, F64", flush=True)
    n = 32
    s, axes, vals = _duffing(gpu_engine, n)
    g = ucuda.run_basins(s, axes, [0.0, 0.0], vals, n_sections=200)
    funcs = oracle.SystemFuncs(s.source(1))
    bj = oracle.make_basins_job(2, axes, [0.0, 0.0], vals, 0.01, 1e-9, 1e-9, 2 * math.pi, 200)
    ref = oracle.basins_cells(funcs, bj, range(n * n), threads=8)
    per_g = g["periods"].cpu().numpy()
    lab_g = g["labels"].cpu().numpy()
    att = g["attractors"]
    print("attractors:", att)
    assert len(att) >= 2
    lab_r = _oracle_labels(ref, att, 1e-3)
    assert (lab_r != -99).all()
    inner = _interior(lab_r.reshape(n, n)).ravel()
    assert inner.sum() > n * n // 4
    np.testing.assert_array_equal(lab_g[inner], lab_r[inner])
    # periodic cells: canonical points agree to the integration tolerance scale
    both = (per_g > 0) & (ref["period"] > 0) & inner
    pts = g["points"].cpu().numpy()
    assert np.abs(pts[:, both] - ref["point"][:, both]).max() < 1e-6
    # the accepted-step counts are of the same size (identical controller)
    steps = g["steps"].cpu().numpy()
    assert abs(steps[0].mean() / ref["steps"][0].mean() - 1.0) < 0.05


def test_basins_duffing_f32c_matches_f64(gpu_engine):
    print("case: forced Duffing basins 48x48, F32 compensated vs F64", flush=True)
    n = 48
    s, axes, vals = _duffing(gpu_engine, n)
    g64 = ucuda.run_basins(s, axes, [0.0, 0.0], vals, n_sections=200)
    g32 = ucuda.run_basins(s, axes, [0.0, 0.0], vals, n_sections=200, precision=ucuda.PREC_F32C)
    l64, l32 = g64["labels"].cpu().numpy(), g32["labels"].cpu().numpy()
    # map catalogues by (period, point)
    m = {}
    for a in g32["attractors"]:
        for b in g64["attractors"]:
            if a["period"] == b["period"] and (a["period"] == 0 or
                                               (abs(a["point"][0] - b["point"][0]) <= 3e-3 and
                                                abs(a["point"][1] - b["point"][1]) <= 3e-3)):
                m[a["label"]] = b["label"]
    l32m = np.array([m.get(v, 0 if v == 0 else -99) for v in l32], dtype=np.int32)
    inner = _interior(l64.reshape(n, n)).ravel()
    assert inner.sum() > n * n // 4
    np.testing.assert_array_equal(l32m[inner], l64[inner])


def test_basins_unforced_two_wells(gpu_engine):
    print("case: unforced damped Duffing (gamma=0): equilibria (+-1, 0)", flush=True)
    n = 24
    s, axes, vals = _duffing(gpu_engine, n, gamma=0.0)
    g = ucuda.run_basins(s, axes, [0.0, 0.0], vals, n_sections=60)
    att = g["attractors"]
    pts = sorted((round(a["point"][0], 1), round(a["point"][1], 1)) for a in att if a["period"] == 1)
    assert pts == [(-1.0, 0.0), (1.0, 0.0)], att
    lab = g["labels"].cpu().numpy().reshape(n, n)
    # point symmetry (x, v) -> (-x, -v) of the unforced system swaps the wells
    flipped = lab[::-1, ::-1]
    inner = _interior(lab).ravel() & _interior(flipped).ravel()
    l1, l2 = lab.ravel()[inner], flipped.ravel()[inner]
    assert ((l1 != l2) | (l1 == 0)).all()


def _map_labels(g_from, g_to, tol=3e-3):
    m = {}
    for a in g_from["attractors"]:
        for b in g_to["attractors"]:
            if a["period"] == b["period"] and (a["period"] == 0 or (abs(a["point"][0] - b["point"][0]) <= tol and
                                                                   abs(a["point"][1] - b["point"][1]) <= tol)):
                m[a["label"]] = b["label"]
    return np.array([m.get(v, 0 if v == 0 else -99) for v in g_from["labels"].cpu().numpy()], dtype=np.int32)


@pytest.mark.parametrize("rtol,h_max", [(1e-9, None), (1e-4, 2 * math.pi)])
def test_basins_harmonic_forcing_fast_vs_strict(gpu_engine, rtol, h_max):
    """Two harmonics with a phase, g cos(w t) + 0.05 sin(2 w t + f) (d = 0.3,
    g = 0.2, w = 1, f = 0.7; ICs in [-1.5, 1.5]^2): the fast
    modules rotate their values to the stage times, strict evaluates cos /
    sin like the oracle. At 1e-4 with steps up to a whole period the
    rotation angles exceed 0.8 and take the reduction branch."""
    print("case: two-harmonic forced Duffing 48x48, fast rotation vs strict calls, rtol %g" % rtol, flush=True)
    n = 48
    s = ucuda.System(gpu_engine, ["x", "v"], ["d", "g", "w", "f"],
                     ["v", "-d*v + x - x^3 + g*cos(w*t) + 0.05*sin(2*w*t + f)"], time_symbol="t")
    assert "#define UC_HARMONICS 2" in s.source(0)
    axes = [ucuda.axis(ucuda.AXIS_VAR, 0, n, -1.5, 1.5), ucuda.axis(ucuda.AXIS_VAR, 1, n, -1.5, 1.5)]
    vals = [0.5, 0.3, 0.2, 1.0, 0.7]  # two period-1 attractors, one per well
    kw = dict(rtol=rtol, atol=rtol, n_sections=120, h_max=h_max)
    gs = ucuda.run_basins(s, axes, [0.0, 0.0], vals, strict=True, **kw)
    ls = gs["labels"].cpu().numpy()
    inner = _interior(ls.reshape(n, n)).ravel()
    assert inner.sum() > n * n // 4 and len(gs["attractors"]) >= 2
    for prec in (ucuda.PREC_F64, ucuda.PREC_F32C):
        g = ucuda.run_basins(s, axes, [0.0, 0.0], vals, precision=prec, **kw)
        lf = _map_labels(g, gs)
        agree = (lf[inner] == ls[inner]).mean()
        print("precision %d: interior agreement %.4f, attractors %d / %d" % (prec, agree, len(g["attractors"]),
                                                                           len(gs["attractors"])))
        if prec == ucuda.PREC_F64 and rtol == 1e-9:
            np.testing.assert_array_equal(lf[inner], ls[inner])
        else:
            assert agree >= 0.99
        if prec == ucuda.PREC_F64:  # same tolerance, same controller: the same accepted-step counts
            st, ss = g["steps"].cpu().numpy(), gs["steps"].cpu().numpy()
            assert abs(st[0].mean() / ss[0].mean() - 1.0) < 0.05
