"""The oracle's time-step functions (segment ends, moves, centres, vorticity
change, Gaussian smoothing; VortexSimulationSystemImp.cpp:90-125 and
GPU_kernel.cu:3118-3254, 3327-3438) pinned by what the mathematics fixes:
closed forms of the segment geometry, the round trip vortex -> segment ->
vortex, exact cases, and for the smoothing a brute-force restatement over
all pairs (no hash tables), constants and linearity."""
import numpy as np
import pytest


def _cent_omega(n, h, seed):
    rng = np.random.default_rng(seed)
    cent = np.zeros((n, 4), np.float32)
    cent[:, :3] = rng.random((n, 3)).astype(np.float32)
    cent[:, 3] = rng.random(n).astype(np.float32) + 3.0
    dirs = rng.standard_normal((3, n))
    dirs /= np.linalg.norm(dirs, axis=0)
    # magnitudes on both sides of h, and exactly h
    mag = h * np.exp(rng.uniform(-3, 3, n))
    mag[:5] = h
    return cent, dirs * mag


def test_segments_geometry_and_round_trip(oracle):
    h = 1.0 / 64
    cent, w = _cent_omega(5000, h, 1)
    enda, endb, kappa = oracle.segments_from_vortices(cent, w, h)
    l = np.linalg.norm(w, axis=0)
    # kappa = max(|w| / h, 1)
    np.testing.assert_allclose(kappa, np.maximum(l / h, 1.0), rtol=1e-15)
    ulp = np.float32(2.0 ** -23) * 2.0
    s = (endb[:, :3].astype(np.float64) - enda[:, :3]).T
    # the segment is w scaled to length min(|w|, h), centred on the vortex
    s_exp = w * np.minimum(1.0, h / l)
    assert np.abs(s - s_exp).max() <= 2 * ulp
    mid = 0.5 * (endb[:, :3].astype(np.float64) + enda[:, :3])
    assert np.abs(mid - cent[:, :3]).max() <= ulp
    # w fields untouched
    np.testing.assert_array_equal(enda[:, 3], cent[:, 3])
    np.testing.assert_array_equal(endb[:, 3], cent[:, 3])
    # vortex -> segment -> vortex: the change is position rounding only
    dw = oracle.vortices_from_segments(enda, endb, w, kappa)
    assert np.abs(dw).max() <= (kappa * 2 * ulp).max()
    assert np.abs(dw).max() < 1e-4 * np.abs(w).max()


def test_vortex_change_exact_case(oracle):
    # ends c -+ s/2 exactly representable: dw = kappa s - w exactly
    n = 64
    rng = np.random.default_rng(2)
    c = np.round(rng.random((n, 3)) * 1024) / 1024
    s = (rng.integers(-64, 64, (n, 3)) / 1024.0)
    a = np.zeros((n, 4), np.float32)
    b = np.zeros((n, 4), np.float32)
    a[:, :3] = c - s / 2
    b[:, :3] = c + s / 2
    w = rng.standard_normal((3, n))
    kappa = rng.random(n) * 4 + 1
    dw = oracle.vortices_from_segments(a, b, w, kappa)
    np.testing.assert_array_equal(dw, kappa * s.T - w)
    cen = oracle.segment_centres(a, b, np.full((n, 4), 7.0, np.float32))
    np.testing.assert_array_equal(cen[:, :3], c.astype(np.float32))
    np.testing.assert_array_equal(cen[:, 3], 7.0)


def test_move_points_closed_form(oracle):
    n = 1000
    rng = np.random.default_rng(3)
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = np.round(rng.random((n, 3)) * 256) / 256
    p[:, 3] = 5.0
    u = np.zeros((n, 4))
    u[:, :3] = rng.integers(-8, 8, (n, 3))
    u[:, 3] = 99.0
    q = oracle.move_points(p, 1.0 / 64, u)
    np.testing.assert_array_equal(q[:, :3], p[:, :3] + u[:, :3] / 64)
    np.testing.assert_array_equal(q[:, 3], 0.0)
    np.testing.assert_array_equal(oracle.move_points(p, 0.0, u)[:, :3], p[:, :3])


def _brute_smooth(oracle, params, pos, v, nu):
    """All pairs: p counts for t when p's hash cell is within +-2 of t's
    floor cell on every axis."""
    g = params.g
    key, _ = oracle.hash_keys(pos, g, params.hf, list(params.originf))
    key = key.astype(np.int64)
    hx, hy, hz = key % g, (key // g) % g, key // (g * g)
    sh = np.stack([(pos[:, c].astype(np.float64) - params.origin[c]).astype(np.float32) for c in range(3)], 1)
    own = np.floor(sh.astype(np.float64) * (1.0 / float(np.float32(params.hf)))).astype(np.int64)
    out = np.zeros((3, pos.shape[0]))
    for t in range(pos.shape[0]):
        sel = ((np.abs(hx - own[t, 0]) <= 2) & (np.abs(hy - own[t, 1]) <= 2) & (np.abs(hz - own[t, 2]) <= 2))
        d = (sh[sel] - sh[t]).astype(np.float64)
        wgt = np.exp(-(d * d).sum(1) / (2 * nu * nu))
        tw = wgt.sum()
        if tw < 1e-12:
            tw = 1e24
        out[:, t] = (v[:, sel] * wgt).sum(1) / tw
    return out


@pytest.mark.parametrize("nu", [0.1, 0.004])
def test_gaussian_smooth_brute_force(oracle, nu):
    g = 16
    rng = np.random.default_rng(4)
    n = 1500
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (0.5 + 0.12 * rng.standard_normal((n, 3))).astype(np.float32)
    o, L = oracle.reference_domain(pos, g)
    h = L / g
    params = oracle.make_params(g, h, o)
    v = rng.standard_normal((3, n))
    got = oracle.gaussian_smooth(params, pos, v, nu)
    ref = _brute_smooth(oracle, params, pos, v, nu)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


def test_gaussian_smooth_constant_and_linear(oracle):
    g = 16
    rng = np.random.default_rng(5)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.random((n, 3)).astype(np.float32)
    o, L = oracle.reference_domain(pos, g)
    h = L / g
    params = oracle.make_params(g, h, o)
    c = np.ones((3, n)) * np.array([[1.5], [-2.0], [0.25]])
    np.testing.assert_allclose(oracle.gaussian_smooth(params, pos, c), c, rtol=1e-14)
    v1, v2 = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    lhs = oracle.gaussian_smooth(params, pos, 2.0 * v1 - 3.0 * v2)
    rhs = 2.0 * oracle.gaussian_smooth(params, pos, v1) - 3.0 * oracle.gaussian_smooth(params, pos, v2)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-13)
