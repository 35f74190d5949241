"""The oracle's baroclinic source (addBaroclinicForce,
VortexSimulationSystemImp.cpp:320-465; GPU_kernel.cu:3655-3898) pinned by
closed forms and brute force: the density domain rule, the density grid
against an all-pairs sum with no hash tables, central differences of a
linear field (with the periodic wrap), and the vorticity change under a
constant gradient."""
import ctypes as C

import numpy as np
import pytest


def _smoke(n, seed, lo=0.3, hi=0.7):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    rho = rng.random(n).astype(np.float32)
    return pos, rho


def _domain(oracle, pos, cell=0.2, max_dim=200):
    o = (C.c_double * 3)()
    h = C.c_double(0)
    dim = C.c_uint32(0)
    p = np.ascontiguousarray(pos, np.float32)
    oracle.lib().orc_density_domain(p.ctypes.data_as(C.c_void_p), C.c_uint32(p.shape[0]), C.c_double(cell),
                                    C.c_uint32(max_dim), o, C.byref(h), C.byref(dim))
    return [o[0], o[1], o[2]], h.value, dim.value


def test_density_domain_closed_form(oracle):
    pos, _ = _smoke(1000, 1)
    pos[0, :3] = [-1.0, 0.5, 0.5]
    pos[1, :3] = [3.0, 0.5, 0.5]
    o, h, dim = _domain(oracle, pos, 0.2, 200)
    ext = pos[:, :3].astype(np.float64).max(0) - pos[:, :3].astype(np.float64).min(0)
    L = 1.2 * ext.max()
    centre = pos[:, :3].astype(np.float64).min(0) + 0.5 * ext
    np.testing.assert_allclose(o, centre - 0.5 * L, rtol=0, atol=1e-15)
    assert dim == int(np.floor(L / 0.2)) + 1 and abs(h - L / dim) < 1e-15
    assert _domain(oracle, pos, 0.001, 200)[2] == 200


def test_assign_density_brute_force(oracle):
    pos, rho = _smoke(3000, 2)
    cent = np.zeros((1, 4), np.float32)
    _, st = oracle.baroclinic(pos, rho, cent, np.zeros((3, 1)), 0.0, cell_target=0.02)
    dim, h, o = st["dim"], st["h"], st["origin"]
    assert dim > 8
    sh = np.stack([(pos[:, c].astype(np.float64) - o[c]).astype(np.float32) for c in range(3)], 1)
    q = sh.astype(np.float64) * (1.0 / h)
    # hash cell of each particle (the oracle's keys, pinned elsewhere)
    key, _ = oracle.hash_keys(pos, dim, float(np.float32(h)), [float(np.float32(v)) for v in o])
    occupied = np.zeros(dim ** 3, bool)
    occupied[key] = True
    ref = np.zeros(dim ** 3)
    for c in range(dim ** 3):
        i, j, k = c % dim, (c // dim) % dim, c // (dim * dim)
        d2 = ((q - np.array([i + 0.5, j + 0.5, k + 0.5])) ** 2).sum(1)
        w = np.where(d2 < 1, (1 - d2) ** 4, 0.0)
        tw = w.sum() + (0.0 if occupied[c] else 1.0)
        ref[c] = (w * rho).sum() / (tw if tw >= 1e-12 else 1e24)
    np.testing.assert_allclose(st["rho"].astype(np.float64), ref.astype(np.float32), rtol=2e-7, atol=1e-30)


def test_grad_density_linear_and_periodic(oracle):
    dim, h = 12, 0.25
    i = np.arange(dim)
    phi = np.broadcast_to((2.0 * i)[None, None, :] + (3.0 * i)[None, :, None] - (1.0 * i)[:, None, None],
                          (dim, dim, dim)).astype(np.float32).copy()
    grad = np.empty(3 * dim ** 3, np.float32)
    oracle.lib().orc_grad_density(phi.ctypes.data_as(C.c_void_p), C.c_uint32(dim), C.c_double(h),
                                  grad.ctypes.data_as(C.c_void_p))
    g = grad.reshape(dim, dim, dim, 3)
    inner = slice(1, dim - 1)
    np.testing.assert_array_equal(g[inner, inner, inner, 0], 2.0 / h)
    np.testing.assert_array_equal(g[inner, inner, inner, 1], 3.0 / h)
    np.testing.assert_array_equal(g[inner, inner, inner, 2], -1.0 / h)
    # periodic wrap at x = 0: (phi[1] - phi[dim-1]) / 2h
    assert g[5, 5, 0, 0] == np.float32((2.0 * 1 - 2.0 * (dim - 1)) * 0.5 / h)


def test_baroclinic_constant_gradient(oracle):
    pos, _ = _smoke(2000, 3)
    rho = (2.0 * pos[:, 0] + 0.5 * pos[:, 2]).astype(np.float32)
    rng = np.random.default_rng(4)
    cent = np.zeros((50, 4), np.float32)
    cent[:, :3] = (0.45 + 0.1 * rng.random((50, 3))).astype(np.float32)
    w0 = rng.standard_normal((3, 50))
    _, st = oracle.baroclinic(pos, rho, cent, w0, 0.0, cell_target=0.05)
    # a constant gradient field: interior vortices see it exactly (hat
    # weights sum to one), so dw = dt_beta 0.1 (G x g)
    dim, h, o = st["dim"], st["h"], st["origin"]
    Gv = np.array([0.7, -1.3, 2.1], np.float32)
    grad = np.tile(Gv, dim ** 3).astype(np.float32)
    w = [np.array(w0[c]) for c in range(3)]
    gv = (C.c_double * 3)(0.0, 0.0, -1.0)
    oc = (C.c_double * 3)(*o)
    dtb = 0.37
    oracle.lib().orc_baroclinic(cent.ctypes.data_as(C.c_void_p), *[a.ctypes.data_as(C.c_void_p) for a in w],
                                grad.ctypes.data_as(C.c_void_p), C.c_double(dtb), C.c_uint32(50), C.c_uint32(dim),
                                C.c_double(1.0 / float(np.float32(h))), oc, gv)
    dw = np.stack(w) - w0
    expect = dtb * 0.1 * np.cross(Gv.astype(np.float64), [0.0, 0.0, -1.0])
    np.testing.assert_allclose(dw, np.repeat(expect[:, None], 50, 1), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("seed", [5, 6])
def test_baroclinic_buoyancy_direction(oracle, seed):
    """Density rising in +x with g = (0, 10, 0): grad rho x g points along +z."""
    pos, _ = _smoke(4000, seed)
    rho = (pos[:, 0] * 3.0).astype(np.float32)
    cent = np.zeros((10, 4), np.float32)
    cent[:, :3] = 0.5
    w, _ = oracle.baroclinic(pos, rho, cent, np.zeros((3, 10)), 1.0, cell_target=0.05)
    # x, y components come from the sampling noise of d rho/dz alone
    assert (w[2] > 0).all() and np.abs(w[:2]).max() < 5e-2 * w[2].min()
