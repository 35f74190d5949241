"""The oracle's potential-flow gradient path (PotentialFieldSolver.cpp:441-739;
GPU_kernel.cu:560-610, 750-912, 964-1200, 1520-1594, 1678-1718, 2977-3117)
pinned by closed forms: exact central differences and the periodic wrap, the
monopole boundary values face by face, the far-field correction of a single
occupied cell from the (separately pinned) direction table, the two-mass
near field in and inside the core, and the whole path against the far field
of a compact cloud, -M d / (4 pi r^3)."""
import ctypes as C

import numpy as np

P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def test_pot_gradient_linear_periodic(oracle):
    n, h = 10, 0.3
    i = np.arange(n)
    phi = (1.5 * i[None, None, :] - 2.0 * i[None, :, None] + 0.5 * i[:, None, None]).astype(np.float64)
    phi = np.ascontiguousarray(np.broadcast_to(phi, (n, n, n)))
    d = np.empty(3 * n ** 3)
    oracle.lib().orc_pot_gradient(P(phi), P(d), C.c_double(h), C.c_double(-1.0), C.c_int(n))
    d = d.reshape(n, n, n, 3)
    s = slice(1, n - 1)
    np.testing.assert_allclose(d[s, s, s], np.broadcast_to([-1.5 / h, 2.0 / h, -0.5 / h], d[s, s, s].shape),
                               rtol=1e-14)
    assert d[4, 4, 0, 0] == -1.0 * (1.5 * 1 - 1.5 * (n - 1)) * 0.5 / h


def test_pot_dirichlet_faces(oracle):
    n, h, M = 4, 0.25, 2.0
    cen, corner = [0.3, 0.6, 0.4], [-0.1, 0.05, 0.0]
    rhs = np.zeros(n ** 3)
    oracle.lib().orc_pot_dirichlet(P(rhs), (C.c_double * 3)(*cen), C.c_double(M), (C.c_double * 3)(*corner),
                                   C.c_double(h), C.c_int(n))
    ref = np.zeros(n ** 3)
    for c in range(n ** 3):
        i, j, k = c % n, (c // n) % n, c // (n * n)
        for axis, pos in ((0, i), (1, j), (2, k)):
            for side, at in ((-1, 0), (1, n - 1)):
                if pos == at:
                    g = np.array([i + 0.5, j + 0.5, k + 0.5])
                    g[axis] += side
                    ref[c] -= M / (12.5663706144 * np.linalg.norm(g * h + corner - cen))
    np.testing.assert_allclose(rhs, ref, rtol=1e-14)
    assert rhs[0] != 0 and rhs[n + n * n + 1] == 0  # a corner and an interior cell


def test_pot_far_field_single_cell(oracle):
    n, K, h, G = 9, 2, 0.1, 1.0
    rho = np.zeros(n ** 3)
    c0 = (4 * n + 4) * n + 4
    rho[c0] = 3.0
    gf = np.zeros(3 * n ** 3)
    phi, far = np.empty(n ** 3), np.empty(3 * n ** 3)
    oracle.lib().orc_pot_far_field(P(rho), P(phi), P(gf), P(far), C.c_double(h), C.c_double(G), C.c_double(1.0),
                                   C.c_int(n), C.c_int(K))
    far = far.reshape(n, n, n, 3)
    dirt = oracle.far_dir_table(K, h).reshape(2 * K + 1, 2 * K + 1, 2 * K + 1, 3)
    m = G * 3.0 * h ** 3
    # a target at (4 + 1, 4, 4): the source is at D = (-1, 0, 0) from it; the
    # diagonal term's gradient there is -(phi_diag at the source) / 2h along x
    diag = 0.250580605913965 * 3.0 * h * h * G
    np.testing.assert_allclose(far[4, 4, 5], -dirt[K, K, K - 1] * m - np.array([-diag * 0.5 / h, 0, 0]),
                               rtol=1e-13, atol=1e-15)
    # beyond K cells the correction is zero
    np.testing.assert_array_equal(far[4, 4, 4 + K + 2], 0.0)


def test_pot_near_two_masses(oracle):
    g, K, h = 8, 1, 0.125
    ev = np.zeros((3, 4), np.float32)
    ev[:, :3] = [[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.52, 0.5, 0.5]]
    mp = np.zeros((1, 4), np.float32)
    mp[0, :3] = [0.625, 0.5, 0.5]
    key, idx = oracle.hash_keys(mp, g, h, [0.0, 0.0, 0.0])
    st, en = oracle.cell_tables(*oracle.sort_pairs(key, idx)[:1], g ** 3)
    pf = np.zeros(9)
    mass = np.array([2.0])
    oracle.lib().orc_pot_near(P(st), P(en), P(ev), P(mp), P(mass), P(pf), C.c_double(h), C.c_double(1.0),
                              C.c_int(g), C.c_int(K), C.c_uint32(3), C.c_uint32(1), (C.c_double * 3)(0, 0, 0))
    pf = pf.reshape(3, 3)
    d = np.float64(np.float32(0.625) - np.float32(0.5))
    np.testing.assert_allclose(pf[0], [2.0 * d / (12.5663706144 * d ** 3), 0, 0], rtol=1e-14)
    d2 = np.float64(np.float32(0.625) - np.float32(0.52))
    np.testing.assert_allclose(pf[2], [2.0 * d2 / (12.5663706144 * d2 ** 3), 0, 0], rtol=1e-14)


def test_potential_gradient_far_from_cloud(oracle):
    rng = np.random.default_rng(1)
    n = 4000
    mp = np.zeros((n, 4), np.float32)
    mp[:, :3] = (0.5 + 0.04 * rng.standard_normal((n, 3))).astype(np.float32)
    m = rng.random(n) / n
    ev = np.zeros((6, 4), np.float32)
    ev[:, :3] = [[0.72, 0.72, 0.72], [0.28, 0.3, 0.3], [0.7, 0.3, 0.7], [0.75, 0.5, 0.5], [0.5, 0.25, 0.5],
                 [0.5, 0.5, 0.75]]
    gr, st = oracle.potential_gradient(mp, m, ev, 32, K=2)
    M, c = m.sum(), np.array(st["center"])
    d = ev[:, :3] - c
    r = np.linalg.norm(d, axis=1)
    ref = -M * d / (4 * np.pi * r[:, None] ** 3)
    np.testing.assert_allclose(gr, ref, rtol=0.03, atol=0.03 * np.abs(ref).max())


def test_pot_scalar_far_single_cell(oracle):
    n, K, h, G = 9, 2, 0.1, 1.0
    rho = np.zeros(n ** 3)
    c0 = (4 * n + 4) * n + 4
    rho[c0] = 2.0
    phi = np.zeros(n ** 3)
    oracle.lib().orc_pot_scalar_far(P(rho), P(phi), C.c_double(h), C.c_double(G), C.c_int(n), C.c_int(K))
    phi = phi.reshape(n, n, n)
    m = G * 2.0 * h * h
    assert phi[4, 4, 4] == -0.25 * m
    np.testing.assert_allclose(phi[4, 5, 6], -m / (12.5663706144 * np.sqrt(5.0)), rtol=1e-15)
    assert phi[4, 4, 4 + K + 1] == 0.0


def test_pot_near_scalar_two_masses(oracle):
    g, K, h = 8, 1, 0.125
    ev = np.zeros((2, 4), np.float32)
    ev[:, :3] = [[0.52, 0.5, 0.5], [0.45, 0.5, 0.5]]
    mp = np.zeros((1, 4), np.float32)
    mp[0, :3] = [0.6, 0.5, 0.5]
    key, idx = oracle.hash_keys(mp, g, h, [0.0, 0.0, 0.0])
    st, en = oracle.cell_tables(key, g ** 3)
    pf = np.zeros(2)
    oracle.lib().orc_pot_near_scalar(P(st), P(en), P(ev), P(mp), P(np.array([3.0])), P(pf), C.c_double(h),
                                     C.c_double(1.0), C.c_int(g), C.c_int(K), C.c_uint32(2), C.c_uint32(1),
                                     (C.c_double * 3)(0, 0, 0))
    # 0.08 away: inside the core a = 0.125, 3 / (4 pi a); 0.15 away: 3 / (4 pi r)
    d1 = np.float64(np.float32(0.6) - np.float32(0.45))
    np.testing.assert_allclose(pf, [3.0 / (12.5663706144 * 0.125), 3.0 / (12.5663706144 * d1)], rtol=1e-15)
