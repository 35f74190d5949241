"""Baroclinic source on the GPU against the oracle (addBaroclinicForce,
VortexSimulationSystemImp.cpp:369-465; GPU_kernel.cu:3655-3898).  The
density grid, its gradient and the vorticity update follow the reference's
arithmetic with the same summation order (the smoke sorts stably, the +-1
cells the kernels visit hold every non-zero term of the reference's +-2), so
they are checked bit for bit -- except on grids whose cells hold 64 or more
smoke particles, where a workgroup sums each cell as a tree (one float
rounding of the density at most)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "dropin", "_build", "bc_main")


def _scene(nh, n, seed):
    """Plume-like smoke (x, y ~ N(0.5, 0.08), z ~ Exp(3)) with values that
    fall with height, and vortices inside it and a few outside."""
    rng = np.random.default_rng(seed)
    heat = np.zeros((nh, 4), np.float32)
    heat[:, :2] = (0.5 + 0.08 * rng.standard_normal((nh, 2))).astype(np.float32)
    heat[:, 2] = rng.exponential(1 / 3, nh).astype(np.float32)
    heat[:, 3] = 1.0
    rho = (np.exp(-3 * heat[:, 2]) * rng.random(nh)).astype(np.float32)
    cent = np.zeros((n, 4), np.float32)
    cent[:, :2] = (0.5 + 0.1 * rng.standard_normal((n, 2))).astype(np.float32)
    cent[:, 2] = rng.exponential(0.4, n).astype(np.float32)
    cent[:4, :3] = [[-5.0, 0.5, 0.5], [9.0, 9.0, 9.0], [0.5, 0.5, -3.0], [0.5, 30.0, 0.5]]
    w = rng.standard_normal((3, n))
    return heat, rho, cent, w


@pytest.mark.parametrize("cell,max_dim,g", [(0.2, 200, (0.0, 10.0, 0.0)), (0.01, 64, (0.0, 0.0, -1.0)),
                                             (0.004, 200, (0.0, 0.0, -1.0)), (0.2, 8, (0.0, 0.0, -1.0))],
                         ids=["ref", "cap64", "fine", "dense"])
def test_add_baroclinic_against_oracle(native, oracle, cell, max_dim, g):
    DB = native.DeviceBuffer
    nh, n, dt_beta = 200_000, 50_000, 5 * 0.004
    heat, rho, cent, w = _scene(nh, n, 7)
    print(f"[baroclinic] nheat={nh} n={n} cell={cell} max_dim={max_dim}", flush=True)
    d = native.Density()
    d_heat, d_rho, d_cent = DB.from_array(heat), DB.from_array(rho), DB.from_array(cent)
    d_w = [DB.from_array(np.ascontiguousarray(w[c])) for c in range(3)]
    d.add_baroclinic(d_heat.ptr, d_rho.ptr, nh, d_cent.ptr, [q.ptr for q in d_w], n, dt_beta, g, cell, max_dim)
    native.synchronize()
    rho_g, grad_g, dim, origin, h = d.grid()
    w_ref, st = oracle.baroclinic(heat, rho, cent, w, dt_beta, g, cell, max_dim)
    print(f"  dim {dim} h {h:.4g}", flush=True)
    assert dim == st["dim"] and h == st["h"] and origin == st["origin"]
    w_got = np.stack([q.to_array(np.float64) for q in d_w])
    assert np.abs(w_got - w).max() > 0
    if nh < 64 * dim ** 3:
        # a thread per cell: the reference's order, bit for bit
        np.testing.assert_array_equal(rho_g, st["rho"])
        np.testing.assert_array_equal(grad_g, st["grad"])
        np.testing.assert_array_equal(w_got, w_ref)
    else:
        # a workgroup per cell (>= 64 smoke per cell): a tree sum, one float
        # rounding of the density at most, carried through the gradient
        ref = st["rho"].astype(np.float64)
        assert (np.abs(rho_g - ref) <= np.spacing(np.abs(st["rho"])).astype(np.float64)).all()
        gscale = np.abs(st["grad"]).max()
        assert np.abs(grad_g.astype(np.float64) - st["grad"]).max() <= 1e-6 * gscale
        np.testing.assert_allclose(w_got - w, w_ref - w, rtol=0, atol=1e-6 * np.abs(w_ref - w).max())


def test_add_baroclinic_twice_and_empty(native, oracle):
    """The object's buffers are reused and grow; nheat = 0 leaves w alone."""
    DB = native.DeviceBuffer
    d = native.Density()
    for nh, n, seed in ((20_000, 5_000, 8), (120_000, 30_000, 9)):
        heat, rho, cent, w = _scene(nh, n, seed)
        print(f"[baroclinic] reuse nheat={nh} n={n}", flush=True)
        d_heat, d_rho, d_cent = DB.from_array(heat), DB.from_array(rho), DB.from_array(cent)
        d_w = [DB.from_array(np.ascontiguousarray(w[c])) for c in range(3)]
        d.add_baroclinic(d_heat.ptr, d_rho.ptr, nh, d_cent.ptr, [q.ptr for q in d_w], n, 0.02, cell_target=0.02,
                         max_dim=100)
        native.synchronize()
        w_ref, _ = oracle.baroclinic(heat, rho, cent, w, 0.02, cell_target=0.02, max_dim=100)
        np.testing.assert_array_equal(np.stack([q.to_array(np.float64) for q in d_w]), w_ref)
    d.add_baroclinic(None, None, 0, d_cent.ptr, [q.ptr for q in d_w], n, 0.02)
    native.synchronize()
    np.testing.assert_array_equal(np.stack([q.to_array(np.float64) for q in d_w]), w_ref)


def test_reference_named_baroclinic(native, oracle, tmp_path):
    """tests/dropin/bc_main.cpp: AssignDensity, GradDensity and
    ParticleGetBaroclinic over the reference's hash names, against the oracle
    on the same grid."""
    assert os.path.exists(EXE), "build() first"
    nh, n, dtb = 100_000, 20_000, 0.02
    heat, rho, cent, w = _scene(nh, n, 10)
    w_ref, st = oracle.baroclinic(heat, rho, cent, w, dtb, cell_target=0.02, max_dim=100)
    dim, h, o = st["dim"], st["h"], st["origin"]
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([nh, n, dim], np.uint32).tobytes())
        f.write(np.array(list(o) + [h, dtb], np.float64).tobytes())
        for a in (heat, rho, cent, w):
            f.write(np.ascontiguousarray(a).tobytes())
    print("[baroclinic] running", EXE, flush=True)
    r = subprocess.run([EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    G = dim ** 3
    grid = np.frombuffer(raw, np.float32, G, 0)
    grad = np.frombuffer(raw, np.float32, 3 * G, 4 * G).reshape(G, 3)
    w_got = np.frombuffer(raw, np.float64, 3 * n, 16 * G).reshape(3, n)
    np.testing.assert_array_equal(grid, st["rho"])
    np.testing.assert_array_equal(grad, st["grad"])
    np.testing.assert_array_equal(w_got, w_ref)
