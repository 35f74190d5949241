"""GPU multigrid against the oracle and against manufactured solutions
(multigrid/multigrid/MultiGrid.cpp:128-250, MultiGrid_kernel.cu:422-934)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _manufactured(oracle, dims, seed):
    xs = np.random.default_rng(seed).standard_normal(int(np.prod(dims)))
    b = -oracle.mg_residual(xs, np.zeros_like(xs), dims)  # b = A x*
    return xs, b


@pytest.mark.parametrize("dims", [(32, 32, 32), (64, 64, 64), (40, 24, 16), (8, 8, 8)])
def test_mg_vcycle_fp64_manufactured(native, oracle, dims):
    print(f"[mg] fp64 dims={dims}", flush=True)
    xs, b = _manufactured(oracle, dims, 1)
    mg = native.MG(dims, native.F64)
    assert mg.levels() == oracle.mg_levels(*dims)
    bx, bb = native.DeviceBuffer(8 * xs.size), native.DeviceBuffer.from_array(b)
    bx.fill_zero()
    cyc, rel, code = mg.solve(bx.ptr, bb.ptr, native.MG_VCYCLE, 2, 2, 1.0, 1e-11, 60)
    x = bx.to_array(np.float64, xs.size)
    assert code == 0 and rel <= 1e-11, (cyc, rel)
    assert cyc <= 16, f"V(2,2) needed {cyc} cycles"
    assert np.abs(x - xs).max() < 1e-8 * np.abs(xs).max() * 100
    # same discrete solution as the oracle
    xo, _, _ = oracle.mg_solve(b, dims, tol=1e-12, max_cycles=60)
    assert rel_l2(x, xo) < 1e-9


def test_mg_ref_fmg_matches_oracle(native, oracle):
    """The reference schedule (one FMG pass + V(4,4) SOR 1.43f, constant
    prolongation) gives the oracle's result to rounding."""
    dims = (32, 32, 32)
    _, b = _manufactured(oracle, dims, 2)
    omega = float(np.float32(1.43))
    mg = native.MG(dims, native.F64)
    bx, bb = native.DeviceBuffer(8 * b.size), native.DeviceBuffer.from_array(b)
    bx.fill_zero()
    cyc, rel, code = mg.solve(bx.ptr, bb.ptr, native.MG_REF_FMG, 4, 4, omega, 0.0, 1)
    x = bx.to_array(np.float64, b.size)
    xo, _, relo = oracle.mg_solve(b, dims, schedule=oracle.MG_REF_FMG, nu1=4, nu2=4, omega=omega)
    assert rel_l2(x, xo) < 1e-10
    assert abs(rel - relo) <= 1e-6 * max(relo, 1e-30) + 1e-14


def test_mg_fp32_reaches_working_precision(native, oracle):
    dims = (64, 64, 64)
    xs, b = _manufactured(oracle, dims, 3)
    mg = native.MG(dims, native.F32)
    bx, bb = native.DeviceBuffer(4 * xs.size), native.DeviceBuffer.from_array(b.astype(np.float32))
    bx.fill_zero()
    cyc, rel, code = mg.solve(bx.ptr, bb.ptr, native.MG_VCYCLE, 2, 2, 1.0, 1e-5, 40)
    assert rel <= 1e-5, (cyc, rel)
    x = bx.to_array(np.float32, xs.size)
    assert rel_l2(x, xs) < 1e-4
