"""Frame output (newTest.cpp:23-37 output()): raw float4 per particle with
w = 1 in Particle_data%04d.bin, the format the reference renderers read
(read_particles.cpp:55-70 takes w as the particle's density)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 1000, (1 << 20) + 5])
def test_frame_bytes(native, tmp_path, n):
    rng = np.random.default_rng(n)
    pos = rng.random((n, 4), dtype=np.float32)
    pos[:, 3] = 7.0  # the writer replaces w
    dens = rng.random(n, dtype=np.float32)
    bp = native.DeviceBuffer.from_array(pos)
    bd = native.DeviceBuffer.from_array(dens)
    print(f"[frames] n={n}", flush=True)
    native.write_frame(tmp_path, 7, bp.ptr, n)
    f = tmp_path / "Particle_data0007.bin"
    assert f.stat().st_size == 16 * n
    expect = pos.copy()
    expect[:, 3] = 1.0
    assert native.read_frame(f).tobytes() == expect.tobytes()
    native.write_frame(tmp_path, 12345, bp.ptr, n, w_ptr=bd.ptr)
    expect[:, 3] = dens
    assert native.read_frame(tmp_path / "Particle_data12345.bin").tobytes() == expect.tobytes()


def test_frame_errors(native, tmp_path):
    pos = native.DeviceBuffer.from_array(np.zeros((4, 4), np.float32))
    with pytest.raises(native.PPPMError):
        native.write_frame(tmp_path / "missing_dir", 0, pos.ptr, 4)
    with pytest.raises(native.PPPMError):
        native.write_frame(tmp_path, -1, pos.ptr, 4)
