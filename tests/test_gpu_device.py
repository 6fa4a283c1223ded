"""Device plumbing (tests/cuda/test-device.cu:26-138 restated for HIP/gfx950)."""
import numpy as np
import pytest

import fehip

pytestmark = pytest.mark.gpu


def test_device_limits(gpu):
    info = fehip.device_info(0)
    assert info.name or info.arch.startswith(b"gfx950")  # the name can be blank on virtualized boxes
    assert info.cu_count >= 1
    assert info.wavefront_size == 64  # CDNA wave64 (the reference asserts 32 for NVIDIA)
    assert info.max_threads_per_block >= 1024
    assert info.lds_per_block >= 64 * 1024
    assert 0 < info.free_global_mem <= info.total_global_mem
    with pytest.raises(fehip.InvalidArgument):
        fehip.device_info(-1)
    with pytest.raises(fehip.InvalidArgument):
        fehip.device_info(fehip.device_count())


def test_device_array_round_trip(gpu):
    host = np.array([1.5, -2.0, 3.25, 0.0, 1e-300, 7.0])
    d = fehip.DeviceArray.from_host(host)
    assert d.size == 6 and d.nbytes == 48
    assert d.to_host().tobytes() == host.tobytes()
    d.zero()
    assert (d.to_host() == 0).all()
    with pytest.raises(fehip.InvalidArgument):
        d.copy_from_host(np.zeros(4))
    d.resize(8)
    d.copy_from_host(np.ones(8))
    assert (d.to_host() == 1).all()
    ints = np.array([0, 1, 1, 2], np.int32)
    assert (fehip.DeviceArray.from_host(ints).to_host() == ints).all()
    assert fehip.DeviceArray(0).to_host().size == 0


def test_errors_forced_from_host(gpu):
    """An oversize allocation and a 4096-thread block fail cleanly, and the error
    does not leak into the next call (tests/cuda/test-device.cu:104-122)."""
    with pytest.raises(fehip.FemHipError) as e:
        fehip.DeviceArray(1 << 50)
    assert e.value.code == fehip.ERR_OUT_OF_MEMORY
    assert "hipMalloc" in str(e.value)
    fehip.call("fem_hip_probe_launch", 1, 64)  # no stale error
    with pytest.raises(fehip.FemHipError):
        fehip.call("fem_hip_probe_launch", 1, 4096)
    fehip.call("fem_hip_probe_launch", 64, 256)


@pytest.mark.parametrize("seed,n,lo,hi", [(1, 0, -1.0, 1.0), (0, 1, 0.0, 1.0), (2**64 - 1, 257, -1.0, 1.0),
                                          (12345, 100003, 2.5, -3.0), (7, 5_000_000, -1.0, 1.0)])
def test_device_rng_equals_host_stream(gpu, seed, n, lo, hi):
    """The device splitmix64 stream is the host one bit for bit (5M exceeds one grid pass)."""
    d = fehip.rng_uniform_device(seed, n, lo, hi)
    fehip.synchronize()
    assert d.to_host().tobytes() == fehip.rng_uniform(seed, n, lo, hi).tobytes()


def test_stream_copy(gpu):
    # bit-equal copy, a size that is not a whole grid, and the 16-byte rule
    host = fehip.rng_uniform(7, 1_000_001 * 2 + 6)
    src = fehip.DeviceArray.from_host(host)
    dst = fehip.DeviceArray(host.size)
    s = fehip.Stream()
    fehip.stream_copy(dst, src, s)
    s.synchronize()
    assert np.array_equal(dst.to_host().view(np.uint64), host.view(np.uint64))
    odd_src, odd_dst = fehip.DeviceArray(3), fehip.DeviceArray(3)
    with pytest.raises(fehip.InvalidArgument):
        fehip.stream_copy(odd_dst, odd_src, s)
    with pytest.raises(fehip.InvalidArgument):
        fehip.stream_copy(dst, odd_src, s)


@pytest.mark.parametrize("n3,P,stride", [((3, 4, 5), 2, 1), ((4, 4, 4), 5, 1), ((2, 3, 2), 8, 3), ((7, 1, 2), 3, 2)])
def test_box_offsets_device_equal_host(gpu, n3, P, stride):
    """The device box offsets (the serving loop's meshes) equal the host generator's entry for entry."""
    d = fehip.box_offsets_device(*n3, P, stride)
    assert d.to_host().tobytes() == fehip.box_offsets(*n3, P, stride).tobytes()


@pytest.mark.parametrize("n,jitter,seed", [(1, 0.1, 0), (5, 0.1, 3), (16, 0.1, 0), (9, 0.3, 1234567)])
def test_box_coords_device_equal_host(gpu, n, jitter, seed):
    """The device vertex coordinates equal the host's bit for bit (the same
    splitmix64 draws, in counter form, and the same roundings)."""
    d = fehip.box_coords_device(n, jitter, seed)
    assert d.to_host().tobytes() == fehip.box_coords(n, jitter, seed).tobytes()
