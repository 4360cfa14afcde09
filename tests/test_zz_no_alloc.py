"""The allocation-free momentum exchange entry point
(d3q19_compute_momentum_exchange_forces_noalloc, device pointers, partials in
a static device array) against the staged export it is a variant of
(d3q19_compute_momentum_exchange_forces, forces_momentum_exchange.ispc:71):
the same bits, on two calls in a row, the second on a smaller grid."""
import ctypes

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float


def _entry():
    import wake_amd

    L = wake_amd._lib.lib()
    if L.wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    fn = L.d3q19_compute_momentum_exchange_forces_noalloc
    fn.argtypes = [P, P, I, I, I, F, F, F, F, P, P]
    fn.restype = I
    return wake_amd, fn


def _case(rng, nx, ny, nz, R):
    """Random populations around the weights; a solid sphere and a moving
    wall plane (both count as walls for the links)."""
    N = nx * ny * nz
    f = (ob.D3Q19_W[:, None] * (1.0 + 0.05 * rng.standard_normal((19, N)))).astype(np.float32).ravel()
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = (nx * 0.4, ny * 0.5, nz * 0.5)
    d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    ct = np.where(d2 <= R * R, 1, 0).astype(np.uint8)
    ct[:, ny - 1, :] = 6
    return f, ct.ravel(), c


def test_noalloc_matches_staged_export_two_calls():
    import torch

    wa, fn = _entry()
    rng = np.random.default_rng(11)
    stream = torch.cuda.current_stream().cuda_stream
    for nx, ny, nz, R in [(48, 32, 24, 6.0), (20, 12, 10, 3.0)]:
        f, ct, c = _case(rng, nx, ny, nz, R)
        ref = [np.zeros(1, np.float32) for _ in range(3)]
        wa._lib.kernel("d3q19_compute_momentum_exchange_forces", f, ct, nx, ny, nz, *c, 2 * R, *ref)
        ref = np.concatenate(ref)
        df = torch.from_numpy(f).cuda()
        dct = torch.from_numpy(ct).cuda()
        out = torch.full((3,), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = fn(df.data_ptr(), dct.data_ptr(), nx, ny, nz, *c, 2 * R, out.data_ptr(), stream)
        assert rc == 0, wa._lib.last_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print(f"{nx}x{ny}x{nz}: noalloc {got}, staged export {ref}")
        assert np.all(np.isfinite(got)) and abs(got[0]) > 0.0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_rejects_bad_arguments():
    import torch

    wa, fn = _entry()
    buf = torch.zeros(64, dtype=torch.float32, device="cuda")
    p = buf.data_ptr()
    for args in [(None, p, 4, 4, 4), (p, p, 0, 4, 4), (p, p, 2048, 2048, 1024)]:
        rc = fn(args[0], args[1], args[2], args[3], args[4], 0.0, 0.0, 0.0, 1.0, p, None)
        assert rc != 0
        assert "noalloc" in wa._lib.last_error()
        wa._lib.lib().wake_hip_clear_error()
    torch.cuda.synchronize()
