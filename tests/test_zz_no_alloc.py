"""lbmc_get_f_noalloc (no device allocation, fixed device scratch) against
lbmc_get_f, both called through the C ABI: bit-identical populations on two
calls in a row, the first on a lattice that spans three scratch chunks (the
last one partial), the second on a smaller lattice of another lattice type
and storage format, so the second call reuses the scratch that still holds
the first call's data."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lbmc = pytest.importorskip("lbmc")

CHUNK = 1 << 20  # slots per scratch chunk (8 MiB of doubles)


def _read_both(lat):
    lib = lbmc.lib()
    ref = [np.full(lat.n, np.nan) for _ in range(lat.q)]
    out = [np.full(lat.n, np.nan) for _ in range(lat.q)]
    ref_p = (ct.POINTER(ct.c_double) * 15)(*[a.ctypes.data_as(ct.POINTER(ct.c_double)) for a in ref])
    out_p = (ct.POINTER(ct.c_double) * 15)(*[a.ctypes.data_as(ct.POINTER(ct.c_double)) for a in out])
    assert lib.lbmc_get_f(lat._h, ref_p) == 0, lib.lbmc_last_error(lat._h).decode()
    assert lib.lbmc_get_f_noalloc(lat._h, out_p) == 0, lib.lbmc_last_error(lat._h).decode()
    return np.stack(ref), np.stack(out)


def _turbulent_d3q15(dims, storage, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    lat = lbmc.Lattice(lbmc.D3Q15, dims, lbmc.BGK, tau=0.55, c_smag=0.17, storage=storage)
    lat.set_fields(geometry=np.zeros(n))
    u = [0.03 * rng.standard_normal(n) for _ in range(3)]
    lat.init_equilibrium(1.0 + 1e-3 * rng.standard_normal(n), u)
    lat.run(0, 7, 0)
    return lat


def _channel_d2q9(dims, storage, seed):
    rng = np.random.default_rng(seed)
    Lx, Ly = dims
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1.0
    geo[-1, :] = 1.0
    geo[rng.random((Ly, Lx)) < 0.05] = 1.0
    lat = lbmc.Lattice(lbmc.D2Q9, dims, lbmc.BGK, forcing=True, tau=0.7, storage=storage)
    lat.set_fields(geometry=geo.reshape(-1), force=[np.full(Lx * Ly, 1e-5), np.zeros(Lx * Ly)])
    lat.init_static()
    lat.run(0, 9, 0)
    return lat


def test_noalloc_matches_get_f_on_two_calls():
    # call 1: 2,129,920 nodes = two full chunks and a partial third
    dims1 = (128, 128, 130)
    print(f"[no-alloc] call 1: D3Q15 FP32 {dims1}", flush=True)
    lat = _turbulent_d3q15(dims1, lbmc.FP32, 11)
    assert lat.n > 2 * CHUNK and lat.n % CHUNK != 0
    ref, out = _read_both(lat)
    lat.close()
    assert np.all(np.isfinite(ref))
    assert np.array_equal(ref, out), f"differs at {np.argwhere(ref != out)[:5]}"

    # call 2: a smaller lattice (one partial chunk), other type and storage
    dims2 = (200, 37)
    print(f"[no-alloc] call 2: D2Q9 FP16S {dims2}", flush=True)
    lat = _channel_d2q9(dims2, lbmc.FP16S, 12)
    assert lat.n < CHUNK
    ref, out = _read_both(lat)
    lat.close()
    assert np.all(np.isfinite(ref))
    assert np.array_equal(ref, out), f"differs at {np.argwhere(ref != out)[:5]}"


@pytest.mark.parametrize("storage,dims", [(lbmc.FP64, (64, 48, 20)), (lbmc.BF16S, (64, 48, 20)),
                                          (lbmc.FP16S, (128, 128, 130))],
                         ids=["fp64", "bf16s", "fp16s-3chunks"])
def test_noalloc_other_storages(storage, dims):
    # the 16-bit D3Q15 lattices (even last axis) are in the half-interleaved layout
    print(f"[no-alloc] D3Q15 {lbmc.STORAGE_NAMES[storage]} {dims}", flush=True)
    lat = _turbulent_d3q15(dims, storage, 13)
    ref, out = _read_both(lat)
    lat.close()
    assert np.array_equal(ref, out)


def test_noalloc_null_arguments():
    lib = lbmc.lib()
    assert lib.lbmc_get_f_noalloc(None, None) == 1
    lat = _channel_d2q9((16, 8), lbmc.FP32, 14)
    assert lib.lbmc_get_f_noalloc(lat._h, None) == 1
    lat.close()
