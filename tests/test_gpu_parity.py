"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

Bars:
  * FP64 storage: populations, density and velocity bit-identical to the
    oracle after every case's steps (same operation order, contraction off).
  * FP32 storage: with eps = 2^-24 and K = 64 * eps * sqrt(steps):
    max|df| <= K * max|f|, max|drho| <= K * max|rho|, max|du| <= K (velocity
    in lattice units, where the populations and rho are O(1): u is a
    difference of O(0.1) populations, so its FP32 error is absolute, not
    relative to |u|).
  * FP16S / BF16S shifted storage: stated per test below.
  * node masks, class tables and bounceback link sets: bit-exact.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

lbmc = pytest.importorskip("lbmc")
import pyoracle  # noqa: E402


def run_pair(case, storage, steps=None, store_every=1):
    steps = case["steps"] if steps is None else steps
    o = cases.build(pyoracle.Oracle, case)
    g = cases.build(lbmc.Lattice, case, storage=storage)
    o.run(0, steps, store_every)
    g.run(0, steps, store_every)
    return o, g


CASES = cases.all_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fp64_bitwise(case):
    print(f"[fp64 bitwise] {case['name']}", flush=True)
    o, g = run_pair(case, lbmc.FP64)
    fo, fg = o.get_f(), g.get_f()
    ro, uo = o.get_macros()
    rg, ug = g.get_macros()
    assert np.array_equal(fo, fg, equal_nan=True), \
        f"f differs at {np.argwhere(fo != fg)[:5]} max {np.nanmax(np.abs(fo - fg))}"
    assert np.array_equal(ro, rg, equal_nan=True)
    assert np.array_equal(uo, ug, equal_nan=True)


EPS32 = 2.0 ** -24


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fp32_tolerance(case):
    steps = case["steps"]
    o, g = run_pair(case, lbmc.FP32)
    fo, fg = o.get_f(), g.get_f()
    ro, uo = o.get_macros()
    rg, ug = g.get_macros()
    assert np.all(np.isfinite(fo)), "oracle diverged: case is not a parity case"
    K = 64 * EPS32 * np.sqrt(steps)
    df = np.max(np.abs(fg - fo)) / np.max(np.abs(fo))
    dr = np.max(np.abs(rg - ro)) / max(np.max(np.abs(ro)), 1e-300)
    du = np.max(np.abs(ug - uo))
    umax = np.max(np.abs(uo))
    print(f"[fp32] {case['name']}: steps {steps} df/max|f| {df:.2e} drho/max|rho| {dr:.2e} "
          f"max|du| {du:.2e} (max|u| {umax:.2e}) bound {K:.2e}", flush=True)
    assert df <= K
    assert dr <= K
    assert du <= K


@pytest.mark.parametrize("case", [cases.turbulence_d3q15(L=10, steps=12), cases.pressure_d2q9(steps=50),
                                  cases.faces_d3q15(L=(64, 6, 5), steps=12),
                                  cases.pressure_d2q9(Lx=64, Ly=24, steps=50)],
                         ids=["turb", "pressure", "faces_lx64", "pressure_lx64"])
def test_wide_offset_kernels(case, monkeypatch):
    """The 64-bit-offset kernels (used when a direction array spans 4 GiB or
    more) forced on a small lattice: FP64 bit-exact, FP32 equal to the
    32-bit-offset kernels bit for bit.  With Lx % 64 == 0 (the lx64 cases)
    they address each direction's slot row from a wave-uniform base."""
    o, g32 = run_pair(case, lbmc.FP32)
    monkeypatch.setenv("LBMC_FORCE_WIDE", "1")
    g64 = cases.build(lbmc.Lattice, case, storage=lbmc.FP64)
    g64.run(0, case["steps"], 1)
    assert np.array_equal(o.get_f(), g64.get_f())
    gw = cases.build(lbmc.Lattice, case, storage=lbmc.FP32)
    gw.run(0, case["steps"], 1)
    assert np.array_equal(g32.get_f(), gw.get_f())


def test_node_types_and_links_bitexact():
    for case in CASES:
        print(f"[mask] {case['name']}", flush=True)
        o = cases.build(pyoracle.Oracle, case)
        g = cases.build(lbmc.Lattice, case, storage=lbmc.FP32)
        mo, to = o.node_types()
        mg, tg = g.node_types()
        assert np.array_equal(to, tg), case["name"]
        assert np.array_equal(mo, mg), case["name"]
        assert np.array_equal(o.links(), g.links()), case["name"]


@pytest.mark.parametrize("storage,tol", [(2, 2e-3), (3, 2e-2)], ids=["fp16s", "bf16s"])
def test_shifted_storage_turbulence(storage, tol):
    """Shifted half-precision storage on the decaying-turbulence case: velocity
    error relative to u_max after 40 steps."""
    case = cases.turbulence_d3q15(L=16, steps=40)
    o, g = run_pair(case, storage)
    ro, uo = o.get_macros()
    rg, ug = g.get_macros()
    umax = np.max(np.abs(uo))
    err = np.max(np.abs(ug - uo)) / umax
    print(f"shifted storage {storage}: max |du|/umax = {err:.3e}", flush=True)
    assert err <= tol
    assert np.max(np.abs(rg - ro)) <= tol * 0.1


def test_store_macros_flag():
    """Macros are written only on store calls; prescribed values persist."""
    case = cases.macro_codes_d2q9()
    o, g = run_pair(case, lbmc.FP64, steps=7, store_every=3)
    assert np.array_equal(o.get_macros()[0], g.get_macros()[0])
    assert np.array_equal(o.get_macros()[1], g.get_macros()[1])


def test_time_index_wraps():
    """Large time indices (t mod L per axis) give the same slots as the oracle."""
    case = cases.turbulence_d3q15(L=6, steps=3)
    o = cases.build(pyoracle.Oracle, case)
    g = cases.build(lbmc.Lattice, case, storage=lbmc.FP64)
    for t in (0, 1, 5, 6, 7, 1_000_003, 2**31 + 5):
        o.step(t, True)
        g.step(t, True)
    assert np.array_equal(o.get_f(), g.get_f())


def test_kinetic_energy_matches_oracle():
    case = cases.turbulence_d3q15(L=12, steps=5)
    o, g = run_pair(case, lbmc.FP64)
    eo, eg = o.kinetic_energy(), g.kinetic_energy()
    assert abs(eg - eo) <= 1e-12 * abs(eo)


def test_get_node_and_set_f_roundtrip():
    case = cases.channel_d2q9(steps=10)
    o, g = run_pair(case, lbmc.FP64)
    rho, u = g.get_macros()
    for idx in (0, 37, g.n - 1):
        r1, u1 = g.get_node(idx)
        assert r1 == rho[idx] and np.array_equal(u1, u[:, idx])
    f = g.get_f()
    g.set_f(f * 1.0)
    assert np.array_equal(g.get_f(), f)


def test_invalid_codes_rejected():
    n = 8 * 6
    lat = lbmc.Lattice(lbmc.D2Q9, (8, 6), micro_bc=True, macro_bc=True)
    micro = np.zeros(n, dtype=np.int32)
    micro[3] = 3  # a NULL slot of the reference's D2Q9 table
    with pytest.raises(lbmc.LbmcError) as e:
        lat.set_fields(geometry=np.zeros(n), micro_bc=micro)
    assert e.value.status == 1
    macro = np.zeros(n, dtype=np.int32)
    macro[5] = 16
    with pytest.raises(lbmc.LbmcError):
        lat.set_fields(geometry=np.zeros(n), macro_bc=macro)
    with pytest.raises(lbmc.LbmcError) as e:
        lbmc.Lattice(lbmc.D2Q9, (0, 6))
    assert e.value.status == 1


def test_step_before_fields_is_state_error():
    lat = lbmc.Lattice(lbmc.D2Q9, (8, 6))
    with pytest.raises(lbmc.LbmcError) as e:
        lat.step(0)
    assert e.value.status == 4


@pytest.mark.parametrize("cap", ["1", "2", "16"])
def test_residency_cap_leaves_results_unchanged(cap, monkeypatch):
    """The residency cap (LBMC_STEP_BLOCKS_PER_CU, dynamic LDS the step
    kernels do not use) only changes how many blocks run at once: FP32 and
    FP16S results are bit-identical to the uncapped run."""
    case = cases.turbulence_d3q15(L=16, steps=9)
    for storage in (lbmc.FP32, lbmc.FP16S):
        out = []
        for c in ("0", cap):
            monkeypatch.setenv("LBMC_STEP_BLOCKS_PER_CU", c)
            g = cases.build(lbmc.Lattice, case, storage=storage)
            g.run(0, case["steps"], 3)
            out.append(g.get_f())
            g.close()
        assert np.array_equal(out[0], out[1]), (cap, storage)
