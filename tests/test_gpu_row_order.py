"""Block order of the one-block-per-chunk step kernels (32- and 64-bit
offsets) on lattices whose x-rows hold R = 2 or 4 chunks of 256 nodes: the
library gives each XCD min(R, 4) consecutive chunks (node_index.hpp
row_chunk), so that the cache
line a rotated row splits at a chunk boundary stays in one L2.  The order
changes which block updates a node, not the update: the FP64 step stays
bit-identical to the oracle on such lattices, and every storage, the
half-interleaved 16-bit kernel and the closure's reuse kernel give the same
bits as the plain order (LBMC_ROW_ORDER=0, read at create)."""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

lbmc = pytest.importorskip("lbmc")
import pyoracle  # noqa: E402


def box_d3q15(dims, seed=11, collision=cases.BGK, c_smag=0.17, steps=7):
    """Periodic D3Q15 box of any shape from a seeded random velocity field."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    u = [0.05 * rng.standard_normal(n) for _ in range(3)]
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    return dict(
        name=f"box_d3q15_{'x'.join(map(str, dims))}_c{collision}",
        desc=dict(lattice=cases.D3Q15, length=dims, collision_type=collision, tau=0.51,
                  tau_mrt=cases.d3q15_rates(0.51), c_smag=c_smag),
        fields=dict(geometry=np.zeros(n)),
        init=("equilibrium", rho, u), steps=steps)


# x-rows of 512 (2 chunks) and 1024 nodes (4 chunks); 30 and 15 rows, so the
# last group of 8 rows is partial
ROW_CASES = [
    box_d3q15((512, 6, 5)),
    box_d3q15((1024, 5, 3), collision=cases.MRT, c_smag=0.0),
    cases.faces_d3q15(L=(512, 6, 4), steps=6),
    cases.channel_d2q9(Lx=1024, Ly=40, steps=9),
    cases.pressure_d2q9(Lx=512, Ly=24, steps=9),
]


def _build(case, storage, env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return cases.build(lbmc.Lattice, case, storage=storage)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("wide", [False, True], ids=["narrow", "wide"])
@pytest.mark.parametrize("case", ROW_CASES, ids=[c["name"] for c in ROW_CASES])
def test_row_order_fp64_matches_oracle(case, wide):
    """Narrow kernels, and the 64-bit-offset row kernels forced on the same
    lattices (LBMC_FORCE_WIDE=1), in the row order."""
    print(f"[row order fp64] {case['name']} wide={wide}", flush=True)
    o = cases.build(pyoracle.Oracle, case)
    g = _build(case, lbmc.FP64, {"LBMC_FORCE_WIDE": "1"} if wide else {})
    for t0 in (0, 2 ** 31 + 5):
        o.run(t0, case["steps"], 3)
        g.run(t0, case["steps"], 3)
        assert np.array_equal(o.get_f(), g.get_f(), equal_nan=True), t0
        ro, uo = o.get_macros()
        rg, ug = g.get_macros()
        assert np.array_equal(ro, rg, equal_nan=True) and np.array_equal(uo, ug, equal_nan=True), t0
    g.close()


@pytest.mark.parametrize("storage", [lbmc.FP32, lbmc.FP16S, lbmc.BF16S], ids=["fp32", "fp16s", "bf16s"])
def test_row_order_matches_plain_order_bitwise(storage):
    for case in ROW_CASES:
        for t0 in (0, 1_000_003):
            print(f"[row order] {case['name']} storage={storage} t0={t0}", flush=True)
            out = []
            for env in ({}, {"LBMC_ROW_ORDER": "0"}, {"LBMC_FORCE_WIDE": "1"},
                        {"LBMC_FORCE_WIDE": "1", "LBMC_ROW_ORDER": "0"}):
                lat = _build(case, storage, env)
                lat.run(t0, case["steps"], 4)
                out.append((lat.get_f(), lat.get_macros()))
                lat.close()
            fa, (ra, ua) = out[0]
            for k, (fb, (rb, ub)) in enumerate(out[1:]):
                assert np.array_equal(fa, fb), (case["name"], k)
                assert np.array_equal(ra, rb) and np.array_equal(ua, ub), (case["name"], k)


def test_row_order_closure_reuse_bitwise():
    """The closure's reuse calls (recorded tau between evaluations) on a
    512-wide box, interval 3: same bits in either block order."""
    import closure

    case = box_d3q15((512, 4, 4), steps=10)
    mu = np.zeros(7, dtype=np.float32)
    mu[6] = 1.0
    isd = np.full(7, 1e3, dtype=np.float32)
    isd[6] = 1e2
    model = lbmc.Mlp(closure.init_params(5), mu=mu, isd=isd, out_a=0.005, out_b=0.515)
    out = []
    for env in ({}, {"LBMC_ROW_ORDER": "0"}):
        lat = _build(case, lbmc.FP32, env)
        lat.set_closure(model)
        lat.set_closure_interval(3)
        lat.run(0, case["steps"], 5)
        out.append((lat.get_f(), lat.get_macros()))
        lat.close()
    model.close()
    (fa, (ra, ua)), (fb, (rb, ub)) = out
    assert np.array_equal(fa, fb)
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
