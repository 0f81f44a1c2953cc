"""Batched stepping (lbmc_batch_*): every member of a batch must end where
the same lattice stepped on its own ends -- bit for bit against the FP64
oracle with FP64 storage (each member a different collision code, boundary
mix and size), and against the single-lattice kernels with FP32 storage;
also from a large t0 (index-shift wrap), across graph replays, and with the
autotuned grid."""
import os

import numpy as np
import pytest

import cases
import lbmc
import pyoracle

pytestmark = pytest.mark.gpu


def _members(lattice):
    out = []
    for c in cases.all_cases():
        if c["desc"]["lattice"] != lattice or "sf" in c["name"]:
            continue
        out.append(c)
    return out


def _run_batch(group, storage, t0, steps, store_every=1, shape=None):
    lats = [cases.build(lbmc.Lattice, c, storage=storage) for c in group]
    b = lbmc.Batch(lats)
    if shape is not None:
        b.set_shape(*shape)
    b.run(t0, steps, store_every)
    out = [(lat.get_f(), lat.get_macros()) for lat in lats]
    b.close()
    for lat in lats:
        lat.close()
    return out


def _oracle(group, t0, steps, store_every=1):
    out = []
    for c in group:
        o = cases.build(pyoracle.Oracle, c)
        o.run(t0, steps, store_every)
        out.append((o.get_f(), o.get_macros()))
        o.close()
    return out


@pytest.mark.parametrize("lattice", [lbmc.D2Q9, lbmc.D3Q15])
@pytest.mark.parametrize("t0", [0, 2 ** 31 + 5])
def test_batch_fp64_matches_oracle_bitwise(lattice, t0):
    group = _members(lattice)
    print(f"[batch] FP64 lattice={lattice} t0={t0} members={len(group)}", flush=True)
    steps = 45  # one graph replay of 32 calls plus 13 single launches
    got = _run_batch(group, lbmc.FP64, t0, steps)
    ref = _oracle(group, t0, steps)
    for c, (fg, (rg, ug)), (fr, (rr, ur)) in zip(group, got, ref):
        assert np.array_equal(fg, fr), c["name"]
        assert np.array_equal(rg, rr), c["name"]
        assert np.array_equal(ug, ur), c["name"]


@pytest.mark.parametrize("storage", [lbmc.FP32, lbmc.FP16S, lbmc.BF16S], ids=["fp32", "fp16s", "bf16s"])
def test_batch_matches_single_lattices(storage):
    """Same node arithmetic and, for the 16-bit formats, the same stochastic
    rounding words (seeded by node and the call's shift counters): a batch
    ends where its members end on their own."""
    group = _members(lbmc.D2Q9)
    print(f"[batch] storage={storage} members={len(group)}", flush=True)
    steps = 70
    got = _run_batch(group, storage, 0, steps, store_every=7)
    for c, (fg, (rg, ug)) in zip(group, got):
        lat = cases.build(lbmc.Lattice, c, storage=storage)
        lat.run(0, steps, 7)
        fs = lat.get_f()
        rs, us = lat.get_macros()
        lat.close()
        assert np.max(np.abs(fg - fs)) <= 1e-6, c["name"]
        assert np.max(np.abs(ug - us)) <= 1e-6, c["name"]


@pytest.mark.parametrize("t0", [0, 2 ** 31 + 5])
@pytest.mark.parametrize("storage", [lbmc.FP16S, lbmc.BF16S], ids=["fp16s", "bf16s"])
@pytest.mark.parametrize("lattice", [lbmc.D2Q9, lbmc.D3Q15], ids=["d2q9", "d3q15"])
def test_pair_steps_match_one_node_steps_bitwise(lattice, storage, t0):
    """A 16-bit lattice stepped on its own takes two nodes per lane: on every
    call in the half-interleaved layout (last axis even), else on calls with
    an even x shift (every case here has an even Lx); a batch member takes
    one node per lane on every call, with the same per-node update and
    stochastic-rounding words: populations and macros bit-identical.  So
    is the same lattice with LBMC_STEP_PAIRS=0 (read at create: one node per
    lane on every call)."""
    group = _members(lattice)
    print(f"[pairs] lattice={lattice} storage={storage} t0={t0} members={len(group)}", flush=True)
    assert all(c["desc"]["length"][0] % 2 == 0 for c in group)
    steps = 21
    got = _run_batch(group, storage, t0, steps, store_every=7)
    for c, (fg, (rg, ug)) in zip(group, got):
        lat = cases.build(lbmc.Lattice, c, storage=storage)
        lat.run(t0, steps, 7)
        fs = lat.get_f()
        rs, us = lat.get_macros()
        lat.close()
        assert np.array_equal(fg, fs), c["name"]
        assert np.array_equal(rg, rs) and np.array_equal(ug, us), c["name"]
        os.environ["LBMC_STEP_PAIRS"] = "0"
        try:
            lat = cases.build(lbmc.Lattice, c, storage=storage)
        finally:
            del os.environ["LBMC_STEP_PAIRS"]
        lat.run(t0, steps, 7)
        assert np.array_equal(lat.get_f(), fs), c["name"]
        lat.close()


def _run_env(c, storage, t0, steps, env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        lat = cases.build(lbmc.Lattice, c, storage=storage)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    lat.run(t0, steps, 7)
    out = lat.get_f(), lat.get_macros()
    lat.close()
    return out


@pytest.mark.parametrize("storage", [lbmc.FP16S, lbmc.BF16S], ids=["fp16s", "bf16s"])
def test_half_layout_matches_plain_layout_bitwise(storage):
    """16-bit lattices with an even last axis in the half-interleaved layout
    (LBMC_HALF_LAYOUT=1, read at create; the default on D3Q15) keep slots k
    and k + N/2 in one 32-bit word and step two nodes per lane on every call;
    with LBMC_HALF_LAYOUT=0 slot k stays at element k and pairs run on even
    x shifts only; with LBMC_STEP_PAIRS=0 as well, one node per lane.  Every
    case (sf faces, boundary codes and odd last axes included), from t0 = 0
    and from a large t0: populations (read back in slot order) and macros
    bit-identical, and the default equals them."""
    for c in cases.all_cases():
        for t0 in (0, 2 ** 31 + 5):
            steps = min(c["steps"], 23)
            print(f"[half layout] {c['name']} storage={storage} t0={t0}", flush=True)
            fh, (rh, uh) = _run_env(c, storage, t0, steps, {"LBMC_HALF_LAYOUT": "1"})
            for env in ({}, {"LBMC_HALF_LAYOUT": "0"}, {"LBMC_HALF_LAYOUT": "1", "LBMC_STEP_PAIRS": "0"},
                        {"LBMC_HALF_LAYOUT": "0", "LBMC_STEP_PAIRS": "0"}):
                fp, (rp, up) = _run_env(c, storage, t0, steps, env)
                assert np.array_equal(fh, fp), (c["name"], env)
                assert np.array_equal(rh, rp) and np.array_equal(uh, up), (c["name"], env)


def test_half_layout_set_f_round_trip():
    """set_f then get_f returns the same populations in either 16-bit layout
    (values that the storage holds exactly), and a half-layout lattice read
    back equals a plain-layout lattice given the same populations."""
    rng = np.random.default_rng(3)
    for lattice, dims in ((lbmc.D3Q15, (12, 10, 8)), (lbmc.D2Q9, (30, 14))):
        q = 15 if lattice == lbmc.D3Q15 else 9
        n = int(np.prod(dims))
        w = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8) if q == 15 else \
            np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
        g = np.round(rng.standard_normal((q, n)) * 64) / 2 ** 18   # exact in FP16S (2^12 g) and BF16S
        f = w[:, None] + g
        got = []
        for flag in ("1", "0"):
            saved = os.environ.get("LBMC_HALF_LAYOUT")
            os.environ["LBMC_HALF_LAYOUT"] = flag
            try:
                lat = lbmc.Lattice(lattice, dims, lbmc.BGK, tau=0.8, storage=lbmc.FP16S)
            finally:
                if saved is None:
                    del os.environ["LBMC_HALF_LAYOUT"]
                else:
                    os.environ["LBMC_HALF_LAYOUT"] = saved
            lat.set_fields(geometry=np.zeros(n))
            lat.set_f(f)
            got.append(lat.get_f())
            lat.close()
        assert np.array_equal(got[0], f) and np.array_equal(got[1], f)


def test_batch_store_schedule():
    # store_every = 10 over 25 calls: the stored macros are those of call 19
    group = _members(lbmc.D2Q9)[:3]
    got = _run_batch(group, lbmc.FP64, 0, 25, store_every=10)
    for c, (_, (rg, ug)) in zip(group, got):
        o = cases.build(pyoracle.Oracle, c)
        o.run(0, 20, 10)
        rr, ur = o.get_macros()
        o.close()
        assert np.array_equal(rg, rr) and np.array_equal(ug, ur), c["name"]


def test_autotune_then_run_matches_oracle():
    group = _members(lbmc.D2Q9)
    lats = [cases.build(lbmc.Lattice, c, storage=lbmc.FP64) for c in group]
    b = lbmc.Batch(lats)
    info = b.info()
    grid, ms = b.autotune(0, steps_per_trial=8)
    shape = b.shape()
    assert shape[0] == grid and ms > 0
    assert 1 <= grid <= b.info()["chunks"] <= info["chunks"]
    assert shape[1] in (256, 512, 1024, 2048) and shape[2] in (0, 4, 6)
    b.close()
    for lat in lats:
        lat.close()
    got = _run_batch(group, lbmc.FP64, 0, 40, shape=shape)
    ref = _oracle(group, 0, 40)
    for c, (fg, _), (fr, _) in zip(group, got, ref):
        assert np.array_equal(fg, fr), c["name"]


@pytest.mark.parametrize("shape", [(1 << 20, 256, 3), (1 << 20, 1024, 5), (7, 512, 0), (1 << 20, 2048, 0)],
                         ids=["cap3", "chunk1024_cap5", "grid7_chunk512", "chunk2048"])
def test_launch_shapes_match_oracle(shape):
    """Residency caps, multi-tile chunks (members smaller than a chunk, and
    chunks ending past a member's last node) and a small strided grid."""
    group = _members(lbmc.D3Q15)
    got = _run_batch(group, lbmc.FP64, 3, 37, shape=shape)
    ref = _oracle(group, 3, 37)
    for c, (fg, _), (fr, _) in zip(group, got, ref):
        assert np.array_equal(fg, fr), c["name"]


def test_batch_set_shape_rejects_bad_arguments():
    lats = [cases.build(lbmc.Lattice, c, storage=lbmc.FP32) for c in _members(lbmc.D2Q9)[:2]]
    b = lbmc.Batch(lats)
    for shape in [(0, 256, 0), (8, 100, 0), (8, 0, 0), (8, 256 * 65, 0), (8, 256, -1)]:
        with pytest.raises(lbmc.LbmcError) as e:
            b.set_shape(*shape)
        assert e.value.status == 1, shape
    b.set_shape(3, 768, 4)
    info = b.info()
    assert info["chunks"] == sum(-(-lat.n // 768) for lat in lats)
    assert b.shape() == (min(3, info["chunks"]), 768, 4)  # the grid never exceeds the chunks
    b.close()
    for lat in lats:
        lat.close()


def test_batch_rejects_mixed_members():
    a = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.BGK, tau=0.8, storage=lbmc.FP32)
    b = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.BGK, tau=0.8, storage=lbmc.FP64)
    a.set_fields()
    b.set_fields()
    with pytest.raises(lbmc.LbmcError):
        lbmc.Batch([a, b])
    c = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.BGK, tau=0.8, storage=lbmc.FP32)
    with pytest.raises(lbmc.LbmcError) as e:
        lbmc.Batch([a, c])  # c has no fields yet
    assert e.value.status == 4
    for x in (a, b, c):
        x.close()


@pytest.mark.parametrize("every,fp16", [(1, False), (3, False), (1, True)], ids=["every1", "every3", "every1-fp16"])
def test_batch_with_closure_members_matches_single_lattices(every, fp16):
    """D2Q9 members with the learned closure (one shared network, bf16 or
    fp16 operands; with and without Guo forcing; the network evaluated every
    call or every 3rd, with the recorded tau in between) beside an MRT
    member: each ends where it ends when stepped on its own, over two runs
    (the second starting from a valid record at a call that is not a
    multiple of 3)."""
    import pymlp
    import workloads

    print("[batch] closure members", flush=True)
    rng = np.random.default_rng(40)
    p = pymlp.init_params(40)
    p["b3"] = 0.1
    mu = np.zeros(7, dtype=np.float32)
    mu[6] = 1.0
    isd = np.full(7, 300.0, dtype=np.float32)
    isd[6] = 100.0
    model = lbmc.Mlp(pymlp.flatten(p), mu=mu, isd=isd, out_a=0.02, out_b=0.62)
    if fp16:
        model.set_eval_fp16(True)
    specs = []
    for k, (Lx, Ly) in enumerate([(40, 24), (72, 33), (17, 50)]):
        geo = workloads.channel(Lx, Ly)
        geo[Ly // 2, Lx // 3] = 1.0
        n = Lx * Ly
        u = [0.02 * rng.standard_normal(n), 0.02 * rng.standard_normal(n)]
        specs.append(dict(L=(Lx, Ly), geo=geo.reshape(-1), u=u, closure=k != 2, guo=k == 1, mrt=k == 2))

    def build(spec):
        lat = lbmc.Lattice(lbmc.D2Q9, spec["L"], lbmc.MRT if spec["mrt"] else lbmc.BGK,
                           forcing=spec["guo"] or spec["mrt"], tau=0.6, tau_mrt=cases.d2q9_rates(0.6),
                           c_smag=0.1, storage=lbmc.FP32)
        n = spec["L"][0] * spec["L"][1]
        force = [np.full(n, 1e-5), np.zeros(n)] if (spec["guo"] or spec["mrt"]) else None
        lat.set_fields(geometry=spec["geo"], force=force)
        lat.init_equilibrium(np.ones(n), spec["u"])
        if spec["closure"]:
            lat.set_closure(model, 0.6, 0.7)
            lat.set_closure_interval(every)
        return lat

    runs = [(0, 50), (50, 23)]
    lats = [build(s) for s in specs]
    b = lbmc.Batch(lats)
    for t0, n in runs:
        b.run(t0, n, 1)
    got = [(lat.get_f(), lat.get_macros()[1]) for lat in lats]
    b.close()
    for lat in lats:
        lat.close()
    for spec, (fg, ug) in zip(specs, got):
        lat = build(spec)
        for t0, n in runs:
            lat.run(t0, n, 1)
        fs = lat.get_f()
        us = lat.get_macros()[1]
        lat.close()
        assert np.max(np.abs(fg - fs)) <= 1e-6, spec["L"]
        assert np.max(np.abs(ug - us)) <= 1e-6, spec["L"]
    model.close()


def test_batch_accepts_members_on_the_64bit_offset_kernels(monkeypatch):
    """A member whose own steps take the 64-bit-offset kernels by choice
    (forced here; the library picks them for large D3Q15 FP32 lattices)
    while its arrays stay below 4 GiB joins a batch, which steps it with
    32-bit offsets: the same bits as a batch of ordinary members, and as the
    member stepped on its own."""
    group = _members(lbmc.D3Q15)[:3]
    print(f"[batch] forced-wide members={len(group)}", flush=True)
    ref = _run_batch(group, lbmc.FP32, 0, 9, store_every=3)
    monkeypatch.setenv("LBMC_FORCE_WIDE", "1")
    got = _run_batch(group, lbmc.FP32, 0, 9, store_every=3)
    for c, (fa, (ra, ua)), (fb, (rb, ub)) in zip(group, ref, got):
        assert np.array_equal(fa, fb), c["name"]
        assert np.array_equal(ra, rb) and np.array_equal(ua, ub), c["name"]
        lat = cases.build(lbmc.Lattice, c, storage=lbmc.FP32)
        lat.run(0, 9, 3)
        assert np.array_equal(lat.get_f(), fa), c["name"]
        lat.close()
