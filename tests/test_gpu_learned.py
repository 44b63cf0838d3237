"""GPU tests of the learned warm start: MFMA forward/backward against the
oracle's MLP (with the same BF16 operand rounding), features bit-exact
against the oracle, recording, training and the warm-started step.

Tolerances: the MFMA accumulation order inside a 16x16x16 product and
expf differ from the oracle's FP64 sums, so network outputs and gradients
are compared to 1e-3 of their scale (outputs) and 1e-2 in norm (gradients);
features are plain FP32 neighbour sums in the canonical order: bit-exact.
"""
import numpy as np
import pytest

from psb_project_amd import learned, scenes
from psb_project_amd._lib import WARM_MLP, WARM_NONE

pytestmark = pytest.mark.gpu


def he_params(seed=0):
    from test_mlp_oracle import he_params as hp
    return hp(seed)


def test_mfma_forward_matches_oracle(ctx, oracle):
    rng = np.random.default_rng(7)
    n = 1000                                   # not a multiple of 16: ragged last tile
    feat = rng.normal(size=(8, n)).astype(np.float32)
    w = he_params(8)
    mean = rng.normal(scale=0.2, size=8).astype(np.float32)
    std = rng.uniform(0.5, 2.0, 8).astype(np.float32)
    m = learned.PressureMLP(ctx, seed=3)
    m.set_params(w)
    m.set_normalization(mean, std, 1.0)
    assert np.array_equal(m.params(), w)
    dfeat = ctx.to_device(feat)
    dout = ctx.array(4 * n)
    m.forward(dfeat.ptr, n, dout.ptr)
    got = dout.download(np.float32, n)
    ref = oracle.mlp_forward(w, mean, std, feat, emulate_bf16=True)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-3 * scale


def test_mfma_gradients_match_oracle(ctx, oracle):
    rng = np.random.default_rng(9)
    ns, batch = 3000, 500
    feat = rng.normal(size=(8, ns)).astype(np.float32)
    target = rng.normal(size=ns).astype(np.float32)
    idx = rng.choice(ns, batch, replace=False).astype(np.uint32)
    w = he_params(10)
    mean, std = np.zeros(8, np.float32), np.ones(8, np.float32)
    m = learned.PressureMLP(ctx)
    m.set_params(w)
    m.set_normalization(mean, std, 1.0)
    df, dt, di = ctx.to_device(feat), ctx.to_device(target), ctx.to_device(idx)
    loss = m.train_step(df.ptr, dt.ptr, ns, di.ptr, batch, apply_adam=False)
    g = m.grad().astype(np.float64)
    gref, lref = oracle.mlp_grad(w, mean, std, feat, target, idx, emulate_bf16=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    for name, a, b in (("W1", 0, 512), ("b1", 512, 576), ("W2", 576, 4672),
                       ("b2", 4672, 4736), ("W3", 4736, 4800), ("b3", 4800, 4801)):
        err = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        assert err < 1e-2, f"{name}: relative error {err}"


def test_training_is_deterministic(ctx):
    """Every wave stores its gradient sums in its own row and one kernel adds
    the rows in wave order, so two training runs on the same data from the
    same initial network end bit for bit equal (the loss too)."""
    rng = np.random.default_rng(11)
    ns, batch = 40000, 8192
    feat = rng.normal(size=(8, ns)).astype(np.float32)
    target = rng.normal(size=ns).astype(np.float32)
    df, dt = ctx.to_device(feat), ctx.to_device(target)
    idx = [ctx.to_device(np.random.default_rng(b).choice(ns, batch, replace=False)
                         .astype(np.uint32)) for b in range(6)]
    runs = []
    for _ in range(2):
        m = learned.PressureMLP(ctx, seed=4)
        losses = [m.train_step(df.ptr, dt.ptr, ns, di.ptr, batch) for di in idx]
        runs.append((m.params(), losses))
        m.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
    assert not np.array_equal(runs[0][0], learned.PressureMLP(ctx, seed=4).params())


def test_features_and_recording_bitwise(ctx, oracle):
    from psb_project_amd.system import FluidSystem
    sc = scenes.pcisph_dam_break(n_side=16)
    g = FluidSystem(ctx, sc.params, sc.n)
    o = oracle.OracleSystem(sc.params, sc.n)
    ds = learned.Dataset(ctx, 10 * sc.n)
    g.set_particles(sc.pos4)
    o.set_particles(sc.pos4)
    o.enable_features(True)
    g.set_recorder(ds, stride=3)
    for step in range(4):
        g.update(1)
        o.update(1)
        fg, fo = g.features(), o.features()
        assert np.array_equal(fg, fo), f"step {step}: max diff {np.abs(fg - fo).max()}"
        assert g.stats()["iterations"] == o.stats()["iterations"]
    gs = g.state()
    feat, tgt = ds.download()
    n = sc.n
    off = 3 % 3                               # offset of the last recorded step (step index 3)
    sel = np.arange(off, n, 3)
    last = slice(feat.shape[1] - sel.size, feat.shape[1])
    assert np.array_equal(feat[:, last], fg[:, sel])
    assert np.allclose(tgt[last], gs["pressure"][sel] / np.float32(g.pressure_scale), rtol=1e-6)
    assert g.pressure_scale == pytest.approx(o.pressure_scale, rel=1e-6)


def test_recording_windows_start_where_asked(ctx):
    """record() with one warm-up count per scene: the second scene's samples
    are exactly those of a system stepped 5 steps unrecorded, then 3
    recorded; the first scene's start at step 0."""
    from psb_project_amd.system import FluidSystem
    sc = scenes.pcisph_dam_break(n_side=16)
    stride, steps = 4, 3
    per = -(-sc.n // stride) * steps
    ds = learned.Dataset(ctx, 2 * per)
    learned.record(ctx, [sc, sc], steps, ds, stride=stride, warmup=[0, 5])
    assert ds.count == 2 * per
    feat, tgt = ds.download()
    for warm, block in ((0, slice(0, per)), (5, slice(per, 2 * per))):
        ref = learned.Dataset(ctx, per)
        g = FluidSystem(ctx, sc.params, sc.n)
        g.set_particles(sc.pos4, sc.vel4, sc.vh4)
        if warm:
            g.update(warm)
        g.set_recorder(ref, stride=stride)
        g.update(steps)
        g.close()
        rf, rt = ref.download()
        assert np.array_equal(feat[:, block], rf) and np.array_equal(tgt[block], rt), warm
        ref.close()
    ds.close()


def test_training_reduces_loss(ctx):
    sc = scenes.pcisph_dam_break(n_side=16)
    ds = learned.Dataset(ctx, 400_000)
    learned.record(ctx, [sc], steps=40, dataset=ds, stride=2)
    assert ds.count > 50_000
    from psb_project_amd.system import FluidSystem
    ps = FluidSystem(ctx, sc.params, sc.n).pressure_scale
    model, losses = learned.train(ctx, ds, ps, epochs=4, batch=4096, lr=3e-3)
    assert losses[-1] < 0.5 * losses[0]


def test_mlp_warm_start_step_matches_oracle(ctx, oracle):
    from psb_project_amd.system import FluidSystem
    sc = scenes.pcisph_dam_break(n_side=16, warm_start=WARM_MLP)
    w = he_params(11)
    w[4736:4801] *= 0.05                       # small pressures: a well-posed start
    mean, std = np.zeros(8, np.float32), np.ones(8, np.float32)
    m = learned.PressureMLP(ctx)
    m.set_params(w)
    g = FluidSystem(ctx, sc.params, sc.n)
    m.set_normalization(mean, std, g.pressure_scale)
    g.set_mlp(m)
    o = oracle.OracleSystem(sc.params, sc.n)
    o.set_mlp(w, mean, std)
    g.set_particles(sc.pos4)
    o.set_particles(sc.pos4)
    g.update(1)
    o.update(1)
    # the network outputs differ at ~1e-3 (MFMA vs FP64 sums), so the
    # corrector may cross the tolerance one iteration apart
    gi, oi = g.stats()["iterations"], o.stats()["iterations"]
    assert abs(gi - oi) <= 1
    gs, os_ = g.state(), o.state()
    assert np.array_equal(gs["ids"], os_["ids"])
    if gi == oi:
        assert np.abs(gs["pos"][:, :3] - os_["pos"][:, :3]).max() < 1e-5
    assert g.stats()["max_error"] <= 0.01


def test_learned_warm_start_cuts_iterations(ctx):
    """The warm-start protocol at half size: train on plain-PCISPH converged
    pressures of 8 seeded 64^3 dam breaks (200 steps each, every 16th
    particle, 4 epochs), then run 2 held-out seeds: the learned start needs at
    least 25 % fewer corrector iterations than the zero start, every step
    still within the 1 % tolerance (the corrector is unchanged).  This covers
    200-step held-out horizons at half the benchmark's training data; the
    benchmark's full protocol reaches about 40 % there (41-46 % over its last
    runs) and about 21 % over the full 2000-step dam break, whose far-wall
    impact lies outside the training horizon."""
    def make(i):
        return scenes.pcisph_dam_break(scene_id=200 + i)
    model, summary = learned.train_and_evaluate(ctx, make, range(8), range(8, 10), steps=200,
                                                stride=16, epochs=4)
    ev = summary["eval"]
    plain, lrn = ev["plain"], ev["learned"]
    assert plain["steps_within_tolerance"] == 1.0 and lrn["steps_within_tolerance"] == 1.0
    assert plain["iters_per_step"] > 5
    assert lrn["iters_per_step"] <= 0.75 * plain["iters_per_step"], ev
    model.close()
