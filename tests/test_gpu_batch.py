"""Batched scenes on the GPU: every scene of a batch must evolve exactly as the
same scene run alone through psb_system (bitwise: same kernels, same
per-scene corrector), whatever the block sizes, with or without HIP graphs."""
import numpy as np
import pytest

from psb_project_amd import learned, scenes
from psb_project_amd._lib import WARM_MLP, WARM_NONE, WARM_PREVIOUS, PsbError
from psb_project_amd.batch import BatchSystem
from psb_project_amd.system import FluidSystem

pytestmark = pytest.mark.gpu

KEYS = ("pos", "vel", "vh", "density", "pressure", "ids")


def small_batch(warm=WARM_NONE, walls=False, n=5, seed=77):
    return scenes.batched_scenes(num_scenes=n, min_count=300, max_count=5000, seed=seed,
                                 warm_start=warm, lattice_walls=walls)


def assert_same(a, b, what=""):
    for k in KEYS:
        x, y = a[k], b[k]
        if x.ndim == 2:
            x, y = x[:, :3], y[:, :3]
        assert np.array_equal(x, y), f"{what} {k}: max diff {np.abs(x.astype(np.float64) - y).max()}"


def single_runs(ctx, scs, steps, model=None):
    out = []
    for sc in scs:
        fs = FluidSystem(ctx, sc.params, sc.n)
        if model is not None:
            fs.set_mlp(model)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        fs.update(steps)
        out.append((fs.state(), fs.stats()))
        fs.close()
    return out


@pytest.mark.parametrize("warm,walls", [(WARM_NONE, False), (WARM_PREVIOUS, False),
                                        (WARM_NONE, True)])
def test_batch_matches_single_systems(ctx, warm, walls):
    scs = small_batch(warm, walls)
    b = BatchSystem(ctx, scs)
    b.set_all()
    b.update(30)
    ref = single_runs(ctx, scs, 30)
    for i, (st, stats) in enumerate(ref):
        got = b.state(i)
        assert_same(got, st, f"scene {i}")
        bs = b.stats(i)
        assert bs["iterations"] == stats["iterations"]
        assert bs["total_iterations"] == stats["total_iterations"]
        assert bs["max_error"] == stats["max_error"]
    # batch-wide largest neighbour count over the update = max of the scenes'
    assert b.stats(0)["max_neighbors_seen"] == max(x[1]["max_neighbors_seen"] for x in ref) > 0
    b.close()


def test_batch_graph_replays_match_eager(ctx):
    scs = small_batch(WARM_NONE, False, n=4, seed=78)
    eager = BatchSystem(ctx, scs)
    graph = BatchSystem(ctx, scs)
    eager.set_all()
    graph.set_all()
    graph.set_graphs(True)
    for step in range(6):                 # every replay checked against the eager step
        eager.update(1)
        graph.update(1)
        for i in range(len(scs)):
            assert_same(graph.state(i), eager.state(i), f"step {step} scene {i}")
    eager.close()
    graph.close()


def test_autotune_keeps_results(ctx):
    scs = small_batch(WARM_NONE, False, n=4, seed=79)
    a = BatchSystem(ctx, scs)
    b = BatchSystem(ctx, scs)
    a.set_all()
    b.set_all()
    choice = b.autotune(steps=1)
    assert set(choice) == {"hash_count", "neighbors_density", "density_epilogue",
                           "pci_nonpressure", "pci_density", "pci_force"}
    assert all(abs(v) in (64, 128, 256, 512) for v in choice.values())
    assert all(v > 0 for k, v in choice.items() if k not in ("pci_density", "pci_force"))
    a.update(8)
    b.update(8)
    for i in range(len(scs)):
        assert_same(b.state(i), a.state(i), f"scene {i}")
    a.close()
    b.close()


@pytest.mark.parametrize("dens,force", [(-128, -512), (-512, -256), (64, -128)])
def test_two_threads_per_particle_keeps_results(ctx, dens, force):
    """The corrector passes with two threads per particle (even and odd
    neighbour positions, added in that order) equal the one-thread runs."""
    scs = small_batch(WARM_NONE, False, n=4, seed=81)
    a = BatchSystem(ctx, scs)
    b = BatchSystem(ctx, scs)
    a.set_all()
    b.set_all()
    sizes = dict.fromkeys(("hash_count", "neighbors_density", "density_epilogue",
                           "pci_nonpressure"), 128)
    a.set_block_sizes(dict(sizes, pci_density=256, pci_force=256))
    b.set_block_sizes(dict(sizes, pci_density=dens, pci_force=force))
    a.update(6)
    b.update(6)
    for i in range(len(scs)):
        assert_same(b.state(i), a.state(i), f"scene {i}")
        assert b.stats(i)["total_iterations"] == a.stats(i)["total_iterations"]
    with pytest.raises(PsbError):
        b.set_block_sizes(dict(sizes, pci_density=-64, pci_force=256))
    with pytest.raises(PsbError):
        b.set_block_sizes(dict(sizes, hash_count=-128, pci_density=256, pci_force=256))
    a.close()
    b.close()


def test_batch_wide_offsets_match_single(ctx):
    """One scene whose neighbour offsets overflow the 16-bit list (a thin
    180 x 100 x 4 slab) beside one where they fit: each equals its single
    run with the 32-bit list."""
    slab = scenes.pcisph_dam_break(n_side=4)
    slab = scenes.Scene("slab", slab.params, scenes.lattice((180, 100, 4), 0.005, seed=11))
    small = scenes.pcisph_dam_break(n_side=12)
    scs = [slab, small]
    b = BatchSystem(ctx, scs)
    b.set_all()
    b.update(3)
    for i, sc in enumerate(scs):
        fs = FluidSystem(ctx, sc.params, sc.n)
        fs.set_compact_list(False)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        fs.update(3)
        assert_same(b.state(i), fs.state(), f"scene {i}")
        assert b.stats(i)["total_iterations"] == fs.stats()["total_iterations"]
        fs.close()
    b.close()


def test_batch_learned_warm_start_matches_single(ctx):
    scs = small_batch(WARM_MLP, False, n=3, seed=80)
    rng = np.random.default_rng(3)
    w = rng.uniform(-0.3, 0.3, learned.num_params()).astype(np.float32)
    w[4736:4801] *= 0.05
    m = learned.PressureMLP(ctx)
    m.set_params(w)
    probe = FluidSystem(ctx, scs[0].params, scs[0].n)
    m.set_normalization(np.zeros(8), np.ones(8), probe.pressure_scale)
    probe.close()
    b = BatchSystem(ctx, scs)
    b.set_mlp(m)
    b.set_all()
    b.update(5)
    ref = single_runs(ctx, scs, 5, model=m)
    for i, (st, stats) in enumerate(ref):
        assert_same(b.state(i), st, f"scene {i}")
        assert b.stats(i)["total_iterations"] == stats["total_iterations"]
    b.close()
    m.close()


def test_batch_latency_samples(ctx):
    scs = small_batch(WARM_NONE, False, n=4, seed=81)
    b = BatchSystem(ctx, scs)
    ref = BatchSystem(ctx, scs)
    b.set_all()
    ref.set_all()
    b.set_latency(True)
    b.update(5)
    lat = b.latency()
    assert lat["samples"] == 5 * len(scs)
    assert 0 < lat["p50_ms"] <= lat["p99_ms"] <= lat["max_ms"]
    # every 3rd step sampled (steps 5..11 of the batch: 6 and 9); the others
    # run without events, with the same results
    b.set_latency(True, every=3)
    b.update(7)
    assert b.latency()["samples"] == 2 * len(scs)
    ref.update(12)
    for i in range(len(scs)):
        assert_same(b.state(i), ref.state(i), f"scene {i}")
    b.close()
    ref.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_retries_match_single_systems(ctx, graphs):
    """A 0.2 % tolerance with a 3-iteration cap makes scenes fail their
    steps (four of the five here); the batch redoes each failed scene alone as substeps (its other
    scenes untouched), so every scene still equals the single system with
    the same retries, bit for bit, with its retry counts; every scene-step
    gives one latency sample."""
    scs = scenes.batched_scenes(num_scenes=5, min_count=300, max_count=5000, seed=83,
                                tolerance=0.002, max_it=3)
    b = BatchSystem(ctx, scs)
    b.set_all()
    b.set_graphs(graphs)
    b.set_latency(True)
    b.update(40)
    ref = single_runs(ctx, scs, 40)
    retried = 0
    for i, (st, stats) in enumerate(ref):
        assert_same(b.state(i), st, f"scene {i}")
        bs = b.stats(i)
        for k in ("iterations", "total_iterations", "retried_steps", "substeps"):
            assert bs[k] == stats[k], (i, k, bs[k], stats[k])
        assert bs["max_error"] == stats["max_error"]
        assert bs["worst_error"] == stats["worst_error"]
        retried += bs["retried_steps"]
    assert retried > 0
    lat = b.latency()
    assert lat["samples"] == 40 * len(scs)
    b.close()


def test_batch_errors(ctx):
    scs = small_batch(WARM_NONE, False, n=2, seed=82)
    b = BatchSystem(ctx, scs)
    with pytest.raises(PsbError):
        b.update(1)                        # particles not set
    with pytest.raises(PsbError):
        b.set_particles(5, scs[0].pos4)    # scene out of range
    b.close()


def test_batch_time_corrector_keeps_the_run(ctx):
    """The corrector replay over every scene times both passes and rewrites
    only scratch: the next steps equal those of a batch never replayed."""
    scs = small_batch(WARM_NONE, False, n=3, seed=84)
    a = BatchSystem(ctx, scs)
    b = BatchSystem(ctx, scs)
    a.set_all()
    b.set_all()
    with pytest.raises(PsbError):
        b.time_corrector(reps=2)          # before a step
    a.update(4)
    b.update(4)
    t = b.time_corrector(reps=5)
    assert t["pci_density"] > 0 and t["pci_force"] > 0
    a.update(4)
    b.update(4)
    for i in range(len(scs)):
        assert_same(b.state(i), a.state(i), f"scene {i}")
    a.close()
    b.close()
