"""Batched persistent evaluation of many independent rings: parity against the
oracle, bitwise equality with each instance evaluated alone, the per-instance
launch baseline, HIP graph replays and the autotuner."""
import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from omega3d_amd import _lib, batch, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib, batch, influence


def small_instances(k=12, seed=scenes.SEED_BASE + 3):
    counts = scenes.log_uniform_counts(k, 100, 3000, seed)
    prm = scenes.Rng(seed, 9)
    Rs = 0.5 + 1.5 * prm.uniform(k)
    sig = 0.01 + 0.04 * prm.uniform(k)
    return [scenes.ring(int(counts[i]), R=float(Rs[i]), sigma=float(sig[i]), seed=seed, stream=i + 1)
            for i in range(k)]


def test_batch_matches_oracle_and_alone(mods):
    _lib, batch, infl = mods
    inst = small_instances()
    cores = ["rm", "wl", "exp", "v2"] * 3
    precs = ["comp", "fp32"] * 6
    b = batch.Batch(inst, cores=cores, precisions=precs, tile=256, tpt=1)
    b.run()
    res = b.results()
    for i, p in enumerate(inst):
        v, g = res[i]
        ov, og = oracle.p2p(cores[i], p.pos, p.str, p.rad, p.pos, p.rad, True, "f64")
        tol = 2e-6 if precs[i] == "comp" else 2e-5   # small random rings: per-pair floor
        assert rel_errors(v, ov)[0] <= tol and rel_errors(g, og)[0] <= tol, i
        # alone, through the single-collection entry point with the same tile: bitwise
        va, ga = infl.p2p(cores[i], p.pos, p.str, p.rad, p.pos, p.rad, True, precs[i], tile=256, tpt=1)
        np.testing.assert_array_equal(v, va)
        np.testing.assert_array_equal(g, ga)


@pytest.mark.parametrize("tpt", [1, 2, 4])
def test_batch_tpt_invariance_and_separate(mods, tpt):
    """A target's sums do not depend on targets-per-thread, on the batch or on the
    launch shape (persistent vs one launch per instance)."""
    _lib, batch, infl = mods
    inst = small_instances(9)
    b = batch.Batch(inst, core="exp", precision="comp", tile=128, tpt=1)
    b.run()
    ref = b.results()
    b2 = batch.Batch(inst, core="exp", precision="comp", tile=128, tpt=tpt)
    b2.run()
    for (v, g), (v2, g2) in zip(ref, b2.results()):
        np.testing.assert_array_equal(v, v2)
        np.testing.assert_array_equal(g, g2)
    b2.run_separate()
    for (v, g), (v2, g2) in zip(ref, b2.results()):
        np.testing.assert_array_equal(v, v2)
        np.testing.assert_array_equal(g, g2)


def test_batch_graph_replays(mods):
    """The captured graph (zero fill, queue reset, persistent kernel) replayed several
    times, with eager work and synchronisations between replays, gives the eager
    result every time."""
    _lib, batch, infl = mods
    inst = small_instances(10)
    b = batch.Batch(inst, core="rm", precision="comp", tile=256, tpt=2)
    s = batch.stream_create()
    try:
        b.run(s)
        _lib.check(_lib.lib.o3d_stream_synchronize(s))
        eager = b.results()
        b.capture(s)
        for rep in range(4):
            b.replay(s)
            _lib.check(_lib.lib.o3d_stream_synchronize(s))
            got = b.results()
            for (v, g), (v2, g2) in zip(eager, got):
                np.testing.assert_array_equal(v, v2)
                np.testing.assert_array_equal(g, g2)
            # eager work in between: an unrelated evaluation
            p = inst[rep]
            infl.p2p("wl", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp")
    finally:
        b.close()
        batch.stream_destroy(s)


def test_batch_autotuner(mods):
    _lib, batch, infl = mods
    inst = small_instances(16)
    b = batch.Batch(inst, core="rm", precision="comp")
    bt, bp, ms = b.tune(reps=1)
    assert bt in (128, 256, 512) and bp in (1, 2, 4) and ms > 0
    cfg = b.config()
    assert cfg["tile"] == bt and cfg["tpt"] == bp and cfg["grid"] > 0
    b.run()
    for (v, g), p in zip(b.results(), inst):
        ov, og = oracle.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "f64")
        assert rel_errors(v, ov)[0] <= 2e-6 and rel_errors(g, og)[0] <= 2e-6


def test_batch_empty_and_tiny(mods):
    _lib, batch, infl = mods
    inst = [scenes.ring(1, seed=1), scenes.ring(2, seed=2), scenes.ring(0, seed=3), scenes.ring(300, seed=4)]
    b = batch.Batch(inst, core="wl", precision="comp")
    b.run()
    for (v, g), p in zip(b.results(), inst):
        if p.n == 0:
            continue
        ov, og = oracle.p2p("wl", p.pos, p.str, p.rad, p.pos, p.rad, True, "f64")
        assert rel_errors(v, ov)[0] <= 2e-6 and rel_errors(g, og)[0] <= 2e-6


def test_batch_factored_instances(mods):
    """Instances of >= 32,768 particles leave the queue for the factored method (AUTO):
    each is bitwise the single-collection factored call, the small ones bitwise the
    direct call; run_separate and graph replays (with eager work between them) give
    the same bits; the large instances meet the compensated bar against the oracle."""
    _lib, batch, infl = mods
    big = [scenes.ring(70001, R=1.3, sigma=0.02, seed=11), scenes.ring(100000, R=0.8, sigma=0.03, seed=12)]
    inst = small_instances(6)[:3] + big[:1] + small_instances(6)[3:] + big[1:]
    b = batch.Batch(inst, core="exp", precision="comp", tile=256, tpt=1)
    s = batch.stream_create()
    try:
        cfg = b.config()
        assert cfg["factored_instances"] == 2 and cfg["factored_scratch_bytes"] > 0
        b.run(s)
        _lib.check(_lib.lib.o3d_stream_synchronize(s))
        eager = b.results()
        for (v, g), p in zip(eager, inst):
            m = "factored" if p.n >= 32768 else "direct"
            va, ga = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp", tile=256 if m == "direct" else 0,
                              tpt=1 if m == "direct" else 0, method=m)
            np.testing.assert_array_equal(v, va)
            np.testing.assert_array_equal(g, ga)
            if p.n >= 32768:
                sel = np.arange(0, p.n, p.n // 512)[:512]
                ov, og = oracle.p2p("exp", p.pos, p.str, p.rad, p.pos[:, sel], p.rad[sel], True, "f64")
                assert rel_errors(v[:, sel], ov)[0] <= 1e-6 and rel_errors(g[:, sel], og)[0] <= 1e-6
        b.run_separate(s)
        _lib.check(_lib.lib.o3d_stream_synchronize(s))
        for (v, g), (v2, g2) in zip(eager, b.results()):
            np.testing.assert_array_equal(v, v2)
            np.testing.assert_array_equal(g, g2)
        b.capture(s)
        for rep in range(3):
            b.replay(s)
            _lib.check(_lib.lib.o3d_stream_synchronize(s))
            for (v, g), (v2, g2) in zip(eager, b.results()):
                np.testing.assert_array_equal(v, v2)
                np.testing.assert_array_equal(g, g2)
            p = inst[rep]
            infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp")
    finally:
        b.close()
        batch.stream_destroy(s)


def test_batch_method_choice(mods):
    _lib, batch, infl = mods
    inst = small_instances(4) + [scenes.ring(66000, seed=13)]
    b = batch.Batch(inst, core="rm", precision="fp32", method="direct")
    assert b.config()["factored_instances"] == 0
    b.close()
    b = batch.Batch(inst, core="rm", precision="fp32", method="factored")
    assert b.config()["factored_instances"] == len(inst)
    b.run()
    for (v, g), p in zip(b.results(), inst):
        va, ga = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "fp32", method="factored")
        np.testing.assert_array_equal(v, va)
        np.testing.assert_array_equal(g, ga)
    b.close()
    with pytest.raises(_lib.O3DError):
        batch.Batch(inst, cores=["rm", "wl", "rm", "rm", "rm"], precision="comp", method="factored")
