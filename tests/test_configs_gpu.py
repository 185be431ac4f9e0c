"""The batched and served workloads at their configured shapes.

Batched: 256 independent rings, N log-uniform on [2,048, 262,144], R ~ U[0.5, 2],
sigma ~ U[0.01, 0.05], RM core, compensated FP32, method AUTO (instances of
>= 32,768 particles take the factored method on the batch's lanes), through the
persistent launch and replays of the captured graph.  Every instance is bitwise
its own single-collection call at the same tile and method; 128 spread targets of
8 instances (the smallest, the largest, three factored, three direct) are within
1e-6 of the CPU FP64 oracle.

Served: the served workload's 16 sessions plus a 1,048,576-particle EXP ring pair,
under one burst of mixed plain / compensated RK2 step requests.  Every session
ends bitwise in the state the same request sequence gives alone; the ring pair's
stage-1 sums (the single-collection call on its initial state, the pipeline the
server runs for it) are within 1e-6 of the CPU FP64 oracle on 256 targets, and
its first step agrees with the oracle's RK2 restatement driven by those sums.
"""
import numpy as np
import pytest

import oracle
from oracle import rk2
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors

pytestmark = pytest.mark.gpu
# method AUTO in a batch: the factored method from this many particles (o3d_biot.h)
BATCH_FACTORED_MIN_N = 32768


@pytest.fixture(scope="module")
def mods():
    from omega3d_amd import _lib, batch, influence, serve, workloads
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return _lib, batch, influence, serve, workloads


def test_batched_256_rings(mods):
    _lib, batch, infl, _, _ = mods
    inst = scenes.batch_instances()
    n = np.array([p.n for p in inst])
    assert len(inst) == 256 and n.min() >= 2048 and n.max() <= 262144
    print(f"batched: 256 rings, {n.sum()} particles", flush=True)
    b = batch.Batch(inst, core="rm", precision="comp", tile=256, tpt=1)
    s = batch.stream_create()
    try:
        cfg = b.config()
        nf = int((n >= BATCH_FACTORED_MIN_N).sum())
        assert cfg["factored_instances"] == nf and nf > 0
        b.run(s)
        _lib.check(_lib.lib.o3d_stream_synchronize(s))
        eager = b.results()
        print("batched: persistent launch done; capturing the whole run", flush=True)
        b.capture(s)
        for rep in range(2):
            b.replay(s)
            _lib.check(_lib.lib.o3d_stream_synchronize(s))
            for i, ((v, g), (v2, g2)) in enumerate(zip(eager, b.results())):
                assert np.array_equal(v, v2) and np.array_equal(g, g2), (rep, i)
            p = inst[rep]
            infl.p2p("wl", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp")   # eager work between replays
        print("batched: graph replays bitwise; single-collection calls", flush=True)
        for i, p in enumerate(inst):
            fac = p.n >= BATCH_FACTORED_MIN_N
            va, ga = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp", tile=0 if fac else 256,
                              tpt=0 if fac else 1, method="factored" if fac else "direct")
            v, g = eager[i]
            assert np.array_equal(v, va) and np.array_equal(g, ga), (i, p.n)
    finally:
        b.close()
        batch.stream_destroy(s)
    fac_idx = [i for i in np.argsort(n) if n[i] >= BATCH_FACTORED_MIN_N]
    dir_idx = [i for i in np.argsort(n) if n[i] < BATCH_FACTORED_MIN_N]
    pick = {int(np.argmin(n)), int(np.argmax(n))}
    pick |= {int(fac_idx[k]) for k in np.linspace(0, len(fac_idx) - 1, 3).astype(int)}
    pick |= {int(dir_idx[k]) for k in np.linspace(0, len(dir_idx) - 1, 3).astype(int)}
    for i in sorted(pick):
        p = inst[i]
        sel = np.linspace(0, p.n - 1, 128).astype(np.int64)
        ov, og = oracle.p2p("rm", p.pos, p.str, p.rad, p.pos[:, sel], p.rad[sel], True, "f64")
        v, g = eager[i]
        ev, eg = rel_errors(v[:, sel], ov), rel_errors(g[:, sel], og)
        print(f"batched: instance {i} n={p.n} vs FP64 oracle: vel {ev} grad {eg}", flush=True)
        assert ev[0] <= 1e-6 and eg[0] <= 1e-6, (i, p.n)


def _burst(sessions, seed=77, n_req=40):
    """(session index, dt, precision) of a burst: every session at least once, the
    1M ring pair (last) three times, the rest drawn at random."""
    rng = np.random.default_rng(seed)
    k = len(sessions)
    picks = list(range(k)) + [k - 1, k - 1] + list(rng.integers(0, k - 1, n_req - k - 2))
    rng.shuffle(picks)
    return [(int(j), float(rng.choice([5e-4, 1e-3])), str(rng.choice(["comp", "fp32"]))) for j in picks]


def _server(serve):
    return serve.Server(hbm_budget=24 << 30, max_active=64, tile=256, tpt=1)


def test_served_sessions_alone_vs_burst(mods):
    _lib, _, infl, serve, workloads = mods
    sessions = workloads.served_sessions() + [(scenes.colliding_rings(n_each=524288, seed=scenes.SEED_BASE + 44),
                                               "exp")]
    assert sessions[-1][0].n == 1048576
    reqs = _burst(sessions)
    print(f"served: {len(sessions)} sessions, {sum(p.n for p, _ in sessions)} particles, {len(reqs)} requests",
          flush=True)
    srv = _server(serve)
    try:
        sids = [srv.open_session(p, c) for p, c in sessions]
        rids = [srv.submit(sids[j], dt, prec) for j, dt, prec in reqs]
        for r in rids:
            info = srv.wait(r)
            assert info["state"] == "done" and info["rc"] == 0, info
        mixed = [srv.read_session(sid) for sid in sids]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["factored_steps"] > 0
    finally:
        srv.close()
    print("served: burst done; each session alone", flush=True)
    srv = _server(serve)
    try:
        for j, (p, c) in enumerate(sessions):
            sid = srv.open_session(p, c)
            for jj, dt, prec in reqs:
                if jj == j:
                    info = srv.wait(srv.submit(sid, dt, prec))
                    assert info["state"] == "done", info
            alone = srv.read_session(sid)
            srv.close_session(sid)
            for a, b in zip(alone, mixed[j]):
                assert np.array_equal(a, b), (j, p.n, c)
    finally:
        srv.close()


def test_served_1m_session_against_oracle(mods):
    _lib, _, infl, serve, _ = mods
    p = scenes.colliding_rings(n_each=524288, seed=scenes.SEED_BASE + 44)
    dt = 1e-3
    # stage 1: the factored pipeline the server runs for this session, on the initial state
    v1, g1 = infl.p2p("exp", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp", method="factored")
    sel = np.linspace(0, p.n - 1, 256).astype(np.int64)
    ov, og = oracle.p2p("exp", p.pos, p.str, p.rad, p.pos[:, sel], p.rad[sel], True, "f64")
    ev, eg = rel_errors(v1[:, sel], ov), rel_errors(g1[:, sel], og)
    print(f"served 1M: stage-1 sums vs FP64 oracle: vel {ev} grad {eg}", flush=True)
    assert ev[0] <= 1e-6 and eg[0] <= 1e-6
    srv = _server(serve)
    try:
        sid = srv.open_session(p, "exp")
        info = srv.wait(srv.submit(sid, dt, "comp"))
        assert info["state"] == "done", info
        x, s, e = srv.read_session(sid)
    finally:
        srv.close()
    gpu_pair = lambda core, pos, st, rad: infl.p2p(core, pos, st, rad, pos, rad, True, "comp", method="factored")
    x32, s32, e32 = rk2.rk2_step("exp", p.pos, p.str, p.rad, np.ones(p.n, np.float32), dt, (0, 0, 0), "f32d",
                                 pair=gpu_pair)
    ex = rel_errors(x - p.pos, x32 - p.pos)[0]
    es = rel_errors(s - p.str, s32 - p.str)[0]
    print(f"served 1M: step increments vs the RK2 restatement: dx {ex:.2e} ds {es:.2e}", flush=True)
    # float storage: each entry within 4 ulps of the restatement plus 2e-5 of the largest increment
    for a, b, a0 in ((x, x32, p.pos), (s, s32, p.str)):
        tol = 4 * np.spacing(np.abs(b)) + 2e-5 * np.abs(b.astype(np.float64) - a0).max()
        assert (np.abs(a.astype(np.float64) - b) <= tol).all()
    assert np.abs(e.astype(np.float64) - e32).max() <= 4 * np.spacing(np.float32(1.0)) + 2e-5 * np.abs(e32 - 1).max()
