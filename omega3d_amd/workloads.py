"""Measurement drivers for the batched, served and panel workloads (used by bench.py).

  batched(k=256)   many independent rings, N log-uniform on [2,048, 262,144], RM core,
                   R ~ U[0.5, 2], sigma ~ U[0.01, 0.05], compensated FP32: the
                   persistent batched launch (autotuned tile / targets per thread),
                   its HIP-graph replay, and one ordinary launch per instance.
  served(...)      Poisson stream of Ralston RK2 step requests on device-resident
                   sessions, N log-uniform on [1,024, 1,048,576], single rings (RM,
                   sigma 0.05) or ring pairs (EXP, sigma 0.02), plain and compensated
                   requests mixed; admission by HBM footprint; p50/p99 step latency
                   and aggregate interactions/s.
  panels()         20,480-panel icosphere -> 524,288 shell targets, recursive
                   centroid rule, compensated FP32.
  bem_matrix_run() the BEM influence matrix of the same icosphere on itself.
  bem_rhs_run()    the BEM right-hand side of the icosphere from the shell points as
                   vortex particles.

    python -m omega3d_amd.workloads batched|served|panels
"""
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib, scenes
from .metrics import field_errors


def _fp64_self(core, p, sel):
    """FP64 kernel: targets p[:, sel] against all of p (double device arrays)."""
    L = _lib.lib
    d64 = lambda a: _lib.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))
    sp, ss, sr = d64(p.pos), d64(p.str), d64(p.rad)
    tp, tr = d64(p.pos[:, sel]), d64(p.rad[sel])
    nt = len(sel)
    v = _lib.DeviceArray.from_host(np.zeros((3, nt)))
    g = _lib.DeviceArray.from_host(np.zeros((9, nt)))
    s, t = _lib.Points64(), _lib.Points64()
    s.n, t.n = p.n, nt
    for d in range(3):
        s.x[d], s.s[d], t.x[d], t.u[d] = sp.row(d), ss.row(d), tp.row(d), v.row(d)
    s.r, t.r = sr.ptr, tr.ptr
    for k in range(9):
        t.ug[k] = g.row(k)
    o = _lib.make_opts(core, True, "fp64")
    _lib.check(L.o3d_points_affect_points64_dev(C.byref(s), C.byref(t), C.byref(o), None))
    _lib.synchronize()
    return v.to_host(), g.to_host()


def batched(k=256, lo=2048, hi=262144, reps=2, tune=True, log=print):
    from .batch import Batch, stream_create, stream_destroy
    t0 = time.perf_counter()
    inst = scenes.batch_instances(k, lo, hi)
    n = np.array([p.n for p in inst])
    pairs = float((n.astype(np.float64) ** 2).sum())
    log(f"batched: {k} rings, {n.sum()} particles, {pairs:.4g} pairs, generated in {time.perf_counter() - t0:.1f}s")
    b = Batch(inst, core="rm", precision="comp")
    s = stream_create()
    try:
        ev0, ev1 = _lib.Event(), _lib.Event()
        tuned = None
        if tune:
            bt, bp, ms = b.tune(stream=s, reps=1)
            tuned = {"tile": bt, "tpt": bp, "ms": round(ms, 3)}
            log(f"batched: autotuner picked tile {bt}, tpt {bp} ({ms:.1f} ms)")

        def timed(fn):
            fn()
            _lib.check(_lib.lib.o3d_stream_synchronize(s))
            ev0.record(s)
            for _ in range(reps):
                fn()
            ev1.record(s)
            _lib.check(_lib.lib.o3d_stream_synchronize(s))
            return ev0.elapsed_ms(ev1) / reps

        ms_batch = timed(lambda: b.run(s))
        res = b.results()
        b.capture(s)
        ms_graph = timed(lambda: b.replay(s))
        ms_sep = timed(lambda: b.run_separate(s))
        # the same instances as ordinary library calls, one after another on the
        # batch's stream: each a view of the batch's device arrays, method and
        # configuration left to the library (AUTO: the source split for small
        # instances, the factored method from 65,536 particles), no host
        # synchronisation between calls; the outputs accumulate (not zeroed: the
        # batch's own zero fill of 15M particles takes ~0.2 ms)
        views = []
        for off, cnt in zip(b.offsets.tolist(), b.counts.tolist()):
            v = _lib.Points()
            v.n = cnt
            for d in range(3):
                v.x[d] = b.all.x[d] + 4 * off
                v.s[d] = b.all.s[d] + 4 * off
                v.u[d] = b.all.u[d] + 4 * off
            v.r = b.all.r + 4 * off
            for kk in range(9):
                v.ug[kk] = b.all.ug[kk] + 4 * off
            views.append(v)
        o_auto = _lib.make_opts("rm", True, "comp", stream=s)

        def plain_calls():
            for v in views:
                _lib.check(_lib.lib.o3d_points_affect_points_dev(C.byref(v), C.byref(v), C.byref(o_auto), None))
        ms_calls = timed(plain_calls)
        # accuracy: every instance on 128 spread targets, and the largest and the
        # smallest on up to 32,768 targets, against the FP64 kernel
        errs = {}
        per = []
        for i, p in enumerate(inst):
            sel = np.linspace(0, p.n - 1, 128).astype(np.int64)
            vr, gr = _fp64_self("rm", p, sel)
            v, g = res[i]
            e = field_errors(v[:, sel], g[:, sel], vr, gr)
            per.append((e["vel_max_rel"], e["grad_max_rel"]))
        per = np.array(per)
        worst = int(np.argmax(per.max(axis=1)))
        errs["every_instance"] = {"instances": len(inst), "targets_per_instance": 128,
                                  "vel_max_rel_max": float(per[:, 0].max()),
                                  "grad_max_rel_max": float(per[:, 1].max()),
                                  "vel_max_rel_median": float(np.median(per[:, 0])),
                                  "grad_max_rel_median": float(np.median(per[:, 1])),
                                  "within_1e-6": int((per.max(axis=1) <= 1e-6).sum()),
                                  "worst_instance": {"index": worst, "n": int(n[worst]),
                                                     "factored": bool(n[worst] >= 32768)}}
        for i in (int(np.argmax(n)), int(np.argmin(n))):
            p = inst[i]
            sel = np.arange(p.n) if p.n <= 32768 else np.linspace(0, p.n - 1, 32768).astype(np.int64)
            vr, gr = _fp64_self("rm", p, sel)
            v, g = res[i]
            errs[f"instance_{i}_n{p.n}"] = field_errors(v[:, sel], g[:, sel], vr, gr)
        cfg = b.config()
    finally:
        b.close()
        stream_destroy(s)
    gi = lambda ms: pairs / (ms / 1e3) / 1e9
    out = {"workload": f"{k} independent rings, N log-uniform [{lo}, {hi}], RM core, compensated FP32",
           "pairs": pairs, "particles": int(n.sum()), "config": cfg, "autotune": tuned,
           "persistent_ms": round(ms_batch, 3), "persistent_ginter_s": round(gi(ms_batch), 2),
           "graph_ms": round(ms_graph, 3), "graph_ginter_s": round(gi(ms_graph), 2),
           "per_instance_ms": round(ms_calls, 3), "per_instance_ginter_s": round(gi(ms_calls), 2),
           "speedup_vs_per_instance": round(ms_calls / ms_graph, 3),
           "per_instance_form": "each instance an ordinary o3d_points_affect_points_dev call (method and "
                                "configuration left to the library), one after another on one stream",
           "separate_launches_ms": round(ms_sep, 3),
           "separate_launches_form": "o3d_batch_run_separate: each instance launched alone at the batch's tile, lane "
                                     "count and method (the bitwise reference of the batched launch)",
           "errors_vs_fp64": errs}
    log("batched: " + json.dumps({kk: vv for kk, vv in out.items() if kk != "errors_vs_fp64"}))
    return out


def served_sessions(sessions=16, lo=1024, hi=1048576, seed=scenes.SEED_BASE + 4):
    """The served workload's sessions: N log-uniform on [lo, hi]; each a single ring
    (RM core, sigma 0.05) or a ring pair (EXP core, sigma 0.02) with equal odds.
    Returns [(particles, core)]."""
    rng = scenes.Rng(seed, 11)
    counts = scenes.log_uniform_counts(sessions, lo, hi, seed)
    kinds = rng.uniform(sessions) < 0.5
    out = []
    for i in range(sessions):
        if kinds[i]:
            half = max(1, int(counts[i]) // 2)
            out.append((scenes.colliding_rings(n_each=half, seed=seed + 100 + i), "exp"))
        else:
            out.append((scenes.ring(int(counts[i]), sigma=0.05, seed=seed, stream=100 + i), "rm"))
    return out


def _latency_stats(lat, ns):
    """p50/p99 (ms) of request latencies, overall and per particle-count bucket"""
    lat, ns = np.asarray(lat, np.float64), np.asarray(ns, np.float64)
    pct = lambda a: {"p50_ms": round(float(np.percentile(a, 50)) * 1e3, 3),
                     "p99_ms": round(float(np.percentile(a, 99)) * 1e3, 3), "requests": int(a.size)}
    out = {"all": pct(lat)}
    for name, lo, hi in (("n<16384", 0, 16384), ("16384<=n<131072", 16384, 131072), ("n>=131072", 131072, 1 << 62)):
        m = (ns >= lo) & (ns < hi)
        if m.any():
            out[name] = pct(lat[m])
    return out


def served(duration=20.0, sessions=16, lo=1024, hi=1048576, load=0.5, rate=None, seed=scenes.SEED_BASE + 4,
           log=print):
    """The served workload, four ways on one request set:
      poisson      arrivals at a fixed rate (load x the device's rate), p50/p99 per N bucket;
      saturated    the same requests submitted at once, so the queue never drains:
                   the served capacity in aggregate interactions/s;
      per_instance the same requests as plain library calls, two evaluations per step
                   (o3d_points_affect_points_dev, method AUTO) per request, one after
                   another on one stream with no host synchronisation in between --
                   the most favourable form of per-instance launches;
      fp64_spot    the largest session's stage-1 sums, by the method the server takes
                   for it, against the FP64 kernel on 1,024 targets, both precisions."""
    from .serve import Server
    t0 = time.perf_counter()
    sess = served_sessions(sessions, lo, hi, seed)
    ns = np.array([p.n for p, _ in sess], np.float64)
    mean_pairs = float(np.mean(2.0 * ns * ns))
    if rate is None:
        # offered load relative to ~1.1e12 pairs/s (compensated FP32, mixed cores)
        rate = load * 1.1e12 / mean_pairs
    arr = scenes.Rng(seed, 12)
    n_req = int(duration * rate) + 1
    gaps = -np.log1p(-arr.uniform(n_req)) / rate
    pick = (arr.uniform(n_req) * sessions).astype(int)
    prec = np.where(arr.uniform(n_req) < 0.5, "comp", "fp32")
    inter = float((2.0 * ns[pick] ** 2).sum())

    def open_all(srv):
        return [srv.open_session(p, core) for p, core in sess]

    out = {"workload": f"RK2 step requests on {sessions} resident sessions, N log-uniform [{lo}, {hi}], "
                       f"rings (RM) and ring pairs (EXP), plain and compensated mixed",
           "sessions": sessions, "particles": int(ns.sum()), "requests": n_req,
           "interactions": inter}
    # ---- Poisson arrivals ----------------------------------------------------------
    srv = Server(hbm_budget=16 << 30, max_active=64)
    try:
        sids = open_all(srv)
        log(f"served: {sessions} sessions, {int(ns.sum())} particles resident, set up in "
            f"{time.perf_counter() - t0:.1f}s")
        srv.wait(srv.submit(sids[int(np.argmin(ns))], 1e-3, "comp"))     # warm-up
        srv.reset_stats()
        rids = []
        t_start = time.perf_counter()
        t_next = t_start
        for j in range(n_req):
            t_next += gaps[j]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            rids.append(srv.submit(sids[pick[j]], 1e-3, str(prec[j])))
        infos = [srv.wait(r) for r in rids]
        st = srv.stats()
        wall = time.perf_counter() - t_start
    finally:
        srv.close()
    lat = [i["latency_s"] for i in infos]
    out["poisson"] = {"rate_per_s": round(rate, 3), "completed": st["completed"], "failed": st["failed"],
                      "p50_s": round(st["p50_s"], 4), "p99_s": round(st["p99_s"], 4),
                      "mean_s": round(st["mean_s"], 4), "ginter_per_s": round(st["ginter_per_s"], 2),
                      "wall_s": round(wall, 2), "ticks": st["ticks"], "hbm_peak_bytes": st["hbm_peak"],
                      "latency_by_n": _latency_stats(lat, ns[pick])}
    log("served poisson: " + json.dumps(out["poisson"]))
    # ---- saturation: every request at once ----------------------------------------------
    srv = Server(hbm_budget=16 << 30, max_active=64)
    try:
        sids = open_all(srv)
        srv.wait(srv.submit(sids[int(np.argmin(ns))], 1e-3, "comp"))
        srv.reset_stats()
        t_start = time.perf_counter()
        rids = [srv.submit(sids[pick[j]], 1e-3, str(prec[j])) for j in range(n_req)]
        infos = [srv.wait(r) for r in rids]
        wall = time.perf_counter() - t_start
        st = srv.stats()
    finally:
        srv.close()
    lat = [i["latency_s"] for i in infos]
    out["saturated"] = {"completed": st["completed"], "failed": st["failed"], "wall_s": round(wall, 3),
                        "ginter_per_s": round(inter / wall / 1e9, 2), "ticks": st["ticks"],
                        "factored_steps": st["factored_steps"], "latency_by_n": _latency_stats(lat, ns[pick])}
    log("served saturated: " + json.dumps(out["saturated"]))
    # ---- the same requests as plain per-instance calls ------------------------------------
    L = _lib.lib
    D = _lib.DeviceArray.from_host
    dev = []
    for p, core in sess:
        dpos, dstr, drad = D(p.pos), D(p.str), D(p.rad)
        dv, dg = _lib.DeviceArray((3, p.n)), _lib.DeviceArray((9, p.n))
        pt = _lib.Points()
        pt.n = p.n
        for d in range(3):
            pt.x[d], pt.s[d], pt.u[d] = dpos.row(d), dstr.row(d), dv.row(d)
        pt.r = drad.ptr
        for k in range(9):
            pt.ug[k] = dg.row(k)
        dev.append((pt, core, (dpos, dstr, drad, dv, dg)))
    opts = {(core, pr): _lib.make_opts(core, True, pr) for _, core in sess for pr in ("comp", "fp32")}

    def plain(j):
        pt, core, _ = dev[pick[j]]
        o = opts[(core, str(prec[j]))]
        for _ in range(2):
            _lib.check(L.o3d_points_affect_points_dev(C.byref(pt), C.byref(pt), C.byref(o), None))

    for j in sorted({int(np.argmax(pick == k)) for k in set(pick.tolist())}):
        plain(j)                 # every session once: workspaces grown before the clock starts
    _lib.synchronize()
    t_start = time.perf_counter()
    for j in range(n_req):
        plain(j)
    _lib.synchronize()
    wall = time.perf_counter() - t_start
    out["per_instance"] = {"wall_s": round(wall, 3), "ginter_per_s": round(inter / wall / 1e9, 2)}
    out["served_vs_per_instance"] = round(out["saturated"]["ginter_per_s"] / out["per_instance"]["ginter_per_s"], 3)
    log("served per-instance: " + json.dumps(out["per_instance"]))
    del dev
    # ---- FP64 spot errors of the largest session ------------------------------------------
    big = int(np.argmax(ns))
    p, core = sess[big]
    from . import influence
    sel = np.linspace(0, p.n - 1, 1024).astype(np.int64)
    vr, gr = _fp64_self(core, p, sel)
    spot = {"session": big, "n": p.n, "core": core, "targets": len(sel),
            "method": "factored" if core in ("rm", "exp") and p.n >= 32768 else "direct"}
    for pr in ("comp", "fp32"):
        v, g = influence.p2p(core, p.pos, p.str, p.rad, p.pos, p.rad, True, pr)
        spot[pr] = field_errors(v[:, sel], g[:, sel], vr, gr)
    out["fp64_spot"] = spot
    log("served fp64 spot: " + json.dumps(spot))
    # headline figures of the Poisson run, as before
    out.update({"rate_per_s": out["poisson"]["rate_per_s"], "p50_s": out["poisson"]["p50_s"],
                "p99_s": out["poisson"]["p99_s"], "ginter_per_s": out["poisson"]["ginter_per_s"]})
    return out


def single_ring_run(n=8192, reps=20, log=print):
    """The 8,192-particle single ring (torus R=1, a=0.1, sigma 0.05, |alpha| = 1/N), RM
    core, velocity + gradient at every particle with the self term excluded, plain
    FP32: kernel time and the error against the FP64 kernel on all targets."""
    L = _lib.lib
    p = scenes.ring(n, sigma=0.05)
    D = _lib.DeviceArray.from_host
    dpos, dstr, drad = D(p.pos), D(p.str), D(p.rad)
    # velocity rows 0-2 and gradient rows 3-11 of one array, zeroed by one fill per step
    dout = _lib.DeviceArray((12, n))
    s = _lib.Points()
    s.n = n
    for d in range(3):
        s.x[d], s.s[d], s.u[d] = dpos.row(d), dstr.row(d), dout.row(d)
    s.r = drad.ptr
    for k in range(9):
        s.ug[k] = dout.row(3 + k)
    o = _lib.make_opts("rm", True, "fp32", self_term=1)
    e0, e1 = _lib.Event(), _lib.Event()
    dout.zero()
    _lib.check(L.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None))
    _lib.synchronize()
    e0.record()
    for _ in range(reps):
        dout.zero()
        _lib.check(L.o3d_points_affect_points_dev(C.byref(s), C.byref(s), C.byref(o), None))
    e1.record()
    _lib.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    out_h = dout.to_host()
    vel, grad = out_h[:3], out_h[3:]
    d64 = lambda a: _lib.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))
    p64, s64_, r64 = d64(p.pos), d64(p.str), d64(p.rad)
    v64, g64 = d64(np.zeros((3, n))), d64(np.zeros((9, n)))
    q = _lib.Points64()
    q.n = n
    for d in range(3):
        q.x[d], q.s[d], q.u[d] = p64.row(d), s64_.row(d), v64.row(d)
    q.r = r64.ptr
    for k in range(9):
        q.ug[k] = g64.row(k)
    o64 = _lib.make_opts("rm", True, "fp64", self_term=1)
    _lib.check(L.o3d_points_affect_points64_dev(C.byref(q), C.byref(q), C.byref(o64), None))
    _lib.synchronize()
    out = {"workload": f"single ring, N={n}, RM core, sigma 0.05, velocity + gradient, self term excluded, plain FP32",
           "gpu_ms": round(ms, 4), "gpu_ginter_s": round(float(n) * n / ms / 1e6, 2),
           "errors_vs_fp64": dict(field_errors(vel, grad, v64.to_host(), g64.to_host()), targets=n)}
    log("single_ring: " + json.dumps(out))
    return out


def bem_matrix_run(levels=5, nunk=3, reps=3, sample_sources=32, log=print):
    """The BEM influence matrix of the 20,480-panel icosphere on itself
    (panels_on_panels_coeff, Coefficients.h:169-444; 3 unknowns per panel, maxlev 3):
    a 61,440 x 61,440 float block written to HBM.  Stage times (root decisions and
    count, one-leaf pairs written, recursed pairs) and the write rate of the whole
    call and of the one-leaf kernel.  Returns (results, sample): sample holds the
    surface and the columns of sample_sources spread source panels, for a check on
    the host."""
    L = _lib.lib
    pnl = scenes.icosphere_panels(levels)
    nodes, area = pnl.nodes, pnl.area
    idx = pnl.idx.reshape(-1, 3)
    P = idx.shape[0]
    R = nunk * P
    D = _lib.DeviceArray.from_host
    dn, di, da = D(nodes), D(pnl.idx.view(np.float32)), D(area)
    pan = _lib.Panels()
    pan.np, pan.nn = P, nodes.shape[1]
    for d in range(3):
        pan.x[d] = dn.row(d)
    pan.idx, pan.area = di.ptr, da.ptr
    out = _lib.DeviceArray((R * R,))
    profs = []
    for _ in range(reps + 1):
        prof = _lib.BemMatrixProfile()
        _lib.check(L.o3d_bem_matrix_profile_dev(C.byref(pan), C.byref(pan), nunk, 1, -1, C.c_void_p(out.ptr), None,
                                                C.byref(prof)))
        _lib.synchronize()
        profs.append(prof)
    e0, e1 = _lib.Event(), _lib.Event()
    e0.record()
    for _ in range(reps):
        _lib.check(L.o3d_bem_matrix_dev(C.byref(pan), C.byref(pan), nunk, 1, -1, C.c_void_p(out.ptr), None))
    e1.record()
    _lib.synchronize()
    call_ms = e0.elapsed_ms(e1) / reps
    best = min(profs[1:], key=lambda q: q.total_ms)
    gb = 4.0 * R * R / 1e9
    js = np.linspace(0, P - 1, sample_sources).astype(np.int64)
    cols = np.empty((R, nunk * len(js)), np.float32)
    col = np.empty(R, np.float32)
    for a, j in enumerate(js):
        for k in range(nunk):
            _lib.check(L.o3d_memcpy_dtoh(col.ctypes.data, out.ptr + 4 * (int(j) * nunk + k) * R, 4 * R))
            cols[:, a * nunk + k] = col
    # the same block written by the runtime's own fill (hipMemset) with no arithmetic:
    # the write rate this GPU reaches for these bytes, the practical ceiling of k_bm_far
    fill = []
    for _ in range(reps + 1):
        _lib.synchronize()
        t = time.perf_counter()
        out.zero()
        _lib.synchronize()
        fill.append((time.perf_counter() - t) * 1e3)
    fill_ms = min(fill[1:])
    out.free()
    res = {"workload": f"BEM influence matrix of the {P}-panel icosphere on itself, {nunk} unknowns per panel, "
                       f"maxlev 3: {R} x {R} floats ({gb:.2f} GB)",
           "call_ms": round(call_ms, 3), "count_ms": round(best.count_ms, 3), "far_ms": round(best.far_ms, 3),
           "near_ms": round(best.near_ms, 3), "recursed_pairs": int(best.near_pairs), "panel_pairs": P * P,
           "matrix_gb": round(gb, 3), "call_write_gbs": round(gb / call_ms * 1e3, 1),
           "far_write_gbs": round(gb / best.far_ms * 1e3, 1),
           "fill_ms": round(fill_ms, 3), "fill_write_gbs": round(gb / fill_ms * 1e3, 1),
           "far_write_vs_fill": round(fill_ms / best.far_ms, 3)}
    log("bem_matrix: " + json.dumps(res))
    sample = {"nodes": nodes, "idx": idx, "area": area, "sources": js, "columns": cols, "nunk": nunk}
    return res, sample


def bem_rhs_run(n_particles=524288, levels=5, nunk=3, reps=5, sample_panels=128, log=print):
    """The BEM right-hand side of the 20,480-panel icosphere (BEMHelper.h:83-117) from
    the 524,288 shell points of the panel workload taken as vortex particles
    (strengths uniform in [-1, 1]^3 / N), compensated FP32, a unit freestream along x
    and no body motion: zero the panel velocities, points_affect_panels, finalize,
    project.  Returns (results, sample) with the velocities and rhs rows of
    sample_panels spread panels, for a check on the host."""
    L = _lib.lib
    pnl = scenes.icosphere_panels(levels)
    P = pnl.area.shape[0]
    pos = scenes.shell_points(n_particles)
    st = (scenes.Rng(scenes.SEED_BASE + 5, 1).uniform(3 * n_particles).reshape(3, -1) * 2.0 - 1.0) / n_particles
    st = st.astype(np.float32)
    D = _lib.DeviceArray.from_host
    dn, di, da = D(pnl.nodes), D(pnl.idx.view(np.float32)), D(pnl.area)
    bd = _lib.Panels()
    bd.np, bd.nn = P, pnl.nodes.shape[1]
    for d in range(3):
        bd.x[d] = dn.row(d)
    bd.idx, bd.area = di.ptr, da.ptr
    dp, ds = D(pos), D(st)
    pts = (_lib.Points * 1)()
    pts[0].n = n_particles
    for d in range(3):
        pts[0].x[d], pts[0].s[d] = dp.row(d), ds.row(d)
    dpu, drhs = _lib.DeviceArray((3, P)), _lib.DeviceArray((nunk * P,))
    pu = (C.c_void_p * 3)(*[dpu.row(d) for d in range(3)])
    o = _lib.make_opts("rm", False, "comp")
    fs = (1.0, 0.0, 0.0)
    b = _lib.BemRhsOpts()
    for d in range(3):
        b.fs[d] = fs[d]
    b.nunk = nunk
    call = lambda: _lib.check(L.o3d_bem_rhs_dev(pts, 1, C.byref(bd), C.byref(b), C.byref(o), pu,
                                                C.c_void_p(drhs.ptr)))
    call()
    _lib.synchronize()
    e0, e1 = _lib.Event(), _lib.Event()
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    _lib.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    pairs = float(P) * n_particles
    res = {"workload": f"BEM right-hand side of the {P}-panel icosphere from {n_particles} vortex particles in "
                       f"1.02<r<1.5 (points_affect_panels, recursive rule maxlev 3, compensated FP32; finalize "
                       f"with a unit freestream; {nunk} unknowns per panel)",
           "ms": round(ms, 3), "pairs": pairs, "g_pairs_per_s": round(pairs / ms / 1e6, 2)}
    log("bem_rhs: " + json.dumps(res))
    js = np.linspace(0, P - 1, sample_panels).astype(np.int64)
    vel, rhs = dpu.to_host(), drhs.to_host().reshape(P, nunk)
    sample = {"nodes": pnl.nodes, "idx": pnl.idx.reshape(-1, 3), "area": pnl.area, "pos": pos, "str": st,
              "panels": js, "pu": vel[:, js], "rhs": rhs[js], "fs": fs, "nunk": nunk}
    return res, sample


def panels_run(n_targets=524288, reps=2, check_targets=4096, log=print):
    """The icosphere workload on the GPU: timing, and the error of a target subset
    against the FP64 panel kernel."""
    L = _lib.lib
    pan = scenes.icosphere_panels()
    tp = scenes.shell_points(n_targets)
    D = _lib.DeviceArray.from_host
    dn, di, dts, dar = D(pan.nodes), D(pan.idx), D(pan.ts), D(pan.area)
    dtp = D(tp)
    dv, dg = _lib.DeviceArray((3, n_targets)), _lib.DeviceArray((9, n_targets))
    p = _lib.Panels()
    p.np, p.nn = pan.np, pan.nodes.shape[1]
    for d in range(3):
        p.x[d], p.ts[d] = dn.row(d), dts.row(d)
    p.idx, p.area, p.ss = di.ptr, dar.ptr, None
    t = _lib.Points()
    t.n = n_targets
    for d in range(3):
        t.x[d], t.u[d] = dtp.row(d), dv.row(d)
    for k in range(9):
        t.ug[k] = dg.row(k)
    o = _lib.make_opts("rm", True, "comp")
    e0, e1 = _lib.Event(), _lib.Event()
    dv.zero(); dg.zero()
    _lib.check(L.o3d_panels_affect_points_dev(C.byref(p), C.byref(t), C.byref(o), None))
    _lib.synchronize()
    e0.record()
    for _ in range(reps):
        dv.zero(); dg.zero()
        _lib.check(L.o3d_panels_affect_points_dev(C.byref(p), C.byref(t), C.byref(o), None))
    e1.record()
    _lib.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    pairs = float(pan.np) * n_targets
    vel, grad = dv.to_host(), dg.to_host()
    from .influence import pan as pan_host
    sel = np.linspace(0, n_targets - 1, check_targets).astype(np.int64)
    vr = np.zeros((3, len(sel)), np.float32)
    gr = np.zeros((9, len(sel)), np.float32)
    pan_host("rm", pan.nodes, pan.idx, pan.ts, pan.area, None, tp[:, sel], True, "fp64", vel=vr, grad=gr)
    out = {"workload": f"{pan.np}-panel icosphere -> {n_targets} targets in 1.02<r<1.5, recursive centroid rule "
                       f"(maxlev 3), compensated FP32", "ms": round(ms, 3),
           "g_panel_target_pairs_per_s": round(pairs / (ms / 1e3) / 1e9, 3),
           "errors_vs_fp64": dict(field_errors(vel[:, sel], grad[:, sel], vr, gr), targets=len(sel))}
    log("panels: " + json.dumps(out))

    # the same workload by the 7-point rule with the closed-form near field
    o7 = _lib.make_opts("rm", True, "comp", method="gauss7")
    prof = _lib.PanelProfile()
    dv.zero(); dg.zero()
    _lib.check(L.o3d_panels_affect_points_dev(C.byref(p), C.byref(t), C.byref(o7), None))   # warm-up
    _lib.synchronize()
    best = None
    for _ in range(reps):
        dv.zero(); dg.zero()
        _lib.check(L.o3d_panels_g7_profile_dev(C.byref(p), C.byref(t), C.byref(o7), C.byref(prof)))
        cur = {k: getattr(prof, k) for k, _ in _lib.PanelProfile._fields_}
        best = cur if best is None or cur["total_ms"] < best["total_ms"] else best
    vel7, grad7 = dv.to_host(), dg.to_host()
    vr7 = np.zeros((3, len(sel)), np.float32)
    gr7 = np.zeros((9, len(sel)), np.float32)
    pan_host("rm", pan.nodes, pan.idx, pan.ts, pan.area, None, tp[:, sel], True, "fp64", vel=vr7, grad=gr7,
             rule="gauss7")
    f_pt = _lib.flops_per_interaction("rm", True, False, True)     # flops_0vs_0pg, a point target
    far_flops = 7.0 * f_pt * best["far_pairs"]
    out7 = {"workload": f"{pan.np}-panel icosphere -> {n_targets} targets in 1.02<r<1.5, 7-point rule per panel, "
                        f"closed-form near field within 2 panel sizes, compensated FP32",
            "ms": round(best["total_ms"], 3), "far_ms": round(best["far_ms"], 3),
            "near_ms": round(best["total_ms"] - best["far_ms"], 3),
            "pairs": best["pairs"], "near_pairs": best["near_pairs"],
            "g_panel_target_pairs_per_s": round(best["pairs"] / (best["total_ms"] / 1e3) / 1e9, 3),
            "far_tflops": round(far_flops / (best["far_ms"] / 1e3) / 1e12, 3),
            "far_flops_per_pair": 7 * f_pt,
            "far_fp32_peak_frac": round(far_flops / (best["far_ms"] / 1e3) / 157.3e12, 4),
            "errors_vs_fp64": dict(field_errors(vel7[:, sel], grad7[:, sel], vr7, gr7), targets=len(sel),
                                   reference="the same rule in FP64 on the GPU")}
    log("panels_gauss7: " + json.dumps(out7))
    out["gauss7"] = out7
    return out


if __name__ == "__main__":
    _lib.check(_lib.lib.o3d_init(0))
    what = sys.argv[1] if len(sys.argv) > 1 else "batched"
    if what == "batched":
        print(json.dumps(batched()))
    elif what == "served":
        print(json.dumps(served()))
    elif what == "panels":
        print(json.dumps(panels_run()))
    elif what == "single_ring":
        print(json.dumps(single_ring_run(reps=50, log=lambda s: None)))
