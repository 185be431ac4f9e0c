"""Headline benchmark: direct Biot-Savart velocity + velocity gradient at N = 1,048,576.

Workload (one "step" = one evaluation of every particle's velocity and 3x3
velocity gradient from every particle, inputs resident in HBM):
  two colliding vortex rings, 524,288 particles each (centres z = +-1, lower ring
  tilted 15 deg about x, circulation +-1, |alpha| jittered +-5%), core radius 0.02,
  compact-exponential core, FP32 pairs with compensated accumulation.

The evaluation uses the library's default method, which at this size is the
factored one (Morton order; well-separated target-block / source-tile pairs
summed through a per-source moment table on the matrix cores; the rest pair by
pair).  Prints ONE JSON line with G interactions/s (value), the plain-FP32 rate
beside it, the direct (pair-by-pair) method's rate, the roofline of the dominant
kernel (the far-field matrix-core kernel: the reference's 68 flops per
interaction times the pairs it takes, over its HIP-event time), max/RMS relative
error against the FP64 kernel on a target subset, step-latency percentiles and the
CPU baseline (the oracle's Vc-style loop on this host's cores, on a bounded
target sample).

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-error]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 peak, vector and matrix alike
FLOPS_PAIR = 68.0          # reference flop count of one EXP vel+grad interaction (Kernels.h)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-each", type=int, default=524288)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--error-targets", type=int, default=16384)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--tpt", type=int, default=0)
    ap.add_argument("--method", default="auto", choices=["auto", "direct", "factored"])
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run the known-traffic copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-instance batched workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the served RK2 workload")
    ap.add_argument("--no-panels", action="store_true", help="skip the icosphere panel workload")
    ap.add_argument("--no-bem", action="store_true", help="skip the BEM influence-matrix workload")
    ap.add_argument("--bem-only", action="store_true", help="run only the BEM workloads (matrix, right-hand side)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--sweep", action="store_true",
                    help="time every (precision, tile, tpt) on the N=1M workload, print them and exit")
    ap.add_argument("--valu-probe", action="store_true",
                    help="measure the FP32 VALU issue cost per wave64 instruction (fmac, add, mul, rsq), print and exit")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bem_leg(args, log):
    """The BEM influence matrix of the icosphere (omega3d_amd.workloads.bem_matrix_run),
    with its CPU baseline and check on the host unless --no-cpu."""
    from omega3d_amd import workloads
    bem, smp = workloads.bem_matrix_run(log=log)
    if not args.no_cpu:
        # the sampled columns through the reference's own recursion on the host
        # cores (oracle/_ref, S = float): the CPU baseline of this block, and the
        # check of the GPU columns against it and the oracle's FP64 tree
        import oracle
        from omega3d_amd.metrics import nerr
        b9, _ = oracle.bem_bases(smp["nodes"], smp["idx"])
        full = (smp["nodes"], smp["idx"], smp["area"], b9)
        js, nu = smp["sources"], smp["nunk"]
        sub = (smp["nodes"], np.ascontiguousarray(smp["idx"][js]), smp["area"][js], np.ascontiguousarray(b9[:, js]))
        keep = np.ones(smp["columns"].shape, bool)
        for a_, j in enumerate(js):    # each panel's own block: constants, not a recursion
            keep[j * nu:(j + 1) * nu, a_ * nu:(a_ + 1) * nu] = False
        want = oracle.bem_matrix("wl", sub, full, nu, False, 3, mode="f64t")
        bem["errors_vs_fp64_tree"] = {"max_rel": nerr(smp["columns"], want, keep), "columns": int(keep.shape[1])}
        try:
            t = time.perf_counter()
            r32 = oracle.ref_bem_matrix("wl", sub, full, nu, False, 3, mode="f32")
            secs = time.perf_counter() - t
            P = smp["idx"].shape[0]
            bem["cpu"] = {"kind": "reference", "cores": oracle.lib().orc_max_threads(),
                          "what": "the reference's rkernel_2vs_2p through panels_on_panels_coeff's loop "
                                  "(oracle/_ref, S = float, OpenMP over source panels)",
                          "sample": f"{len(js)} source panels x all {P} target panels",
                          "sample_s": round(secs, 3), "full_block_s_estimate": round(secs * P / len(js), 1),
                          "reference_float_vs_fp64_tree": nerr(r32, want, keep),
                          "gpu_vs_reference_float": nerr(smp["columns"], r32.astype(np.float64), keep)}
        except FileNotFoundError:
            bem["cpu"] = None
        log(f"bench: bem matrix check {bem.get('errors_vs_fp64_tree')} cpu {bem.get('cpu')}")
    return bem


def bem_rhs_leg(args, log):
    """The BEM right-hand side of the icosphere (omega3d_amd.workloads.bem_rhs_run),
    with its CPU baseline and check on the host unless --no-cpu: the sampled panels
    through the reference's own rkernel_2vs_0p used backwards (oracle/_ref,
    S = float, A = double) and the oracle's finalize and projection."""
    from omega3d_amd import workloads
    res, smp = workloads.bem_rhs_run(log=log)
    if not args.no_cpu:
        import oracle
        from omega3d_amd.metrics import rel_errors
        js = smp["panels"]
        idx = np.ascontiguousarray(smp["idx"][js])
        try:
            t = time.perf_counter()
            pu = oracle.ref_ptpan("rm", smp["pos"], smp["str"], smp["nodes"], idx, smp["area"][js], mode="f32d")
            secs = time.perf_counter() - t
            kind = "reference"
        except FileNotFoundError:
            t = time.perf_counter()
            pu = oracle.ptpan("rm", smp["pos"], smp["str"], smp["nodes"], idx, smp["area"][js], mode="f32d")
            secs = time.perf_counter() - t
            kind = "port"
        pu_fin, rhs = oracle.bem_rhs(smp["nodes"], idx, pu, smp["fs"], nunk=smp["nunk"])
        P = smp["idx"].shape[0]
        res["errors_vs_reference"] = {"pu_max_rel": rel_errors(smp["pu"], pu_fin)[0],
                                      "rhs_max_rel": rel_errors(smp["rhs"].T, rhs.reshape(-1, smp["nunk"]).T)[0],
                                      "panels": int(len(js))}
        res["cpu"] = {"kind": kind, "cores": oracle.lib().orc_max_threads(),
                      "what": "points_affect_panels' loop over the reference's rkernel_2vs_0p used backwards "
                              "(Influence.h:1190-1210; S = float, A = double), OpenMP over panels",
                      "sample": f"{len(js)} panels x all {smp['pos'].shape[1]} particles",
                      "sample_s": round(secs, 3), "full_s_estimate": round(secs * P / len(js), 2)}
        log(f"bench: bem rhs check {res['errors_vs_reference']} cpu {res['cpu']}")
    return res


def scene_io_leg(c, log):
    """Row f3's data formats on the headline workload: the 1,048,576 particles of
    the two rings written as the reference's VTU file (omega3d_amd.scene), timed,
    and beside it the reference's own writer (VtkXmlWriter.h through
    oracle/_ref/libo3dref_vtk.so) on the same arrays, with a byte comparison."""
    import tempfile
    from omega3d_amd import scene
    vel = np.zeros((3, c.n), np.float32)
    t = time.perf_counter()
    data = scene.points_vtu_bytes(c.pos, 0.0, c.str, c.rad, vel)
    secs = time.perf_counter() - t
    out = {"workload": f"the {c.n} particles of the two rings written as a VTU file (position, circulation, "
                       "radius, velocity; base64 binary)", "bytes": len(data), "seconds": round(secs, 3),
           "mb_per_s": round(len(data) / secs / 1e6, 1)}
    ref = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "libo3dref_vtk.so")
    if os.path.exists(ref):
        L = C.CDLL(ref)
        L.ref_write_points_vtk.argtypes = [C.c_char_p, C.c_int64] + [C.c_void_p] * 10 + [C.c_double, C.c_int]
        L.ref_write_points_vtk.restype = C.c_int
        arrs = [np.ascontiguousarray(x, np.float32) for x in (c.pos[0], c.pos[1], c.pos[2], c.str[0], c.str[1],
                                                             c.str[2], c.rad, vel[0], vel[1], vel[2])]
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "ref.vtu")
            t = time.perf_counter()
            L.ref_write_points_vtk(fn.encode(), c.n, *[x.ctypes.data_as(C.c_void_p) for x in arrs], 0.0, 0)
            rsecs = time.perf_counter() - t
            same = open(fn, "rb").read() == data
        out["reference_writer"] = {"kind": "reference", "seconds": round(rsecs, 3), "byte_identical": same}
    log("scene_io: " + json.dumps(out))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding over several GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    from omega3d_amd import _lib, scenes
    from omega3d_amd.metrics import field_errors
    L = _lib.lib
    _lib.check(L.o3d_init(0))
    dev = _lib.device_info()
    log(f"bench: device {dev['name']} CUs {dev['cus']}")
    if args.valu_probe:
        out = {}
        for op, name in enumerate(("v_fmac_f32", "v_add_f32", "v_mul_f32", "v_rsq_f32")):
            ms, wi = C.c_float(), C.c_double()
            _lib.check(L.o3d_valu_probe(op, C.byref(ms), C.byref(wi)))
            simd_cycles = dev["cus"] * 4 * dev["clock_khz"] * 1e3 * ms.value * 1e-3
            out[name] = {"ms": round(ms.value, 4), "wave_instructions": wi.value,
                         "cycles_per_wave_instruction_at_max_clock": round(simd_cycles / wi.value, 3)}
        print(json.dumps({"valu_probe": out, "clock_mhz": dev["clock_khz"] // 1000, "cus": dev["cus"]}))
        return

    if args.bem_only:
        print(json.dumps({"bem_matrix": bem_leg(args, log), "bem_rhs": bem_rhs_leg(args, log)}))
        return

    t0 = time.perf_counter()
    c = scenes.colliding_rings(n_each=args.n_each)
    N = c.n
    log(f"bench: generated two rings N={N} in {time.perf_counter() - t0:.2f}s")
    D = _lib.DeviceArray.from_host
    dpos, dstr, drad = D(c.pos), D(c.str), D(c.rad)
    dvel = _lib.DeviceArray((3, N))
    dgrad = _lib.DeviceArray((9, N))
    src, tgt = _lib.Points(), _lib.Points()
    src.n = tgt.n = N
    for d in range(3):
        src.x[d] = tgt.x[d] = dpos.row(d)
        src.s[d] = dstr.row(d)
        tgt.u[d] = dvel.row(d)
    src.r = tgt.r = drad.ptr
    for k in range(9):
        tgt.ug[k] = dgrad.row(k)

    def run(prec, steps, warmup, method=args.method):
        o = _lib.make_opts("exp", True, prec, 0, args.tile, args.tpt, method=method)
        evs = [(_lib.Event(), _lib.Event()) for _ in range(steps)]
        for _ in range(warmup):
            dvel.zero(); dgrad.zero()
            _lib.check(L.o3d_points_affect_points_dev(C.byref(src), C.byref(tgt), C.byref(o), None))
        _lib.synchronize()
        t = time.perf_counter()
        for k in range(steps):
            dvel.zero(); dgrad.zero()
            evs[k][0].record()
            _lib.check(L.o3d_points_affect_points_dev(C.byref(src), C.byref(tgt), C.byref(o), None))
            evs[k][1].record()
        _lib.synchronize()
        wall = time.perf_counter() - t
        kms = np.array([a.elapsed_ms(b) for a, b in evs])
        return wall, kms

    if args.sweep:
        res = []
        for prec in ("comp", "fp32"):
            for tile in (128, 256, 512, 1024):
                for tpt in (1, 2, 4):
                    o = _lib.make_opts("exp", True, prec, 0, tile, tpt, method="direct")
                    e0, e1 = _lib.Event(), _lib.Event()
                    _lib.check(L.o3d_points_affect_points_dev(C.byref(src), C.byref(tgt), C.byref(o), None))
                    e0.record()
                    _lib.check(L.o3d_points_affect_points_dev(C.byref(src), C.byref(tgt), C.byref(o), None))
                    e1.record()
                    _lib.synchronize()
                    ms = e0.elapsed_ms(e1)
                    res.append({"precision": prec, "tile": tile, "tpt": tpt, "ms": round(ms, 2),
                                "ginter_s": round(float(N) * N / ms / 1e6, 1)})
                    log(f"sweep: {res[-1]}")
        print(json.dumps({"sweep": res}))
        return

    if args.copy_probe_mb:
        nb = args.copy_probe_mb << 20
        a_, b_ = _lib.DeviceArray(nb // 4), _lib.DeviceArray(nb // 4)
        a_.zero()
        for _ in range(3):
            _lib.check(L.o3d_copy_probe(b_.ptr, a_.ptr, nb, None))
        _lib.synchronize()
        log(f"bench: copy probe {args.copy_probe_mb} MiB x3 done")
        del a_, b_

    log(f"bench: compensated FP32, {args.warmup} warmup + {args.steps} timed steps")
    wall, kms = run("comp", args.steps, args.warmup)
    vel_c, grad_c = dvel.to_host(), dgrad.to_host()
    inter = float(N) * float(N)
    value = inter * args.steps / wall / 1e9
    kavg = kms.mean() / 1e3
    flops = N * (12.0 + FLOPS_PAIR * N)
    call_tflops = flops / kavg / 1e12
    log(f"bench: comp {value:.1f} G int/s, call avg {kms.mean():.1f} ms, {call_tflops:.2f} TFLOP/s")

    # stage times of the factored path, HIP events on the evaluation's stream
    factored = args.method == "factored" or (args.method == "auto" and args.tile == 0 and args.tpt == 0)
    prof = None
    if factored:
        o = _lib.make_opts("exp", True, "comp", 0, args.tile, 0, method="factored")
        recs = []
        for _ in range(max(1, min(args.steps, 3))):
            dvel.zero(); dgrad.zero()
            pr = _lib.FactoredProfile()
            _lib.check(L.o3d_factored_profile_dev(C.byref(src), C.byref(tgt), C.byref(o), C.byref(pr)))
            recs.append({k: getattr(pr, k) for k, _ in _lib.FactoredProfile._fields_})
        prof = {k: float(np.mean([r[k] for r in recs])) for k in recs[0]}
        prof["steps"] = len(recs)
        log(f"bench: factored stages {prof}")

    psteps = max(1, min(args.steps, 3))
    pwall, pkms = run("fp32", psteps, 1)
    vel_p, grad_p = dvel.to_host(), dgrad.to_host()
    plain_value = inter * psteps / pwall / 1e9
    log(f"bench: plain fp32 {plain_value:.1f} G int/s, kernel avg {pkms.mean():.1f} ms")

    direct = None
    if factored:
        dwall, dkms = run("comp", 1, 0, method="direct")
        direct = {"value": round(inter / dwall / 1e9, 3), "unit": "G interactions/s",
                  "kernel_ms": round(float(dkms.mean()), 3), "kernel": "k_p2p<EXP,thick,grad,compensated>",
                  "tflops": round(flops / (dkms.mean() / 1e3) / 1e12, 3)}
        log(f"bench: direct method {direct}")

    errors = None
    if not args.no_error:
        nt = min(args.error_targets, N)
        sel = np.linspace(0, N - 1, nt).astype(np.int64)
        t = time.perf_counter()
        tp = np.ascontiguousarray(c.pos[:, sel], np.float64)
        tr = np.ascontiguousarray(c.rad[sel], np.float64)
        d64 = lambda a: _lib.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))
        spos64, sstr64, srad64 = d64(c.pos), d64(c.str), d64(c.rad)
        tpos64, trad64 = d64(tp), d64(tr)
        v64 = _lib.DeviceArray.from_host(np.zeros((3, nt)))
        g64 = _lib.DeviceArray.from_host(np.zeros((9, nt)))
        s64, t64 = _lib.Points64(), _lib.Points64()
        s64.n, t64.n = N, nt
        for d in range(3):
            s64.x[d], s64.s[d] = spos64.row(d), sstr64.row(d)
            t64.x[d], t64.u[d] = tpos64.row(d), v64.row(d)
        s64.r, t64.r = srad64.ptr, trad64.ptr
        for k in range(9):
            t64.ug[k] = g64.row(k)
        o = _lib.make_opts("exp", True, "fp64")
        _lib.check(L.o3d_points_affect_points64_dev(C.byref(s64), C.byref(t64), C.byref(o), None))
        _lib.synchronize()
        vref, gref = v64.to_host(), g64.to_host()
        errors = {"reference": "FP64 kernel", "targets": nt,
                  "compensated": field_errors(vel_c[:, sel], grad_c[:, sel], vref, gref),
                  "plain_fp32": field_errors(vel_p[:, sel], grad_p[:, sel], vref, gref)}
        log(f"bench: FP64 subset ({nt} targets) in {time.perf_counter() - t:.1f}s: {errors}")

    cpu = None
    if not args.no_cpu:
        import oracle
        nth = oracle.lib().orc_max_threads()
        if oracle.lib().orc_have_avx2():
            # calibrate on a few targets, then run ~cpu_seconds worth
            t = time.perf_counter()
            oracle.p2p_vcstyle("exp", c.pos, c.str, c.rad, c.pos, c.rad, 0, 4 * nth, nth)
            dt = time.perf_counter() - t
            ntc = int(min(N, max(4 * nth, 4 * nth * args.cpu_seconds / max(dt, 1e-3))))
            t = time.perf_counter()
            oracle.p2p_vcstyle("exp", c.pos, c.str, c.rad, c.pos, c.rad, 0, ntc, nth)
            dt = time.perf_counter() - t
            cpu = {"value": round(ntc * float(N) / dt / 1e9, 4), "unit": "G interactions/s", "cores": nth,
                   "kind": "port",
                   "sample": f"Vc-style AVX2 restatement of Influence.h:388-442 (8 float lanes, rsqrt/rcp "
                             f"estimates, OpenMP over targets), first {ntc} targets x all {N} sources, "
                             f"{dt:.1f} s"}
            # beside it, the reference's own scalar kernels (oracle/_ref: Kernels.h and
            # CoreFunc.h compiled in place, EXP core, S = float, A = double: the
            # non-Vc build) on a few targets of the same workload
            try:
                nr = 16 * nth
                t = time.perf_counter()
                oracle.ref_p2p("exp", c.pos, c.str, c.rad, c.pos[:, :nr], c.rad[:nr], True, "f32d")
                rdt = time.perf_counter() - t
                cpu["reference_scalar"] = {
                    "value": round(nr * float(N) / rdt / 1e9, 4), "unit": "G interactions/s", "cores": nth,
                    "kind": "reference", "sample": f"first {nr} targets x all {N} sources, {rdt:.1f} s"}
            except FileNotFoundError:
                cpu["reference_scalar"] = None
            log(f"bench: cpu baseline {cpu}")

    from omega3d_amd import workloads
    single = workloads.single_ring_run(log=log)
    if not args.no_cpu:
        # the same single ring on the host cores: the reference's scalar loop restated
        # (float storage and accumulation) and the Vc-style loop
        import oracle
        p0 = scenes.ring(8192, sigma=0.05)
        nth = oracle.lib().orc_max_threads()
        t = time.perf_counter()
        oracle.p2p("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, True, "f32f", nthreads=nth)
        t_exact = time.perf_counter() - t
        t_vc = None
        if oracle.lib().orc_have_avx2():
            t = time.perf_counter()
            oracle.p2p_vcstyle("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, 0, p0.n, nth)
            t_vc = time.perf_counter() - t
        single["cpu"] = {"cores": nth, "scalar_float_ms": round(t_exact * 1e3, 2),
                         "vc_style_ms": None if t_vc is None else round(t_vc * 1e3, 2),
                         "kind": "port", "what": "the oracle's restatement of Influence.h:202-545 on the host "
                                                 "cores (self pair included), full workload"}
        log(f"bench: single ring on the CPU {single['cpu']}")
    batched = None if args.no_batch else workloads.batched(log=log)
    served = None if args.no_serve else workloads.served(duration=args.serve_seconds, log=log)
    panels = None if args.no_panels else workloads.panels_run(log=log)
    if panels is not None and not args.no_cpu:
        # the c2 workload's recursive rule through the reference's own rkernel_2vs_0pg
        # (oracle/_ref, S = float, A = double) on a spread sample of the targets
        import oracle
        pan = scenes.icosphere_panels()
        tp = scenes.shell_points(524288)
        sel = np.linspace(0, tp.shape[1] - 1, 1024).astype(np.int64)
        try:
            t = time.perf_counter()
            oracle.ref_pan("rm", pan.nodes, pan.idx.reshape(-1, 3), pan.ts, pan.area, pan.ss, tp[:, sel], True, "f32d")
            secs = time.perf_counter() - t
            panels["cpu"] = {"kind": "reference", "cores": oracle.lib().orc_max_threads(),
                             "what": "panels_affect_points' loop over the reference's rkernel_2vs_0pg "
                                     "(Influence.h:729-760; S = float, A = double), OpenMP over targets",
                             "sample": f"{len(sel)} spread targets x all {pan.np} panels", "sample_s": round(secs, 3),
                             "full_s_estimate": round(secs * tp.shape[1] / len(sel), 2)}
        except FileNotFoundError:
            panels["cpu"] = None
        log(f"bench: panels cpu {panels['cpu']}")
    scene_io = None if args.no_cpu else scene_io_leg(c, log)
    bem = None if args.no_bem else bem_leg(args, log)
    bem_rhs = None if args.no_bem else bem_rhs_leg(args, log)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "ff_far_exp_n1m_traffic.json" if factored else "p2p_exp_n1m_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("traffic_bytes")
    executed = None
    cfile = os.path.join(ROOT, "profiles", "ff_far_exp_n1m_counters.json")
    if factored and os.path.exists(cfile):
        executed = json.load(open(cfile))
    if factored:
        far_flops = FLOPS_PAIR * prof["far_pairs"]
        achieved = far_flops / (prof["far_ms"] / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_ff_far<EXP,thick,compensated>", "kernel_ms": round(prof["far_ms"], 3),
                "flops_per_launch": far_flops, "far_pair_fraction": round(prof["far_tile_fraction"], 5),
                "stages_ms": {k: round(prof[k], 3) for k in ("prep_ms", "far_ms", "near_ms", "total_ms")},
                "whole_call": {"achieved": round(call_tflops, 3), "frac": round(call_tflops / FP32_PEAK_TFLOPS, 4),
                               "what": "reference flops of all N^2 pairs over the whole evaluation (sort, far, "
                                       "near)"},
                "note": "achieved counts the reference's 68 FP32 flops per far pair; the kernel executes an exact "
                        "factorisation instead (about 8 VALU operations per pair plus BF16/FP16 split matrix products), "
                        "so it can exceed the FP32 peak; 'executed' gives the utilisation of the units it runs on",
                "executed": executed}
    else:
        roof = {"bound": "valu", "achieved": round(call_tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(call_tflops / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_p2p<EXP,thick,grad,compensated>", "kernel_ms": round(float(kms.mean()), 3),
                "flops_per_launch": flops}

    out = {
        "metric": "G interactions/s (vel+grad) at N=1M vs plain FP32; p50/p99 latency, mixed N",
        "value": round(value, 3),
        "unit": "G interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "two colliding vortex rings, N=1,048,576 (524,288 each), compact-exponential core, "
                               "sigma 0.02, velocity + gradient at every particle, compensated FP32",
                   "n": N, "core": "exp", "precision": "compensated", "self_term": "included",
                   "method": "factored" if factored else "direct"},
        "plain_fp32": {"value": round(plain_value, 3), "unit": "G interactions/s",
                       "kernel_ms": round(float(pkms.mean()), 3)},
        "latency_ms": ({"p50": round(served["p50_s"] * 1e3, 3), "p99": round(served["p99_s"] * 1e3, 3),
                        "what": "served RK2 step latency, mixed N (see served)"} if served else
                       {"p50": round(float(np.percentile(kms, 50)), 3),
                        "p99": round(float(np.percentile(kms, 99)), 3),
                        "what": "per-step kernel time of this workload"}),
        "roofline": roof,
        "direct_method": direct,
        "errors_vs_fp64": errors,
        "cpu_baseline": cpu,
        "single_ring": single,
        "batched": batched,
        "served": served,
        "panels": panels,
        "scene_io": scene_io,
        "bem_matrix": bem,
        "bem_rhs": bem_rhs,
        "device": dev["name"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
