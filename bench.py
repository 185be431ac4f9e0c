"""Headline benchmark: direct Biot-Savart velocity + velocity gradient at N = 1,048,576.

Workload (one "step" = zeroing the outputs and one evaluation of every particle's
velocity and 3x3 velocity gradient from every particle, inputs resident in HBM):
  two colliding vortex rings, 524,288 particles each (centres z = +-1, lower ring
  tilted 15 deg about x, circulation +-1, |alpha| jittered +-5%), core radius 0.02,
  compact-exponential core, FP32 pairs with compensated accumulation.

The headline is the direct pair-by-pair sum on the vector ALUs (k_p2p, method
DIRECT): `value` is G interactions/s over the timed steps, `roofline` its FP32
VALU rate in the reference's flop count (68 per interaction + 12 per target) over
the call's HIP-event time, against the 157.3 TFLOP/s FP32 vector peak.  Beside it:
the plain-FP32 direct rate, max/RMS relative error of both against the FP64
kernel on a target subset, the factored method (the library's AUTO choice at this
size; matrix-core arithmetic, reported under its own key), and the CPU baseline:
the reference's own kernels (oracle/_ref) on this host's cores on a bounded
target sample, with the GPU result compared against them.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu] [--no-error]
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="timed steps of the plain-FP32 and factored legs (each after one warm-up)")
    ap.add_argument("--n-each", type=int, default=524288)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--error-targets", type=int, default=16384)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--tpt", type=int, default=0)
    ap.add_argument("--method", default="direct", choices=["auto", "direct", "factored"],
                    help="method of the headline leg (the north-star kernel is the direct one)")
    ap.add_argument("--no-factored", action="store_true", help="skip the factored-method leg")
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


def fp64_subset(L, _lib, c, sel):
    """FP64 kernel (k_p2p64): targets c[:, sel] against all of c, double device arrays."""
    nt = len(sel)
    d64 = lambda a: _lib.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))
    spos64, sstr64, srad64 = d64(c.pos), d64(c.str), d64(c.rad)
    tpos64, trad64 = d64(c.pos[:, sel]), d64(c.rad[sel])
    v64 = _lib.DeviceArray.from_host(np.zeros((3, nt)))
    g64 = _lib.DeviceArray.from_host(np.zeros((9, nt)))
    s64, t64 = _lib.Points64(), _lib.Points64()
    s64.n, t64.n = c.n, nt
    for d in range(3):
        s64.x[d], s64.s[d] = spos64.row(d), sstr64.row(d)
        t64.x[d], t64.u[d] = tpos64.row(d), v64.row(d)
    s64.r, t64.r = srad64.ptr, trad64.ptr
    for k in range(9):
        t64.ug[k] = g64.row(k)
    o = _lib.make_opts("exp", True, "fp64")
    _lib.check(L.o3d_points_affect_points64_dev(C.byref(s64), C.byref(t64), C.byref(o), None))
    _lib.synchronize()
    return v64.to_host(), g64.to_host()


def traffic_of(kernel_file):
    """HBM bytes per launch of a kernel from a committed rocprofv3 --pmc summary
    (profiles/*.json), with the command, commit and date it records: the counters
    cannot be read from inside this process."""
    f = os.path.join(ROOT, "profiles", kernel_file)
    if not os.path.exists(f):
        return None, None
    d = json.load(open(f))
    src = {"file": "profiles/" + kernel_file, "commit": d.get("commit"), "date": d.get("date"),
           "command": d.get("command"), "measured_kernel_ms": d.get("kernel_ms")}
    return d.get("traffic_bytes"), src


def cpu_leg(args, c, vel_c, grad_c, log):
    """CPU baseline of the headline workload on this host's cores: the reference's
    own EXP kernels (oracle/_ref: Kernels.h / CoreFunc.h compiled in place, the
    non-Vc build, S = float, A = double, OpenMP over targets; Influence.h:405-440)
    on a spread sample of targets sized for about --cpu-seconds, with the GPU
    result on the same targets compared against it; beside it the oracle's
    Vc-style AVX2 restatement of the Vc loop.  Without oracle/_ref the port is the
    baseline."""
    import oracle
    from omega3d_amd.metrics import field_errors
    N = c.n
    nth = oracle.lib().orc_max_threads()
    out = None
    if oracle.ref_lib("exp") is not None:
        nr0 = 8 * nth
        sel0 = np.linspace(0, N - 1, nr0).astype(np.int64)
        t = time.perf_counter()
        oracle.ref_p2p("exp", c.pos, c.str, c.rad, c.pos[:, sel0], c.rad[sel0], True, "f32d")
        dt0 = time.perf_counter() - t
        ntc = int(min(N, max(1024, nr0 * args.cpu_seconds / max(dt0, 1e-3))))
        sel = np.linspace(0, N - 1, ntc).astype(np.int64)
        t = time.perf_counter()
        rv, rg = oracle.ref_p2p("exp", c.pos, c.str, c.rad, c.pos[:, sel], c.rad[sel], True, "f32d")
        dt = time.perf_counter() - t
        out = {"value": round(ntc * float(N) / dt / 1e9, 4), "unit": "G interactions/s", "cores": nth,
               "kind": "reference",
               "sample": f"the reference's kernel_0v_0bg / EXP core_func (oracle/_ref, S = float, A = double, "
                         f"OpenMP over targets) for {ntc} evenly spread targets x all {N} sources, {dt:.1f} s",
               "gpu_vs_reference": dict(field_errors(vel_c[:, sel], grad_c[:, sel], rv, rg), targets=ntc,
                                        what="the headline's compensated direct result against the reference's "
                                             "own kernels on the same targets (bar 1e-5)")}
    if oracle.lib().orc_have_avx2():
        t = time.perf_counter()
        oracle.p2p_vcstyle("exp", c.pos, c.str, c.rad, c.pos, c.rad, 0, 4 * nth, nth)
        dt = time.perf_counter() - t
        ntc = int(min(N, max(4 * nth, 4 * nth * args.cpu_seconds / 2 / max(dt, 1e-3))))
        t = time.perf_counter()
        oracle.p2p_vcstyle("exp", c.pos, c.str, c.rad, c.pos, c.rad, 0, ntc, nth)
        dt = time.perf_counter() - t
        port = {"value": round(ntc * float(N) / dt / 1e9, 4), "unit": "G interactions/s", "cores": nth,
                "kind": "port",
                "sample": f"Vc-style AVX2 restatement of Influence.h:388-442 (8 float lanes, rsqrt/rcp estimates, "
                          f"OpenMP over targets), first {ntc} targets x all {N} sources, {dt:.1f} s"}
        if out is None:
            out = port
        else:
            out["vc_style_port"] = port
    log(f"bench: cpu baseline {out}")
    return out


def single_ring_cpu(single, log):
    """c0 on the host cores: the reference's own RM kernels (oracle/_ref, S = float,
    A = double) over the full 8,192 x 8,192 workload, repeated for at least 5 s, with
    the GPU's self-included FP32 and compensated results compared against them; and
    the oracle's Vc-style loop."""
    import oracle
    from omega3d_amd import influence, scenes
    from omega3d_amd.metrics import field_errors
    p0 = scenes.ring(8192, sigma=0.05)
    nth = oracle.lib().orc_max_threads()
    out = {"cores": nth}
    if oracle.ref_lib("rm") is not None:
        reps, t = 0, time.perf_counter()
        while True:
            rv, rg = oracle.ref_p2p("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, True, "f32d")
            reps += 1
            if time.perf_counter() - t >= 5.0:
                break
        secs = (time.perf_counter() - t) / reps
        gv, gg = influence.p2p("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, True, "fp32")
        cv, cg = influence.p2p("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, True, "comp")
        out["reference"] = {"kind": "reference", "ms": round(secs * 1e3, 3), "reps": reps,
                            "ginter_s": round(p0.n * float(p0.n) / secs / 1e9, 3),
                            "what": "the reference's kernel_0v_0bg / RM core_func (oracle/_ref, S = float, "
                                    "A = double, OpenMP over targets), every target, self pair included",
                            "gpu_fp32_vs_reference": field_errors(gv, gg, rv, rg),
                            "gpu_compensated_vs_reference": field_errors(cv, cg, rv, rg)}
    if oracle.lib().orc_have_avx2():
        t = time.perf_counter()
        oracle.p2p_vcstyle("rm", p0.pos, p0.str, p0.rad, p0.pos, p0.rad, 0, p0.n, nth)
        out["vc_style_port_ms"] = round((time.perf_counter() - t) * 1e3, 3)
    log(f"bench: single ring on the CPU {out}")
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
        for op, name in enumerate(("v_fmac_f32 (scalar-register operand)", "v_add_f32", "v_mul_f32", "v_rsq_f32",
                                   "v_fmac_f32 (vector-register operands)")):
            ms, wi = C.c_float(), C.c_double()
            _lib.check(L.o3d_valu_probe(op, C.byref(ms), C.byref(wi)))
            simd_cycles = dev["cus"] * 4 * dev["clock_khz"] * 1e3 * ms.value * 1e-3
            out[name] = {"ms": round(ms.value, 4), "wave_instructions": wi.value,
                         "cycles_per_wave_instruction_at_max_clock": round(simd_cycles / wi.value, 3)}
        # one wave per SIMD: FMAs alone, MFMAs alone, both in the same waves, and one
        # MFMA wave beside one FMA wave (two per SIMD); each probe iteration of a wave
        # is 64 FMAs and/or 4 v_mfma_f32_32x32x16_f16
        mix = {}
        for op, name in ((5, "fma waves, 1 per SIMD"), (6, "mfma waves, 1 per SIMD"),
                         (7, "mfma and fma in the same wave, 1 per SIMD"),
                         (8, "1 mfma wave + 1 fma wave per SIMD")):
            ms, wi = C.c_float(), C.c_double()
            _lib.check(L.o3d_valu_probe(op, C.byref(ms), C.byref(wi)))
            waves = dev["cus"] * 4
            iters = wi.value / waves / (64.0 if op == 5 else 4.0)
            mix[name] = {"ms": round(ms.value, 4), "ns_per_iteration_per_wave": round(ms.value * 1e6 / iters, 2)}
        print(json.dumps({"valu_probe": out, "mfma_valu_overlap": mix, "clock_mhz": dev["clock_khz"] // 1000,
                          "cus": dev["cus"]}))
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

    def run(prec, steps, warmup, method):
        """warmup + steps calls; each step zeroes the outputs and makes one call, the
        call bracketed by HIP events on the stream it runs on (the null stream)"""
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

    inter = float(N) * float(N)
    flops = N * (12.0 + FLOPS_PAIR * N)       # Influence.h:475 with flops_0v_0bg<EXP> = 68

    # ---- headline: the direct kernel, compensated --------------------------------
    log(f"bench: direct, compensated FP32, {args.warmup} warmup + {args.steps} timed steps")
    wall, kms = run("comp", args.steps, args.warmup, args.method)
    vel_c, grad_c = dvel.to_host(), dgrad.to_host()
    value = inter * args.steps / wall / 1e9
    kavg = float(kms.mean())
    achieved = flops / (kavg / 1e3) / 1e12
    log(f"bench: direct comp {value:.1f} G int/s, call avg {kavg:.2f} ms, {achieved:.2f} TFLOP/s")

    psteps = max(1, min(args.steps, args.plain_steps))
    pwall, pkms = run("fp32", psteps, 1, args.method)
    vel_p, grad_p = dvel.to_host(), dgrad.to_host()
    plain = {"value": round(inter * psteps / pwall / 1e9, 3), "unit": "G interactions/s",
             "call_ms": round(float(pkms.mean()), 3), "steps": psteps,
             "tflops": round(flops / (pkms.mean() / 1e3) / 1e12, 3),
             "kernel": "k_p2p<EXP,thick,grad,fp32>"}
    log(f"bench: direct plain fp32 {plain}")

    # ---- the factored method (AUTO's choice at this size), its own arithmetic -------
    factored = None
    if not args.no_factored:
        fsteps = max(1, min(args.steps, args.plain_steps))
        fwall, fkms = run("comp", fsteps, 1, "factored")
        vel_f, grad_f = dvel.to_host(), dgrad.to_host()
        o = _lib.make_opts("exp", True, "comp", 0, 0, 0, method="factored")
        recs = []
        for _ in range(3):
            dvel.zero(); dgrad.zero()
            pr = _lib.FactoredProfile()
            _lib.check(L.o3d_factored_profile_dev(C.byref(src), C.byref(tgt), C.byref(o), C.byref(pr)))
            recs.append({k: getattr(pr, k) for k, _ in _lib.FactoredProfile._fields_})
        prof = {k: float(np.mean([r[k] for r in recs])) for k in recs[0]}
        # matrix-core work of the far kernel, from its structure: per 32 sources x 32
        # targets, two 32x32x16 BF16 products (r^2) and six 32x32x16 FP16 products
        # (three split products x two 16-column halves of the moment table)
        mfma_flops = prof["far_pairs"] / 1024.0 * 8 * (32 * 32 * 16 * 2)
        factored = {"value": round(inter * fsteps / fwall / 1e9, 3), "unit": "G interactions/s",
                    "call_ms": round(float(fkms.mean()), 3), "steps": fsteps,
                    "dtype": "far pairs: r^2 as BF16x3 split products and weights x 29 moment columns as FP16x2 "
                             "split products on the matrix cores, FP32 VALU weights, FP64 epilogue folds; near "
                             "pairs: FP32 pairs with compensated accumulation, as the direct kernel",
                    "stages_ms": {k: round(prof[k], 3) for k in ("prep_ms", "far_ms", "near_ms", "total_ms")},
                    "far_pair_fraction": round(prof["far_tile_fraction"], 5),
                    "far_kernel": {"kernel": "k_ff_far<EXP,thick,compensated>", "ms": round(prof["far_ms"], 3),
                                   "matrix_flops_per_launch": mfma_flops,
                                   "matrix_tflops": round(mfma_flops / (prof["far_ms"] / 1e3) / 1e12, 2),
                                   "matrix_peak_frac": round(mfma_flops / (prof["far_ms"] / 1e3) / 2.5e15, 4),
                                   "what": "dense FP16/BF16 matrix-core rate of the far kernel's products "
                                           "(counted from the kernel's structure) over its HIP-event time, "
                                           "against the 2.5 PFLOP/s dense peak; unit utilisation from counters "
                                           "is in profiles/ff_far_exp_n1m_counters.md"}}
        log(f"bench: factored {factored}")

    errors = None
    if not args.no_error:
        nt = min(args.error_targets, N)
        sel = np.linspace(0, N - 1, nt).astype(np.int64)
        t = time.perf_counter()
        vref, gref = fp64_subset(L, _lib, c, sel)
        errors = {"reference": "FP64 kernel (k_p2p64, the reference's formulas in double)", "targets": nt,
                  "compensated": field_errors(vel_c[:, sel], grad_c[:, sel], vref, gref),
                  "plain_fp32": field_errors(vel_p[:, sel], grad_p[:, sel], vref, gref)}
        if factored is not None:
            factored["errors_vs_fp64"] = dict(field_errors(vel_f[:, sel], grad_f[:, sel], vref, gref), targets=nt)
        plain["errors_vs_fp64"] = dict(errors["plain_fp32"], targets=nt)
        log(f"bench: FP64 subset ({nt} targets) in {time.perf_counter() - t:.1f}s: {errors}")

    cpu = None if args.no_cpu else cpu_leg(args, c, vel_c, grad_c, log)

    from omega3d_amd import workloads
    single = workloads.single_ring_run(log=log)
    if not args.no_cpu:
        single["cpu"] = single_ring_cpu(single, log)
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

    traffic, tsrc = traffic_of("p2p_exp_n1m_traffic.json")
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic, "traffic_source": tsrc,
            "kernel": "k_p2p<EXP,thick,grad,compensated>", "kernel_ms": round(kavg, 3),
            "flops_per_launch": flops,
            "what": "the reference's count, N*(12 + 68 N) flops per launch (Influence.h:475, flops_0v_0bg of the "
                    "EXP core), over the mean HIP-event time of the call on its stream (the call also sorts the "
                    "targets into Morton order, about 0.1 ms); FP32 vector peak with packed FP32",
            "counters": "profiles/p2p_exp_n1m_counters.md"}

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
                               "sigma 0.02, velocity + gradient at every particle, FP32 pairs with compensated "
                               "accumulation, direct pair-by-pair sum",
                   "n": N, "core": "exp", "precision": "compensated", "self_term": "included",
                   "method": args.method},
        "plain_fp32": plain,
        "latency_ms": ({"p50": round(served["p50_s"] * 1e3, 3), "p99": round(served["p99_s"] * 1e3, 3),
                        "what": "served RK2 step latency, mixed N (see served)"} if served else
                       {"p50": round(float(np.percentile(kms, 50)), 3),
                        "p99": round(float(np.percentile(kms, 99)), 3),
                        "what": "per-step call time of this workload"}),
        "roofline": roof,
        "errors_vs_fp64": errors,
        "factored": factored,
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
