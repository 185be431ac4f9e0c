"""Writes the golden fixtures in this directory from the REFERENCE'S OWN kernels.

Run in a container that has the reference checkout (oracle/build_ref.sh builds
oracle/_ref/libo3dref_<core>.so from Omega3D's Kernels.h / CoreFunc.h /
MathHelper.h).  The fixtures hold seeded inputs and the reference's outputs:

  p2p_<core>.npz   points_affect_points scalar loop (Influence.h:405-440 and
                   :280-307) with kernel_0v_0bg / 0v_0pg / 0v_0b / 0v_0p, for
                   accumulators A = float, A = double (S = float) and S = A = double.
  pan_<core>.npz   panels_affect_points scalar loop (Influence.h:729-775) with
                   rkernel_2vs_0pg / rkernel_2vs_0p, maxlev = 3, covering every
                   recursion depth, with and without a source strength.
  ptpan_<core>.npz points_affect_panels scalar loop (Influence.h:1190-1210): the
                   panel kernel run backwards from particles onto the same panels,
                   particles placed so that every recursion depth is used.
  flops.npz        the reference's flops_0v_0bg .. flops_0vs_0p for each core.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

CORES = ["rm", "wl", "exp", "v2"]
MODES = ["f32f", "f32d", "f64"]


def p2p_inputs(rng):
    """257 sources (one coincident pair, radii spread across the EXP branches) and
    64 targets, 16 of which coincide with sources."""
    ns, nt = 257, 64
    spos = rng.uniform(-0.3, 0.3, (3, ns)).astype(np.float32)
    spos[:, 1] = spos[:, 0]
    sstr = rng.normal(size=(3, ns)).astype(np.float32) / ns
    srad = rng.uniform(0.02, 0.2, ns).astype(np.float32)
    tpos = rng.uniform(-0.35, 0.35, (3, nt)).astype(np.float32)
    tpos[:, :16] = spos[:, 100:116]
    trad = rng.uniform(0.02, 0.2, nt).astype(np.float32)
    return spos, sstr, srad, tpos, trad


def panel_inputs(rng):
    """An 80-panel icosphere (radius 0.5) plus one big panel; targets on a line through
    the big panel's plane and around the sphere so that every depth 0..3 is used."""
    from omega3d_amd import scenes
    verts, faces = scenes.icosphere(1, 0.5)
    big = np.array([[1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.5, 0.8, 0.0]])
    nodes = np.concatenate([verts, big]).T.astype(np.float32)
    faces = np.concatenate([faces, [[len(verts), len(verts) + 1, len(verts) + 2]]])
    p0, p1, p2 = (nodes.T[faces[:, k]].astype(np.float64) for k in range(3))
    area = (0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)).astype(np.float32)
    ts = (rng.normal(size=(3, len(faces))) * area).astype(np.float32)
    ss = rng.normal(size=len(faces)).astype(np.float32)
    line = np.stack([np.linspace(1.0, 2.0, 40), np.full(40, 0.3), np.linspace(0.001, 0.6, 40)])
    sph = rng.normal(size=(3, 60))
    sph = sph / np.linalg.norm(sph, axis=0) * rng.uniform(0.45, 0.9, 60)
    tpos = np.concatenate([line, sph], axis=1).astype(np.float32)
    return nodes, faces.reshape(-1).astype(np.uint32), ts, area, ss, tpos


def main():
    for core in CORES:
        if oracle.ref_lib(core) is None:
            raise SystemExit(f"{oracle.ref_path(core)} missing: run sh oracle/build_ref.sh first")
    rng = np.random.default_rng(20240607)
    flops = {}
    for core in CORES:
        spos, sstr, srad, tpos, trad = p2p_inputs(rng)
        out = dict(spos=spos, sstr=sstr, srad=srad, tpos=tpos, trad=trad)
        for mode in MODES:
            for thick in (True, False):
                for grads in (True, False):
                    v, g = oracle.ref_p2p(core, spos, sstr, srad, tpos, trad if thick else None, grads, mode)
                    key = f"{mode}_{'tv' if thick else 'tp'}_{'g' if grads else 'v'}"
                    out["vel_" + key] = v
                    if grads:
                        out["grad_" + key] = g
        np.savez_compressed(os.path.join(HERE, f"p2p_{core}.npz"), **out)

        nodes, idx, ts, area, ss, ptgt = panel_inputs(rng)
        out = dict(nodes=nodes, idx=idx, ts=ts, area=area, ss=ss, tpos=ptgt)
        for mode in MODES:
            for with_ss in (True, False):
                for grads in (True, False):
                    v, g = oracle.ref_pan(core, nodes, idx, ts, area, ss if with_ss else None, ptgt, grads, mode)
                    key = f"{mode}_{'vs' if with_ss else 'v'}_{'g' if grads else 'v'}"
                    out["vel_" + key] = v
                    if grads:
                        out["grad_" + key] = g
        np.savez_compressed(os.path.join(HERE, f"pan_{core}.npz"), **out)
        flops[core] = np.array([oracle.ref_flops(core, k) for k in range(6)], np.int32)
    np.savez_compressed(os.path.join(HERE, "flops.npz"), **flops)
    rng = np.random.default_rng(20241202)
    for core in CORES:
        nodes, idx, _, area, _, ppos = panel_inputs(rng)
        pstr = (rng.normal(size=ppos.shape) * 0.05).astype(np.float32)
        out = dict(nodes=nodes, idx=idx, area=area, spos=ppos, sstr=pstr)
        for mode in MODES:
            out["vel_" + mode] = oracle.ref_ptpan(core, ppos, pstr, nodes, idx, area, mode)
        np.savez_compressed(os.path.join(HERE, f"ptpan_{core}.npz"), **out)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
