"""Scene input and particle output (omega3d_amd/scene.py).

VTU output of particles and of panels is checked byte for byte against files the
reference's own writer produced (tests/golden/vtk_*.vtu from VtkXmlWriter.h, see
make_vtk_golden.py), and against that writer live when oracle/_ref holds it.  The JSON scene generators
cannot be built from the reference (FlowFeature.h needs Eigen3), so they are
pinned by geometry: ring particles on the circle, in the plane normal to the axis,
tangent strengths summing to circulation x circumference; the reference's counts
(1 + 2 pi R / ips stations, layered disks); blob and block strengths summing to the
given strength; the example scenes of the reference parse with the expected counts.
Their bodies' meshes (sphere, box, quads) are pinned the same way: closed and
consistently wound, on the sphere or the box faces, the reference's panel counts,
areas, normals toward the fluid, boundary velocities.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from omega3d_amd import scene

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REFVTK = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libo3dref_vtk.so")


@pytest.mark.parametrize("name", ["vtk_part", "vtk_fldpt", "vtk_one"])
def test_vtu_bytes_match_the_reference_writer(name):
    g = np.load(os.path.join(GOLD, "vtk_inputs.npz"))
    inert = bool(g[f"{name}_inert"])
    got = scene.points_vtu_bytes(g[f"{name}_pos"], float(g[f"{name}_time"]),
                                 None if inert else g[f"{name}_str"], None if inert else g[f"{name}_rad"],
                                 g[f"{name}_vel"])
    want = open(os.path.join(GOLD, name + ".vtu"), "rb").read()
    assert got == want


@pytest.mark.skipif(not os.path.exists(REFVTK), reason="oracle/_ref/libo3dref_vtk.so not built")
def test_vtu_bytes_match_the_live_reference_writer(tmp_path):
    L = C.CDLL(REFVTK)
    L.ref_write_points_vtk.argtypes = [C.c_char_p, C.c_int64] + [C.c_void_p] * 10 + [C.c_double, C.c_int]
    rng = np.random.default_rng(3)
    for n, inert in ((1000, 0), (333, 1)):
        a = [np.ascontiguousarray(rng.normal(size=n).astype(np.float32)) for _ in range(10)]
        fn = str(tmp_path / f"r{n}.vtu")
        L.ref_write_points_vtk(fn.encode(), n, *[x.ctypes.data_as(C.c_void_p) for x in a], 2.5, inert)
        got = scene.points_vtu_bytes(np.stack(a[0:3]), 2.5, None if inert else np.stack(a[3:6]),
                                     None if inert else a[6], np.stack(a[7:10]))
        assert got == open(fn, "rb").read(), (n, inert)



@pytest.mark.parametrize("name", ["vtk_pan_src", "vtk_pan", "vtk_pan_inert"])
def test_panel_vtu_bytes_match_the_reference_writer(name):
    g = np.load(os.path.join(GOLD, "vtk_inputs.npz"))
    got = scene.panels_vtu_bytes(g[f"{name}_nodes"], g[f"{name}_idx"], float(g[f"{name}_time"]),
                                 g[f"{name}_vs"] if g[f"{name}_vort"] else None,
                                 g[f"{name}_ss"] if g[f"{name}_src"] else None)
    assert got == open(os.path.join(GOLD, name + ".vtu"), "rb").read()


@pytest.mark.skipif(not os.path.exists(REFVTK), reason="oracle/_ref/libo3dref_vtk.so not built")
def test_panel_vtu_bytes_match_the_live_reference_writer(tmp_path):
    """the meshed bodies of the example scenes, with random sheet strengths, through
    the reference's writer and this one"""
    L = C.CDLL(REFVTK)
    L.ref_write_panels_vtk.argtypes = [C.c_char_p, C.c_int64] + [C.c_void_p] * 3 + [C.c_int64] + [C.c_void_p] * 5 + [
        C.c_double]
    rng = np.random.default_rng(4)
    for f in ("flow_over_sphere.json", "driven_cavity.json"):
        pan, _ = scene.Scene(os.path.join(GOLD, "scenes", f)).surfaces()
        vs = rng.normal(size=(3, pan.np)).astype(np.float32)
        ss = rng.normal(size=pan.np).astype(np.float32)
        for src in (True, False):
            keep = [np.ascontiguousarray(a) for a in (pan.nodes[0], pan.nodes[1], pan.nodes[2], pan.idx,
                                                      vs[0], vs[1], vs[2], ss)]
            ptr = [a.ctypes.data_as(C.c_void_p) for a in keep]
            fn = str(tmp_path / scene.panel_vtu_name(0, 7))
            L.ref_write_panels_vtk(fn.encode(), pan.nodes.shape[1], ptr[0], ptr[1], ptr[2], pan.np, ptr[3],
                                   ptr[4], ptr[5], ptr[6], ptr[7] if src else None, 0.75)
            got = scene.panels_vtu_bytes(pan.nodes, pan.idx, 0.75, vs, ss if src else None)
            assert got == open(fn, "rb").read(), (f, src)

def test_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    pos, st, vel = (rng.normal(size=(3, 50)).astype(np.float32) for _ in range(3))
    rad = rng.uniform(0.01, 0.1, 50).astype(np.float32)
    fn = scene.write_points_vtu(str(tmp_path / scene.vtu_name(0, 12)), pos, 0.5, st, rad, vel)
    assert os.path.basename(fn) == "part_00_00012.vtu"
    d = scene.read_points_vtu(fn)
    np.testing.assert_array_equal(d["position"], pos)
    np.testing.assert_array_equal(d["circulation"], st)
    np.testing.assert_array_equal(d["radius"], rad)
    np.testing.assert_array_equal(d["velocity"], vel)
    assert d["TimeValue"][0] == 0.5 and (d["connectivity"] == np.arange(50)).all()


def _ring_checks(pos, st, c, nrm, R, circ, ips):
    n = np.asarray(nrm, float) / np.linalg.norm(nrm)
    d = pos.astype(np.float64) - np.asarray(c, float)[:, None]
    assert np.abs(np.linalg.norm(d, axis=0) - R).max() < 2e-6 * max(1, R)
    assert np.abs(n @ d).max() < 2e-6
    a = st.astype(np.float64)
    assert np.abs((a * d).sum(0)).max() < 1e-6 * np.abs(a).max() * R and np.abs(n @ a).max() < 1e-6
    nd = int(1 + 2 * math.pi * R / ips)
    assert pos.shape[1] == nd
    assert abs(np.linalg.norm(a, axis=0).sum() - abs(circ) * 2 * math.pi * R) < 1e-5 * abs(circ) * 2 * math.pi * R
    # circulation sense: alpha along n x d
    assert (np.sign(circ) * (np.cross(n, d.T) * a.T).sum(1) > 0).all()


def test_example_scenes_parse():
    counts = {}
    for f in sorted(os.listdir(os.path.join(GOLD, "scenes"))):
        s = scene.Scene(os.path.join(GOLD, "scenes", f))
        counts[f] = s.n
        assert s.rad.shape == (s.n,) and np.all(s.rad == s.vdelta)
    assert counts["colliding_vortex_rings_nv.json"] == 2 * 158
    assert counts["flow_over_sphere.json"] == 0           # a body and no flow structures


def test_singular_rings_of_an_example():
    doc = json.load(open(os.path.join(GOLD, "scenes", "colliding_vortex_rings_nv.json")))
    s = scene.Scene(doc)
    assert abs(float(s.ips) - math.sqrt(8) * math.sqrt(0.002 / 40.000003814697266)) < 1e-7
    off = 0
    for ff in doc["flowstructures"]:
        nd = int(1 + 2 * math.pi * ff["major radius"] / float(s.ips))
        _ring_checks(s.pos[:, off:off + nd], s.str[:, off:off + nd], ff["center"], ff["normal"],
                     ff["major radius"], ff["circulation"], float(s.ips))
        off += nd


def test_thick_ring_blob_block_and_rules():
    doc = {"simparams": {"nominalDt": 0.01}, "flowparams": {"Re": 200.0},
           "flowstructures": [
               {"type": "thick ring", "center": [0.1, 0.2, -0.3], "normal": [0.0, 0.0, 1.0], "major radius": 1.0,
                "minor radius": 0.05, "circulation": 2.0},
               {"type": "vortex blob", "center": [2, 0, 0], "strength": [0.0, 0.5, 1.0], "radius": 0.1,
                "softness": 0.04},
               {"type": "uniform block", "center": [-2, 0, 0], "size": [0.2, 0.1, 0.3], "strength": [1, 2, 3]},
               {"type": "single particle", "center": [5, 5, 5], "strength": [0, 0, 1], "enabled": False},
               {"type": "particle emitter", "center": [0, 0, 5], "strength": [0, 0, 1]},
               {"type": "no such feature", "center": [0, 0, 0]}]}
    s = scene.Scene(doc)
    assert [t for t, _ in s.features] == ["thick ring", "vortex blob", "uniform block", "particle emitter"]
    ips = float(s.ips)
    # thick ring: stations x disk particles, all within the minor radius of the core circle
    pos, st = scene.thick_ring([0.1, 0.2, -0.3], [0, 0, 1], 1.0, 0.05, 2.0, s.ips)
    nlay = int(1 + 0.05 / ips)
    ndisk = 1 + sum(int(1 + 2 * math.pi * (l * ips) / ips) for l in range(1, nlay))
    nst = int(1 + 2 * math.pi / ips)
    assert pos.shape[1] == nst * ndisk
    d = pos.astype(np.float64) - np.array([[0.1], [0.2], [-0.3]])
    rho = np.hypot(d[0], d[1])
    assert np.max(np.hypot(rho - 1.0, d[2])) <= 0.05 + 1e-6
    # per station the disk's strengths sum to this_ips * circulation * (mean of (R + x)/R)
    tang = np.stack([-d[1] / rho, d[0] / rho, 0 * rho])
    total = (st.astype(np.float64) * tang).sum()
    assert abs(total - 2.0 * 2 * math.pi * 1.0) < 2e-3 * 2.0 * 2 * math.pi
    # blob: weights normalised, so the strengths sum to the given vector
    pb, sb = scene.vortex_blob([2, 0, 0], [0.0, 0.5, 1.0], 0.1, 0.04, s.ips)
    np.testing.assert_allclose(sb.astype(np.float64).sum(1), [0.0, 0.5, 1.0], atol=1e-6)
    assert np.linalg.norm(pb - np.array([[2], [0], [0]]), axis=0).max() < 0.1 + 0.02 + 1e-6
    # block: the third count follows the y size, as in the reference; strengths sum
    pk, sk = scene.uniform_block([-2, 0, 0], [0.2, 0.1, 0.3], [1, 2, 3], s.ips)
    ni, nj = int(1 + 0.2 / ips), int(1 + 0.1 / ips)
    assert pk.shape[1] == ni * nj * nj
    np.testing.assert_allclose(sk.astype(np.float64).sum(1), [1, 2, 3], rtol=1e-5)
    assert s.n == pos.shape[1] + pb.shape[1] + pk.shape[1] and len(s.emitters) == 1


def test_omega2d_files_are_rejected():
    with pytest.raises(ValueError):
        scene.Scene({"version": {"Omega2D": 1}})


# --- boundary bodies ---------------------------------------------------------
# The reference's meshers need libigl (igl::upsample) and Eigen, so they cannot be
# built here; the meshes are pinned by geometry and by the reference's own counts.

def _closed_and_oriented(faces):
    """every directed edge once and its reverse once: a closed, consistently wound surface"""
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    fwd = set(map(tuple, e.tolist()))
    assert len(fwd) == len(e)
    assert all((b, a) in fwd for a, b in fwd)
    return len(fwd) // 2


def _passes(start, ips):
    """refinement passes of the sphere and box meshers (GeomHelper.cpp:57-74, 110-113)"""
    k, m = 0, float(np.float32(start))
    while m > 1.2 * float(ips):
        k, m = k + 1, float(np.float32(m * 0.5))
    return k


def test_upsample_splits_each_triangle_in_four():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    x2, f2 = scene.upsample(x, np.array([[0, 1, 2]]))
    assert x2.shape == (6, 3) and f2.shape == (4, 3)
    assert np.array_equal(x2[:3], x)
    a, n = scene.panel_records(x2, f2)
    assert np.allclose(a, 0.125) and np.allclose(n, [0, 0, 1])


def test_sphere_body_of_an_example():
    s = scene.Scene(os.path.join(GOLD, "scenes", "flow_over_sphere.json"))
    assert s.n == 0 and s.unmeshed == [] and len(s.boundaries) == 1
    b = s.boundaries[0]
    k = _passes(0.35, s.ips)
    x, f = b["nodes"].astype(np.float64), b["faces"]
    assert k == 2 and len(f) == 20 * 4 ** k and len(x) == 10 * 4 ** k + 2
    assert np.allclose(np.linalg.norm(x, axis=1), 0.5, rtol=0, atol=1e-6)
    ne = _closed_and_oriented(f)
    assert len(x) - ne + len(f) == 2
    a, n = scene.panel_records(b["nodes"], f)
    cen = x[f].mean(axis=1)
    assert ((cen * n).sum(1) > 0).all()                      # external: normals into the fluid
    assert 0.95 * math.pi < a.sum() < math.pi                # inscribed in the sphere of area pi
    assert np.all(b["bc"] == 0)


def test_box_and_cavity_bodies_of_the_examples():
    s = scene.Scene(os.path.join(GOLD, "scenes", "flow_over_cube.json"))
    b = s.boundaries[0]
    k = _passes(0.7, s.ips)
    x, f = b["nodes"].astype(np.float64), b["faces"]
    assert k == 4 and len(f) == 12 * 4 ** k
    assert np.all(np.abs(x) <= 0.5 + 1e-7)
    tri = x[f]
    on_face = [np.all(np.abs(np.abs(tri[:, :, c]) - 0.5) < 1e-7, axis=1) for c in range(3)]
    assert np.all(np.sum(on_face, axis=0) == 1)               # each panel in exactly one face plane
    ne = _closed_and_oriented(f)
    assert len(np.unique(x, axis=0)) - ne + len(f) == 2
    a, n = scene.panel_records(b["nodes"], f)
    assert abs(a.sum() - 6.0) < 1e-9
    assert ((tri.mean(axis=1) * n).sum(1) > 0).all()

    doc = json.load(open(os.path.join(GOLD, "scenes", "driven_cavity.json")))
    s = scene.Scene(doc)
    ips = float(s.ips)
    meshes = doc["bodies"][0]["meshes"]
    assert len(s.boundaries) == len(meshes) == 6
    for m, b in zip(meshes, s.boundaries):
        p0, p1, p2, p3 = (np.array(m[k], np.float64) for k in ("startpt", "p1", "p2", "p3"))
        n1 = max(1, int(0.5 * (np.linalg.norm(p1 - p0) + np.linalg.norm(p3 - p2)) / ips))
        n2 = max(1, int(0.5 * (np.linalg.norm(p2 - p1) + np.linalg.norm(p0 - p3)) / ips))
        assert len(b["faces"]) == 2 * n1 * n2 == 450
        a, n = scene.panel_records(b["nodes"], b["faces"])
        assert abs(a.sum() - 1.0) < 1e-5
        cen = b["nodes"].astype(np.float64)[b["faces"]].mean(axis=1)
        assert ((cen - 0.5) * n).sum(1).max() < 0              # into the cavity, where the fluid is
        assert np.all(b["bc"] == np.array(m["vel"], np.float32))
    pan, bc = s.surfaces()
    assert pan.np == 2700 and bc.shape == (3, 2700) and abs(float(pan.area.sum()) - 6.0) < 1e-4


def test_body_options_and_unmeshed_geometries():
    base = {"simparams": {"nominalDt": 0.05}, "flowparams": {"Re": 50.0}}
    doc = dict(base, bodies=[{"name": "b", "translation": [0.0, 0.0, 1.0], "rotation": [0.0, 0.0, 0.0], "meshes": [
        {"geometry": "sphere", "scale": 2.0, "translation": [1.0, 2.0, 3.0], "external": False},
        {"geometry": "disk", "translation": [0, 0, 0]},
        {"geometry": "wing.obj", "translation": [0, 0, 0]}]},
        {"name": "spun", "translation": [0, 0, 0], "rotation": [0.0, 10.0, 0.0],
         "meshes": [{"geometry": "rect", "translation": [0, 0, 0]}]},
        {"name": "moving", "translation": ["sin(t)", 0, 0], "rotation": [0, 0, 0],
         "meshes": [{"geometry": "rect", "translation": [0, 0, 0]}]}])
    s = scene.Scene(doc)
    assert len(s.boundaries) == 1
    b = s.boundaries[0]
    x = b["nodes"].astype(np.float64) - [1.0, 2.0, 4.0]
    assert np.allclose(np.linalg.norm(x, axis=1), 1.0, atol=2e-6)
    assert len(b["faces"]) == 20 * 4 ** _passes(0.7, s.ips)  # 0.35 x the largest scale
    a, n = scene.panel_records(b["nodes"], b["faces"])
    assert ((x[b["faces"]].mean(axis=1) * n).sum(1) < 0).all()   # "external": false turns them inward
    reasons = {(u[0], u[1]): u[2] for u in s.unmeshed}
    assert set(reasons) == {("b", "disk"), ("b", "wing.obj"), ("spun", "rect"), ("moving", "rect")}
    assert "decimate" in reasons[("b", "disk")] and "libigl" in reasons[("b", "wing.obj")]


# --- measurement features ----------------------------------------------------

def test_measurement_features():
    doc = {"simparams": {"nominalDt": 0.01}, "flowparams": {"Re": 100.0}, "drawparams": {"tracerScale": 0.2},
           "measurements": [
               {"type": "point", "center": [0.1, 0.2, 0.3]},
               {"type": "measurement line", "center": [0, 0, 0], "end": [0.3, 0.4, 0], "lagrangian": True},
               {"type": "measurement quad", "start": [-1, -1, 0.5], "axis1": [2, 0, 0], "axis2": [0, 1, 0],
                "dx": [0.1, 0.25]},
               {"type": "tracer blob", "center": [0, 0, 1], "rad": 0.03},
               {"type": "tracer emitter", "center": [0, 1, 0]},
               {"type": "no such feature", "center": [0, 0, 0]},
               {"type": "point", "center": [9, 9, 9], "enabled": False}]}
    s = scene.Scene(doc)
    tips = float(s.tips)
    assert abs(tips - 0.2 * float(s.ips)) < 1e-9
    assert [m[0] for m in s.measures] == ["point", "measurement line", "measurement quad", "tracer blob",
                                          "tracer emitter"]
    assert [m[2] for m in s.measures] == [False, True, False, False, True]     # the emitter starts lagrangian
    x = s.fieldpts.astype(np.float64)
    off = 0
    assert np.allclose(x[:, 0], [0.1, 0.2, 0.3]) and np.all(x[:, 0] == np.float32([0.1, 0.2, 0.3]))
    off += 1
    nl = int(np.float32(1.0) + np.float32(0.5) / np.float32(tips))             # 1 + length / tips
    line = x[:, off:off + nl]
    assert np.allclose(line[:, 0], 0) and np.allclose(line[:, -1], [0.3, 0.4, 0])
    assert np.allclose(np.diff(line, axis=1), (line[:, -1:] - line[:, :1]) / (nl - 1), atol=1e-7)
    off += nl
    quad = x[:, off:off + 20 * 4]                                                # 2 / 0.1 x 1 / 0.25 cells
    assert np.allclose(quad[:, 0], [-1 + 0.05, -1 + 0.125, 0.5])
    assert np.all((quad[0] > -1) & (quad[0] < 1) & (quad[1] > -1) & (quad[1] < 0) & (quad[2] == 0.5))
    assert len(np.unique(quad[0].round(6))) == 20 and len(np.unique(quad[1].round(6))) == 4
    off += 80
    ijk = np.array([(i, j, k) for i in range(-9, 10) for j in range(-9, 10) for k in range(-9, 10)], np.float32)
    inside = np.sqrt((ijk ** 2).sum(1)) * np.float32(tips) < np.float32(0.03)
    nb = int(inside.sum())
    blob = x[:, off:off + nb] - np.array([[0], [0], [1.0]])
    site = ijk[inside].T * tips
    assert np.all(np.abs(blob - site) <= 0.5 * tips + 1e-7)                     # each point jittered about its site
    off += nb
    assert np.allclose(x[:, off], [0, 1, 0]) and x.shape[1] == off + 1
