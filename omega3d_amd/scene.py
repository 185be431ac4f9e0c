"""Scene input and particle output in the reference's formats.

JSON scenes (JsonHelper.cpp:55-190, FlowFeature.cpp:43-67 and each feature's
from_json / init_elements):
  * ``simparams.nominalDt``, ``flowparams.Re`` and ``flowparams.Uinf`` set the time
    step, Reynolds number and freestream; the particle spacing is
    ips = sqrt(8) * sqrt(dt / Re) and the core radius vdelta = 1.5 * ips
    (Simulation.cpp:57-58, 77-79; Points.h:102-103);
  * ``flowstructures`` of type "single particle", "vortex blob", "uniform block",
    "block of random", "singular ring", "thick ring" and "particle emitter" become
    particle collections by the reference's generators, restated here in their own
    float/double mixing (FlowFeature.cpp:205-1019).  "block of random" is unseeded
    in the reference (std::random_device, :561-562); here it takes a seed.  A
    particle emitter adds one particle per step (step_elements, :702-712) and none
    at start.  Disabled features are skipped;
  * ``measurements`` ("tracer", "point", "tracer emitter", "tracer blob", "tracer
    line", "measurement line", "measurement quad") become inert field points
    (``fieldpts``) at the tracer spacing tracerScale x ips (MeasureFeature.cpp);
  * ``bodies``: each mesh of a body is made into flat triangles at the particle
    spacing by the reference's meshers (BoundaryFeature.cpp:39-60, GeomHelper.cpp:
    41-127): "sphere" (the 12-node icosahedron of radius 0.5, split 1 -> 4 and
    projected back to the sphere while the estimated triangle size exceeds 1.2 ips,
    then scaled), "rect" (the 12-triangle unit cube, split likewise, then scaled) and
    "quad" (a bilinear grid, BoundaryFeature.cpp:695-760); "disk" and mesh files need
    libigl and are listed in ``unmeshed``.

VTU output of a particle collection (Points.h:851-1039 through VtkXmlWriter.h) and
of a panel set (Surfaces.h:1301-1450):
byte for byte what the reference writes -- the same element and attribute order
(tinyxml2's printer, attributes in std::map order), base64 data whose UInt32 length
header holds the length of the base64 text (the reference's own convention).

The reference's generators cannot be compiled here (FlowFeature.h needs Eigen3
through Body.h), so their restatement is pinned by geometry in tests/test_scene.py;
the VTU writer is pinned byte for byte against the reference's own writer.
"""
import base64
import json
import math
import struct

import numpy as np

f32 = np.float32


def _f(x):
    return f32(x)


# ---- particle output ---------------------------------------------------------------

def _b64_block(raw: bytes) -> str:
    enc = base64.b64encode(raw).decode("ascii")
    hdr = base64.b64encode(struct.pack("<I", len(enc))).decode("ascii")
    return " " + hdr + enc + " "


def _data_array(indent, attrs, raw):
    a = " ".join(f'{k}="{v}"' for k, v in sorted(attrs.items()))
    return f'{indent}<DataArray {a} format="binary">{_b64_block(raw)}</DataArray>\n'


def points_vtu_bytes(pos, time, strength=None, radius=None, vel=None):
    """The VTU file Points::write_vtk writes for n particles: pos (3, n), strength (3, n)
    and radius (n,) for an active collection (None for field points), vel (3, n)."""
    pos = np.asarray(pos, f32)
    n = pos.shape[1]
    vel = np.zeros((3, n), f32) if vel is None else np.asarray(vel, f32)
    inter = lambda a: np.ascontiguousarray(np.asarray(a, f32).T).tobytes()
    I1, I2, I3, I4 = " " * 4, " " * 8, " " * 12, " " * 16
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n',
           f"{I1}<UnstructuredGrid>\n", f"{I2}<FieldData>\n",
           _data_array(I3, {"Name": "TimeValue", "NumberOfTuples": "1", "type": "Float64"}, struct.pack("<d", time)),
           f"{I2}</FieldData>\n",
           f'{I2}<Piece NumberOfCells="{n}" NumberOfPoints="{n}">\n',
           f"{I3}<Points>\n",
           _data_array(I4, {"Name": "position", "NumberOfComponents": "3", "type": "Float32"}, inter(pos)),
           f"{I3}</Points>\n", f"{I3}<Cells>\n",
           _data_array(I4, {"Name": "connectivity", "type": "Int32"}, np.arange(n, dtype="<i4").tobytes()),
           _data_array(I4, {"Name": "offsets", "type": "Int32"}, np.arange(1, n + 1, dtype="<i4").tobytes()),
           _data_array(I4, {"Name": "types", "type": "UInt8"}, np.ones(n, np.uint8).tobytes()),
           f"{I3}</Cells>\n"]
    pd = {"Vectors": "velocity"}
    if radius is not None:
        pd["Scalars"] = "radius"
    out.append(f"{I3}<PointData " + " ".join(f'{k}="{v}"' for k, v in sorted(pd.items())) + ">\n")
    if strength is not None:
        out.append(_data_array(I4, {"Name": "circulation", "NumberOfComponents": "3", "type": "Float32"},
                               inter(strength)))
    if radius is not None:
        out.append(_data_array(I4, {"Name": "radius", "type": "Float32"}, np.asarray(radius, "<f4").tobytes()))
    out.append(_data_array(I4, {"Name": "velocity", "NumberOfComponents": "3", "type": "Float32"}, inter(vel)))
    out += [f"{I3}</PointData>\n", f"{I2}</Piece>\n", f"{I1}</UnstructuredGrid>\n", "</VTKFile>\n"]
    return "".join(out).encode("ascii")


def panels_vtu_bytes(nodes, idx, time, sheet=None, source=None):
    """The VTU file Surfaces::write_vtk (Surfaces.h:1301-1450) writes for a panel
    set: nodes (3, nn), idx the 3 node indices of each panel (flat or (np, 3)),
    sheet (3, np) the vortex sheet strength vector of each panel (ps0 x1 + ps1 x2,
    vortex_sheet_to_vector, Surfaces.h:340-361) for an active or reactive surface
    (None: inert, no cell data), source (np,) its source sheet strengths when it
    has them.  The reference never writes point data for panels."""
    nodes = np.asarray(nodes, f32)
    idx = np.asarray(idx).reshape(-1)
    nn, npn = nodes.shape[1], idx.size // 3
    inter = lambda a: np.ascontiguousarray(np.asarray(a, f32).T).tobytes()
    I1, I2, I3, I4 = " " * 4, " " * 8, " " * 12, " " * 16
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n',
           f"{I1}<UnstructuredGrid>\n", f"{I2}<FieldData>\n",
           _data_array(I3, {"Name": "TimeValue", "NumberOfTuples": "1", "type": "Float64"}, struct.pack("<d", time)),
           f"{I2}</FieldData>\n",
           f'{I2}<Piece NumberOfCells="{npn}" NumberOfPoints="{nn}">\n',
           f"{I3}<Points>\n",
           _data_array(I4, {"Name": "position", "NumberOfComponents": "3", "type": "Float32"}, inter(nodes)),
           f"{I3}</Points>\n", f"{I3}<Cells>\n",
           _data_array(I4, {"Name": "connectivity", "type": "Int32"}, idx.astype("<i4").tobytes()),
           _data_array(I4, {"Name": "offsets", "type": "Int32"}, (3 * np.arange(1, npn + 1)).astype("<i4").tobytes()),
           _data_array(I4, {"Name": "types", "type": "UInt8"}, np.full(npn, 5, np.uint8).tobytes()),
           f"{I3}</Cells>\n"]
    if sheet is not None:
        src = source is not None
        out.append(f'{I3}<CellData Scalars="source sheet strength">\n' if src else f"{I3}<CellData>\n")
        out.append(_data_array(I4, {"Name": "vortex sheet strength", "NumberOfComponents": "3", "type": "Float32"},
                               inter(sheet)))
        if src:
            out.append(_data_array(I4, {"Name": "source sheet strength", "type": "Float32"},
                                   np.asarray(source, "<f4").tobytes()))
        out.append(f"{I3}</CellData>\n")
    out += [f"{I2}</Piece>\n", f"{I1}</UnstructuredGrid>\n", "</VTKFile>\n"]
    return "".join(out).encode("ascii")


def panel_vtu_name(index, frame):
    """panel_II_FFFFF.vtu (Surfaces.h:1315-1316)"""
    return f"panel_{index:02d}_{frame:05d}.vtu"


def write_panels_vtu(path, nodes, idx, time, sheet=None, source=None):
    data = panels_vtu_bytes(nodes, idx, time, sheet, source)
    with open(path, "wb") as fh:
        fh.write(data)
    return path


def vtu_name(index, frame, inert=False):
    """part_II_FFFFF.vtu / fldpt_II_FFFFF.vtu (Points.h:866-867)"""
    return f"{'fldpt_' if inert else 'part_'}{index:02d}_{frame:05d}.vtu"


def write_points_vtu(path, pos, time, strength=None, radius=None, vel=None):
    with open(path, "wb") as fh:
        fh.write(points_vtu_bytes(pos, time, strength, radius, vel))
    return path


def read_points_vtu(path):
    """The arrays of a particle VTU written in this format: {name: array}, plus
    TimeValue; vectors come back as (3, n)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    types = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}
    out = {}
    for da in root.iter("DataArray"):
        text = da.text.strip()
        hlen = 8                                       # base64 of the 4-byte header
        raw = base64.b64decode(text[hlen:])
        a = np.frombuffer(raw, types[da.get("type")])
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = a.reshape(-1, nc).T.copy() if nc > 1 else a.copy()
    return out


# ---- JSON scenes ---------------------------------------------------------------------

def spacing(dt, re):
    """(ips, vdelta) of Simulation::get_ips / get_vdelta (Simulation.cpp:57-58, 77-79)"""
    core_size_ratio = _f(math.sqrt(8.0))
    hnu = np.sqrt(_f(dt) / _f(re))                   # float dt, re (Simulation.h:170-171)
    ips = _f(core_size_ratio * hnu)
    return ips, _f(_f(1.5) * ips)


def _normalize(v):
    """normalizeVec (MathHelper.h:185-192) in float"""
    v = [_f(x) for x in v]
    s = _f(_f(v[0] * v[0]) + _f(v[1] * v[1]))
    s = _f(s + _f(v[2] * v[2]))
    ln = _f(1.0 / float(np.sqrt(s)))
    return [_f(x * ln) for x in v]


def _onb(n):
    """branchlessONB (MathHelper.h:214-226) in float"""
    sign = _f(math.copysign(1.0, n[2]))
    a = _f(-1.0 / float(_f(sign + n[2])))
    b = _f(_f(n[0] * n[1]) * a)
    b1 = [_f(1.0 + float(_f(_f(_f(sign * n[0]) * n[0]) * a))), _f(sign * b), _f(-sign * n[0])]
    b2 = [b, _f(sign + _f(_f(n[1] * n[1]) * a)), _f(-n[1])]
    return b1, b2


def _cs(theta):
    t = float(theta)
    return _f(math.cos(t)), _f(math.sin(t))


def singular_ring(center, normal, majrad, circ, ips):
    """SingularRing::init_elements (FlowFeature.cpp:788-831)"""
    majrad, circ, ips = _f(majrad), _f(circ), _f(ips)
    ndiam = int(1 + (2.0 * math.pi * float(majrad)) / float(ips))
    this_ips = _f((2.0 * math.pi * float(majrad)) / float(_f(ndiam)))
    n = _normalize(normal)
    b1, b2 = _onb(n)
    pos = np.zeros((3, ndiam), f32)
    st = np.zeros((3, ndiam), f32)
    for i in range(ndiam):
        theta = _f(2.0 * math.pi * float(_f(i)) / float(_f(ndiam)))
        c, s = _cs(theta)
        for d in range(3):
            pos[d, i] = _f(_f(center[d]) + _f(majrad * _f(_f(b1[d] * c) + _f(b2[d] * s))))
            st[d, i] = _f(_f(this_ips * circ) * _f(_f(b2[d] * c) - _f(b1[d] * s)))
    return pos, st


def thick_ring(center, normal, majrad, minrad, circ, ips):
    """ThickRing::init_elements (FlowFeature.cpp:940-1019)"""
    majrad, minrad, circ, ips = _f(majrad), _f(minrad), _f(circ), _f(ips)
    nlayers = int(1 + float(minrad) / float(ips))
    disk = [(_f(0.0), _f(0.0), _f(1.0))]
    for l in range(1, nlayers):
        thisrad = _f(_f(l) * ips)
        nthislayer = int(1 + (2.0 * math.pi * float(thisrad)) / float(ips))
        for i in range(nthislayer):
            phi = _f(2.0 * math.pi * float(_f(i)) / float(_f(nthislayer)))
            c, s = _cs(phi)
            disk.append((_f(thisrad * c), _f(thisrad * s), _f(_f(majrad + _f(thisrad * c)) / majrad)))
    nd = len(disk)
    ndiam = int(1 + (2.0 * math.pi * float(majrad)) / float(ips))
    this_ips = _f((2.0 * math.pi * float(majrad)) / float(_f(ndiam)))
    n = _normalize(normal)
    b1, b2 = _onb(n)
    pos = np.zeros((3, ndiam * nd), f32)
    st = np.zeros((3, ndiam * nd), f32)
    k = 0
    for i in range(ndiam):
        theta = _f(2.0 * math.pi * float(_f(i)) / float(_f(ndiam)))
        ct, stt = _cs(theta)
        for (dx, dy, dl) in disk:
            sscale = _f(_f(_f(dl * this_ips) * circ) / _f(nd))
            for d in range(3):
                pos[d, k] = _f(_f(_f(center[d]) + _f(_f(majrad + dx) * _f(_f(b1[d] * ct) + _f(b2[d] * stt))))
                               + _f(dy * n[d]))
                st[d, k] = _f(sscale * _f(_f(b2[d] * ct) - _f(b1[d] * stt)))
            k += 1
    return pos, st


def vortex_blob(center, strength, rad, softness, ips):
    """VortexBlob::init_elements (FlowFeature.cpp:290-343)"""
    rad, soft, ips = _f(rad), _f(softness), _f(ips)
    irad = int(1 + (float(rad) + 0.5 * float(soft)) / float(ips))
    xs, ws = [], []
    for i in range(-irad, irad + 1):
        for j in range(-irad, irad + 1):
            for k in range(-irad, irad + 1):
                dr = _f(_f(np.sqrt(_f(i * i + j * j + k * k))) * ips)
                if float(dr) < float(rad) + 0.5 * float(soft):
                    xs.append([_f(_f(center[0]) + _f(ips * _f(i))), _f(_f(center[1]) + _f(ips * _f(j))),
                               _f(_f(center[2]) + _f(ips * _f(k)))])
                    w = 1.0
                    if float(dr) > float(rad) - 0.5 * float(soft):
                        w = 0.5 - 0.5 * math.sin(math.pi * (float(dr) - float(rad)) / float(soft))
                    ws.append(w)
    tot = 0.0
    for w in ws:                                     # sequential double sum, as the reference
        tot += w
    scale = 1.0 / tot
    pos = np.array(xs, f32).T.copy() if xs else np.zeros((3, 0), f32)
    st = np.zeros((3, len(ws)), f32)
    for i, w in enumerate(ws):
        for d in range(3):
            st[d, i] = _f(float(_f(_f(strength[d]) * _f(w))) * scale)
    return pos, st


def uniform_block(center, size, strength, ips):
    """UniformBlock::init_elements (FlowFeature.cpp:408-440); as in the reference, the
    third count is taken from the y size"""
    ips = _f(ips)
    isz = int(1 + float(_f(size[0])) / float(ips))
    jsz = int(1 + float(_f(size[1])) / float(ips))
    ksz = int(1 + float(_f(size[1])) / float(ips))
    tot = isz * jsz * ksz
    each = _f(1.0 / float(_f(tot)))
    pos = np.zeros((3, tot), f32)
    st = np.zeros((3, tot), f32)
    m = 0
    for i in range(isz):
        for j in range(jsz):
            for k in range(ksz):
                for d, (q, nq) in enumerate(((i, isz), (j, jsz), (k, ksz))):
                    pos[d, m] = _f(float(_f(center[d])) + float(_f(size[d])) * ((float(q) + 0.5) / float(nq) - 0.5))
                    st[d, m] = _f(_f(strength[d]) * each)
                m += 1
    return pos, st


def block_of_random(center, size, maxstr, num, seed=0):
    """BlockOfRandom::init_elements (FlowFeature.cpp:552-590) with a seeded generator
    (the reference draws from an unseeded std::random_device)"""
    rng = np.random.default_rng(seed)
    pos = np.array([_f(center[d]) + _f(size[d]) * rng.uniform(-0.5, 0.5, num) for d in range(3)], f32)
    st = (np.float32(maxstr) * rng.uniform(-0.5, 0.5, (3, num)) / np.float32(num)).astype(f32)
    return pos, st



# ---------------------------------------------------------------------------
# Boundary bodies: the meshes of a scene's "bodies" (BoundaryFeature.cpp:39-60
# dispatch; each geometry's from_json and init_elements), as flat triangles with
# a boundary-condition velocity per panel.
# ---------------------------------------------------------------------------

# the icosahedron of radius 0.5 that the sphere mesher starts from (Icosahedron.h
# ico0): its 12 vertices are the cyclic permutations of (0, +-A, +-B), in the
# order of _ico0_nodes, and its 20 faces wound outward
_ICO_A, _ICO_B = 0.2628655, 0.4253255
_ICO0_FACES = [(0, 1, 2), (3, 4, 5), (4, 3, 6), (7, 4, 8), (9, 3, 10), (9, 6, 3), (0, 9, 10), (7, 2, 1),
               (2, 8, 11), (3, 5, 10), (10, 5, 1), (6, 9, 11), (8, 4, 6), (0, 10, 1), (11, 0, 2), (7, 5, 4),
               (7, 1, 5), (6, 11, 8), (9, 0, 11), (2, 7, 8)]
# the unit cube of the box mesher (Cube.h cube0): node i at (i >> 2, (i >> 1) & 1, i & 1),
# two triangles per side, wound outward
_CUBE0_FACES = [(0, 4, 5), (0, 5, 1), (4, 6, 7), (4, 7, 5), (1, 5, 7), (1, 7, 3), (6, 2, 3), (6, 3, 7),
                (2, 0, 1), (2, 1, 3), (2, 6, 4), (2, 4, 0)]


def _ico0_nodes():
    """ico0's 12 vertices, (nn, 3) float32"""
    rows = [(_ICO_B, 0.0, _ICO_A), (0.0, _ICO_A, _ICO_B), (0.0, -_ICO_A, _ICO_B), (0.0, _ICO_A, -_ICO_B),
            (-_ICO_B, 0.0, -_ICO_A), (-_ICO_A, _ICO_B, 0.0), (0.0, -_ICO_A, -_ICO_B), (-_ICO_B, 0.0, _ICO_A),
            (-_ICO_A, -_ICO_B, 0.0), (_ICO_B, 0.0, -_ICO_A), (_ICO_A, _ICO_B, 0.0), (_ICO_A, -_ICO_B, 0.0)]
    return np.array(rows, np.float32)


def upsample(nodes, faces):
    """One in-plane subdivision (IglRefine.cpp:32-75, igl::upsample): a node at the
    midpoint of every edge, each triangle split into four of the same winding; the
    original nodes keep their indices and the midpoints follow in order of first
    appearance.  nodes (nn, 3) float32, faces (nf, 3) int; midpoints in double."""
    nn = nodes.shape[0]
    f = np.asarray(faces, np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])      # edges ab, bc, ca of every face
    key = np.minimum(e[:, 0], e[:, 1]) * nn + np.maximum(e[:, 0], e[:, 1])
    nf = f.shape[0]
    order = np.stack([key[:nf], key[nf:2 * nf], key[2 * nf:]], axis=1).reshape(-1)   # face-major order
    uniq, first = np.unique(order, return_index=True)
    rank = np.empty(len(uniq), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    mid_of = dict(zip(uniq.tolist(), (nn + rank).tolist()))
    a, b = uniq // nn, uniq % nn
    mids = np.empty((len(uniq), 3), np.float64)
    mids[rank] = 0.5 * (nodes[a].astype(np.float64) + nodes[b].astype(np.float64))
    ab = np.array([mid_of[k] for k in key[:nf].tolist()], np.int64)
    bc = np.array([mid_of[k] for k in key[nf:2 * nf].tolist()], np.int64)
    ca = np.array([mid_of[k] for k in key[2 * nf:].tolist()], np.int64)
    out = np.stack([np.stack([f[:, 0], ab, ca], 1), np.stack([ab, f[:, 1], bc], 1),
                    np.stack([ca, bc, f[:, 2]], 1), np.stack([ab, bc, ca], 1)], axis=1).reshape(-1, 3)
    return np.concatenate([nodes, mids.astype(np.float32)]), out


def _ovoid(sx, sy, sz, ips):
    """generate_ovoid (GeomHelper.cpp:41-91): ico0, refined while the estimated
    triangle size 0.35 max(scale) (halved per pass) exceeds 1.2 ips, each pass
    projected back to radius 0.5, then scaled per axis"""
    x, f = _ico0_nodes(), np.array(_ICO0_FACES, np.int64)
    meansize = _f(0.35 * float(max(sx, sy, sz)))
    while float(meansize) > 1.2 * float(ips):
        x, f = upsample(x, f)
        rad = np.sqrt((x.astype(np.float64) ** 2).sum(axis=1)).astype(np.float32)
        x = (x * (0.5 / rad.astype(np.float64)).astype(np.float32)[:, None]).astype(np.float32)
        meansize = _f(meansize * _f(0.5))
    return (x * np.array([sx, sy, sz], np.float32)).astype(np.float32), f


def _cuboid(dx, dy, dz, ips):
    """generate_cuboid (GeomHelper.cpp:94-127): cube0 refined while 0.7 max(scale)
    (halved per pass) exceeds 1.2 ips, then scaled per axis"""
    x = np.array([((i >> 2) & 1, (i >> 1) & 1, i & 1) for i in range(8)], np.float32)
    f = np.array(_CUBE0_FACES, np.int64)
    meansize = _f(0.7 * float(max(dx, dy, dz)))
    while float(meansize) > 1.2 * float(ips):
        x, f = upsample(x, f)
        meansize = _f(meansize * _f(0.5))
    return (x * np.array([dx, dy, dz], np.float32)).astype(np.float32), f


def _quad(p0, p1, p2, p3, ips):
    """BoundaryQuad::init_elements (BoundaryFeature.cpp:695-760): num1 x num2 cells
    of a bilinear quad (num = int(mean opposite side / ips), at least 1), two
    triangles per cell, in the reference's float/double mixing"""
    p0, p1, p2, p3 = (np.array(p, np.float32) for p in (p0, p1, p2, p3))
    d = lambda a, b: math.sqrt(sum(float(a[k] - b[k]) ** 2 for k in range(3)))
    side1 = _f(0.5 * (d(p1, p0) + d(p3, p2)))
    side2 = _f(0.5 * (d(p2, p1) + d(p0, p3)))
    num1 = max(1, int(side1 / ips))
    num2 = max(1, int(side2 / ips))
    x = np.empty(((num1 + 1) * (num2 + 1), 3), np.float32)
    k = 0
    for i in range(num1 + 1):
        s = float(_f(_f(i) / _f(num1)))
        for j in range(num2 + 1):
            t = float(_f(_f(j) / _f(num2)))
            w = [_f((1.0 - s) * (1.0 - t)), _f(s * (1.0 - t)), _f(s * t), _f((1.0 - s) * t)]
            for c in range(3):
                acc = _f(w[0] * p0[c])
                acc = _f(acc + _f(w[1] * p1[c]))
                acc = _f(acc + _f(w[2] * p2[c]))
                x[k, c] = _f(acc + _f(w[3] * p3[c]))
            k += 1
    f = []
    for i in range(num1):
        i1, i2 = (num2 + 1) * i, (num2 + 1) * (i + 1)
        for j in range(num2):
            f.append((i1 + j, i2 + j, i2 + j + 1))
            f.append((i1 + j, i2 + j + 1, i1 + j + 1))
    return x, np.array(f, np.int64)


def _scale3(m):
    sc = m.get("scale", 1.0)
    if isinstance(sc, (list, tuple)):
        return [_f(v) for v in sc[:3]]
    return [_f(sc)] * 3


def mesh_boundary(m, ips):
    """One mesh entry of a body (parse_boundary_json, BoundaryFeature.cpp:39-60).
    Returns (nodes (nn, 3) float32, faces (np, 3) int64, bc (np, 3) float32), or None
    with the reason for geometries not meshed here: "disk" (igl::decimate) and mesh
    files (ExteriorFromFile, read through libigl).  "sphere" and "rect" are reactive
    zero-velocity panels, reversed when "external" is false (ElementPacket.h:140-154
    swaps each triangle's first two nodes); a "quad" carries its "vel" and is always
    external."""
    g = m.get("geometry")
    if g in ("sphere", "rect"):
        sx, sy, sz = _scale3(m)
        x, f = (_ovoid if g == "sphere" else _cuboid)(sx, sy, sz, ips)
        x = (x + np.array(m["translation"], np.float32)).astype(np.float32)
        if not m.get("external", True):
            f = f[:, [1, 0, 2]]
        return x, f, np.zeros((f.shape[0], 3), np.float32), None
    if g == "quad":
        x, f = _quad(m["startpt"], m["p1"], m["p2"], m["p3"], ips)
        bc = np.tile(np.array(m["vel"], np.float32), (f.shape[0], 1))
        return x, f, bc, None
    if g == "disk":
        return None, None, None, "the disk mesher simplifies its triangles with igl::decimate (GeomHelper.cpp:207-330)"
    return None, None, None, "mesh files are read through libigl (ExteriorFromFile)"


def panel_records(nodes, faces):
    """area and in-plane basis of flat triangles (Surfaces.h:766-816: x1 along
    node0 -> node1, x2 its in-plane normal toward node 2, n = x1 x x2, area =
    base * height / 2), in double; returns area (np,) and n (np, 3)"""
    p = nodes.astype(np.float64)
    p0, p1, p2 = p[faces[:, 0]], p[faces[:, 1]], p[faces[:, 2]]
    e1 = p1 - p0
    base = np.linalg.norm(e1, axis=1)
    x1 = e1 / base[:, None]
    e2 = p2 - p0
    e2 = e2 - (e2 * x1).sum(1)[:, None] * x1
    height = np.linalg.norm(e2, axis=1)
    n = np.cross(x1, e2 / height[:, None])
    return 0.5 * base * height, n



# ---------------------------------------------------------------------------
# Measurement features: inert field points (MeasureFeature.cpp:44-68 dispatch,
# each type's from_json and init_elements), at the tracer spacing tracerScale x ips
# (main_batch.cpp:95; tracerScale from "drawparams", 0.15 by default).
# ---------------------------------------------------------------------------

def measure_points(m, tips, seed=0):
    """Field points (3, n) float32 of one measurement feature, or None for a type
    that names none (MeasureFeature.cpp:58-61: ignored).  "tracer blob" jitters its
    lattice points by a uniform [-0.5, 0.5) x tips per axis; the reference draws that
    from an unseeded generator (:276-278), here it takes a seed."""
    t = m.get("type")
    if t in ("tracer", "point", "tracer emitter"):                 # SinglePoint::init_elements (:164-176)
        return np.array(m["center"], f32).reshape(3, 1)
    if t == "tracer blob":                                        # MeasurementBlob::init_elements (:273-320)
        c, rad = np.array([_f(v) for v in m["center"]], f32), _f(m["rad"])
        irad = int(_f(1.0) + _f(rad / tips))
        r = np.arange(-irad, irad + 1)
        ijk = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)   # i, then j, then k
        dr = np.sqrt((ijk ** 2).sum(1).astype(f32)) * f32(tips)
        ijk = ijk[dr < rad].astype(f32)
        u = np.random.default_rng(seed).uniform(-0.5, 0.5, ijk.shape).astype(f32)      # three per point, in order
        return (c[None, :] + f32(tips) * (ijk + u)).astype(f32).T.copy()
    if t in ("tracer line", "measurement line"):                  # MeasurementLine::init_elements (:415-445)
        a, b = [_f(v) for v in m["center"]], [_f(v) for v in m["end"]]
        llen = _f(math.sqrt(sum((float(b[d]) - float(a[d])) ** 2 for d in range(3))))
        ilen = int(_f(_f(1.0) + _f(llen / tips)))
        with np.errstate(invalid="ignore", divide="ignore"):           # ilen = 1: 0 / 0, as in the reference
            frac = np.arange(ilen, dtype=f32) / f32(ilen - 1)
        a64, b32 = np.array(a, np.float64), np.array(b, f32)
        # (1.0 - frac) * a in double, frac * b in float, the sum rounded to float
        out = (1.0 - frac.astype(np.float64))[None, :] * a64[:, None] + (frac[None, :] * b32[:, None]).astype(np.float64)
        return out.astype(f32)
    if t == "measurement quad":                                   # Grid2dPoints::init_elements (:552-583)
        x0, s1, s2 = [np.array([_f(v) for v in m[k]], f32) for k in ("start", "axis1", "axis2")]
        ds, df = _f(m["dx"][0]), _f(m["dx"][1])

        def steps(d, dist):
            # the loop variable of the reference: a float, started at 0.5 d / dist
            # (in double) and stepped by d / dist (in float) while below 1 + 0.01 d / dist
            out, v = [], _f(0.5 * float(d) / float(dist))
            while float(v) < 1.0 + 0.01 * float(d) / float(dist):
                out.append(v)
                v = _f(v + _f(d / dist))
            return np.array(out, f32)
        sp = steps(ds, _f(math.sqrt(sum(float(v) ** 2 for v in s1))))
        fp = steps(df, _f(math.sqrt(sum(float(v) ** 2 for v in s2))))
        pts = (x0[None, None, :] + s1[None, None, :] * sp[:, None, None]) + s2[None, None, :] * fp[None, :, None]
        return pts.astype(f32).reshape(-1, 3).T.copy()
    return None


# (moves, emits) each type starts with before its "lagrangian" / "emits" keys
# (MeasureFeature.cpp:52-58 with the constructors' positional arguments: "tracer
# emitter" and "tracer line" pass (0, 0, false, true), i.e. z = 0 and moves = true)
_MEASURE_DEFAULTS = {"tracer": (False, False), "point": (False, False), "tracer emitter": (True, False),
                     "tracer blob": (False, False), "tracer line": (True, False),
                     "measurement line": (False, False), "measurement quad": (False, False)}

class Scene:
    """A parsed JSON scene: time step, Reynolds number, freestream, particle spacing
    and core radius, and the particle collections of its enabled flow structures."""

    def __init__(self, doc, seed=0):
        if isinstance(doc, (str, bytes)) and not str(doc).lstrip().startswith("{"):
            with open(doc) as fh:
                doc = json.load(fh)
        elif isinstance(doc, (str, bytes)):
            doc = json.loads(doc)
        ver = doc.get("version", {})
        if "Omega2D" in ver:
            raise ValueError("input file is an Omega2D file (JsonHelper.cpp:69-72)")
        self.description = doc.get("description", "")
        sp, fp = doc.get("simparams", {}), doc.get("flowparams", {})
        self.dt = float(sp.get("nominalDt", 0.01))
        self.viscous = sp.get("viscous", "none")
        self.re = float(fp.get("Re", 100.0))
        self.fs = [float(x) for x in fp.get("Uinf", [0.0, 0.0, 0.0])]
        self.ips, self.vdelta = spacing(self.dt, self.re)
        self.bodies = doc.get("bodies", [])
        self.features = []       # (type, dict) of the enabled flow structures
        self.emitters = []       # (position, strength) added each step
        pos, st = [], []
        for k, ff in enumerate(doc.get("flowstructures", [])):
            if "type" not in ff or not ff.get("enabled", True):
                continue
            t = ff["type"]
            c = ff.get("center", [0.0, 0.0, 0.0])
            if t == "single particle":
                p, s = np.array(c, f32).reshape(3, 1), np.array(ff["strength"], f32).reshape(3, 1)
            elif t == "vortex blob":
                p, s = vortex_blob(c, ff["strength"], ff["radius"], ff["softness"], self.ips)
            elif t == "uniform block":
                p, s = uniform_block(c, ff["size"], ff["strength"], self.ips)
            elif t == "block of random":
                p, s = block_of_random(c, ff["size"], ff["max strength"], int(ff["num"]), seed + k)
            elif t == "singular ring":
                p, s = singular_ring(c, ff["normal"], ff["major radius"], ff["circulation"], self.ips)
            elif t == "thick ring":
                p, s = thick_ring(c, ff["normal"], ff["major radius"], ff["minor radius"], ff["circulation"],
                                  self.ips)
            elif t == "particle emitter":
                self.emitters.append((np.array(c, f32), np.array(ff["strength"], f32)))
                self.features.append((t, ff))
                continue
            else:
                continue                  # FlowFeature.cpp:58-61: not an available flow feature
            self.features.append((t, ff))
            pos.append(p)
            st.append(s)
        self.pos = np.concatenate(pos, axis=1) if pos else np.zeros((3, 0), f32)
        self.str = np.concatenate(st, axis=1) if st else np.zeros((3, 0), f32)
        self.rad = np.full(self.pos.shape[1], self.vdelta, f32)
        # measurement features: inert field points at the tracer spacing
        dp = doc.get("drawparams", {})
        self.tracer_scale = _f(dp.get("tracerScale", 0.15))
        self.tips = _f(self.tracer_scale * self.ips)
        self.measures = []       # (type, dict, moves, emits) of the enabled measurement features
        fp_list = []
        for k, m in enumerate(doc.get("measurements", [])):
            if "type" not in m or not m.get("enabled", True) or m["type"] not in _MEASURE_DEFAULTS:
                continue
            x = measure_points(m, self.tips, seed + 1000 + k)
            moves, emits = _MEASURE_DEFAULTS[m["type"]]
            self.measures.append((m["type"], m, bool(m.get("lagrangian", moves)), bool(m.get("emits", emits))))
            fp_list.append(x)
        self.fieldpts = np.concatenate(fp_list, axis=1) if fp_list else np.zeros((3, 0), f32)
        # bodies: each mesh at the particle spacing (main_batch.cpp:85), placed by the
        # body's translation at t = 0; bodies that move by formula or start rotated
        # are listed, not meshed
        self.boundaries = []     # dict(body, geometry, nodes (nn, 3), faces (np, 3), bc (np, 3))
        self.unmeshed = []       # (body, geometry, reason)
        for b in self.bodies:
            name = b.get("name", "")
            tr, rot = b.get("translation", [0.0, 0.0, 0.0]), b.get("rotation", [0.0, 0.0, 0.0])
            fixed = all(isinstance(v, (int, float)) for v in list(tr) + list(rot))
            for m in b.get("meshes", []):
                g = m.get("geometry")
                if not fixed or any(float(v) != 0.0 for v in rot):
                    self.unmeshed.append((name, g, "the body moves by formula or starts rotated"))
                    continue
                x, f, bc, why = mesh_boundary(m, self.ips)
                if why:
                    self.unmeshed.append((name, g, why))
                    continue
                x = (x + np.array(tr, f32)).astype(f32)
                self.boundaries.append(dict(body=name, geometry=g, nodes=x, faces=f, bc=bc))

    def surfaces(self):
        """The meshed bodies as one panel set (Surfaces<float> layout, zero sheet
        strengths, as the reference creates reactive panels) and their boundary
        velocities (3, np)."""
        from .scenes import Panels
        xs, fs, bcs, off = [], [], [], 0
        for bd in self.boundaries:
            xs.append(bd["nodes"])
            fs.append(bd["faces"] + off)
            bcs.append(bd["bc"])
            off += bd["nodes"].shape[0]
        x = np.concatenate(xs) if xs else np.zeros((0, 3), f32)
        f = np.concatenate(fs) if fs else np.zeros((0, 3), np.int64)
        area, _ = panel_records(x, f)
        npn = f.shape[0]
        pan = Panels(np.ascontiguousarray(x.T, dtype=f32), f.reshape(-1).astype(np.uint32),
                     np.zeros((3, npn), f32), area.astype(f32), np.zeros(npn, f32))
        bc = np.ascontiguousarray((np.concatenate(bcs) if bcs else np.zeros((0, 3), f32)).T, dtype=f32)
        return pan, bc

    @property
    def n(self):
        return self.pos.shape[1]
