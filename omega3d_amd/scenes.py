"""Seeded, portable synthetic inputs for the direct Biot-Savart workloads.

Every draw comes from a counter-based SplitMix64 stream (Steele, Lea & Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014), so the same
seed gives the same particles on any host: floats take the top 24 bits x 2^-24,
doubles the top 53 bits x 2^-53, Gaussians use Box-Muller in double.

Workloads (sizes are the defaults; all arrays are float32 SoA):
  * ring(n, ...)            single vortex ring: centres uniform in a torus of major
                            radius R, tube radius a, strengths tangent to the ring
                            with |alpha| = gamma/n, one core radius.
  * colliding_rings(...)    two rings of 524,288 particles, centres z=+-1, the
                            lower one tilted 15 deg about x, circulation +-1,
                            |alpha| jittered +-5%, core radius 0.02.
  * icosphere_panels(...)   unit sphere as an icosahedron split 5 times (20,480 flat
                            triangles) carrying Gaussian vortex-sheet strength
                            projected into each panel plane, plus targets uniform in
                            the shell 1.02 < r < 1.5.
  * batch_instances(...)    many independent rings with log-uniform particle counts.

The reference builds its rings from deterministic lattices (FlowFeature.cpp:940-1019)
and its random blocks from an unseeded random_device (FlowFeature.cpp:561-562); these
generators are new and seeded so parity runs are repeatable.
"""
from dataclasses import dataclass
import math

import numpy as np

SEED_BASE = 0x0314D0000
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based SplitMix64 stream: draw k of stream s is mix(seed' + (k+1)*golden)
    where seed' = mix(seed ^ s*golden)."""

    def __init__(self, seed, stream=0):
        with np.errstate(over="ignore"):
            base = np.array([np.uint64(seed) ^ (np.uint64(stream) * _GOLDEN)], dtype=np.uint64)
            self._base = _mix(base)[0]
        self._ctr = 0

    def bits(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(self._ctr + 1, self._ctr + 1 + n, dtype=np.uint64)
            out = _mix(self._base + k * _GOLDEN)
        self._ctr += n
        return out

    def uniform(self, n):
        """doubles in [0, 1) from the top 53 bits"""
        return (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform32(self, n):
        """float32 in [0, 1) from the top 24 bits"""
        return ((self.bits(n) >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)

    def normal(self, n):
        """standard normals, Box-Muller in double"""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


@dataclass
class Particles:
    """SoA vortex particles (the Points<float> layout of Points.h / ElementBase.h)."""
    pos: np.ndarray      # (3, n) float32
    str: np.ndarray      # (3, n) float32, absolute vortex strength alpha
    rad: np.ndarray      # (n,)   float32 core radius

    @property
    def n(self):
        return self.pos.shape[1]


def _torus_points(rng, n, R, a):
    """n points exactly uniform in volume inside a torus around the z axis:
    theta ~ U[0,2pi), cross-section (r, phi) with r = a sqrt(u), accepted with
    probability (R + r cos phi)/(R + a)."""
    out_t = [np.zeros(0)]
    out_r = [np.zeros(0)]
    out_p = [np.zeros(0)]
    have = 0
    while have < n:
        m = int((n - have) * 1.25) + 64
        th = 2 * np.pi * rng.uniform(m)
        r = a * np.sqrt(rng.uniform(m))
        ph = 2 * np.pi * rng.uniform(m)
        acc = rng.uniform(m) * (R + a) < (R + r * np.cos(ph))
        out_t.append(th[acc]); out_r.append(r[acc]); out_p.append(ph[acc])
        have += int(acc.sum())
    th = np.concatenate(out_t)[:n]
    r = np.concatenate(out_r)[:n]
    ph = np.concatenate(out_p)[:n]
    rho = R + r * np.cos(ph)
    pos = np.stack([rho * np.cos(th), rho * np.sin(th), r * np.sin(ph)])
    tang = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)])
    return pos, tang


def ring(n=8192, R=1.0, a=0.1, sigma=0.05, gamma=1.0, seed=SEED_BASE + 0, stream=0, jitter=0.0):
    """Single vortex ring in the xy-plane; |alpha| = gamma/n (total circulation gamma)."""
    rng = Rng(seed, stream)
    pos, tang = _torus_points(rng, n, R, a)
    mag = np.full(n, gamma / max(n, 1))
    if jitter:
        mag *= 1.0 + jitter * (2.0 * rng.uniform(n) - 1.0)
    return Particles(pos.astype(np.float32), (tang * mag).astype(np.float32),
                     np.full(n, sigma, np.float32))


def _rot_x(v, deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.stack([v[0], c * v[1] - s * v[2], s * v[1] + c * v[2]])


def colliding_rings(n_each=524288, R=1.0, a=0.1, sigma=0.02, tilt_deg=15.0, jitter=0.05,
                    seed=SEED_BASE + 1):
    """Two rings: ring 1 at z=+1 with circulation +1, ring 2 rotated tilt_deg about x
    and moved to z=-1 with circulation -1; |alpha| = (1/n_each)(1 + jitter U[-1,1])."""
    rng = Rng(seed, 0)
    p1, t1 = _torus_points(rng, n_each, R, a)
    m1 = (1.0 / n_each) * (1.0 + jitter * (2.0 * rng.uniform(n_each) - 1.0))
    p2, t2 = _torus_points(rng, n_each, R, a)
    m2 = (1.0 / n_each) * (1.0 + jitter * (2.0 * rng.uniform(n_each) - 1.0))
    p1 = p1 + np.array([[0.0], [0.0], [1.0]])
    p2 = _rot_x(p2, tilt_deg) + np.array([[0.0], [0.0], [-1.0]])
    s1 = t1 * m1
    s2 = -_rot_x(t2, tilt_deg) * m2
    pos = np.concatenate([p1, p2], axis=1).astype(np.float32)
    st = np.concatenate([s1, s2], axis=1).astype(np.float32)
    return Particles(pos, st, np.full(2 * n_each, sigma, np.float32))


@dataclass
class Panels:
    """SoA flat triangular panels (the Surfaces<float> layout of Surfaces.h)."""
    nodes: np.ndarray    # (3, nn) float32
    idx: np.ndarray      # (3*np,) uint32
    ts: np.ndarray       # (3, np) float32 absolute vortex strength = sheet strength x area
    area: np.ndarray     # (np,) float32
    ss: np.ndarray       # (np,) float32 sheet source strength (zeros for vortex sheets)

    @property
    def np(self):
        return self.area.shape[0]


def icosphere(levels=5, radius=1.0):
    """Icosahedron split `levels` times by edge midpoints re-projected onto the sphere.
    Returns nodes (nn, 3) float64 and faces (nf, 3) int, faces wound outward."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts) * radius, np.array(faces, dtype=np.int64)


def icosphere_panels(levels=5, radius=1.0, seed=SEED_BASE + 2):
    """Panels on an icosphere with Gaussian vortex-sheet strength projected into each
    panel plane.  Area and the in-plane basis follow Surfaces.h:766-816 (x1 along
    node0->node1, x2 its in-plane normal toward node 2, area = base*height/2); the
    absolute strength is ts = gamma * area (Surfaces.h:309-335)."""
    rng = Rng(seed, 0)
    verts, faces = icosphere(levels, radius)
    p0, p1, p2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    e1 = p1 - p0
    base = np.linalg.norm(e1, axis=1)
    x1 = e1 / base[:, None]
    e2 = p2 - p0
    e2 = e2 - (e2 * x1).sum(1)[:, None] * x1
    height = np.linalg.norm(e2, axis=1)
    x2 = e2 / height[:, None]
    n = np.cross(x1, x2)
    area = 0.5 * base * height
    g = rng.normal(3 * len(faces)).reshape(len(faces), 3)
    gam = g - (g * n).sum(1)[:, None] * n
    ts = gam * area[:, None]
    return Panels(np.ascontiguousarray(verts.T, dtype=np.float32),
                  faces.reshape(-1).astype(np.uint32),
                  np.ascontiguousarray(ts.T, dtype=np.float32),
                  area.astype(np.float32), np.zeros(len(faces), np.float32))


def shell_points(n=524288, r0=1.02, r1=1.5, seed=SEED_BASE + 2, stream=1):
    """n points uniform in volume in the spherical shell r0 < r < r1."""
    rng = Rng(seed, stream)
    r = np.cbrt(r0 ** 3 + rng.uniform(n) * (r1 ** 3 - r0 ** 3))
    z = 2.0 * rng.uniform(n) - 1.0
    ph = 2 * np.pi * rng.uniform(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * s * np.cos(ph), r * s * np.sin(ph), r * z]).astype(np.float32)


def log_uniform_counts(k, lo, hi, seed, stream=7):
    rng = Rng(seed, stream)
    u = rng.uniform(k)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch_instances(k=256, lo=2048, hi=262144, seed=SEED_BASE + 3):
    """k independent single rings: n log-uniform on [lo, hi], R ~ U[0.5, 2.0],
    sigma ~ U[0.01, 0.05]; instance i uses sub-stream i + 1."""
    counts = log_uniform_counts(k, lo, hi, seed)
    prm = Rng(seed, 9)
    Rs = 0.5 + 1.5 * prm.uniform(k)
    sig = 0.01 + 0.04 * prm.uniform(k)
    return [ring(int(counts[i]), R=float(Rs[i]), a=0.1, sigma=float(sig[i]), seed=seed, stream=i + 1)
            for i in range(k)]
