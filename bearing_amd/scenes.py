"""Seeded synthetic scenes: the workloads the contact pipeline is measured and checked on.

Each scene is a list of body specs, dicts with the keys the reference's
``world.addBody`` options use (``src/world.js:189-218`` DEFAULTS, ``docs/API.md``):
``type``, ``shape``, ``size`` / ``radius`` / ``height``, ``position``, ``rotation``,
``velocity``, ``angularVelocity``, ``friction``, ``restitution``, ``density``, ``mass``.
A key left out takes the reference's default. Every number is a double produced here
from a splitmix64 stream (doubles as ``(x >> 11) * 2**-53``, Gaussians by the Marsaglia
polar method), so the same seed gives the same scene bit for bit in any language that
reads the JSON this module writes.

Scenes:
  pyramid      210 unit boxes in 20 rows on a static 40x1x40 ground (test/contact.test.js:45-64)
  rain         spheres raining into a walled 20x20 floor
  tower        a box tower, each level turned about y, on a 10x1x10 ground
  pile_world   one world of the batched workload: a tower or a sphere pile
  throw        a thrown box landing and hitting a 6x6 wall of unit boxes
  capsules     capsules, boxes and balls dropped in a heap on a floor
  terrain      the mixed pile small: 400 bodies on a 48 x 48 heightfield
  mixed_pile   20,000 boxes, balls, capsules and hulls dropped onto a 256 x 256 heightfield
  rocks        hulls of random points with boxes, balls and capsules dropped in a heap, and
               small balls fired through two static hull boulders
  darts        small balls and rods fired at 20-200 m/s at a 2 cm static plate and the
               floor: bodies the step would put through them, which the continuous catch
               puts back (src/world.js:1636-1681)
"""
from __future__ import annotations

import json
import math
import sys

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): the generator every scene draws from."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """A double in [0, 1): the top 53 bits."""
        return (self.next_u64() >> 11) * (2.0 ** -53)

    def between(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def gauss(self, sigma: float = 1.0) -> float:
        """Marsaglia's polar method; the second value of each pair is dropped."""
        while True:
            u = 2.0 * self.uniform() - 1.0
            v = 2.0 * self.uniform() - 1.0
            s = u * u + v * v
            if 0.0 < s < 1.0:
                return sigma * u * math.sqrt(-2.0 * math.log(s) / s)


def _ground(size, top=0.0, friction=None):
    spec = {"type": "static", "shape": "box", "size": list(size), "position": [0.0, top - size[1] * 0.5, 0.0]}
    if friction is not None:
        spec["friction"] = friction
    return spec


def pyramid(rows: int = 20):
    """Rows of unit boxes, row k holding rows-k boxes 1 m apart, centred, at y = 0.5 + k.

    Boxes have friction 0.6, density 1000, restitution 0; the ground is a static
    40x1x40 box whose top is at y = 0.
    """
    bodies = [_ground((40.0, 1.0, 40.0))]
    for row in range(rows):
        for k in range(rows - row):
            bodies.append({"shape": "box", "size": [1.0, 1.0, 1.0], "friction": 0.6,
                           "position": [k - (rows - 1 - row) / 2, 0.5 + row, 0.0]})
    return bodies


def _walls(half: float, height: float):
    """Four static walls 1 m thick and `height` tall around a floor 2*half across."""
    h = height * 0.5
    return [
        {"type": "static", "shape": "box", "size": [1.0, height, 2 * half], "position": [half + 0.5, h, 0.0]},
        {"type": "static", "shape": "box", "size": [1.0, height, 2 * half], "position": [-half - 0.5, h, 0.0]},
        {"type": "static", "shape": "box", "size": [2 * half, height, 1.0], "position": [0.0, h, half + 0.5]},
        {"type": "static", "shape": "box", "size": [2 * half, height, 1.0], "position": [0.0, h, -half - 0.5]},
    ]


def rain(count: int = 2000, seed: int = 2, spread: float = 5.0, lo: float = 2.0, hi: float = 42.0):
    """Spheres of radius U[0.1, 0.3], restitution U[0, 0.7], friction 0.5, density 1000,
    centres uniform in x, z in [-spread, spread], y in [lo, hi], at rest, over a static
    20x1x20 floor walled by four static 1x6x20 boxes."""
    rng = SplitMix64(seed)
    bodies = [_ground((20.0, 1.0, 20.0))] + _walls(10.0, 6.0)
    for _ in range(count):
        r = rng.between(0.1, 0.3)
        e = rng.between(0.0, 0.7)
        x = rng.between(-spread, spread)
        y = rng.between(lo, hi)
        z = rng.between(-spread, spread)
        bodies.append({"shape": "sphere", "radius": r, "restitution": e, "friction": 0.5, "position": [x, y, z]})
    return bodies


def _yaw(theta: float):
    """The unit quaternion [x, y, z, w] turning by theta about +y."""
    return [0.0, math.sin(theta * 0.5), 0.0, math.cos(theta * 0.5)]


def tower(height: int, rng: SplitMix64, ground=(10.0, 1.0, 10.0)):
    """`height` unit boxes stacked, each level turned U[0, 15] degrees about y."""
    bodies = [_ground(ground)]
    for k in range(height):
        theta = math.radians(rng.between(0.0, 15.0))
        bodies.append({"shape": "box", "size": [1.0, 1.0, 1.0], "position": [0.0, 0.5 + k, 0.0],
                       "rotation": _yaw(theta)})
    return bodies


def sphere_pile(count: int, rng: SplitMix64):
    """`count` spheres as in rain(), over a static 10x1x10 ground, dropped from y in [1, 1 + count/25]."""
    bodies = [_ground((10.0, 1.0, 10.0))]
    top = 1.0 + count / 25.0
    for _ in range(count):
        r = rng.between(0.1, 0.3)
        e = rng.between(0.0, 0.7)
        x = rng.between(-2.5, 2.5)
        y = rng.between(1.0, top)
        z = rng.between(-2.5, 2.5)
        bodies.append({"shape": "sphere", "radius": r, "restitution": e, "friction": 0.5, "position": [x, y, z]})
    return bodies


def batched(worlds: int = 4096, seed: int = 4, lo: int = 8, hi: int = 512):
    """Independent worlds, body counts log-uniform in [lo, hi]; even worlds box towers of
    height U{4..20}, odd worlds sphere piles of the world's count. Returns a list of scenes."""
    rng = SplitMix64(seed)
    out = []
    for w in range(worlds):
        n = int(math.floor(math.exp(rng.between(math.log(lo), math.log(hi + 1)))))
        if w % 2 == 0:
            height = 4 + int(rng.uniform() * 17)
            out.append(tower(height, rng))
        else:
            out.append(sphere_pile(n, rng))
    return out


def throw(rng: SplitMix64, dt_wall: float = 4.0):
    """A box of half-extents U[0.2, 0.5]^3 and mass U[1, 10] kg thrown with horizontal
    velocity N(0, 2^2), upward 3 m/s and spin N(0, 1)^3, over a ground of friction
    U[0.2, 0.9] and restitution U[0, 0.5], towards a 6x6 wall of unit boxes 4 m away."""
    half = [rng.between(0.2, 0.5) for _ in range(3)]
    mass = rng.between(1.0, 10.0)
    mu = rng.between(0.2, 0.9)
    e = rng.between(0.0, 0.5)
    vx, vz = rng.gauss(2.0), rng.gauss(2.0)
    w = [rng.gauss(1.0) for _ in range(3)]
    ground = {"type": "static", "shape": "box", "size": [20.0, 1.0, 20.0], "position": [0.0, -0.5, 0.0],
              "friction": mu, "restitution": e}
    bodies = [ground]
    for row in range(6):
        for col in range(6):
            bodies.append({"shape": "box", "size": [1.0, 1.0, 1.0], "position": [dt_wall, 0.5 + row, col - 2.5]})
    bodies.append({"shape": "box", "size": [2 * h for h in half], "mass": mass, "position": [0.0, 1.0, 0.0],
                   "velocity": [vx, 3.0, vz], "angularVelocity": w})
    return bodies


def darts(rng: SplitMix64, count: int = 64):
    """`count` small bodies, balls of radius U[0.05, 0.15] and rods of 0.06 x 0.06 x
    U[0.2, 0.6], fired from x in [-2, 0] at U[20, 200] m/s towards a static plate 2 cm thick,
    4 m tall and 6 m wide at x = 3, every fourth one down at the floor instead; rods turned
    any way and spinning N(0, 60). A static 20x1x20 floor. Restitution U[0, 0.5]."""
    bodies = [_ground((20.0, 1.0, 20.0)),
              {"type": "static", "shape": "box", "size": [0.02, 4.0, 6.0], "position": [3.0, 2.0, 0.0]}]
    for k in range(count):
        speed = rng.between(20.0, 200.0)
        pos = [rng.between(-2.0, 0.0), rng.between(0.8, 3.2), rng.between(-2.0, 2.0)]
        if k % 4 == 3:
            vel = [rng.gauss(1.0), -speed, rng.gauss(1.0)]
        else:
            vel = [speed, rng.gauss(1.0), rng.gauss(1.0)]
        e = rng.between(0.0, 0.5)
        if k % 2 == 0:
            bodies.append({"shape": "sphere", "radius": rng.between(0.05, 0.15), "position": pos, "velocity": vel,
                           "restitution": e})
        else:
            q = [rng.gauss(1.0) for _ in range(4)]
            n = math.sqrt(sum(x * x for x in q))
            bodies.append({"shape": "box", "size": [0.06, 0.06, rng.between(0.2, 0.6)], "position": pos,
                           "rotation": [x / n for x in q], "velocity": vel,
                           "angularVelocity": [rng.gauss(60.0) for _ in range(3)], "restitution": e})
    return bodies


def capsule_heap(rng: SplitMix64, count: int = 48):
    """`count` capsules (radius U[0.05, 0.2], half-length U[0.1, 0.5], turned any way), with
    a box of sides U[0.2, 0.8] and a ball of radius U[0.1, 0.3] after every sixth, dropped at
    rest from x, z in [-2, 2], y in [0.5, 6] onto a static 20x1x20 floor."""
    bodies = [_ground((20.0, 1.0, 20.0))]
    for k in range(count):
        pos = [rng.between(-2.0, 2.0), rng.between(0.5, 6.0), rng.between(-2.0, 2.0)]
        q = [rng.gauss(1.0) for _ in range(4)]
        n = math.sqrt(sum(x * x for x in q))
        r = rng.between(0.05, 0.2)
        bodies.append({"shape": "capsule", "radius": r, "height": 2 * r + 2 * rng.between(0.1, 0.5),
                       "position": pos, "rotation": [x / n for x in q]})
        if k % 6 == 5:
            bodies.append({"shape": "box", "size": [rng.between(0.2, 0.8) for _ in range(3)],
                           "position": [rng.between(-2.0, 2.0), rng.between(0.5, 6.0), rng.between(-2.0, 2.0)]})
            bodies.append({"shape": "sphere", "radius": rng.between(0.1, 0.3),
                           "position": [rng.between(-2.0, 2.0), rng.between(0.5, 6.0), rng.between(-2.0, 2.0)]})
    return bodies


def value_noise(rng: SplitMix64, columns: int, rows: int, spacing: float = 1.0, amplitude: float = 2.0,
                cell: float = 8.0):
    """Heights for a heightfield, row by row: bilinear value noise, lattice values
    U[-1, 1] * amplitude every `cell` metres (row-major lattice, drawn in order)."""
    nx = int(math.ceil((columns - 1) * spacing / cell)) + 1
    nz = int(math.ceil((rows - 1) * spacing / cell)) + 1
    lattice = [[rng.between(-1.0, 1.0) * amplitude for _ in range(nx)] for _ in range(nz)]
    heights = []
    for j in range(rows):
        z = j * spacing / cell
        j0 = min(int(math.floor(z)), nz - 2)
        fz = z - j0
        for i in range(columns):
            x = i * spacing / cell
            i0 = min(int(math.floor(x)), nx - 2)
            fx = x - i0
            a, b = lattice[j0][i0], lattice[j0][i0 + 1]
            c, d = lattice[j0 + 1][i0], lattice[j0 + 1][i0 + 1]
            heights.append((a * (1.0 - fx) + b * fx) * (1.0 - fz) + (c * (1.0 - fx) + d * fx) * fz)
    return heights


def _cloud(rng: SplitMix64, points: int, sigma: float):
    """A hull's vertices: `points` Gaussian points (sigma per axis), flat x, y, z."""
    return [rng.gauss(sigma) for _ in range(3 * points)]


def rocks(rng: SplitMix64, count: int = 40):
    """`count` rocks, each the hull of 12 Gaussian points (sigma U[0.15, 0.35]), turned any
    way and dropped at rest from x, z in [-2.5, 2.5], y in [0.5, 7]; a box of sides
    U[0.2, 0.6], a ball of radius U[0.1, 0.25] and a capsule after every fifth; two static
    boulders (hulls of 16 points, sigma 0.5) on a static 20x1x20 floor, and four balls of
    radius 0.03 fired at each boulder at 40-70 m/s, which the step would put through it."""
    bodies = [_ground((20.0, 1.0, 20.0))]
    boulders = [(-3.5, 0.5, 0.0), (3.5, 0.5, 0.0)]
    for c in boulders:
        bodies.append({"type": "static", "shape": "hull", "vertices": _cloud(rng, 16, 0.5), "position": list(c)})
    for k in range(count):
        pos = [rng.between(-2.5, 2.5), rng.between(0.5, 7.0), rng.between(-2.5, 2.5)]
        q = [rng.gauss(1.0) for _ in range(4)]
        n = math.sqrt(sum(x * x for x in q))
        bodies.append({"shape": "hull", "vertices": _cloud(rng, 12, rng.between(0.15, 0.35)), "position": pos,
                       "rotation": [x / n for x in q]})
        if k % 5 == 4:
            bodies.append({"shape": "box", "size": [rng.between(0.2, 0.6) for _ in range(3)],
                           "position": [rng.between(-2.5, 2.5), rng.between(0.5, 7.0), rng.between(-2.5, 2.5)]})
            bodies.append({"shape": "sphere", "radius": rng.between(0.1, 0.25),
                           "position": [rng.between(-2.5, 2.5), rng.between(0.5, 7.0), rng.between(-2.5, 2.5)]})
            r = rng.between(0.05, 0.15)
            bodies.append({"shape": "capsule", "radius": r, "height": 2 * r + 2 * rng.between(0.1, 0.4),
                           "position": [rng.between(-2.5, 2.5), rng.between(0.5, 7.0), rng.between(-2.5, 2.5)]})
    for bx, by, bz in boulders:
        side = -1.0 if bx < 0 else 1.0
        for j in range(4):
            y = by + rng.between(-0.2, 0.3)
            z = bz + rng.between(-0.25, 0.25)
            speed = rng.between(40.0, 70.0)
            bodies.append({"shape": "sphere", "radius": 0.03, "position": [bx + side * 3.0, y, z],
                           "velocity": [-side * speed, 0.0, 0.0]})
    return bodies


def mixed_pile(rng: SplitMix64, count: int = 20000, columns: int = 256, volume=(60.0, 80.0, 60.0), low: float = 2.0,
               cell: float = 8.0, amplitude: float = 2.0):
    """The mixed pile: a static heightfield of `columns` x `columns` samples 1 m apart
    (value_noise, amplitude 2 m, 8 m cells) and `count` bodies dropped at rest onto it from
    a volume centred over it (y from `low` up): 40% boxes (half-extents U[0.1, 0.5]),
    25% balls (radius U[0.1, 0.4]), 20% capsules (radius U[0.05, 0.2], half-length
    U[0.1, 0.5]), 15% hulls of 16 points N(0, 0.3^2); every turn normalised 4D Gaussian."""
    heights = value_noise(rng, columns, columns, 1.0, amplitude, cell)
    bodies = [{"type": "static", "shape": "heightfield", "heights": heights, "columns": columns, "spacing": 1.0}]
    vx, vy, vz = volume
    for _ in range(count):
        u = rng.uniform()
        pos = [rng.between(-vx / 2, vx / 2), rng.between(low, low + vy), rng.between(-vz / 2, vz / 2)]
        q = [rng.gauss(1.0) for _ in range(4)]
        n = math.sqrt(sum(x * x for x in q))
        rot = [x / n for x in q]
        if u < 0.4:
            body = {"shape": "box", "size": [2 * rng.between(0.1, 0.5) for _ in range(3)]}
        elif u < 0.65:
            body = {"shape": "sphere", "radius": rng.between(0.1, 0.4)}
        elif u < 0.85:
            r = rng.between(0.05, 0.2)
            body = {"shape": "capsule", "radius": r, "height": 2 * r + 2 * rng.between(0.1, 0.5)}
        else:
            body = {"shape": "hull", "vertices": _cloud(rng, 16, 0.3)}
        body["position"] = pos
        body["rotation"] = rot
        bodies.append(body)
    return bodies


def _unit(q):
    """A rotation as the reference keeps it (src/core/check.js:105-116 quat): scaled by its
    largest magnitude, then to unit length."""
    q = [float(x) for x in q]
    m = max(abs(q[0]), abs(q[1]), abs(q[2]), abs(q[3]))
    if not m > 0:
        raise ValueError("resolve: rotation must be 4 finite numbers, not all zero")
    x, y, z, w = q[0] / m, q[1] / m, q[2] / m, q[3] / m
    s = 1 / math.sqrt(x * x + y * y + z * z + w * w)
    return [x * s, y * s, z * s, w * s]


def resolve(bodies):
    """addBody's resolved columns (bm_bodies, as native.World.add takes them) for scenes of
    boxes and spheres, with the reference's defaults and arithmetic (src/world.js:2150-2245
    and src/shapes.js:14-56: density 1000 unless a mass is given, friction 0.5, restitution
    0; the thinnest size and the reach; inverses where positive; rotations normalised)."""
    n = len(bodies)
    cols = {name: [] for name in ("type", "shape", "lock", "position", "rotation", "velocity", "angular", "dims",
                                  "inertia", "inv_inertia", "mass", "inv_mass", "gravity_scale", "damping",
                                  "angular_damping", "size", "reach", "friction", "restitution")}
    for b in bodies:
        dynamic = b.get("type", "dynamic") == "dynamic"
        shape = b["shape"]
        if shape == "box":
            w, h, d = (float(x) for x in b["size"])
            dims = [w, h, d]
            volume = w * h * d
        elif shape == "sphere":
            r = float(b["radius"])
            dims = [r, 0.0, 0.0]
            volume = (4 / 3) * math.pi * r * r * r
        else:
            raise ValueError(f"resolve: a {shape} (boxes and spheres only)")
        mass = float(b["mass"]) if "mass" in b else float(b.get("density", 1000.0)) * volume
        if shape == "box":
            inertia = [mass * (h * h + d * d) / 12, mass * (w * w + d * d) / 12, mass * (w * w + h * h) / 12]
            size, reach = min(w, h, d), 0.5 * math.sqrt(w * w + h * h + d * d)
        else:
            inertia = [0.4 * mass * r * r] * 3
            size, reach = 2 * r, r
        cols["type"].append(0 if dynamic else 1)
        cols["shape"].append(0 if shape == "box" else 1)
        cols["lock"].append(0)
        cols["position"] += [float(x) for x in b.get("position", [0.0, 0.0, 0.0])]
        cols["rotation"] += _unit(b.get("rotation", [0.0, 0.0, 0.0, 1.0]))
        cols["velocity"] += [float(x) for x in b.get("velocity", [0.0, 0.0, 0.0])]
        cols["angular"] += [float(x) for x in b.get("angularVelocity", [0.0, 0.0, 0.0])]
        cols["dims"] += dims
        cols["inertia"] += inertia
        cols["inv_inertia"] += [1 / i if i > 0 else 0.0 for i in inertia]
        cols["mass"].append(mass)
        cols["inv_mass"].append(1 / mass if dynamic else 0.0)
        cols["gravity_scale"].append(float(b.get("gravityScale", 1.0)))
        cols["damping"].append(float(b.get("linearDamping", 0.0)))
        cols["angular_damping"].append(float(b.get("angularDamping", 0.0)))
        cols["size"].append(size)
        cols["reach"].append(reach)
        cols["friction"].append(float(b.get("friction", 0.5)))
        cols["restitution"].append(float(b.get("restitution", 0.0)))
    import numpy as np
    from .native import BODY_COLUMNS
    dtypes = {name: dtype for name, _, dtype in BODY_COLUMNS}
    return {k: np.asarray(v, dtypes[k]) for k, v in cols.items()}


def throws(worlds: int = 512, seed: int = 5):
    """The differentiation form's worlds: `worlds` thrown-box scenes (throw()), each from its
    own draws of one seeded stream."""
    rng = SplitMix64(seed)
    return [throw(rng) for _ in range(worlds)]


def throw_targets(worlds: int = 512, seed: int = 5):
    """The differentiation form's loss targets: for each thrown-box world, the point the
    box's final centre is pulled towards (the loss is its squared distance to it), drawn
    from its own seeded stream: x in [0, 4] (from the throw towards the wall), y in
    [0.2, 1.0], z in [-2, 2]."""
    rng = SplitMix64(seed * 1000 + 1)
    return [[rng.between(0.0, 4.0), rng.between(0.2, 1.0), rng.between(-2.0, 2.0)] for _ in range(worlds)]


SCENES = {
    "pyramid": lambda: pyramid(),
    "rain": lambda: rain(),
    "rain_small": lambda: rain(count=150, seed=2, spread=3.0, lo=0.5, hi=6.0),
    "tower": lambda: tower(12, SplitMix64(40)),
    "pile_small": lambda: sphere_pile(60, SplitMix64(41)),
    "throw": lambda: throw(SplitMix64(5)),
    "darts": lambda: darts(SplitMix64(7)),
    "capsules": lambda: capsule_heap(SplitMix64(9)),
    "rocks": lambda: rocks(SplitMix64(11)),
    "terrain": lambda: mixed_pile(SplitMix64(13), count=400, columns=48, volume=(24.0, 30.0, 24.0)),
    "mixed_pile": lambda: mixed_pile(SplitMix64(3)),
}


def main(argv):
    name = argv[1]
    json.dump(SCENES[name](), sys.stdout)
    sys.stdout.write("\n")


if __name__ == "__main__":
    main(sys.argv)
