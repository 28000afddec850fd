"""Regenerates the golden fixtures under tests/golden/data from the reference.

Needs the reference at /root/reference and the Node binary on PATH; run from the
repository root:  python tests/golden/make_fixtures.py [job ...]
Writes each scene's JSON (bearing_amd/scenes.py), then runs make_fixtures.mjs under
the lower.mjs loader once per job. The fixtures are what the tests read; this
script is not run by any test.
"""
import json
import math
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bearing_amd import scenes  # noqa: E402

DATA = os.path.join(HERE, "data")

# name: scene, dt, steps, soft-phase steps, colour steps, trace step, transforms every
JOBS = {
    "pyramid": dict(scene="pyramid", dt=1 / 60, steps=600, soft=[1, 2, 30, 55, 63], colour=[1, 30, 62, 66], trace=3,
                    every=100),
    "rain_small": dict(scene="rain_small", dt=1 / 60, steps=240, soft=[20, 40, 60, 90], colour=[60], trace=45, every=60),
    "tower": dict(scene="tower", dt=1 / 60, steps=120, soft=[1, 10, 40, 64], colour=[10, 63], trace=5, every=30),
    "throw": dict(scene="throw", dt=1 / 120, steps=120, soft=[30, 60, 90, 119], colour=[60], trace=70, every=30),
    "pile_small": dict(scene="pile_small", dt=1 / 60, steps=150, soft=[30, 70, 110], colour=[70], trace=100, every=50),
    "darts": dict(scene="darts", dt=1 / 60, steps=90, soft=[2, 20], colour=[], trace=-1, every=30),
    "capsules": dict(scene="capsules", dt=1 / 60, steps=240, soft=[60, 120], colour=[120], trace=-1, every=60),
    "rocks": dict(scene="rocks", dt=1 / 60, steps=240, soft=[60, 120], colour=[120], trace=-1, every=60),
    "terrain": dict(scene="terrain", dt=1 / 60, steps=300, soft=[150], colour=[150], trace=-1, every=100),
    # The 20,000-body mixed pile on a 256 x 256 heightfield at full size: hashes only.
    "mixed_pile": dict(scene="mixed_pile", dt=1 / 60, steps=600, soft=[], colour=[], trace=-1, every=600),
    # The 2,000-sphere rain at full size: its solve phase mid-landing is the bench's workload.
    "rain": dict(scene="rain", dt=1 / 60, steps=900, soft=[300], colour=[], trace=-1, every=300),
}


def write_job(name):
    """Writes trajectory `name`'s scene file and job file; returns the job file's path."""
    job = dict(JOBS[name])
    os.makedirs(os.path.join(DATA, "scenes"), exist_ok=True)
    scene_path = os.path.join(DATA, "scenes", job["scene"] + ".json")
    with open(scene_path, "w") as f:
        json.dump(scenes.SCENES[job["scene"]](), f)
    job["scene"] = scene_path
    job["out"] = os.path.join(DATA, name)
    job_path = os.path.join(DATA, name + ".job.json")
    with open(job_path, "w") as f:
        json.dump(job, f)
    return job_path


def run(name):
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_fixtures.mjs"), write_job(name)], check=True)


def run_events(name):
    """The reference's contact events after every step of trajectory `name` (make_events.mjs),
    into data/<name>.events.bmfx.gz."""
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_events.mjs"), write_job(name)], check=True)


def collide_pairs(seed, n):
    """Seeded shape pairs in the configurations the narrowphase branches on: spheres near
    and in boxes, boxes stacked flat (exactly and nearly), tilted, turned about the normal,
    edge on edge, corner down, and apart within and beyond the margin. One record a pair:
    per shape kind (0 sphere, 2 box), centre 3, unit quaternion 4, half 3, weight 3; then
    margin, tolerance."""
    import numpy as np
    rng = np.random.default_rng(seed)
    recs = []

    def quat(kind):
        if kind == "id":
            return [0.0, 0.0, 0.0, 1.0]
        if kind == "yaw":
            t = rng.uniform(0, 2 * np.pi)
            return [0.0, np.sin(t / 2), 0.0, np.cos(t / 2)]
        if kind == "tilt":
            ax = rng.normal(size=3)
            ax /= np.linalg.norm(ax)
            t = rng.normal(0, 0.02)
            return list(ax * np.sin(t / 2)) + [np.cos(t / 2)]
        v = rng.normal(size=4)
        return list(v / np.linalg.norm(v))

    def weight():
        return [0.0, -9.81, 0.0] if rng.random() < 0.7 else [0.0, 0.0, 0.0]

    for k in range(n):
        case = k % 8
        margin = float(rng.choice([0.0, rng.uniform(0, 0.05), rng.uniform(0, 0.4)]))
        if case == 0:   # sphere - sphere
            ra, rb = rng.uniform(0.1, 1, 2)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            dist = ra + rb + rng.uniform(-0.3, 0.3) * (ra + rb)
            if rng.random() < 0.05:
                dist = 0.0
            ca = rng.uniform(-3, 3, 3)
            a = [0, *ca, 0, 0, 0, 1, ra, 0, 0, *weight()]
            b = [0, *(ca + d * dist), 0, 0, 0, 1, rb, 0, 0, *weight()]
            tol = min(ra, rb) * 2 / 200
        elif case in (1, 2):   # sphere - box (either order)
            h = rng.uniform(0.1, 1.5, 3)
            r = rng.uniform(0.05, 0.6)
            qb = quat(rng.choice(["id", "yaw", "any"]))
            local = rng.uniform(-1.4, 1.4, 3) * (h + r)
            cb = rng.uniform(-2, 2, 3)
            R = np.array(_axes(qb)).reshape(3, 3).T
            cs = cb + R @ local
            sph = [0, *cs, 0, 0, 0, 1, r, 0, 0, *weight()]
            box = [2, *cb, *qb, *h, *weight()]
            a, b = (sph, box) if case == 1 else (box, sph)
            tol = min(2 * r, 2 * h.min()) / 200
        else:   # box - box
            ha = rng.uniform(0.1, 2.0, 3)
            hb = rng.uniform(0.1, 1.0, 3)
            if case == 3:   # stacked flat, maybe exactly
                qa, qb = [0.0, 0, 0, 1], quat(rng.choice(["id", "yaw"]))
                off = np.array([rng.uniform(-1, 1) * ha[0], 0, rng.uniform(-1, 1) * ha[2]])
                if rng.random() < 0.5:
                    off = np.round(off * 4) / 4
                gap = float(rng.choice([0.0, rng.uniform(-0.02, 0.02), rng.uniform(-0.2, 0.2)]))
                ca = rng.uniform(-2, 2, 3)
                cb = ca + off + np.array([0, ha[1] + hb[1] + gap, 0])
                if rng.random() < 0.3:
                    ha, hb = np.ones(3) * 0.5, np.ones(3) * 0.5
                    cb = ca + np.array([float(rng.choice([0.0, 0.5, 1.0])), 1.0 + gap, 0.0])
            else:
                qa = quat(["tilt", "any", "yaw", "tilt", "any"][case - 4])
                qb = quat(["tilt", "any", "any", "any", "tilt"][case - 4])
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                reach = np.linalg.norm(ha) * 0.8 + np.linalg.norm(hb) * 0.8
                ca = rng.uniform(-2, 2, 3)
                cb = ca + d * reach * rng.uniform(0.5, 1.3)
            a = [2, *ca, *qa, *ha, *weight()]
            b = [2, *cb, *qb, *hb, *weight()]
            tol = 2 * min(ha.min(), hb.min()) / 200
        recs.append([float(x) for x in a] + [float(x) for x in b] + [margin, tol])
    return np.array(recs, np.float64)


def _axes(q):
    x, y, z, w = q
    return [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
            2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
            2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]


def capsule_pairs(seed, n):
    """Seeded pairs with a capsule (kind 1: half = radius, half its segment's length): against
    a sphere (both orders), a capsule (crossing, near-parallel and parallel, end to end) and a
    box (resting on a face, across an edge, through it, off a corner). Records as
    collide_pairs'."""
    import numpy as np
    rng = np.random.default_rng(seed)
    recs = []

    def quat(kind):
        if kind == "id":
            return [0.0, 0.0, 0.0, 1.0]
        if kind == "flat":   # the capsule's axis (its y) turned to lie along x or z
            t = np.pi / 2
            ax = [0.0, 0.0, 1.0] if rng.random() < 0.5 else [1.0, 0.0, 0.0]
            ax = np.array(ax)
            if rng.random() < 0.5:
                t += rng.normal(0, 0.01)
            return list(ax * np.sin(t / 2)) + [np.cos(t / 2)]
        v = rng.normal(size=4)
        return list(v / np.linalg.norm(v))

    def weight():
        return [0.0, -9.81, 0.0] if rng.random() < 0.7 else [0.0, 0.0, 0.0]

    def capsule(centre, q):
        r = rng.uniform(0.05, 0.4)
        hl = rng.uniform(0.0, 0.8) if rng.random() < 0.9 else 0.0
        return [1, *centre, *q, r, hl, 0.0, *weight()], r, hl

    for k in range(n):
        case = k % 6
        margin = float(rng.choice([0.0, rng.uniform(0, 0.05), rng.uniform(0, 0.4)]))
        if case in (0, 1):   # sphere - capsule
            cc = rng.uniform(-2, 2, 3)
            cap, r, hl = capsule(cc, quat(rng.choice(["id", "flat", "any"])))
            rs = rng.uniform(0.05, 0.5)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            cs = cc + d * (r + rs + hl) * rng.uniform(0.3, 1.3)
            sph = [0, *cs, 0, 0, 0, 1, rs, 0, 0, *weight()]
            a, b = (sph, cap) if case == 0 else (cap, sph)
            tol = min(2 * r, 2 * rs) / 200
        elif case == 2:      # capsule - capsule
            ca = rng.uniform(-2, 2, 3)
            qa = quat(rng.choice(["id", "flat", "any"]))
            a, ra, ha = capsule(ca, qa)
            qb = qa if rng.random() < 0.4 else quat(rng.choice(["id", "flat", "any"]))
            if rng.random() < 0.3:   # nearly parallel
                qb = list(np.array(qa) + rng.normal(0, 1e-4, 4))
                qb = list(np.array(qb) / np.linalg.norm(qb))
            off = rng.normal(size=3)
            off /= np.linalg.norm(off)
            b, rb, hb = capsule(ca + off * rng.uniform(0.0, 1.2) * (ra + 0.4 + ha), qb)
            tol = min(2 * ra, 2 * rb) / 200
        else:                # capsule - box
            h = rng.uniform(0.1, 1.5, 3)
            qb = quat(rng.choice(["id", "id", "any"]))
            cb = rng.uniform(-2, 2, 3)
            R = np.array(_axes(qb)).reshape(3, 3).T
            qc = quat(rng.choice(["flat", "any", "id"]))
            local = rng.uniform(-1.3, 1.3, 3) * (h + 0.3)
            if case == 3:    # resting on the top face, lying flat
                local = np.array([rng.uniform(-1, 1) * h[0], h[1] + rng.uniform(-0.05, 0.3), rng.uniform(-1, 1) * h[2]])
                qc = quat("flat")
            cap, r, hl = capsule(cb + R @ local, qc)
            if case == 3:
                cap[2] += r * 0.9
            box = [2, *cb, *qb, *h, *weight()]
            a, b = (cap, box) if rng.random() < 0.8 else (box, cap)
            tol = min(2 * r, 2 * h.min()) / 200
        recs.append(a + b + [margin, tol])
    return np.asarray(recs, np.float64)


def hull_pairs(seed, n):
    """Seeded pairs with a hull (16 points N(0, 0.3^2) each, as the mixed pile's rocks):
    hull on hull, box on hull, sphere on hull and capsule on hull, either order, apart within
    and beyond the margin, touching and overlapping, at any turn. JSON-ready dicts."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def quat():
        if rng.random() < 0.2:
            return [0.0, 0.0, 0.0, 1.0]
        v = rng.normal(size=4)
        return list(v / np.linalg.norm(v))

    def weight():
        return [0.0, -9.81, 0.0] if rng.random() < 0.7 else [0.0, 0.0, 0.0]

    def shape(kind):
        if kind == 3:
            pts = rng.normal(0, 0.3, (16, 3))
            return {"kind": 3, "points": [float(x) for x in pts.reshape(-1)], "reach": float(np.linalg.norm(pts, axis=1).max())}
        if kind == 2:
            h = rng.uniform(0.1, 1.0, 3)
            return {"kind": 2, "half": [float(x) for x in h], "reach": float(np.linalg.norm(h))}
        if kind == 1:
            r, hl = rng.uniform(0.05, 0.3), rng.uniform(0.0, 0.6)
            return {"kind": 1, "half": [r, hl, 0.0], "reach": r + hl}
        r = rng.uniform(0.05, 0.5)
        return {"kind": 0, "half": [r, 0.0, 0.0], "reach": r}

    pairs = []
    kinds = [(3, 3), (3, 3), (2, 3), (3, 2), (0, 3), (3, 0), (1, 3), (3, 1)]
    for k in range(n):
        ka, kb = kinds[k % len(kinds)]
        a, b = shape(ka), shape(kb)
        ca = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        cb = ca + d * (a["reach"] + b["reach"]) * rng.uniform(0.2, 1.0)
        for s, c in ((a, ca), (b, cb)):
            s["center"] = [float(x) for x in c]
            s["q"] = quat() if s["kind"] != 0 else [0.0, 0.0, 0.0, 1.0]
            s["weight"] = weight()
            s.pop("reach")
        margin = float(rng.choice([0.0, rng.uniform(0, 0.05), rng.uniform(0, 0.3)]))
        pairs.append({"a": a, "b": b, "margin": margin})
    return pairs


def run_hulls(seed=2026, n=1200):
    path = os.path.join(DATA, "hulls.in.json")
    with open(path, "w") as f:
        json.dump(hull_pairs(seed, n), f)
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_hulls.mjs"), path, os.path.join(DATA, f"collide_hull.{seed}.bmfx.gz")],
                   check=True)
    os.remove(path)


def level_cases(seed, n):
    """Seeded shapes against heightfields, for touchLevel (src/triangle.js:1078): three fields
    of value noise (24 x 20 samples, 0.75 m apart; a flat plateau for seams, a raised crest
    line for sharp folds), placed at the origin, moved, and moved and turned about y; spheres,
    capsules, boxes and hulls (16 points N(0, 0.3^2)) at any turn, from above the surface to
    sunk into it, each with the box level() would look in (its own box grown by the margin,
    in the field's frame). JSON-ready."""
    import numpy as np
    from bearing_amd import scenes
    rng = np.random.default_rng(seed)
    fields = []
    for f in range(3):
        cols, rows, sp = 24, 20, 0.75
        h = scenes.value_noise(scenes.SplitMix64(100 + f), cols, rows, sp, 1.0, 4.0)
        for j in range(rows):
            for i in range(cols):
                if i < 8 and j < 8:
                    h[j * cols + i] = 0.25
                elif i == 16:
                    h[j * cols + i] += 0.5
        fields.append({"columns": cols, "rows": rows, "spacing": sp, "heights": h})
    poses = [((0.0, 0.0, 0.0), 0.0), ((1.5, -0.2, 2.25), 0.0), ((-0.75, 0.4, 0.5), math.pi / 6)]

    def axes_of(q):
        x, y, z, w = q
        return [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]

    def quat():
        r = rng.random()
        if r < 0.25:
            return [0.0, 0.0, 0.0, 1.0]
        if r < 0.4:   # turned about y only: faces level
            t = rng.uniform(0, math.pi)
            return [0.0, math.sin(t / 2), 0.0, math.cos(t / 2)]
        v = rng.normal(size=4)
        return list(v / np.linalg.norm(v))

    def surface(f, x, z):
        fd = fields[f]
        cols, rows, sp = fd["columns"], fd["rows"], fd["spacing"]
        x0, z0 = -((cols - 1) * sp) / 2, -((rows - 1) * sp) / 2
        i = min(max(int((x - x0) / sp), 0), cols - 1)
        j = min(max(int((z - z0) / sp), 0), rows - 1)
        return fd["heights"][j * cols + i]

    cases = []
    for k in range(n):
        f = k % 3
        pos, yaw = poses[(k // 3) % 3]
        q0 = [0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2)]
        A = axes_of(q0)
        kind = int(rng.integers(0, 4))
        if kind == 0:
            r = rng.uniform(0.1, 0.5)
            shape, reach = {"kind": 0, "half": [r, 0.0, 0.0]}, r
        elif kind == 1:
            r, hl = rng.uniform(0.05, 0.2), rng.uniform(0.1, 0.5)
            shape, reach = {"kind": 1, "half": [r, hl, 0.0]}, r + hl
        elif kind == 2:
            hh = rng.uniform(0.1, 0.6, 3)
            shape, reach = {"kind": 2, "half": [float(x) for x in hh]}, float(np.linalg.norm(hh))
        else:
            pts = rng.normal(0, 0.3, (16, 3))
            shape, reach = {"kind": 3, "points": [float(x) for x in pts.reshape(-1)]}, float(np.linalg.norm(pts, axis=1).max())
        lx, lz = rng.uniform(-7.5, 7.5), rng.uniform(-6.0, 6.0)
        ly = surface(f, lx, lz) + rng.uniform(-0.3, 0.5) * reach
        centre = [pos[0] + A[0] * lx + A[3] * ly + A[6] * lz, pos[1] + A[1] * lx + A[4] * ly + A[7] * lz,
                  pos[2] + A[2] * lx + A[5] * ly + A[8] * lz]
        shape["center"] = centre
        shape["q"] = quat() if kind != 0 else [0.0, 0.0, 0.0, 1.0]
        margin = float(rng.choice([0.0, rng.uniform(0, 0.05), rng.uniform(0, 0.2)]))
        g = reach + margin
        box = [lx - g, ly - g, lz - g, lx + g, ly + g, lz + g]
        cases.append({"field": f, "position": list(pos), "axes": A, "shape": shape, "box": box, "margin": margin})
    return {"fields": fields, "cases": cases}


def run_level(seed=2026, n=2400):
    path = os.path.join(DATA, "level.in.json")
    with open(path, "w") as f:
        json.dump(level_cases(seed, n), f)
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_level.mjs"), path, os.path.join(DATA, f"level.{seed}.bmfx.gz")],
                   check=True)
    os.remove(path)


def run_collide(seed=2026, n=4000, capsules=False):
    pairs = capsule_pairs(seed, n) if capsules else collide_pairs(seed, n)
    name = "collide_capsule" if capsules else "collide"
    path = os.path.join(DATA, name + ".in.bin")
    pairs.tofile(path)
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_collide.mjs"), path, os.path.join(DATA, f"{name}.{seed}.bmfx.gz")],
                   check=True)
    os.remove(path)


def impact_pairs(seed, n):
    """Seeded pairs closing on each other within about a step, for the time of impact:
    spheres and boxes, slow and fast, spinning below and above the pi/dt cap, the second
    body dynamic or static, some already within the tolerance and some parting. One record
    a pair: per body kind (0 sphere, 2 box), centre 3, unit quaternion 4, half 3, velocity 3,
    spin 3, gravity scale, dynamic (1 or 0), reach; then dt, tolerance, gravity 3."""
    import numpy as np
    rng = np.random.default_rng(seed)
    recs = []

    def quat():
        if rng.random() < 0.3:
            return [0.0, 0.0, 0.0, 1.0]
        v = rng.normal(size=4)
        return list(v / np.linalg.norm(v))

    def body(kind, centre, dynamic):
        if kind == 0:
            r = rng.uniform(0.05, 0.6)
            half, reach, size = [r, 0.0, 0.0], r, 2 * r
            q = [0.0, 0.0, 0.0, 1.0]
        else:
            h = rng.uniform(0.05, 1.0, 3)
            half, reach, size = list(h), float(np.linalg.norm(h)), 2 * float(h.min())
            q = quat()
        spin = rng.choice([0.0, 1.0, 20.0, 400.0]) if dynamic else 0.0
        w = list(rng.normal(0, spin, 3)) if spin > 0 else [0.0, 0.0, 0.0]
        return {"kind": kind, "c": np.asarray(centre, float), "q": q, "half": half, "w": w, "reach": reach,
                "size": size, "dyn": dynamic, "gs": float(rng.choice([1.0, 1.0, 0.5])) if dynamic else 0.0}

    for k in range(n):
        kinds = [(0, 0), (0, 2), (2, 0), (2, 2)][k % 4]
        dt = float(rng.choice([1 / 60, 1 / 120]))
        a = body(kinds[0], rng.uniform(-5, 5, 3), True)
        b = body(kinds[1], [0, 0, 0], rng.random() < 0.7)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        speed = float(np.exp(rng.uniform(np.log(0.5), np.log(80))))
        case = rng.random()
        if case < 0.1:
            gap0 = rng.uniform(-0.01, 0.002)     # about touching
        elif case < 0.2:
            gap0 = rng.uniform(2, 4) * speed * dt  # too far to meet
        else:
            gap0 = rng.uniform(-0.5, 1.0) * speed * dt
        b["c"] = a["c"] + d * (a["reach"] + b["reach"] + gap0)
        towards = d + rng.normal(0, 0.1, 3)
        towards /= np.linalg.norm(towards)
        va = towards * speed * (-1 if rng.random() < 0.1 else 1)
        vb = rng.normal(0, 1, 3) if b["dyn"] else np.zeros(3)
        tol = min(a["size"], b["size"]) / 200
        rec = []
        for x, v in ((a, va), (b, vb)):
            rec += [x["kind"], *x["c"], *x["q"], *x["half"], *v, *x["w"], x["gs"], 1.0 if x["dyn"] else 0.0, x["reach"]]
        rec += [dt, tol, 0.0, -9.81, 0.0]
        recs.append(rec)
    return np.asarray(recs, np.float64)


def run_impact(seed=2026, n=2000):
    pairs = impact_pairs(seed, n)
    path = os.path.join(DATA, "impact.in.bin")
    pairs.tofile(path)
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_impact.mjs"), path, os.path.join(DATA, f"impact.{seed}.bmfx.gz")],
                   check=True)
    os.remove(path)


def run_built(name):
    """The reference's built hulls (src/hull.js build()) of scene `name`'s hull bodies, from its
    scene file in data/scenes, into data/<name>.built.bmfx.gz."""
    subprocess.run(["node", "--no-warnings", "--experimental-loader", os.path.join(HERE, "lower.mjs"),
                    os.path.join(HERE, "make_built.mjs"), os.path.join(DATA, "scenes", f"{name}.json"),
                    os.path.join(DATA, f"{name}.built.bmfx.gz")], check=True)


def batched_sample(worlds):
    """The worlds of the batched scenes whose every step is recorded: of the box towers (even
    worlds) and of the sphere piles (odd worlds) each, the smallest, the median and the
    largest by body count (the first of equals), and the first three."""
    picked = []
    for kind in (0, 1):
        idx = list(range(kind, len(worlds), 2))
        by = sorted(idx, key=lambda i: (len(worlds[i]), i))
        for i in (by[0], by[len(by) // 2], by[-1], *idx[:3]):
            if i not in picked:
                picked.append(i)
    return sorted(picked)


def run_batched(worlds=4096, seed=4, steps=300, dt=1 / 60, procs=8):
    """The batched form's worlds (scenes.batched: log-uniform 8..512 bodies, box towers and
    sphere piles, each on its own 10x1x10 ground), each stepped `steps` times by the reference
    on its own: every world's final transforms hash and counts, and the sampled worlds' every
    step. Written to data/batched.traj.bmfx.gz: hashes as 32 bytes a world (u8), counts as i32."""
    import numpy as np
    import fixtures

    import shutil
    import tempfile

    scenes_list = scenes.batched(worlds, seed=seed)
    sample = batched_sample(scenes_list)
    tmp = tempfile.mkdtemp(prefix="bm_batched_")
    scene_path = os.path.join(tmp, "batched.json")
    with open(scene_path, "w") as f:
        json.dump(scenes_list, f)
    # Ranges of about equal body counts, one a process.
    sizes = np.array([len(w) for w in scenes_list])
    cuts = np.searchsorted(np.cumsum(sizes), np.linspace(0, sizes.sum(), procs + 1)[1:-1])
    bounds = [0, *[int(c) + 1 for c in cuts], worlds]
    jobs, running = [], []
    for k in range(procs):
        job = {"scenes": scene_path, "from": bounds[k], "to": bounds[k + 1], "steps": steps, "dt": dt,
               "sample": sample, "out": os.path.join(tmp, f"batched.{k}.out.json")}
        path = os.path.join(tmp, f"batched.{k}.job.json")
        with open(path, "w") as f:
            json.dump(job, f)
        jobs.append(job)
        running.append(subprocess.Popen(["node", "--no-warnings", "--experimental-loader",
                                         os.path.join(HERE, "lower.mjs"), os.path.join(HERE, "make_batched.mjs"), path]))
    for p in running:
        if p.wait() != 0:
            raise RuntimeError("make_batched.mjs failed")
    final, counts, per = [], {k: [] for k in ("awake", "contacts", "frozen", "impacts", "catches")}, {}
    for job in jobs:
        out = json.load(open(job["out"]))
        assert (out["from"], out["to"]) == (job["from"], job["to"])
        final += out["final"]
        for k in counts:
            counts[k] += out[k]
        per.update({int(i): r for i, r in out["sample"].items()})
    shutil.rmtree(tmp)
    arrays = {"final": np.frombuffer(b"".join(bytes.fromhex(h) for h in final), np.uint8).copy(),
              "sample": np.asarray(sample, np.int32)}
    for k, v in counts.items():
        arrays[k] = np.asarray(v, np.int32)
    for i in sample:
        r = per[i]
        arrays[f"s{i}.hashes"] = np.frombuffer(b"".join(bytes.fromhex(h) for h in r["hashes"]), np.uint8).copy()
        for k in ("awake", "contacts", "frozen"):
            arrays[f"s{i}.{k}"] = np.asarray(r[k], np.int32)
    meta = {"worlds": worlds, "seed": seed, "steps": steps, "dt": dt, "bodies": int(sizes.sum())}
    fixtures.save(os.path.join(DATA, "batched.traj.bmfx.gz"), arrays, meta)


if __name__ == "__main__":
    if sys.argv[1:] == ["batched"]:
        run_batched()
        sys.exit(0)
    if sys.argv[1:] == ["collide"]:
        run_collide()
        sys.exit(0)
    if sys.argv[1:] == ["collide_hulls"]:
        run_hulls()
        sys.exit(0)
    if sys.argv[1:] == ["collide_capsules"]:
        run_collide(n=3000, capsules=True)
        sys.exit(0)
    if sys.argv[1:] == ["level"]:
        run_level()
        sys.exit(0)
    if sys.argv[1:] == ["impact"]:
        run_impact()
        sys.exit(0)
    if sys.argv[1:2] == ["events"]:
        for name in sys.argv[2:] or list(JOBS):
            run_events(name)
        sys.exit(0)
    if sys.argv[1:2] == ["built"]:
        for name in sys.argv[2:] or ["rocks", "terrain"]:
            run_built(name)
        sys.exit(0)
    for name in sys.argv[1:] or [j for j in JOBS if j not in ("rain", "mixed_pile")]:
        run(name)
