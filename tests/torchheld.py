"""A plain FP64 PyTorch transcription of the held step, for checking the adjoint kernels'
gradients with torch.autograd: test infrastructure, written from the reference's JavaScript
and not from this library's own code.

The held trajectory (what bm_adjoint_held exports: each step's solved manifolds with their
ends held on their bodies, normals in a's frame, feature ids, carried rows, colours) is stepped
from an initial state with given masses, frictions and restitutions, every number a 0-dim
float64 tensor, every branch taken on the values as the reference takes it, so autograd
differentiates the pass's own branches. A step, for each world:

* the REUSE rebuild of every row from where its bodies are (src/contact.js:468-508), the
  tangent basis (1283-1292), friction and restitution mixed as emit() mixes them (889-921),
  the impulses carried from the row its pair stored last (999-1036);
* the world inverse inertias (1302-1310, 1493-1508) and each awake body's start;
* prepareRows (1337-1445, with lever 1565-1581 and core 1199-1214) and makeSoft (1590-1597);
* each substep: _accelerate (src/world.js:1495-1530), warmRows (src/contact.js:1616-1663),
  solveRows with its bias (1805-2090, pushes 1672-1724), _move (src/world.js:1533-1563,
  advance 2820-2828), solveRows relaxing;
* restitution in the order the rows were found (src/contact.js:2229-2251), and store with
  slipped and turned (1067-1182).

Worlds here hold boxes and balls; a lopsided body that tumbles, capsules and parts are left
out (the scenes this checks have none). Mass enters as the adjoint parametrises it: the inverse
mass and inverse inertias scaled by mass0 / mass, the recorded values at mass0.
"""
import math

import torch

P = 4
STIFFNESS = 1.0 / 8        # src/world.js:115
DAMPING_RATIO = 10.0       # src/world.js:118
SPIN = math.pi             # src/world.js:171: half a turn a step
SPIN_DISC = 2.0 / 3        # src/contact.js:146: a ball's spinning friction
COMPLIANCE = 2.0 ** -20    # src/forest.js:67
STARTS = 35
OVERFLOW = 32


def f(x):
    return torch.as_tensor(float(x), dtype=torch.float64)


def rotate(q, v, inverse):
    """src/core/math.js:211-220: v turned by q (by its conjugate when inverse)."""
    sg = -1.0 if inverse else 1.0
    ux, uy, uz, w = q[0] * sg, q[1] * sg, q[2] * sg, q[3]
    x, y, z = v
    tx, ty, tz = 2 * (uy * z - uz * y), 2 * (uz * x - ux * z), 2 * (ux * y - uy * x)
    return [x + w * tx + (uy * tz - uz * ty), y + w * ty + (uz * tx - ux * tz), z + w * tz + (ux * ty - uy * tx)]


def axes_of(q):
    """src/core/math.js:223-228: the rotation's columns, x, y, z axes one after another."""
    x, y, z, w = q
    return [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
            2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
            2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def basis(n):
    """src/contact.js:1283-1292."""
    nx, ny, nz = n
    ax, ay, az = abs(nx.item()), abs(ny.item()), abs(nz.item())
    zero = nx * 0
    if ax <= ay and ax <= az:
        t = [zero, nz, -ny]
    elif ay <= az:
        t = [-nz, zero, nx]
    else:
        t = [ny, -nx, zero]
    l = 1 / torch.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
    t = [t[0] * l, t[1] * l, t[2] * l]
    return t, [ny * t[2] - nz * t[1], nz * t[0] - nx * t[2], nx * t[1] - ny * t[0]]


def invi_mul(I, l):
    return [I[0] * l[0] + I[1] * l[1] + I[2] * l[2], I[1] * l[0] + I[3] * l[1] + I[4] * l[2],
            I[2] * l[0] + I[4] * l[1] + I[5] * l[2]]


def lever(Ia, Ib, ima, imb, ra, rb, d):
    """src/contact.js:1565-1581: jac (12) and the effective mass along d."""
    ca = [ra[1] * d[2] - ra[2] * d[1], ra[2] * d[0] - ra[0] * d[2], ra[0] * d[1] - ra[1] * d[0]]
    cb = [rb[1] * d[2] - rb[2] * d[1], rb[2] * d[0] - rb[0] * d[2], rb[0] * d[1] - rb[1] * d[0]]
    ia, ib = invi_mul(Ia, ca), invi_mul(Ib, cb)
    k = ima + imb + ca[0] * ia[0] + ca[1] * ia[1] + ca[2] * ia[2] + cb[0] * ib[0] + cb[1] * ib[1] + cb[2] * ib[2]
    return ca + cb + ia + ib, (1 / k if k.item() > 0 else k * 0)


def turning(Ia, Ib, d):
    """src/contact.js:1705-1715: 1 / (d . (Ia^-1 + Ib^-1) d)."""
    def along(I):
        return I[0] * d[0] * d[0] + I[3] * d[1] * d[1] + I[5] * d[2] * d[2] + 2 * (
            I[1] * d[0] * d[1] + I[2] * d[0] * d[2] + I[4] * d[1] * d[2])
    k = along(Ia) + along(Ib)
    return 1 / k if k.item() > 0 else k * 0


class World:
    """One world of a held trajectory: bodies [b0, b1) of the export, its rows step by step."""

    def __init__(self, held, wld, state0, mass, friction, restitution):
        H = held
        self.H = H
        self.wld = wld
        self.b0, self.b1 = int(H["body_offset"][wld]), int(H["body_offset"][wld + 1])
        nb = self.b1 - self.b0
        self.dt = float(H["dt"][wld])
        self.substeps = int(H["substeps"][wld])
        self.g = [float(x) for x in H["gravity"][3 * wld:3 * wld + 3]]
        self.p, self.q, self.v, self.w = [], [], [], []
        self.inv_mass, self.inv_inertia = [], []
        self.fr, self.re = [], []
        for s in range(nb):
            G = self.b0 + s
            st = state0[G]
            self.p.append([st[0], st[1], st[2]])
            self.q.append([st[3], st[4], st[5], st[6]])
            self.v.append([st[7], st[8], st[9]])
            self.w.append([st[10], st[11], st[12]])
            dynamic = H["type"][G] == 0
            if H["tumbles"][G] and dynamic:
                raise NotImplementedError("torchheld: a lopsided body tumbles, which this transcription leaves out")
            if H["shape"][G] not in (0, 1):
                raise NotImplementedError("torchheld: boxes and balls only")
            down = f(H["mass0"][G]) / mass[G] if dynamic else f(1.0)
            self.inv_mass.append(f(H["inv_mass0"][G]) * down if dynamic else f(H["inv_mass0"][G]))
            self.inv_inertia.append([f(H["inv_inertia0"][3 * G + k]) * down if dynamic else f(H["inv_inertia0"][3 * G + k])
                                     for k in range(3)])
            self.fr.append(friction[G])
            self.re.append(restitution[G])
        # What each row stored at its step's end, for the rows that carry from it: by global row.
        self.stored = {}
        first = int(H["first"])
        for r in range(first):
            self.stored[r] = {"points": int(H["count"][r]), "pid": [int(x) for x in H["pid"][4 * r:4 * r + 4]],
                              "impulse": [f(x) for x in H["store_impulse"][4 * r:4 * r + 4]],
                              "friction": [f(x) for x in H["store_friction"][3 * r:3 * r + 3]],
                              "twist": [f(x) for x in H["store_twist"][3 * r:3 * r + 3]],
                              "slip": [f(x) for x in H["store_slip"][3 * r:3 * r + 3]]}

    def dynamic(self, s):
        return self.H["type"][self.b0 + s] == 0

    # ---- a step ----

    def step(self, t):
        H, W = self.H, self.H["worlds"]
        r0, r1 = int(H["row_offset"][t * W + self.wld]), int(H["row_offset"][t * W + self.wld + 1])
        rows = [self.place(r) for r in range(r0, r1)]
        nb = self.b1 - self.b0
        # inertias (src/contact.js:1302-1310, inverseAt 1493-1508) and the step's start
        self.invI = []
        for s in range(nb):
            if self.dynamic(s):
                ax = axes_of(self.q[s])
                i0, i1, i2 = self.inv_inertia[s]
                x, y, z = ax[0:3], ax[3:6], ax[6:9]
                self.invI.append([x[0] * x[0] * i0 + y[0] * y[0] * i1 + z[0] * z[0] * i2,
                                  x[0] * x[1] * i0 + y[0] * y[1] * i1 + z[0] * z[1] * i2,
                                  x[0] * x[2] * i0 + y[0] * y[2] * i1 + z[0] * z[2] * i2,
                                  x[1] * x[1] * i0 + y[1] * y[1] * i1 + z[1] * z[1] * i2,
                                  x[1] * x[2] * i0 + y[1] * y[2] * i1 + z[1] * z[2] * i2,
                                  x[2] * x[2] * i0 + y[2] * y[2] * i1 + z[2] * z[2] * i2])
            else:
                self.invI.append([f(0.0)] * 6)
        self.start = [list(p) for p in self.p]
        self.start_q = [list(q) for q in self.q]
        h = self.dt / self.substeps
        if rows:
            for row in rows:
                self.prepare(row)
            omega = 2 * math.pi * (STIFFNESS / h)
            a1 = 2 * DAMPING_RATIO + h * omega
            a2 = h * omega * a1
            a3 = 1 / (1 + a2)
            self.soft = (omega / a1, a2 * a3, a3)
        starts = [int(x) for x in H["starts"][(t * W + self.wld) * STARTS:(t * W + self.wld + 1) * STARTS]]
        order = [rows[k] for c in range(OVERFLOW + 1) for k in range(starts[c], starts[c + 1])]
        for sub in range(self.substeps):
            self.accelerate(h)
            for row in order:
                self.warm(row)
            for row in order:
                self.solve(row, h, True, sub > 0)
            self.move(h)
            for row in order:
                self.solve(row, h, False, False)
        if rows:
            gravity = math.sqrt(self.g[0] ** 2 + self.g[1] ** 2 + self.g[2] ** 2)
            for k in range(r1 - r0):
                self.bounce(rows[int(H["narrow"][r0 + k])], gravity)
            for row in rows:
                self.store(row)

    def place(self, r):
        """The REUSE rebuild (src/contact.js:468-508) of row r from where its bodies are, its basis,
        the mixes of emit() (889-921) and the impulses carried (999-1036)."""
        H = self.H
        a, b = int(H["a"][r]), int(H["b"][r])
        count = int(H["count"][r])
        Aa, Ab = axes_of(self.q[a]), axes_of(self.q[b])
        lx, ly, lz = (f(x) for x in H["local_normal"][3 * r:3 * r + 3])
        n = [Aa[0] * lx + Aa[3] * ly + Aa[6] * lz, Aa[1] * lx + Aa[4] * ly + Aa[7] * lz,
             Aa[2] * lx + Aa[5] * ly + Aa[8] * lz]
        ra, rb = float(H["rounded"][self.b0 + a]), -float(H["rounded"][self.b0 + b])
        row = {"r": r, "a": a, "b": b, "count": count, "normal": n, "points": [], "sep": [],
               "pid": [int(x) for x in H["pid"][4 * r:4 * r + count]],
               "shift": [[f(x) for x in H["shift"][12 * r + 3 * k:12 * r + 3 * k + 3]] for k in range(count)],
               "tol": f(H["tolerance"][r]), "patch": f(H["patch"][r])}
        pa0, pb0 = self.p[a], self.p[b]
        for k in range(count):
            e = [f(x) for x in H["ends"][24 * r + 6 * k:24 * r + 6 * k + 6]]
            pa = [pa0[c] + Aa[c] * e[0] + Aa[3 + c] * e[1] + Aa[6 + c] * e[2] + n[c] * ra for c in range(3)]
            pb = [pb0[c] + Ab[c] * e[3] + Ab[3 + c] * e[4] + Ab[6 + c] * e[5] + n[c] * rb for c in range(3)]
            s = (pb[0] - pa[0]) * n[0] + (pb[1] - pa[1]) * n[1] + (pb[2] - pa[2]) * n[2]
            half = 0.5 * s
            if rb != 0:
                pt = [pb[c] - n[c] * half for c in range(3)]
            elif ra != 0:
                pt = [pa[c] + n[c] * half for c in range(3)]
            else:
                pt = [(pa[c] + pb[c]) * 0.5 for c in range(3)]
            row["points"].append(pt)
            row["sep"].append(s)
        row["t1"], row["t2"] = basis(n)
        row["mu"] = torch.sqrt(self.fr[a] * self.fr[b])
        row["e"] = torch.maximum(self.re[a], self.re[b])
        zero = f(0.0)
        row.update(impulse=[zero] * count, tangent=[zero, zero], roll=[zero, zero], spin=zero, drift=[zero, zero])
        src = int(H["src"][r])
        kept = self.stored.get(src) if src >= 0 else None
        if kept is not None and kept["points"] > 0:
            t1, t2 = row["t1"], row["t2"]
            L = kept["twist"]
            if count <= 2:
                row["roll"] = [dot(L, t1), dot(L, t2)]
            row["spin"] = dot(L, n)
            F = kept["friction"]
            row["tangent"] = [dot(F, t1), dot(F, t2)]
            same = False
            for k in range(count):
                for j in range(kept["points"]):
                    if kept["pid"][j] == row["pid"][k]:
                        row["impulse"][k] = kept["impulse"][j]
                        same = True
                        break
            if same:
                S = kept["slip"]
                row["drift"] = [dot(S, t1), dot(S, t2)]
        return row

    def core(self, s, r):
        """src/contact.js:1199-1214 for a box or a ball: the point's anchor in the body's frame,
        and where it is now relative to the body's centre."""
        if self.H["shape"][self.b0 + s] == 1:
            zero = f(0.0)
            return [zero, zero, zero], [zero, zero, zero]
        return rotate(self.q[s], r, True), r

    def prepare(self, row):
        """src/contact.js:1337-1445 for one row."""
        a, b, n, t1, t2 = row["a"], row["b"], row["normal"], row["t1"], row["t2"]
        Ia, Ib = self.invI[a], self.invI[b]
        ima, imb = self.inv_mass[a], self.inv_mass[b]
        count = row["count"]
        if count <= 2:
            row["roll_mass"] = [turning(Ia, Ib, t1), turning(Ia, Ib, t2)]
        row["spin_mass"] = turning(Ia, Ib, n)
        tol = row["tol"]
        weight = 0.0
        A = [f(0.0)] * 3
        B = [f(0.0)] * 3
        E = [f(0.0)] * 3
        F = [f(0.0)] * 3
        row.update(ra=[], rb=[], la=[], lb=[], adjusted=[], nmass=[], jac=[], most=[], relvel=[])
        pa, pb, va, vb, wa, wb = self.p[a], self.p[b], self.v[a], self.v[b], self.w[a], self.w[b]
        weight_t = None
        for k in range(count):
            pt, sh = row["points"][k], row["shift"][k]
            ra = [pt[c] - pa[c] for c in range(3)]
            rb = [pt[c] + sh[c] - pb[c] for c in range(3)]
            la, ca = self.core(a, ra)
            lb, cb = self.core(b, rb)
            row["ra"].append(ra); row["rb"].append(rb); row["la"].append(la); row["lb"].append(lb)
            row["adjusted"].append(row["sep"][k] - ((cb[0] - ca[0]) * n[0] + (cb[1] - ca[1]) * n[1] + (cb[2] - ca[2]) * n[2]))
            jac, nm = lever(Ia, Ib, ima, imb, ra, rb, n)
            row["jac"].append(jac); row["nmass"].append(nm)
            row["most"].append(f(0.0))
            s = row["sep"][k] / tol
            sv = s.item()
            near = f(1.0) if sv <= 1 else (f(0.0) if sv >= 2 else 2 - s)
            weight_t = near if weight_t is None else weight_t + near
            weight += near.item()
            A = [A[c] + near * ra[c] for c in range(3)]
            B = [B[c] + near * rb[c] for c in range(3)]
            E = [E[c] + ra[c] for c in range(3)]
            F = [F[c] + rb[c] for c in range(3)]
            dv = [vb[0] + wb[1] * rb[2] - wb[2] * rb[1] - va[0] - wa[1] * ra[2] + wa[2] * ra[1],
                  vb[1] + wb[2] * rb[0] - wb[0] * rb[2] - va[1] - wa[2] * ra[0] + wa[0] * ra[2],
                  vb[2] + wb[0] * rb[1] - wb[1] * rb[0] - va[2] - wa[0] * ra[1] + wa[1] * ra[0]]
            row["relvel"].append(dot(dv, n))
        if count >= 2:
            K = [[None] * count for _ in range(count)]
            for k in range(count):
                jk = row["jac"][k]
                for l in range(k + 1):
                    jl = row["jac"][l]
                    e = ima + imb + jk[0] * jl[6] + jk[1] * jl[7] + jk[2] * jl[8] + jk[3] * jl[9] + jk[4] * jl[10] + jk[5] * jl[11]
                    K[k][l] = e
                    K[l][k] = e
            row["K"] = K
        if weight == 0:
            weight_t, A, B = f(float(count)), E, F
        A = [x / weight_t for x in A]
        B = [x / weight_t for x in B]
        row["anchor_a"], row["anchor_b"] = A, B
        row["lever"] = []
        for k in range(count):
            d = [row["ra"][k][c] - A[c] for c in range(3)]
            al = dot(d, n)
            d = [d[c] - al * n[c] for c in range(3)]
            row["lever"].append(torch.sqrt(dot(d, d)))
        f1, _ = lever(Ia, Ib, ima, imb, A, B, t1)
        f2, _ = lever(Ia, Ib, ima, imb, A, B, t2)
        row["fjac"] = f1 + f2
        k11 = ima + imb + f1[0] * f1[6] + f1[1] * f1[7] + f1[2] * f1[8] + f1[3] * f1[9] + f1[4] * f1[10] + f1[5] * f1[11]
        k22 = ima + imb + f2[0] * f2[6] + f2[1] * f2[7] + f2[2] * f2[8] + f2[3] * f2[9] + f2[4] * f2[10] + f2[5] * f2[11]
        k12 = f2[0] * f1[6] + f2[1] * f1[7] + f2[2] * f1[8] + f2[3] * f1[9] + f2[4] * f1[10] + f2[5] * f1[11]
        det = k11 * k22 - k12 * k12
        row["tm"] = [k22 / det, -k12 / det, k11 / det] if det.item() > 0 else [f(0.0)] * 3
        row["now"] = [None] * count

    # ---- the substep's sweeps ----

    def accelerate(self, h):
        """src/world.js:1495-1530."""
        most = SPIN / (h * self.substeps)
        for s in range(len(self.p)):
            if not self.dynamic(s):
                continue
            G = self.b0 + s
            gs = float(self.H["gravity_scale"][G])
            self.v[s] = [self.v[s][c] + self.g[c] * h * gs for c in range(3)]
            if self.H["damping"][G] != 0:
                k = 1 / (1 + h * float(self.H["damping"][G]))
                self.v[s] = [x * k for x in self.v[s]]
            if self.H["angular_damping"][G] != 0:
                k = 1 / (1 + h * float(self.H["angular_damping"][G]))
                self.w[s] = [x * k for x in self.w[s]]
            if self.H["capped"][G]:
                w = self.w[s]
                w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2]
                if w2.item() > most * most:
                    k = most / torch.sqrt(w2)
                    self.w[s] = [x * k for x in w]

    def move(self, h):
        """src/world.js:1533-1563 and advance (2820-2828) for bodies that do not tumble."""
        half = h * 0.5
        for s in range(len(self.p)):
            if not self.dynamic(s):
                continue
            self.p[s] = [self.p[s][c] + self.v[s][c] * h for c in range(3)]
            wx, wy, wz = self.w[s]
            if wx.item() != 0 or wy.item() != 0 or wz.item() != 0:
                x, y, z, w = self.q[s]
                nx = x + half * (wx * w + wy * z - wz * y)
                ny = y + half * (wy * w + wz * x - wx * z)
                nz = z + half * (wz * w + wx * y - wy * x)
                nw = w - half * (wx * x + wy * y + wz * z)
                l = 1 / torch.sqrt(nx * nx + ny * ny + nz * nz + nw * nw)
                self.q[s] = [nx * l, ny * l, nz * l, nw * l]

    def kick(self, row, va, vb, wa, wb, jac, d, n):
        """A normal impulse d along n at a point with turning terms jac (src/contact.js push)."""
        ima, imb = self.inv_mass[row["a"]], self.inv_mass[row["b"]]
        va = [va[c] - ima * (n[c] * d) for c in range(3)]
        vb = [vb[c] + imb * (n[c] * d) for c in range(3)]
        wa = [wa[c] - jac[6 + c] * d for c in range(3)]
        wb = [wb[c] + jac[9 + c] * d for c in range(3)]
        return va, vb, wa, wb

    def write_back(self, row, va, vb, wa, wb):
        a, b = row["a"], row["b"]
        if self.inv_mass[a].item() != 0:
            self.v[a], self.w[a] = va, wa
        if self.inv_mass[b].item() != 0:
            self.v[b], self.w[b] = vb, wb

    def warm(self, row):
        """src/contact.js:1616-1663 for one row."""
        a, b, n, t1, t2 = row["a"], row["b"], row["normal"], row["t1"], row["t2"]
        ima, imb = self.inv_mass[a], self.inv_mass[b]
        va, vb, wa, wb = self.v[a], self.v[b], self.w[a], self.w[b]
        for k in range(row["count"]):
            va, vb, wa, wb = self.kick(row, va, vb, wa, wb, row["jac"][k], row["impulse"][k], n)
        fj = row["fjac"]
        l1, l2 = row["tangent"]
        p = [t1[c] * l1 + t2[c] * l2 for c in range(3)]
        va = [va[c] - ima * p[c] for c in range(3)]
        vb = [vb[c] + imb * p[c] for c in range(3)]
        wa = [wa[c] - (fj[6 + c] * l1 + fj[18 + c] * l2) for c in range(3)]
        wb = [wb[c] + (fj[9 + c] * l1 + fj[21 + c] * l2) for c in range(3)]
        r1, r2 = row["roll"]
        sp = row["spin"]
        L = [t1[c] * r1 + t2[c] * r2 + n[c] * sp for c in range(3)]
        ia, ib = invi_mul(self.invI[a], L), invi_mul(self.invI[b], L)
        wa = [wa[c] - ia[c] for c in range(3)]
        wb = [wb[c] + ib[c] for c in range(3)]
        self.write_back(row, va, vb, wa, wb)

    def solve(self, row, h, bias, known):
        """src/contact.js:1805-2090 for one row: bias (a solving pass) or not (relaxing)."""
        a, b, n, t1, t2 = row["a"], row["b"], row["normal"], row["t1"], row["t2"]
        ima, imb = self.inv_mass[a], self.inv_mass[b]
        count = row["count"]
        inv = 1 / h
        bias_rate, mass_scale, impulse_scale = self.soft
        springy = impulse_scale / mass_scale + COMPLIANCE
        max_bias = row["tol"] * inv
        if not known:
            d = [(self.p[b][c] - self.start[b][c]) - (self.p[a][c] - self.start[a][c]) for c in range(3)]
            dn = dot(d, n)
            na, nb = rotate(self.q[a], n, True), rotate(self.q[b], n, True)
        target, ms, is_, give = [None] * count, [None] * count, [None] * count, [None] * count
        for k in range(count):
            if known:
                s = row["now"][k]
            else:
                s = dn + dot(row["lb"][k], nb) - dot(row["la"][k], na) + row["adjusted"][k]
                if not bias:
                    row["now"][k] = s
            sv = s.item()
            if sv > 0:
                target[k], ms[k], is_[k], give[k] = s * inv, 1.0, 0.0, COMPLIANCE
            elif bias:
                t = bias_rate * s
                target[k] = torch.maximum(t, -max_bias)
                ms[k], is_[k], give[k] = mass_scale, impulse_scale, springy
            else:
                target[k], ms[k], is_[k], give[k] = f(0.0), 1.0, 0.0, COMPLIANCE
        va, vb, wa, wb = self.v[a], self.v[b], self.w[a], self.w[b]

        def closing(jac):
            return ((vb[0] - va[0]) * n[0] + (vb[1] - va[1]) * n[1] + (vb[2] - va[2]) * n[2]
                    + wb[0] * jac[3] + wb[1] * jac[4] + wb[2] * jac[5] - wa[0] * jac[0] - wa[1] * jac[1] - wa[2] * jac[2])

        solved, now_closing = [None] * count, [None] * count
        K = row.get("K")
        if count == 1 or not K[0][0].item() > 0:
            for k in range(count):
                vn = closing(row["jac"][k])
                old = row["impulse"][k]
                nxt = old - row["nmass"][k] * ms[k] * (vn + target[k]) - is_[k] * old
                if nxt.item() < 0:
                    nxt = nxt * 0
                now_closing[k], solved[k] = vn, nxt
        else:
            for k in range(count):
                now_closing[k] = closing(row["jac"][k])
            wanted = []
            bits = 0
            for k in range(count):
                r = -now_closing[k] - target[k]
                for l in range(count):
                    r = r + K[k][l] * row["impulse"][l]
                wanted.append(r)
                if r.item() > 0:
                    bits |= 1 << k
            rnd = 0
            while True:
                solved = self.pushes(K, count, bits, give, wanted)
                flip = -1
                for k in range(count):
                    if flip >= 0:
                        break
                    if bits >> k & 1:
                        if solved[k].item() < 0:
                            flip = k
                    else:
                        r = -wanted[k]
                        for l in range(count):
                            r = r + K[k][l] * solved[l]
                        if r.item() < 0:
                            flip = k
                if flip < 0 or rnd == 1 << count:
                    break
                bits ^= 1 << flip
                rnd += 1
        for k in range(count):
            old = row["impulse"][k]
            nxt = solved[k] if solved[k].item() > 0 else solved[k] * 0
            d = nxt - old
            row["impulse"][k] = nxt
            if d.item() > 0 and row["most"][k].item() == 0:
                row["relvel"][k] = now_closing[k]
            if d.item() > row["most"][k].item():
                row["most"][k] = d
            va, vb, wa, wb = self.kick(row, va, vb, wa, wb, row["jac"][k], d, n)
        mu = row["mu"]
        pressed = row["impulse"][0]
        twisted = row["lever"][0] * row["impulse"][0]
        for k in range(1, count):
            pressed = pressed + row["impulse"][k]
            twisted = twisted + row["lever"][k] * row["impulse"][k]
        fj = row["fjac"]
        dv = [vb[c] - va[c] for c in range(3)]
        v1 = dot(dv, t1) + wb[0] * fj[3] + wb[1] * fj[4] + wb[2] * fj[5] - wa[0] * fj[0] - wa[1] * fj[1] - wa[2] * fj[2]
        v2 = (dot(dv, t2) + wb[0] * fj[15] + wb[1] * fj[16] + wb[2] * fj[17]
              - wa[0] * fj[12] - wa[1] * fj[13] - wa[2] * fj[14])
        s1, s2 = row["drift"]
        if bias and (s1.item() != 0 or s2.item() != 0):
            b1, b2 = bias_rate * s1, bias_rate * s2
            bb = b1 * b1 + b2 * b2
            if bb.item() > (max_bias * max_bias).item():
                scale = max_bias / torch.sqrt(bb)
                b1, b2 = b1 * scale, b2 * scale
            v1, v2 = v1 + b1, v2 + b2
        tm = row["tm"]
        o1, o2 = row["tangent"]
        n1 = o1 - (tm[0] * v1 + tm[1] * v2)
        n2 = o2 - (tm[1] * v1 + tm[2] * v2)
        limit = mu * pressed
        l2 = n1 * n1 + n2 * n2
        if l2.item() > (limit * limit).item():
            ma, mb, mc = tm
            al, x, y, ll = f(0.0), n1, n2, l2
            while limit.item() > 0:
                ab = al * mb
                d = (1 + al * ma) * (1 + al * mc) - ab * ab
                px, py = ma * x + mb * y, mb * x + mc * y
                wx = ((1 + al * mc) * px - ab * py) / d
                wy = ((1 + al * ma) * py - ab * px) / d
                lw = x * wx + y * wy
                if not lw.item() > 0:
                    break
                nxt = al + (torch.sqrt(ll) / limit - 1) * ll / lw
                if not nxt.item() > al.item():
                    break
                al = nxt
                ab2 = al * mb
                d2 = (1 + al * ma) * (1 + al * mc) - ab2 * ab2
                x = ((1 + al * mc) * n1 - ab2 * n2) / d2
                y = ((1 + al * ma) * n2 - ab2 * n1) / d2
                ll = x * x + y * y
                if not ll.item() > (limit * limit).item():
                    break
            scale = limit / torch.sqrt(ll)
            n1, n2 = x * scale, y * scale
            row["sliding"] = True
        else:
            row["sliding"] = False
        row["tangent"] = [n1, n2]
        d1, d2 = n1 - o1, n2 - o2
        p = [t1[c] * d1 + t2[c] * d2 for c in range(3)]
        va = [va[c] - ima * p[c] for c in range(3)]
        vb = [vb[c] + imb * p[c] for c in range(3)]
        wa = [wa[c] - (fj[6 + c] * d1 + fj[18 + c] * d2) for c in range(3)]
        wb = [wb[c] + (fj[9 + c] * d1 + fj[21 + c] * d2) for c in range(3)]
        Ia, Ib = self.invI[a], self.invI[b]
        if count <= 2:
            most = row["patch"] * pressed
            u = [wb[c] - wa[c] for c in range(3)]
            r1, r2 = row["roll"]
            m1, m2 = row["roll_mass"]
            n1 = r1 - m1 * dot(u, t1)
            n2 = r2 - m2 * dot(u, t2)
            l2 = n1 * n1 + n2 * n2
            if l2.item() > (most * most).item():
                scale = most / torch.sqrt(l2)
                n1, n2 = n1 * scale, n2 * scale
            row["roll"] = [n1, n2]
            d1, d2 = n1 - r1, n2 - r2
            L = [t1[c] * d1 + t2[c] * d2 for c in range(3)]
            ia, ib = invi_mul(Ia, L), invi_mul(Ib, L)
            wa = [wa[c] - ia[c] for c in range(3)]
            wb = [wb[c] + ib[c] for c in range(3)]
        limit = SPIN_DISC * mu * row["patch"] * pressed if count == 1 else mu * twisted
        s0 = row["spin"]
        s1 = s0 - row["spin_mass"] * dot([wb[c] - wa[c] for c in range(3)], n)
        if s1.item() > limit.item():
            s1 = limit
        elif s1.item() < -limit.item():
            s1 = -limit
        row["spin"] = s1
        ds = s1 - s0
        L = [n[c] * ds for c in range(3)]
        ia, ib = invi_mul(Ia, L), invi_mul(Ib, L)
        wa = [wa[c] - ia[c] for c in range(3)]
        wb = [wb[c] + ib[c] for c in range(3)]
        self.write_back(row, va, vb, wa, wb)

    @staticmethod
    def pushes(K, count, bits, give, wanted):
        """src/contact.js:1672-1724: the pushing points' equations, Gaussian elimination with no
        pivoting; the other points' impulses 0."""
        idx = [k for k in range(count) if bits >> k & 1]
        n = len(idx)
        solved = [wanted[0] * 0 for _ in range(count)]
        A = [[K[idx[r]][idx[s]] for s in range(n)] for r in range(n)]
        for r in range(n):
            A[r][r] = A[r][r] * (1 + give[idx[r]])
        R = [wanted[idx[r]] for r in range(n)]
        for col in range(n):
            pivot = A[col][col]
            for r in range(col + 1, n):
                fct = A[r][col] / pivot
                for s in range(col + 1, n):
                    A[r][s] = A[r][s] - fct * A[col][s]
                R[r] = R[r] - fct * R[col]
        X = [None] * n
        for r in range(n - 1, -1, -1):
            x = R[r]
            for s in range(r + 1, n):
                x = x - A[r][s] * X[s]
            X[r] = x / A[r][r]
            solved[idx[r]] = X[r]
        return solved

    # ---- the step's end ----

    def bounce(self, row, gravity):
        """src/contact.js:2229-2251 for one row."""
        e = row["e"]
        if e.item() == 0:
            return
        threshold = math.sqrt(2 * gravity * row["tol"].item())
        a, b, n = row["a"], row["b"], row["normal"]
        for k in range(row["count"]):
            if row["relvel"][k].item() > -threshold or row["most"][k].item() == 0:
                continue
            jac = row["jac"][k]
            va, vb, wa, wb = self.v[a], self.v[b], self.w[a], self.w[b]
            vn = ((vb[0] - va[0]) * n[0] + (vb[1] - va[1]) * n[1] + (vb[2] - va[2]) * n[2]
                  + wb[0] * jac[3] + wb[1] * jac[4] + wb[2] * jac[5] - wa[0] * jac[0] - wa[1] * jac[1] - wa[2] * jac[2])
            old = row["impulse"][k]
            nxt = old - row["nmass"][k] * (vn + e * row["relvel"][k])
            if nxt.item() < 0:
                nxt = nxt * 0
            d = nxt - old
            row["impulse"][k] = nxt
            va, vb, wa, wb = self.kick(row, va, vb, wa, wb, jac, d, n)
            self.write_back(row, va, vb, wa, wb)

    def turned(self, s):
        """src/contact.js:1168-1182: the body's move over the step and its turn vector."""
        q, q0 = self.q[s], self.start_q[s]
        ax, ay, az, aw = q
        bx, by, bz, bw = -q0[0], -q0[1], -q0[2], q0[3]
        ux = aw * bx + ax * bw + ay * bz - az * by
        uy = aw * by + ay * bw + az * bx - ax * bz
        uz = aw * bz + az * bw + ax * by - ay * bx
        uw = aw * bw - ax * bx - ay * by - az * bz
        if uw.item() < 0:
            ux, uy, uz, uw = -ux, -uy, -uz, -uw
        sine = torch.sqrt(ux * ux + uy * uy + uz * uz)
        k = sine * 0 if sine.item() == 0 else 2 * torch.atan2(sine, uw) / sine
        return [self.p[s][c] - self.start[s][c] for c in range(3)] + [ux * k, uy * k, uz * k]

    def store(self, row):
        """src/contact.js:1067-1115 (store, keep) and slipped (1122-1153) for one row."""
        a, b, n, t1, t2 = row["a"], row["b"], row["normal"], row["t1"], row["t2"]
        count = row["count"]
        pressed = row["impulse"][0]
        for k in range(1, count):
            pressed = pressed + row["impulse"][k]
        impact = any(any(x.item() != 0 for x in sh) for sh in row["shift"])
        if row.get("sliding", False) or pressed.item() == 0 or impact:
            row["drift"] = [f(0.0), f(0.0)]
        else:
            d = [f(0.0)] * 3
            if self.dynamic(b):
                T = self.turned(b)
                r = row["anchor_b"]
                d = [T[0] + T[4] * r[2] - T[5] * r[1], T[1] + T[5] * r[0] - T[3] * r[2],
                     T[2] + T[3] * r[1] - T[4] * r[0]]
            if self.dynamic(a):
                T = self.turned(a)
                r = row["anchor_a"]
                d = [d[0] - (T[0] + T[4] * r[2] - T[5] * r[1]), d[1] - (T[1] + T[5] * r[0] - T[3] * r[2]),
                     d[2] - (T[2] + T[3] * r[1] - T[4] * r[0])]
            row["drift"] = [row["drift"][0] + dot(d, t1), row["drift"][1] + dot(d, t2)]
        r1, r2 = row["roll"]
        sp = row["spin"]
        l1, l2 = row["tangent"]
        d1, d2 = row["drift"]
        self.stored[row["r"]] = {
            "points": count, "pid": row["pid"], "impulse": list(row["impulse"]),
            "twist": [t1[c] * r1 + t2[c] * r2 + n[c] * sp for c in range(3)],
            "friction": [t1[c] * l1 + t2[c] * l2 for c in range(3)],
            "slip": [t1[c] * d1 + t2[c] * d2 for c in range(3)]}

    def state(self):
        out = []
        for s in range(len(self.p)):
            out.append(torch.stack(self.p[s] + self.q[s] + self.v[s] + self.w[s]))
        return torch.stack(out)


def forward(held, state0, mass, friction, restitution):
    """The held trajectory from state0 (bodies x 13) with the given parameters (a tensor a body):
    the final state, bodies x 13, differentiable in all four."""
    finals = []
    for wld in range(held["worlds"]):
        rows0 = [[state0[g, k] for k in range(13)] for g in range(state0.shape[0])]
        world = World(held, wld, rows0, [mass[g] for g in range(len(mass))], [friction[g] for g in range(len(friction))],
                      [restitution[g] for g in range(len(restitution))])
        for t in range(held["steps"]):
            world.step(t)
        finals.append(world.state())
    return torch.cat(finals)
