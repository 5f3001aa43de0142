"""Generate element golden vectors with sympy at 40 significant digits.

The energies are the reference's element energies, written here directly from their defining
formulas:
  spring        springEnergy.cpp:20-30        E = 1/2 k (|x0-x1| - L0)^2
  anchor        springAnchorEnergy.cpp:21-26  E = 1/2 k |x - a|^2
  StVK membrane membraneStvkEnergy.cpp:9-28   C = (E00, E11, 2E01), E = 1/2 C^T Keff C
  bending       discreteBendingEnergy.cpp:9-30  E = 1/2 keff (theta - theta0)^2
  SNH           stableNeoHookeanEnergy.cpp:33-61  E = 1/2 lambda (J - alpha)^2 + 1/2 mu (tr F^T F - 3)
and the rest-state preparation comes from each energy's preparePrecomputations
(membraneStvkEnergy.cpp:45-71, discreteBendingEnergy.cpp:48-72, stableNeoHookeanEnergy.cpp:82-108).
Gradients and Hessians are exact symbolic derivatives (sympy.diff), evaluated with mpmath at
40 digits on the exact binary values of the double inputs. The result does not depend on the
oracle or on the product.

Inputs: the canonical tiny elements of apps/energyViewer.cpp:20-137 (at rest and displaced), plus
seeded random elements. Output: tests/golden/element_golden.json (values written as 17-digit
decimal strings of the 40-digit results).

Run:  python tests/golden/gen_golden.py
"""
import json
import os
import random

import mpmath
import sympy as sp

mpmath.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def vec(syms):
    return sp.Matrix(syms)


def build_energy(kind):
    """Return (position symbols, parameter symbols, energy expression)."""
    if kind == "spring":
        X = sp.symbols("x0:6")
        P = sp.symbols("L0 k")
        a, b = vec(X[0:3]), vec(X[3:6])
        d = a - b
        E = sp.Rational(1, 2) * P[1] * (sp.sqrt(d.dot(d)) - P[0]) ** 2
    elif kind == "anchor":
        X = sp.symbols("x0:3")
        P = sp.symbols("ax ay az k")
        d = vec(X) - vec(P[0:3])
        E = sp.Rational(1, 2) * P[3] * d.dot(d)
    elif kind == "stvk":
        X = sp.symbols("x0:9")
        P = sp.symbols("d0:4 K0:9")
        x0, x1, x2 = vec(X[0:3]), vec(X[3:6]), vec(X[6:9])
        Dinv = sp.Matrix(2, 2, P[0:4])
        K = sp.Matrix(3, 3, P[4:13])
        u, v = x1 - x0, x2 - x0
        F = sp.Matrix.hstack(u, v) * Dinv
        G = sp.Rational(1, 2) * (F.T * F - sp.eye(2))
        C = sp.Matrix([G[0, 0], G[1, 1], 2 * G[0, 1]])
        E = sp.Rational(1, 2) * (C.T * K * C)[0, 0]
    elif kind == "bending":
        X = sp.symbols("x0:12")
        P = sp.symbols("theta0 k")
        x0, x1, x2, x3 = vec(X[0:3]), vec(X[3:6]), vec(X[6:9]), vec(X[9:12])
        e0, e3, e4 = x1 - x0, x2 - x1, x3 - x1
        c1, c2 = e0.cross(e3), e0.cross(e4)
        n1 = c1 / sp.sqrt(c1.dot(c1))
        n2 = -c2 / sp.sqrt(c2.dot(c2))
        eh = e0 / sp.sqrt(e0.dot(e0))
        theta = sp.atan2(n1.cross(n2).dot(eh), n1.dot(n2))
        E = sp.Rational(1, 2) * P[1] * (theta - P[0]) ** 2
    elif kind == "snh":
        X = sp.symbols("x0:12")
        P = sp.symbols("d0:9 lam mu alpha")
        x0, x1, x2, x3 = vec(X[0:3]), vec(X[3:6]), vec(X[6:9]), vec(X[9:12])
        Dinv = sp.Matrix(3, 3, P[0:9])
        F = sp.Matrix.hstack(x1 - x0, x2 - x0, x3 - x0) * Dinv
        J = F.det()
        E = sp.Rational(1, 2) * P[9] * (J - P[11]) ** 2 + sp.Rational(1, 2) * P[10] * ((F.T * F).trace() - 3)
    else:
        raise ValueError(kind)
    return X, P, E


def mp(x):
    return mpmath.mpf(float(x))  # exact binary value of the double


def prepare(kind, rest, raw):
    """preparePrecomputations at 40 digits, from rest positions (flat list) and addStencil args."""
    R = [mp(r) for r in rest]
    if kind == "spring":
        return [mp(raw[0]), mp(raw[1])]
    if kind == "anchor":
        return [mp(r) for r in raw]
    if kind == "stvk":
        young, poisson = mp(raw[0]), mp(raw[1])
        yc = young / (1 - poisson * poisson)
        K = [yc, yc * poisson, 0, yc * poisson, yc, 0, 0, 0, yc * (1 - poisson) / 2]
        u = (R[3] - R[0], R[4] - R[1])
        v = (R[6] - R[0], R[7] - R[1])
        det = u[0] * v[1] - v[0] * u[1]
        Dinv = [v[1] / det, -v[0] / det, -u[1] / det, u[0] / det]
        area = det / 2
        assert area > 0
        return Dinv + [k * area for k in K]
    if kind == "bending":
        x = [mpmath.matrix(R[3 * i:3 * i + 3]) for i in range(4)]
        e0, e3, e4 = x[1] - x[0], x[2] - x[1], x[3] - x[1]

        def cross(a, b):
            return mpmath.matrix([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])

        def norm(a):
            return mpmath.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)

        def dot(a, b):
            return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

        factor = 6 * dot(e0, e0) / (mpmath.mpf(1) / 2 * (norm(cross(e0, e3)) + norm(cross(e0, e4))))
        keff = 1 / ((1 / mp(raw[1])) / factor)
        n1 = cross(e0, e3) / norm(cross(e0, e3))
        n2 = -cross(e0, e4) / norm(cross(e0, e4))
        theta0 = mpmath.atan2(dot(cross(n1, n2), e0 / norm(e0)), dot(n1, n2))
        return [theta0, keff]
    if kind == "snh":
        young, poisson = mp(raw[0]), mp(raw[1])
        mu = young / (2 * (1 + poisson))
        lam = mu / (1 - 2 * poisson)
        Dm = mpmath.matrix(3, 3)
        for c in range(3):
            for r in range(3):
                Dm[r, c] = R[3 * (c + 1) + r] - R[r]
        vol = mpmath.det(Dm) / 6
        assert vol > 0
        Di = Dm ** -1
        return [Di[r, c] for r in range(3) for c in range(3)] + [lam * vol, mu * vol, 1 + mu / lam]
    raise ValueError(kind)


def s17(x):
    return mpmath.nstr(x, 17, min_fixed=-3, max_fixed=3, strip_zeros=False)


class Evaluator:
    def __init__(self, kind):
        X, P, E = build_energy(kind)
        g = [sp.diff(E, xi) for xi in X]
        H = [[sp.diff(gi, xj) for xj in X] for gi in g]
        args = list(X) + list(P)
        self.fE = sp.lambdify(args, E, "mpmath")
        self.fg = sp.lambdify(args, g, "mpmath")
        self.fH = sp.lambdify(args, H, "mpmath")
        self.n = len(X)

    def __call__(self, xs, prm):
        a = [mp(v) for v in xs] + list(prm)
        return self.fE(*a), self.fg(*a), self.fH(*a)


def case(kind, ev, xs, rest, raw, name):
    prm = prepare(kind, rest, raw)
    E, g, H = ev(xs, prm)
    return {
        "kind": kind,
        "name": name,
        "x": [float(v) for v in xs],
        "rest": [float(v) for v in rest],
        "raw": [float(v) for v in raw],
        "prepared": [s17(p) for p in prm],
        "energy": s17(E),
        "grad": [s17(v) for v in g],
        "hess": [[s17(v) for v in row] for row in H],
    }


def main():
    rng = random.Random(20240611)
    out = []
    U = lambda a, b: rng.uniform(a, b)  # noqa: E731

    # ---- spring (energyViewer.cpp:101-111) ----
    ev = Evaluator("spring")
    base = [0, 0, 0, 1, 0, 0]
    out.append(case("spring", ev, [0, 0, 0, 1.3, 0.2, -0.1], base, [1, 1], "energyViewer_spring_stretched"))
    out.append(case("spring", ev, [0, 0, 0, 0.6, 0.1, 0.05], base, [1, 1], "energyViewer_spring_compressed"))
    for i in range(8):
        xs = [U(-1, 1) for _ in range(6)]
        out.append(case("spring", ev, xs, xs, [U(0.2, 2.0), 10 ** U(-2, 3)], f"random_{i}"))

    # ---- anchor ----
    ev = Evaluator("anchor")
    for i in range(6):
        xs = [U(-1, 1) for _ in range(3)]
        raw = [U(-1, 1), U(-1, 1), U(-1, 1), 10 ** U(0, 5)]
        out.append(case("anchor", ev, xs, xs, raw, f"random_{i}"))

    # ---- StVK membrane (energyViewer.cpp:20-31) ----
    ev = Evaluator("stvk")
    rest = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    out.append(case("stvk", ev, rest, rest, [1, 0], "energyViewer_stvk_rest"))
    out.append(case("stvk", ev, [0, 0, 0, 1.2, 0.1, 0.3, -0.1, 0.9, 0.2], rest, [1, 0], "energyViewer_stvk_deformed"))
    for i in range(10):
        # CCW rest triangle in the xy plane (z = 0), random current positions
        a = [U(-1, 1), U(-1, 1), 0.0]
        b = [a[0] + U(0.2, 1.0), a[1] + U(-0.3, 0.3), 0.0]
        c = [a[0] + U(-0.3, 0.3), a[1] + U(0.2, 1.0), 0.0]
        rest = a + b + c
        xs = [r + U(-0.3, 0.3) for r in rest]
        out.append(case("stvk", ev, xs, rest, [10 ** U(0, 4), U(0.0, 0.45)], f"random_{i}"))

    # ---- discrete bending (energyViewer.cpp:59-71) ----
    ev = Evaluator("bending")
    rest = [1, -1, 0, 1, 1, 0, 0, 0, 0, 2, 0, 0]
    cur = [1, 0, 1, 1, 0, -1, 0, 0, 0, 2, 0, 0]
    out.append(case("bending", ev, cur, rest, [0, 1], "energyViewer_bending_flat"))
    cur2 = [1, 0, 1, 1, 0, -1, 0, 0.4, 0, 2, 0.1, 0.2]
    out.append(case("bending", ev, cur2, rest, [0, 1], "energyViewer_bending_folded"))
    for i in range(12):
        # rest hinge with random dihedral, current = rest + noise
        e0 = [U(-1, 1), U(-1, 1), 0.0]
        e1 = [e0[0] + U(0.5, 1.5), e0[1] + U(-0.3, 0.3), 0.0]
        w1 = [(e0[0] + e1[0]) / 2 + U(-0.3, 0.3), (e0[1] + e1[1]) / 2 + U(0.3, 1.0), U(-0.5, 0.5)]
        w2 = [(e0[0] + e1[0]) / 2 + U(-0.3, 0.3), (e0[1] + e1[1]) / 2 - U(0.3, 1.0), U(-0.5, 0.5)]
        rest = e0 + e1 + w1 + w2
        xs = [r + U(-0.2, 0.2) for r in rest]
        out.append(case("bending", ev, xs, rest, [0.0, 10 ** U(-3, 1)], f"random_{i}"))

    # ---- Stable Neo-Hookean (energyViewer.cpp:125-137) ----
    ev = Evaluator("snh")
    rest = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]
    out.append(case("snh", ev, rest, rest, [10, 0.2], "energyViewer_snh_rest"))
    out.append(case("snh", ev, [0, 0, 0, 1.1, 0.1, 0, 0, 0.8, 0.1, 0.1, -0.1, 1.2], rest, [10, 0.2],
                    "energyViewer_snh_deformed"))
    out.append(case("snh", ev, [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, -0.5], rest, [10, 0.2], "energyViewer_snh_inverted"))
    for i in range(10):
        while True:
            pts = [U(-1, 1) for _ in range(12)]
            u = [pts[3 + k] - pts[k] for k in range(3)]
            v = [pts[6 + k] - pts[k] for k in range(3)]
            w = [pts[9 + k] - pts[k] for k in range(3)]
            det = (u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0])
                   + u[2] * (v[0] * w[1] - v[1] * w[0]))
            if det > 0.05:
                break
        xs = [p + U(-0.2, 0.2) for p in pts]
        out.append(case("snh", ev, xs, pts, [10 ** U(2, 6), U(0.0, 0.48)], f"random_{i}"))

    path = os.path.join(HERE, "element_golden.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py", "digits": 40, "cases": out}, f, indent=0)
    print(f"wrote {len(out)} cases to {path}")


if __name__ == "__main__":
    main()
