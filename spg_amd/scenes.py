"""Synthetic scenes of the VBD workloads, built the way spg's apps wire scenes
(apps/solverCompare.cpp:332-498 for cloth, 118-189,260-330 for tetrahedra): particles with
positions, rest positions and masses, then energies in addEnergy order with their addStencil
arguments. The same Scene feeds the GPU solver (through the C++ mirror) and the FP64 oracle.

Workloads (all gravity (0, -9.81, 0), dt = 1/60, one substep):
  cloth_config1()  32x32 grid, 1 m square in the xy plane, one diagonal per quad (1922 triangles),
                   mass 1e-3 kg per vertex, the two top corners pinned, StVK (Y t = 1e4 * 1e-3, nu 0.3)
                   + discrete bending (k 1e-4) on interior edges, noise +-1e-4 m. 10 iterations.
  cloth_config2()  the same construction at 512x512 (262,144 vertices, 522,242 triangles,
                   782,341 hinges), mass 1e-3 (32/512)^2, Y log-uniform in [1e3, 1e5], top row pinned.
                   20 iterations.
  cube_config3()   64^3 cells of side 1/64 m, 6 tets per cell (274,625 vertices, 1,572,864 tets), SNH
                   Y 1e5, nu 0.45, density 1000 lumped, bottom face pinned, rest scaled by 1.05 in y.
  cloth_batch()    256 cloths of side U{16..128}, Y log-U[1e3, 1e5], bending log-U[1e-5, 1e-3].
Pins are SpringAnchorEnergy (the reference's only pinning mechanism, springAnchorEnergy.cpp:21-39),
stiffness 1e5 for cloth (solverCompare.cpp:401-409) and 1e4 for tetrahedra (solverCompare.cpp:300-311),
anchored at the initial positions. Energy order: bending, membrane, anchor
(solverCompare.cpp:460,493,496); anchor, SNH for tetrahedra (solverCompare.cpp:326-327).
"""
from dataclasses import dataclass, field

import numpy as np

from . import native

DT = 1.0 / 60.0
GRAVITY = (0.0, -9.81, 0.0)
MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) -> 53-bit uniform doubles in [0, 1)."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniforms(self, n):
        """n uniforms as a float64 array (vectorised splitmix64 with the same sequence)."""
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.s) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.s = (self.s + n * 0x9E3779B97F4A7C15) & MASK64
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Scene:
    x: np.ndarray          # (n, 3) initial positions
    x0: np.ndarray         # (n, 3) rest positions
    m: np.ndarray          # (n,) masses
    energies: list = field(default_factory=list)  # [(kind, stencils, addStencil args)] in addEnergy order
    iterations: int = 1
    dt: float = DT
    gravity: tuple = GRAVITY
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.x.shape[0]

    def particle_group(self, prepare=True):
        """spg::ParticleGroup through the C++ mirror: addParticle, addStencil, preparePrecomputations."""
        pg = native.ParticleGroup()
        pg.add_particles(self.x, self.x0, self.m)
        for kind, st, args in self.energies:
            pg.add_energy(kind, st, args, prepare=prepare and kind in PREPARED_KINDS)
        return pg

    def vbd_object(self):
        """Prepared arrays (spg_vbd_object layout) computed by the C++ mirror's preparePrecomputations."""
        pg = self.particle_group()
        return {"x": self.x, "v": None, "m": self.m, "x0": self.x0,
                "energies": [(kind, st, pg.energy_params(i)) for i, (kind, st, _) in enumerate(self.energies)]}

    def xpbd_object(self):
        """Prepared arrays (spg_xpbd_object layout: effectiveCompliance() form, invMasses()) from the C++
        mirror."""
        pg = self.particle_group()
        return {"x": self.x, "v": None, "w": pg.inv_masses(),
                "energies": [(kind, st, pg.xpbd_params(i)) for i, (kind, st, _) in enumerate(self.energies)]}


# energies with a preparePrecomputations of their own (rest-state matrices, rest angles, volumes)
PREPARED_KINDS = ("stvk", "bending", "snh", "membrane_bw", "membrane_choi", "bending_quadratic", "bending_bw",
                  "stvk_tet", "ssnh")


def grid_cloth_mesh(n):
    """n x n vertices, 1 m square in the xy plane (z = 0), row-major ids (row = y), every quad split
    along the same diagonal into two counter-clockwise triangles."""
    s = 1.0 / (n - 1)
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    x0 = np.stack([j.ravel() * s, i.ravel() * s, np.zeros(n * n)], axis=1)
    a = (np.arange(n - 1)[:, None] * n + np.arange(n - 1)[None, :]).ravel()
    b, c, d = a + 1, a + n + 1, a + n
    faces = np.empty((2 * a.size, 3), dtype=np.int32)
    faces[0::2] = np.stack([a, b, c], axis=1)
    faces[1::2] = np.stack([a, c, d], axis=1)
    return x0, faces


def cloth(n, young, poisson=0.3, thickness=1e-3, bending=1e-4, mass=1e-3, pin="corners", noise=1e-4, seed=1,
          anchor_k=1e5, iterations=10, membrane="stvk", bending_kind="bending", edge_springs=None, edge_k=None,
          stretch=1.0):
    """membrane: "stvk" (MembraneStvkEnergy), "membrane_bw" or "membrane_choi" (kWeft = kWarp = Y t,
    kBias = Y t / (2 (1 + nu))); bending_kind: "bending", "bending_quadratic" or "bending_bw";
    edge_springs: None, or "spring_continuum" / "spring_squared_continuum" over the mesh edges (stiffness
    edge_k, default Y t; rest length from the rest positions), added after the membrane; stretch: the initial
    positions are the rest positions scaled by this factor in the cloth's plane (pre-strain), plus the noise."""
    x0, faces = grid_cloth_mesh(n)
    rng = SplitMix64(seed)
    x = x0 * np.array([stretch, stretch, 1.0]) + (2.0 * rng.uniforms(3 * n * n).reshape(-1, 3) - 1.0) * noise
    m = np.full(n * n, mass)
    hinges = native.trimesh_bending_stencils(n * n, faces)
    if pin == "corners":
        pins = np.array([(n - 1) * n, n * n - 1], dtype=np.int32)
    elif pin == "top_row":
        pins = np.arange((n - 1) * n, n * n, dtype=np.int32)
    else:
        raise ValueError(pin)
    yt = young * thickness
    if membrane == "stvk":
        mem = ("stvk", faces, np.tile([yt, poisson], (len(faces), 1)))
    else:
        mem = (membrane, faces, np.tile([yt, yt, yt / (2 * (1 + poisson))], (len(faces), 1)))
    energies = [
        (bending_kind, hinges, np.column_stack([np.zeros(len(hinges)), np.full(len(hinges), bending)])),
        mem,
    ]
    if edge_springs:
        edges = np.unique(np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), axis=1),
                          axis=0).astype(np.int32)
        L0 = np.linalg.norm(x0[edges[:, 0]] - x0[edges[:, 1]], axis=1)
        energies.append((edge_springs, edges, np.column_stack([L0, np.full(len(edges), yt if edge_k is None else edge_k)])))
    energies.append(("anchor", pins[:, None], np.column_stack([x[pins], np.full(len(pins), anchor_k)])))
    return Scene(x=x, x0=x0, m=m, energies=energies, iterations=iterations,
                 meta={"kind": "cloth", "n": n, "young": young, "bending": bending, "seed": seed,
                       "triangles": len(faces), "hinges": len(hinges), "pins": len(pins)})


def cloth_config1(seed=1):
    return cloth(32, young=1e4, poisson=0.3, thickness=1e-3, bending=1e-4, mass=1e-3, pin="corners", seed=seed,
                 iterations=10)


def config2_young(seed):
    return 10.0 ** (3.0 + 2.0 * SplitMix64(seed ^ 0x5EED).uniform())


def cloth_config2(seed=1, n=512, iterations=20):
    return cloth(n, young=config2_young(seed), poisson=0.3, thickness=1e-3, bending=1e-4,
                 mass=1e-3 * (32.0 / n) ** 2, pin="top_row", seed=seed, iterations=iterations)


def kuhn_cube_mesh(cells, side=1.0):
    """(cells+1)^3 vertices, id = (ix (cells+1) + iy)(cells+1) + iz; each cube cell split into the
    6 tetrahedra of the Kuhn triangulation along its main diagonal, positively oriented."""
    n = cells + 1
    s = side / cells
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x0 = np.stack([ix.ravel() * s, iy.ravel() * s, iz.ravel() * s], axis=1)
    c = np.arange(cells)
    cx, cy, cz = np.meshgrid(c, c, c, indexing="ij")
    base = ((cx * n + cy) * n + cz).ravel()
    step = np.array([n * n, n, 1])
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        v0 = base
        v1 = v0 + step[perm[0]]
        v2 = v1 + step[perm[1]]
        v3 = v2 + step[perm[2]]
        tets.append(np.stack([v0, v1, v2, v3], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    p = x0[tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    neg = vol < 0
    tets[neg, 0], tets[neg, 1] = tets[neg, 1].copy(), tets[neg, 0].copy()
    return x0, tets


def cube(cells, young=1e5, poisson=0.45, density=1000.0, anchor_k=1e4, stretch_y=1.05, iterations=20, side=1.0,
         material="snh"):
    """material: "snh" (StableNeoHookeanEnergy), "stvk_tet" (StvkEnergy) or "ssnh" (StableSquaredNeoHookeanEnergy)."""
    x0, tets = kuhn_cube_mesh(cells, side)
    p = x0[tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6.0
    m = np.zeros(x0.shape[0])
    np.add.at(m, tets.ravel(), np.repeat(vol * density / 4.0, 4))
    x = x0.copy()
    x[:, 1] *= stretch_y
    pins = np.flatnonzero(x0[:, 1] == 0.0).astype(np.int32)
    energies = [
        ("anchor", pins[:, None], np.column_stack([x[pins], np.full(len(pins), anchor_k)])),
        (material, tets, np.tile([young, poisson], (len(tets), 1))),
    ]
    return Scene(x=x, x0=x0, m=m, energies=energies, iterations=iterations,
                 meta={"kind": "cube", "cells": cells, "tets": len(tets), "pins": len(pins)})


def cube_config3():
    return cube(64)


def cloth_batch(count=256, seed=7, lo=16, hi=128, iterations=15):
    """Independent cloths of uneven size (config-4 style): side U{lo..hi}, Y log-U[1e3, 1e5],
    bending log-U[1e-5, 1e-3], top row pinned, mass 1e-3 (32/side)^2 per vertex."""
    rng = SplitMix64(seed)
    scenes = []
    for k in range(count):
        side = lo + int(rng.uniform() * (hi - lo + 1))
        young = 10.0 ** (3.0 + 2.0 * rng.uniform())
        bend = 10.0 ** (-5.0 + 2.0 * rng.uniform())
        scenes.append(cloth(side, young=young, poisson=0.3, thickness=1e-3, bending=bend,
                            mass=1e-3 * (32.0 / side) ** 2, pin="top_row", seed=seed * 1000 + k,
                            iterations=iterations))
    return scenes


# ---- rigid bodies ----

@dataclass
class RigidScene:
    """A RigidBodyGroup: bodies (centre of mass, axis-angle orientation, mass, local inertia) and its
    energies [(kind, stencils, addStencil args)] in addEnergy order."""
    x: np.ndarray          # (nb, 3)
    theta: np.ndarray      # (nb, 3)
    m: np.ndarray          # (nb,)
    inertia: np.ndarray    # (nb, 3, 3) local inertia tensors
    energies: list = field(default_factory=list)
    iterations: int = 10
    nsubsteps: int = 1
    dt: float = DT
    gravity: tuple = GRAVITY
    meta: dict = field(default_factory=dict)

    @property
    def nb(self):
        return self.x.shape[0]


def box_inertia(mass, width, height, depth):
    """Solid box about its centre (apps/solverCompare.cpp:540-543)."""
    return np.diag([mass * (height ** 2 + depth ** 2), mass * (width ** 2 + depth ** 2),
                    mass * (width ** 2 + height ** 2)]) / 12.0


def rb_anchored(mass=1.0, w=1.0, h=1.0, d=1.0, iterations=10):
    """createAnchoredRigidBody (apps/solverCompare.cpp:524-555): one box held by three anchor springs."""
    raw = np.array([[0, h * 0.5, 0, 0, h * 0.5, -1, 100], [w * 0.5, -h * 0.5, 0, w, -h * 0.5, 1, 100],
                    [-w * 0.5, -h * 0.5, 0, -w, -h * 0.5, 1, 100]], dtype=np.float64)
    return RigidScene(x=np.zeros((1, 3)), theta=np.zeros((1, 3)), m=np.array([mass]),
                      inertia=box_inertia(mass, w, h, d)[None], iterations=iterations,
                      energies=[("rb_anchor", np.zeros((3, 1), np.int32), raw)], meta={"kind": "rb_anchored"})


def rb_articulated(n, mass=1.0, w=1.0, h=1.0, d=1.0, iterations=10):
    """createArticulatedRigidBody (apps/solverCompare.cpp:557-603): a row of boxes, the first held by three
    stiff anchors, each next one tied to its predecessor by a zero-length squared spring between facing
    face centres and an orientation alignment."""
    x = np.column_stack([w * np.arange(n), np.zeros(n), np.zeros(n)]).astype(np.float64)
    anchors = np.array([[0, h * 0.5, 0, 0, h * 0.5, 0, 1e10], [0, -h * 0.5, 0, 0, -h * 0.5, 0, 1e10],
                        [-w * 0.5, 0, 0, -w * 0.5, 0, 0, 1e10]], dtype=np.float64)
    energies = [("rb_anchor", np.zeros((3, 1), np.int32), anchors)]
    if n > 1:
        pairs = np.column_stack([np.arange(n - 1), np.arange(1, n)]).astype(np.int32)
        sp = np.tile([w * 0.5, 0, 0, -w * 0.5, 0, 0, 0.0, 1e10], (n - 1, 1))
        energies += [("rb_spring_squared", pairs, sp), ("rb_alignment", pairs, np.full((n - 1, 1), 1e5))]
    return RigidScene(x=x, theta=np.zeros((n, 3)), m=np.full(n, mass),
                      inertia=np.repeat(box_inertia(mass, w, h, d)[None], n, 0), energies=energies,
                      iterations=iterations, meta={"kind": "rb_articulated", "bodies": n})


def rb_chain(n, mass=1.0, w=1.0, h=1.0, d=1.0, iterations=10):
    """createRigidBodyChain (apps/solverCompare.cpp:605-654): boxes on a diagonal, the first anchored at
    a corner, consecutive ones joined corner to corner by springs of rest length sqrt(2 w)."""
    i = np.arange(n, dtype=np.float64)
    x = np.column_stack([-(h + 1) * i, np.zeros(n), -(h + 1) * i])
    energies = [("rb_anchor", np.zeros((1, 1), np.int32),
                 np.array([[w * 0.5, h * 0.5, 0, w * 0.5, h * 0.5, 0, 100]], dtype=np.float64))]
    if n > 1:
        pairs = np.column_stack([np.arange(n - 1), np.arange(1, n)]).astype(np.int32)
        sp = np.tile([-w * 0.5, -h * 0.5, 0, w * 0.5, h * 0.5, 0, np.sqrt(w * 2), 100], (n - 1, 1))
        energies.append(("rb_spring", pairs, sp))
    return RigidScene(x=x, theta=np.zeros((n, 3)), m=np.full(n, mass),
                      inertia=np.repeat(box_inertia(mass, w, h, d)[None], n, 0), energies=energies,
                      iterations=iterations, meta={"kind": "rb_chain", "bodies": n})


def rb_lattice(n, spacing=1.5, size=1.0, mass=1.0, k=1e3, k_align=10.0, tilt=0.05, seed=3, iterations=10):
    """An n x n horizontal lattice of boxes (the rigid-body throughput workload): neighbours in x and z
    joined by springs between facing face centres (rest length spacing - size) and by orientation
    alignments, the row z = 0 held by anchor springs at its top face centres, every body tilted by a
    random rotation of angle up to `tilt`. Energy order: anchors, springs, alignments."""
    rng = SplitMix64(seed)
    ix, iz = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ix, iz = ix.ravel(), iz.ravel()
    nb = n * n
    x = np.column_stack([spacing * ix, np.zeros(nb), spacing * iz]).astype(np.float64)
    u = rng.uniforms(3 * nb).reshape(nb, 3) * 2.0 - 1.0
    theta = u * (tilt / np.sqrt(3.0))
    hs = size * 0.5
    row = np.flatnonzero(iz == 0).astype(np.int32)
    anchors = np.column_stack([np.tile([0, hs, 0], (len(row), 1)), x[row] + [0, hs, 0], np.full(len(row), k * 10)])
    idx = np.arange(nb).reshape(n, n)
    px = np.column_stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()]).astype(np.int32)
    pz = np.column_stack([idx[:-1, :].ravel(), idx[1:, :].ravel()]).astype(np.int32)
    rest = spacing - size
    spx = np.tile([hs, 0, 0, -hs, 0, 0, rest, k], (len(px), 1))
    spz = np.tile([0, 0, hs, 0, 0, -hs, rest, k], (len(pz), 1))
    pairs = np.concatenate([px, pz])
    energies = [("rb_anchor", row[:, None], anchors), ("rb_spring", pairs, np.concatenate([spx, spz])),
                ("rb_alignment", pairs, np.full((len(pairs), 1), k_align))]
    return RigidScene(x=x, theta=theta, m=np.full(nb, mass),
                      inertia=np.repeat(box_inertia(mass, size, size, size)[None], nb, 0), energies=energies,
                      iterations=iterations,
                      meta={"kind": "rb_lattice", "n": n, "bodies": nb, "springs": len(pairs),
                            "alignments": len(pairs), "anchors": len(row)})
