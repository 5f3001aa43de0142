"""CPU-only checks of the native libraries (no GPU compute calls):
  * both shared libraries load and export every function declared in include/*.h;
  * host-side logic against the oracle: greedy colouring (bit-exact, incl. the known class sizes of
    the 512^2 cloth and the 64^3 Kuhn cube), TriangleMesh bending stencils, preparePrecomputations;
  * the reference's setup errors surface with the reference's messages.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
from spg_amd import native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spg_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("spg_vbd.h", native.VBD_LIB), ("spg_vbd_learned.h", native.VBD_LIB),
                                        ("spg_host.h", native.HOST_LIB)])
def test_exports(header, lib):
    native.host_lib()
    L = ctypes.CDLL(lib)
    names = _declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_coloring_matches_oracle_config1_and_config2():
    for s in (scenes.cloth_config1(), scenes.cloth_config2(seed=1)):
        c = native.color_vertices(s.n, [(st, st.shape[1]) for _, st, _ in s.energies])
        o = ol.from_scene(s)
        co, nc = o.colors()
        assert np.array_equal(c, co)
        assert o.coloring_valid(c)
    assert np.bincount(c).tolist() == [65536] * 4


def test_coloring_cube_known_sizes():
    s = scenes.cube_config3()
    c = native.color_vertices(s.n, [(st, st.shape[1]) for _, st, _ in s.energies])
    assert np.bincount(c).tolist() == [35937, 51520, 39813, 48895, 35925, 28263, 17136, 17136]


def test_coloring_random_graph_matches_oracle():
    rng = np.random.default_rng(2)
    nv = 500
    tets = rng.integers(0, nv, size=(700, 4)).astype(np.int32)
    ok = np.array([len(set(t)) == 4 for t in tets])
    tets = tets[ok]
    x = rng.uniform(size=(nv, 3))
    o = ol.Oracle(x, x, np.ones(nv))
    o.add_energy("snh", tets, np.tile([1e3, 0.3], (len(tets), 1)))
    co, _ = o.colors()
    c = native.color_vertices(nv, [(tets, 4)])
    assert np.array_equal(c, co)


def test_bending_stencils_match_oracle():
    for n in (5, 32, 77):
        x0, faces = scenes.grid_cloth_mesh(n)
        assert np.array_equal(native.trimesh_bending_stencils(n * n, faces), ol.bending_stencils(faces, n * n))
    rng = np.random.default_rng(4)
    faces = rng.integers(0, 40, size=(120, 3)).astype(np.int32)
    faces = faces[(faces[:, 0] != faces[:, 1]) & (faces[:, 1] != faces[:, 2]) & (faces[:, 0] != faces[:, 2])]
    assert np.array_equal(native.trimesh_bending_stencils(40, faces), ol.bending_stencils(faces, 40))


@pytest.mark.parametrize("make", [scenes.cloth_config1, lambda: scenes.cube(6)], ids=["cloth", "cube"])
def test_prepared_data_matches_oracle(make):
    s = make()
    pg = s.particle_group()
    o = ol.from_scene(s)
    for i in range(len(s.energies)):
        a, b = pg.energy_params(i), o.prepared(i)
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))


def test_reference_setup_errors():
    pg = native.ParticleGroup()
    pg.add_particles(np.zeros((3, 3)), np.array([[0, 0, 0], [0, 1.0, 0], [1.0, 0, 0]]), np.ones(3))
    with pytest.raises(native.NativeError, match="Zero/negative rest length invalid"):
        pg.add_energy("spring", [[0, 1]], [[0.0, 1.0]])
    with pytest.raises(native.NativeError, match="Rest area of membrane StVK triangle is not positive"):
        pg.add_energy("stvk", [[0, 1, 2]], [[1.0, 0.3]])
    pg2 = native.ParticleGroup()
    pg2.add_particles(np.zeros((3, 3)), np.array([[0, 0, 0.5], [1.0, 0, 0], [0, 1.0, 0]]), np.ones(3))
    with pytest.raises(native.NativeError, match="must have coordinate z=0"):
        pg2.add_energy("stvk", [[0, 1, 2]], [[1.0, 0.3]])
    pg3 = native.ParticleGroup()
    pg3.add_particles(np.zeros((4, 3)), np.zeros((4, 3)), np.ones(4))
    with pytest.raises(native.NativeError, match="Invalid restTheta"):
        pg3.add_energy("bending", [[0, 1, 2, 3]], [[3.5, 1.0]])
    idx = pg3.add_energy("snh", [[0, 1, 2, 3]], [[1.0, 0.3]], prepare=False)
    with pytest.raises(native.NativeError, match="SNH tetrahedron rest volume is not positive"):
        pg3.prepare_energy(idx)
    with pytest.raises(native.NativeError, match="Triangle mesh face contains repeated vertex index"):
        native.trimesh_bending_stencils(4, [[0, 1, 1]])
    with pytest.raises(native.NativeError, match="Triangle mesh face contains index to invalid vertex"):
        native.trimesh_bending_stencils(4, [[0, 1, 9]])


def test_double_prepare_throws():
    s = scenes.cloth(6, young=1e3)
    pg = s.particle_group()
    with pytest.raises(native.NativeError, match="Energy was already prepared"):
        pg.prepare_energy(0)


def test_coloring_bad_index_message():
    with pytest.raises(native.NativeError, match="invalid node indices"):
        native.color_vertices(3, [(np.array([0, 5], dtype=np.int32), 2)])


def _golden(kind):
    import json
    G = json.load(open(os.path.join(ROOT, "tests", "golden", "element_golden.json")))["cases"]
    cs = [c for c in G if c["kind"] == kind]
    ns = native.STENCIL_SIZE[native.KIND[kind]]
    xs = np.array([c["x"] for c in cs])
    prm = np.array([[float(v) for v in c["prepared"]] for c in cs])
    g = np.array([[float(v) for v in c["grad"]] for c in cs]).reshape(len(cs), ns, 3)
    H = np.array([[[float(v) for v in r] for r in c["hess"]] for c in cs])
    hb = np.array([[H[k, 3 * s:3 * s + 3, 3 * s:3 * s + 3] for s in range(ns)] for k in range(len(cs))])
    return xs, prm, g, hb


@pytest.mark.parametrize("kind", ["spring", "anchor", "stvk", "bending", "snh"])
@pytest.mark.parametrize("form", [0, 1, 2], ids=["solve-form", "general-form", "unified-form"])
def test_device_closed_forms_on_host(kind, form):
    """The solver's closed forms (element.h), instantiated on the host, against the 40-digit golden
    derivatives: FP64 within 1e-12 relative, FP32 within 2e-5 relative (per element, max-norm)."""
    xs, prm, g_ref, h_ref = _golden(kind)
    for prec, tol in ((1, 1e-12), (0, 2e-5)):
        g, h = native.element_derivatives(kind, xs, prm, precision=prec, device=-1, form=form)
        for k in range(len(xs)):
            sh = np.max(np.abs(h_ref[k]))
            sg = np.max(np.abs(g_ref[k]))
            assert np.max(np.abs(h[k] - h_ref[k])) <= tol * sh, (kind, k, prec)
            if sg > 1e-9 * sh:
                assert np.max(np.abs(g[k] - g_ref[k])) <= tol * sg, (kind, k, prec)
            else:
                assert np.max(np.abs(g[k])) <= 1e-6 * sh
