"""CPU-only checks of the native libraries (no GPU compute calls):
  * both shared libraries load and export every function declared in include/*.h;
  * XPBD host logic: stencil colouring (bit-exact against the oracle), the compliance-form parameters
    of the C++ mirror, and the projection closed forms instantiated on the host against the oracle's
    dual-number projection;
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
                                        ("spg_xpbd.h", native.VBD_LIB), ("spg_implicit.h", native.VBD_LIB),
                                        ("spg_rigid.h", native.VBD_LIB),
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


# ------------------------------------------------------------------------------------------------
# XPBD host logic
# ------------------------------------------------------------------------------------------------

def test_xpbd_stencil_coloring_matches_oracle():
    for s in (scenes.cloth_config1(), scenes.cloth_config2(seed=1), scenes.cube(8)):
        for kind, st, _ in s.energies:
            a = native.xpbd_color_stencils(st)
            b, nc = ol.color_stencils(st)
            assert np.array_equal(a, b), kind
    # 512^2 cloth: class sizes of the greedy stencil colouring
    s = scenes.cloth_config2(seed=1)
    sizes = {k: np.bincount(native.xpbd_color_stencils(st)).tolist() for k, st, _ in s.energies}
    assert sizes["stvk"] == [87296, 87169, 87039, 86955, 86828, 86701, 254]
    assert sizes["bending"] == [65536] + [65280] * 4 + [65025] * 7 + [255, 255]
    assert sizes["anchor"] == [512]


def test_xpbd_stencil_coloring_random_matches_oracle():
    rng = np.random.default_rng(11)
    for ss in (1, 2, 3, 4):
        st = np.array([rng.choice(300, size=ss, replace=False) for _ in range(900)], dtype=np.int32)
        assert np.array_equal(native.xpbd_color_stencils(st), ol.color_stencils(st)[0])


def test_xpbd_stencil_coloring_bad_index():
    with pytest.raises(native.NativeError, match="invalid node indices"):
        native.xpbd_color_stencils(np.array([[0, -1]], dtype=np.int32))


@pytest.mark.parametrize("make", [scenes.cloth_config1, lambda: scenes.cube(6)], ids=["cloth", "cube"])
def test_xpbd_params_match_oracle(make):
    """effectiveCompliance() form of the C++ mirror against the oracle's compliance (computed from the raw
    addStencil arguments), and the inverse masses."""
    s = make()
    pg = s.particle_group()
    o = ol.from_scene(s)
    assert np.array_equal(pg.inv_masses(), native.inverse_masses(s.m))
    for i, (kind, st, _) in enumerate(s.energies):
        xp = pg.xpbd_params(i)
        cm = o.compliance(i)
        pr = o.prepared(i)
        t = native.KIND[kind]
        if t == native.MEMBRANE_STVK:
            got_c, got_p = xp[:, 4:], xp[:, :4]
            ref_p = pr[:, :4]
        elif t == native.STABLE_NEO_HOOKEAN:
            got_c, got_p = xp[:, 9:13], np.column_stack([xp[:, :9], xp[:, 13]])
            ref_p = np.column_stack([pr[:, :9], pr[:, 11]])
        elif t == native.SPRING_ANCHOR:
            got_c, got_p, ref_p = xp[:, 3:], xp[:, :3], pr[:, :3]
        else:
            got_c, got_p, ref_p = xp[:, 1:], xp[:, :1], pr[:, :1]
        assert np.allclose(got_c, cm, rtol=1e-14, atol=0), kind
        assert np.array_equal(got_p, ref_p) or np.allclose(got_p, ref_p, rtol=1e-15), kind


def _xpbd_layout(kind, p, cm):
    """Oracle prepared params + compliance -> spg_xpbd_energy params of one stencil."""
    if kind == "spring":
        return np.array([p[0], cm[0]])
    if kind == "anchor":
        return np.array([*p[:3], cm[0]])
    if kind == "stvk":
        return np.concatenate([p[:4], cm])
    if kind == "bending":
        return np.array([p[0], cm[0]])
    return np.concatenate([p[:9], cm, [p[11]]])


def _projection_cases():
    import test_xpbd_oracle as T
    rng = np.random.default_rng(5)
    out = []
    for kind, xs, p in T.CASES:
        ns = xs.shape[0]
        w = rng.uniform(0.5, 2.0, size=ns)
        c, J = ol.constraints(kind, xs.ravel(), p)
        for scale in (0.0, 1.0, 1e3):  # stiff, balanced and soft compliance
            base = float(np.mean(np.sum(J * J, axis=0) * np.mean(w))) * (1 / 60) ** 2
            out.append((kind, xs, w, p, T._compliance(kind, scale * base)))
    # SNH near rest (F ~ I), where dC0/dx and dC1/dx are nearly parallel and only alpha~ keeps A regular:
    # a 2 cm tet with Y = 1e5 (det(A) / A00 A11 ~ 2.2e-2: FP32 path), 1e7 and 1e8 (~3e-3, 5e-4: the kernels'
    # FP64 fallback runs), nu = 0.45
    x0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.02
    D = np.linalg.inv((x0[1:] - x0[0]).T)
    vol = 0.02 ** 3 / 6
    for young in (1e5, 1e7, 1e8):
        mu = young / 2.9
        lam = mu / (1 - 0.9)
        p = np.concatenate([D.ravel(), [lam * vol, mu * vol, 1 + mu / lam]])
        cm = np.array([1 / lam, 0, 0, 1 / mu]) / vol
        out.append(("snh", x0 + 2e-4 * rng.normal(size=(4, 3)), np.full(4, 1 / (1000 * vol)), p, cm))
    return out


PROJ = _projection_cases()


@pytest.mark.parametrize("case", range(len(PROJ)))
def test_xpbd_closed_forms_on_host_match_oracle(case):
    """Closed-form projections (xpbd_element.h, host instantiation) against the oracle's generic projection
    with dual-number constraint gradients: FP64 within 1e-12 of the displacement, FP32 (the kernels' path
    with its FP64 fallback) within 1e-5."""
    kind, xs, w, p, cm = PROJ[case]
    dt = 1 / 60
    ref = ol.xpbd_project(kind, xs.ravel(), w, p, cm, dt)
    xp = _xpbd_layout(kind, p, cm)
    scale = max(np.abs(ref).max(), 1e-300)
    for prec, tol in ((1, 1e-12), (0, 1e-5)):
        got = native.xpbd_project_stencils(kind, xs.ravel()[None], w[None], xp[None], dt, precision=prec,
                                           device=-1)[0]
        assert np.abs(got - ref).max() <= tol * scale, (kind, prec, got, ref)


def test_xpbd_closed_form_guards_on_host():
    dt = 0.01
    # spring with both ends immovable, and with coincident ends: no correction
    z = native.xpbd_project_stencils("spring", [[0, 0, 0, 1, 0, 0]], [[0, 0]], [[0.5, 0.0]], dt, 0, -1)
    assert not np.any(z)
    z = native.xpbd_project_stencils("spring", [[0, 0, 0, 0, 0, 0]], [[1, 1]], [[0.5, 0.0]], dt, 0, -1)
    assert not np.any(z)
    # anchor exactly at its anchor point
    z = native.xpbd_project_stencils("anchor", [[1, 2, 3]], [[1.0]], [[1, 2, 3, 1e-3]], dt, 0, -1)
    assert not np.any(z)
    # zero-mass vertices are never moved
    import test_xpbd_oracle as T
    for kind, xs, p in T.CASES[:5]:
        ns = xs.shape[0]
        w = np.ones(ns)
        w[0] = 0.0
        nc = ol.num_constraints(kind)
        xp = _xpbd_layout(kind, p, T._compliance(kind, 1e-3))
        d = native.xpbd_project_stencils(kind, xs.ravel()[None], w[None], xp[None], dt, 0, -1)[0]
        assert not np.any(d[0]), kind


def _element_cases():
    import test_energies_oracle as T
    out = []
    for kind in T.KINDS:
        p = T.prepared(kind)
        for seed, scale in ((0, 0.03), (1, 0.3)):
            out.append((kind, T.deformed(kind, scale=scale, seed=seed), p))
    return out


ELEMENTS = _element_cases()


@pytest.mark.parametrize("case", range(len(ELEMENTS)))
def test_implicit_closed_forms_on_host_match_oracle(case):
    """Full element gradients (implicit_element.h closed forms, host instantiation) and their forward-mode
    Hessians against the oracle's second-order dual numbers, for the further energies: gradient 1e-12 and
    Hessian 1e-10 relative to their max-norms."""
    kind, x, p = ELEMENTS[case]
    g, h = native.implicit_element_derivatives(kind, x.ravel()[None], p[None], device=-1)
    _, gr, _, hr = ol.element(kind, x.ravel(), p, mode=ol.AD, full_hess=True)
    assert np.abs(g[0] - gr).max() <= 1e-12 * np.abs(gr).max(), kind
    assert np.abs(h[0] - hr).max() <= 1e-10 * np.abs(hr).max(), kind


@pytest.mark.parametrize("kind", ["spring_continuum", "spring_squared_continuum", "membrane_bw", "membrane_choi",
                                  "bending_quadratic", "bending_bw", "stvk_tet", "ssnh"])
def test_mirror_further_energies_prepare_like_oracle(kind):
    """The C++ mirror's addStencil + preparePrecomputations of the further energies give the oracle's prepared
    parameters (spg_vbd_energy layout) and its compliance diagonal (spg_xpbd_energy layout) exactly."""
    import test_energies_oracle as T
    x0 = T.REST[kind]
    pg = native.ParticleGroup()
    pg.add_particles(x0, x0, np.ones(len(x0)))
    pg.add_energy(kind, np.arange(len(x0))[None], np.asarray(T.RAW[kind])[None], prepare=True)
    po = T.prepared(kind)
    assert np.array_equal(pg.energy_params(0)[0], po)
    o = ol.Oracle(x0, x0, np.ones(len(x0)))
    o.add_energy(kind, [np.arange(len(x0))], [T.RAW[kind]])
    o.prepare()
    nc = ol.num_constraints(kind)
    cm = o.compliance(0)[0].reshape(nc, nc)
    xp = pg.xpbd_params(0)[0]
    t = native.KIND[kind]
    n_geo = native.XPBD_PARAMS[t] - nc if t not in (native.QUADRATIC_BENDING,) else 4
    assert np.array_equal(xp[:n_geo], po[:n_geo])
    diag = xp[n_geo:]
    assert np.allclose(diag, np.diag(cm)[:len(diag)], rtol=1e-15, atol=0)


@pytest.mark.parametrize("kind,raw,msg", [
    ("spring_continuum", [0.0, 1.0], "Zero/negative rest length invalid"),
    ("bending_quadratic", [0.2, 1.0], "Rest theta != 0 not supported for quadratic bending"),
    ("bending_bw", [-3.2, 1.0], "Invalid restTheta"),
])
def test_mirror_further_energies_validation(kind, raw, msg):
    import test_energies_oracle as T
    x0 = T.REST[kind]
    pg = native.ParticleGroup()
    pg.add_particles(x0, x0, np.ones(len(x0)))
    with pytest.raises(native.NativeError, match=msg):
        pg.add_energy(kind, np.arange(len(x0))[None], np.asarray(raw)[None], prepare=True)


def _further_xpbd_cases():
    import test_energies_oracle as T
    out = []
    for kind in T.KINDS:
        x0 = T.REST[kind]
        o = ol.Oracle(x0, x0, np.ones(len(x0)))
        o.add_energy(kind, [np.arange(len(x0))], [T.RAW[kind]])
        o.prepare()
        p = o.prepared(0)[0]
        nc = ol.num_constraints(kind)
        cm = o.compliance(0)[0].reshape(nc, nc)
        t = native.KIND[kind]
        ngeo = 4 if t == native.QUADRATIC_BENDING else native.XPBD_PARAMS[t] - nc
        diag = np.diag(cm)[:native.XPBD_PARAMS[t] - ngeo]
        for seed, (scale, cscale) in enumerate(((0.05, 1.0), (0.3, 1e-3), (0.3, 1e3))):
            x = T.deformed(kind, scale=scale, seed=seed)
            out.append((kind, x, p, np.diag(np.diag(cm) * cscale), np.concatenate([p[:ngeo], diag * cscale])))
    return out


FURTHER_XPBD = _further_xpbd_cases()


@pytest.mark.parametrize("case", range(len(FURTHER_XPBD)))
def test_xpbd_further_projection_on_host_matches_oracle(case):
    """The generic projection of the further energies (xpbd_element.h, host instantiation) against the oracle's
    projection with dual-number constraint gradients: FP64 within 1e-10 of the displacement; FP32 (with its
    FP64 recomputation) within 1e-5 of the displacement plus 1e-6 of the positions, since a nearly satisfied
    constraint such as |d| / L0 - 1 loses digits to cancellation when it is evaluated from FP32 positions."""
    kind, x, p, cm, xp = FURTHER_XPBD[case]
    dt = 1 / 60
    w = np.linspace(0.5, 2.0, len(x))
    ref = ol.xpbd_project(kind, x.ravel(), w, p, cm, dt)
    scale = max(np.abs(ref).max(), 1e-300)
    for prec, tol in ((1, 1e-10 * scale), (0, 1e-5 * scale + 1e-6 * np.abs(x).max())):
        got = native.xpbd_project_stencils(kind, x.ravel()[None], w[None], xp[None], dt, precision=prec, device=-1)[0]
        assert np.abs(got - ref).max() <= tol, (kind, prec, np.abs(got - ref).max() / scale)
