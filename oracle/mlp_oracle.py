"""CPU restatement of the learned warm start (test infrastructure only: imported by tests/ and by
nothing that ships; the product path is the HIP code in spg_amd/csrc/mlp_*.hip and
warmstart_kernels.hip).

The reference has no learned component (spg's VBD starts every substep from the inertial guess,
spg/source/sim/solver/vbd.cpp:58-83), so nothing here restates reference code. What pins it:
  * forward / loss / gradient: torch FP64 autograd of the same network and central finite
    differences (tests/test_mlp_oracle.py);
  * feature rows: computed from the scene's raw material parameters (Young's modulus, bending
    stiffness) and plain numpy geometry, independently of the solver's recovery of those values
    from the prepared per-element parameters.

Network: 24 -> 128 -> 128 -> 128 -> 3 with SiLU after each hidden layer. Flat parameter layout:
W1 (128x24, out x in), b1, W2 (128x128), b2, W3, b3, W4 (3x128), b4. Inputs are standardised with
(f - mean) / std. Loss L = (1/n) sum_rows sum_k (y_k - t_k)^2.

With emulate=True the values are rounded to bfloat16 (round to nearest even) where the GPU kernels
round them: standardised inputs, weights, hidden activations, and the output / hidden gradients that
feed the gradient products; sums stay in FP64 (the GPU accumulates in FP32).
"""
import numpy as np

SHAPES = [(128, 24), (128, 128), (128, 128), (3, 128)]
NUM_PARAMS = sum(o * i + o for o, i in SHAPES)  # 36611


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p):
    p = np.asarray(p, dtype=np.float64)
    out, k = [], 0
    for o, i in SHAPES:
        w = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((w, b))
    assert k == p.size == NUM_PARAMS
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _inputs(features, mean, std, emulate):
    f = np.asarray(features, dtype=np.float64)
    if emulate:  # the GPU standardises in FP32 with a precomputed reciprocal
        inv = (np.float32(1.0) / np.asarray(std, dtype=np.float32)).astype(np.float32)
        x = ((np.asarray(features, dtype=np.float32) - np.asarray(mean, dtype=np.float32)) * inv).astype(np.float32)
        return bf16(x)
    return (f - np.asarray(mean, dtype=np.float64)) / np.asarray(std, dtype=np.float64)


def forward(params, features, mean=None, std=None, emulate=False, keep=False):
    """Outputs (n, 3); with keep=True also the per-layer (input, pre-activation) pairs."""
    mean = np.zeros(24) if mean is None else mean
    std = np.ones(24) if std is None else std
    r = bf16 if emulate else (lambda a: a)
    layers = unflatten(params)
    h = _inputs(features, mean, std, emulate)
    saved = []
    for l, (w, b) in enumerate(layers):
        z = h @ r(w).T + b
        saved.append((h, z))
        h = r(silu(z)) if l < 3 else z
    return (h, saved) if keep else h


def loss_grad(params, features, targets, mean=None, std=None, emulate=False):
    """Loss and its gradient with respect to the flat parameters."""
    r = bf16 if emulate else (lambda a: a)
    layers = unflatten(params)
    y, saved = forward(params, features, mean, std, emulate, keep=True)
    t = np.asarray(targets, dtype=np.float64)
    n = y.shape[0]
    d = y - t
    loss = float(np.sum(d * d) / n)
    dz = 2.0 * d / n                  # dL/dY
    grads = [None] * 4
    dz_op = r(dz)                     # operand of the gradient products
    for l in range(3, -1, -1):
        h_in, _ = saved[l]
        grads[l] = (dz_op.T @ h_in, dz.sum(axis=0))
        if l == 0:
            break
        w = r(layers[l][0])
        dz = (dz_op @ w) * silu_grad(saved[l - 1][1])
        dz_op = r(dz)
    return loss, flatten(grads)


# ------------------------------------------------------------------------------------------------
# Feature rows (the layout the solver's feature kernel writes)
# ------------------------------------------------------------------------------------------------

def static_tables(scene):
    """One-ring neighbours (sorted, unique) with rest lengths, and per-vertex static features
    (log10 Young's modulus of incident membrane / solid / spring elements, log10 bending stiffness of
    incident hinges, anchored flag), from the scene's raw addStencil arguments."""
    n = scene.n
    x0 = np.asarray(scene.x0, dtype=np.float64)
    nbrs = [set() for _ in range(n)]
    sy, cy = np.zeros(n), np.zeros(n)
    sb, cb = np.zeros(n), np.zeros(n)
    anchored = np.zeros(n)
    size = {"spring": 2, "stvk": 3, "bending": 4, "snh": 4, "anchor": 1}
    for kind, st, args in scene.energies:
        st = np.asarray(st, dtype=np.int64).reshape(-1, size[kind])
        if st.shape[0] == 0:
            continue
        args = np.asarray(args, dtype=np.float64).reshape(st.shape[0], -1)
        if kind in ("stvk", "snh", "spring"):
            logy = np.log10(args[:, 0] if kind != "spring" else args[:, 1])
            for q in range(st.shape[1]):
                np.add.at(sy, st[:, q], logy)
                np.add.at(cy, st[:, q], 1.0)
            for e in st:
                for a in e:
                    for b in e:
                        if a != b:
                            nbrs[a].add(int(b))
        elif kind == "bending":
            logk = np.log10(args[:, 1])
            for q in range(4):
                np.add.at(sb, st[:, q], logk)
                np.add.at(cb, st[:, q], 1.0)
        elif kind == "anchor":
            anchored[st[:, 0]] = 1.0
    nb = [np.array(sorted(s), dtype=np.int64) for s in nbrs]
    rest = [np.linalg.norm(x0[j] - x0[i], axis=1) if len(j) else np.zeros(0) for i, j in enumerate(nb)]
    stat = np.stack([np.where(cy > 0, sy / np.maximum(cy, 1), 0.0), np.where(cb > 0, sb / np.maximum(cb, 1), 0.0),
                     anchored], axis=1)
    return nb, rest, stat


def features(x_tilde, v_kicked, corr, tables, h, g):
    """(n, 24) feature rows from the substep's inertial target, the velocity after the gravity kick,
    and the previous substep's x* - x~ (see include/spg_vbd_learned.h for the layout)."""
    nb, rest, stat = tables
    g = np.asarray(g, dtype=np.float64)
    a = np.linalg.norm(g)
    a = a if a > 0 else 1.0
    s, u = h * h * a, h * a
    v = np.asarray(v_kicked, dtype=np.float64) - h * g
    c = np.asarray(corr, dtype=np.float64)
    xt = np.asarray(x_tilde, dtype=np.float64)
    n = xt.shape[0]
    out = np.zeros((n, 24))
    for i in range(n):
        j = nb[i]
        out[i, 0:3] = c[i] / s
        out[i, 3:6] = v[i] / u
        if len(j):
            out[i, 6:9] = c[j].mean(axis=0) / s
            out[i, 9:12] = np.abs(c[j] - c[i]).max(axis=0) / s
            out[i, 12:15] = v[j].mean(axis=0) / u
            out[i, 15:18] = np.abs(v[j] - v[i]).max(axis=0) / u
            strain = np.linalg.norm(xt[j] - xt[i], axis=1) / rest[i] - 1.0
            out[i, 21:24] = strain.mean(), strain.max(), strain.min()
        out[i, 18:21] = stat[i]
    return out
