"""FP64 CPU oracle of the learned RF surrogate (a path with no counterpart in the reference).

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU leg may import or run
anything under ``oracle/``, and only as the checker: the package ``blochparty_amd`` never imports it.

The surrogate replaces the per-spin composition of a reused pulse's substeps (the reference's ``pulse_maps``,
src/blochparty/engines/gpu/cuda/pulse.cu.in:3-32) by a network. A pulse of K substeps is cut into M blocks; each
block's pure rotation (no relaxation) is a unit quaternion, taken in the frame that turns with the off-resonance
precession (the rotation about z by the block's free-precession angle theta = -2 pi f tau, half before and half
after, is divided out), so that it varies slowly with the offset f. The network maps features of (f, b) (the spin's
offset during the pulse and its B1 scale times the pulse's peak) -> the M block quaternions; the relaxation is put
back exactly, half a block before and after each block's rotation (a Strang split of the pulse into M blocks). What
each function computes:

- ``substep_quat``  the rotation of one substep, w = -2 pi dt (b1 rf_x, b1 rf_y, f) (as core.rf_segment's rotation)
- ``block_quats``   the M interaction-frame block quaternions of a pulse at offsets f and B1 scales b1
- ``block_map``     the affine map m -> m + D m + v of a pulse from its block quaternions (relaxation split around
                    each block), the maps format of the reused-pulse path
- ``first_order`` / ``first_quats``: each block's first-order (small-tip) interaction-frame rotation,
                    sum_k Rz(theta/2 - Phi_k) w_k with w_k = -2 pi dt_k b rf_k (per Hz of b), and its quaternion: the
                    part of a block's quaternion the network does not learn
- ``first_moment`` / ``second_moment``: the time moments of that sum, H = sum_k s_k Rz(theta/2 - Phi_k) w_k (s_k
                    from the block's middle) and G = int v v^T ds - (tau/2) v(tau) v(tau)^T (v(s) the partial sums, linear
                    within a substep), which ``block_map`` uses to correct its relaxation split
- ``features``      the network's inputs: f and b scaled into the box, sin and cos of 2 pi f q S, q = 1 .. Q
- ``mlp_forward`` / ``mlp_grad`` / ``adam``: the network d -> H -> H -> 4 M (tanh), its mean-squared-error gradient,
                    and Adam with bias correction

How each function is pinned (tests/test_surrogate_oracle.py):
- ``substep_quat`` / ``block_quats``: quaternion rotation = core.rotate (a Rodrigues restatement pinned against SciPy)
  to 1e-12 for complex RF and off-resonance; unit norm; a hard pulse on resonance (closed form cos, sin of flip/2);
- ``block_map``: with one substep per block (M = K) the split is the exact substep sequence, so it equals
  core.rf_segment applied K times to 1e-12 (with relaxation); its error at M = 10 against that is under 1e-5
  (the Strang split error, second order: halving short blocks' length divides it by about four);
- ``first_order`` / ``first_quats``: a hard pulse on resonance, where it is exact (rotations about one axis commute);
  the residual block_quats - first_quats of a shaped pulse off resonance falls faster than b (by more than 3.5 per
  halving of b: second order or higher), while a planted mistake (the precession phase's sign flipped) leaves a
  first-order residual (a factor of 2), which this check catches;
- ``block_map`` with moments: on a slice-selective pulse whose blocks precess many turns (spins far from its band),
  the corrected split is at least 10x closer to the substep-exact map than the plain split and at least second order
  in the block length (halving the blocks divides its error by more than 3), while a planted mistake (the
  correction's sign flipped) is further off than the plain split; on a symmetric slab pulse near its band (where the
  plain split's errors cancel between mirrored blocks) it stays within 1e-6 of the substeps;
- ``features``: the box's corners map to +-1; the harmonic structure (feature q at f equals feature 1 at q f) and the
  phases at f = 1 / (4 q S) (a quarter turn: sin 1, cos 0), so a wrong factor of 2 pi or q fails;
- ``mlp_forward`` / ``mlp_grad`` / ``mlp_loss``: torch.nn's Linear / Tanh stack (the library routine) loaded with
  the same flat parameters, outputs and autograd's flat gradient of the mean squared error to 1e-12; central finite
  differences of the loss to 1e-7 relative, with a planted mistake (the tanh derivative dropped) that this check
  catches;
- ``adam``: torch.optim.Adam (the library routine) to 1e-12.
"""

from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def qmul(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Hamilton product p q of (..., 4) quaternions (w, x, y, z): the rotation q, then p."""
    pw, px, py, pz = np.moveaxis(p, -1, 0)
    qw, qx, qy, qz = np.moveaxis(q, -1, 0)
    return np.stack([pw * qw - px * qx - py * qy - pz * qz,
                     pw * qx + px * qw + py * qz - pz * qy,
                     pw * qy - px * qz + py * qw + pz * qx,
                     pw * qz + px * qy - py * qx + pz * qw], -1)


def quat_rotate(q: np.ndarray, m: np.ndarray) -> np.ndarray:
    """The right-handed rotation of unit quaternion q applied to vectors m (..., 3)."""
    return np.einsum("...ij,...j->...i", quat_matrix(q), m)


def quat_matrix(q: np.ndarray) -> np.ndarray:
    """(..., 3, 3) rotation matrix of unit quaternions (..., 4)."""
    w, x, y, z = np.moveaxis(q, -1, 0)
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def rot_quat(wx, wy, wz) -> np.ndarray:
    """Unit quaternion of the right-handed rotation by |w| about w."""
    phi = np.sqrt(wx * wx + wy * wy + wz * wz)
    s = np.where(phi > 0, np.sin(0.5 * phi) / np.where(phi > 0, phi, 1.0), 0.5)
    return np.stack([np.cos(0.5 * phi), s * wx, s * wy, s * wz], -1)


def substep_quat(dt, rf_x, rf_y, f, b1) -> np.ndarray:
    """One RF substep's rotation: w = -2 pi dt (b1 rf_x, b1 rf_y, f) (core.py:124-150's rotation vector with the
    spin's total offset f in Hz)."""
    w = -TWO_PI * dt
    f = np.asarray(f, dtype=np.float64)
    b1 = np.asarray(b1, dtype=np.float64)
    return rot_quat(w * b1 * rf_x + 0 * f, w * b1 * rf_y + 0 * f, w * f + 0 * b1)


def block_bounds(k: int, m: int) -> np.ndarray:
    """The M + 1 substep indices that cut K substeps into M blocks of near-equal length."""
    return np.round(np.linspace(0, k, m + 1)).astype(np.int64)


def block_quats(table: np.ndarray, bounds: np.ndarray, f, b1) -> np.ndarray:
    """(N, M, 4) interaction-frame quaternions of the pulse ``table`` ((K, >=3) rows dt, rf_x, rf_y, and b0_offset
    in column 6 when present) at spin offsets f [Hz] and B1 scales b1 (N,): block j's rotation Q_j (substeps
    bounds[j]..bounds[j+1]-1 composed in time order, substep k's offset f + b0_offset_k) as
    Rz(-theta/2) Q_j Rz(-theta/2), theta = -2 pi sum_k dt_k (f + b0_offset_k) its free precession angle."""
    f = np.asarray(f, dtype=np.float64)
    b1 = np.asarray(b1, dtype=np.float64)
    n, m = len(f), len(bounds) - 1
    b0 = table[:, 6] if table.shape[1] > 6 else np.zeros(len(table))
    out = np.zeros((n, m, 4))
    for j in range(m):
        q = np.zeros((n, 4))
        q[:, 0] = 1.0
        for k in range(bounds[j], bounds[j + 1]):
            q = qmul(substep_quat(table[k, 0], table[k, 1], table[k, 2], f + b0[k], b1), q)
        tau = table[bounds[j] : bounds[j + 1], 0].sum()
        area = (table[bounds[j] : bounds[j + 1], 0] * b0[bounds[j] : bounds[j + 1]]).sum()
        half = rot_quat(0 * f, 0 * f, 0.5 * TWO_PI * (f * tau + area))  # Rz(-theta / 2)
        out[:, j] = qmul(half, qmul(q, half))
    return out


def first_order(table: np.ndarray, bounds: np.ndarray, f, phase_sign: float = 1.0) -> np.ndarray:
    """(N, M, 2) per Hz of b: block j's first-order interaction-frame rotation vector at offsets f,
    sum_k Rz(-phi_k) (-2 pi dt_k rf_k) with phi_k = Phi_k - theta/2, Phi_k = -2 pi (sum_{i<k} dt_i (f + off_i) +
    dt_k (f + off_k) / 2) the z angle up to the middle of substep k and theta the block's (the first Magnus term of
    Rz(-theta/2) Q Rz(-theta/2) with each substep split symmetrically about its z part). ``phase_sign=-1`` plants a
    mistake for the check that pins this function."""
    f = np.asarray(f, dtype=np.float64)
    b0 = table[:, 6] if table.shape[1] > 6 else np.zeros(len(table))
    m = len(bounds) - 1
    out = np.zeros((len(f), m, 2))
    for j in range(m):
        rows = range(bounds[j], bounds[j + 1])
        z_total = sum(table[k, 0] * (f + b0[k]) for k in rows)
        z = np.zeros_like(f)
        for k in rows:
            zk = table[k, 0] * (f + b0[k])
            phi = phase_sign * -TWO_PI * (z + 0.5 * zk - 0.5 * z_total)
            ax, ay = -TWO_PI * table[k, 0] * table[k, 1], -TWO_PI * table[k, 0] * table[k, 2]
            out[:, j, 0] += np.cos(phi) * ax + np.sin(phi) * ay
            out[:, j, 1] += np.cos(phi) * ay - np.sin(phi) * ax
            z = z + zk
    return out


def first_moment(table: np.ndarray, bounds: np.ndarray, f) -> np.ndarray:
    """(N, M, 2) per Hz of b: sum_k s_k Rz(-phi_k) (-2 pi dt_k rf_k), s_k = the middle of substep k minus the block's
    middle (``first_order``'s terms weighted by their time from the block's middle)."""
    f = np.asarray(f, dtype=np.float64)
    b0 = table[:, 6] if table.shape[1] > 6 else np.zeros(len(table))
    m = len(bounds) - 1
    out = np.zeros((len(f), m, 2))
    for j in range(m):
        rows = range(bounds[j], bounds[j + 1])
        z_total = sum(table[k, 0] * (f + b0[k]) for k in rows)
        tau = table[bounds[j] : bounds[j + 1], 0].sum()
        z, t = np.zeros_like(f), 0.0
        for k in rows:
            zk = table[k, 0] * (f + b0[k])
            phi = -TWO_PI * (z + 0.5 * zk - 0.5 * z_total)
            s = t + 0.5 * table[k, 0] - 0.5 * tau
            ax, ay = -TWO_PI * table[k, 0] * table[k, 1], -TWO_PI * table[k, 0] * table[k, 2]
            out[:, j, 0] += s * (np.cos(phi) * ax + np.sin(phi) * ay)
            out[:, j, 1] += s * (np.cos(phi) * ay - np.sin(phi) * ax)
            z, t = z + zk, t + table[k, 0]
    return out


def second_moment(table: np.ndarray, bounds: np.ndarray, f) -> np.ndarray:
    """(N, M, 3) per Hz^2 of b: G = int v v^T ds - (tau/2) v(tau) v(tau)^T as (xx, xy, yy), v(s) the partial sums of
    ``first_order``'s terms, growing linearly within each substep."""
    f = np.asarray(f, dtype=np.float64)
    b0 = table[:, 6] if table.shape[1] > 6 else np.zeros(len(table))
    m = len(bounds) - 1
    out = np.zeros((len(f), m, 3))
    for j in range(m):
        rows = range(bounds[j], bounds[j + 1])
        z_total = sum(table[k, 0] * (f + b0[k]) for k in rows)
        tau = table[bounds[j] : bounds[j + 1], 0].sum()
        z = np.zeros_like(f)
        v = np.zeros((len(f), 2))
        g = np.zeros((len(f), 2, 2))
        for k in rows:
            dt = table[k, 0]
            zk = dt * (f + b0[k])
            phi = -TWO_PI * (z + 0.5 * zk - 0.5 * z_total)
            ax, ay = -TWO_PI * dt * table[k, 1], -TWO_PI * dt * table[k, 2]
            u = np.stack([np.cos(phi) * ax + np.sin(phi) * ay, np.cos(phi) * ay - np.sin(phi) * ax], 1)
            outer = lambda p, q: p[:, :, None] * q[:, None, :]
            g += dt * (outer(v, v) + 0.5 * (outer(v, u) + outer(u, v)) + outer(u, u) / 3.0)
            v = v + u
            z = z + zk
        g -= 0.5 * tau * (v[:, :, None] * v[:, None, :])
        out[:, j] = np.stack([g[:, 0, 0], g[:, 0, 1], g[:, 1, 1]], 1)
    return out


def first_quats(first: np.ndarray, b) -> np.ndarray:
    """(N, M, 4) the quaternions of the rotations b F (first (N, M, 2) per Hz of b, b (N,))."""
    r = first * np.asarray(b, dtype=np.float64)[:, None, None]
    return rot_quat(r[..., 0], r[..., 1], 0.0 * r[..., 0])


def block_map(quats: np.ndarray, taus: np.ndarray, f, pd, r1, r2, areas=None, moments=None,
              correction_sign: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """The affine map of a pulse from its (N, M, 4) interaction-frame block quaternions: D (N, 3, 3) and v (N, 3)
    with m' = m + D m + v; block j is relax(tau_j / 2), Rz(theta/2) Q~_j Rz(theta/2), relax(tau_j / 2),
    theta = -2 pi (f tau_j + areas_j) (areas: the blocks' b0_offset areas, default 0). ``moments``: (H, G) the blocks'
    time moments ((N, M, 2) H_j times b, ``first_moment``; (N, M, 3) G_j times b^2, ``second_moment``): each block's
    map then gets the split's correction (Duhamel's formula: first order in the rates; second order in the RF)
    Rz(theta/2) Q~ (R2 - R1) (S(H) - T(G)) Rz(theta/2) on the linear part and Rz(theta/2) Q~ R1 PD
    (H_y, -H_x, -tr(G) / 2) on the offset, S(H) the symmetric matrix with (0, 2) = (2, 0) = H_y and
    (1, 2) = (2, 1) = -H_x, T(G) = [[-G_yy, G_xy, 0], [G_xy, -G_xx, 0], [0, 0, G_xx + G_yy]].
    ``correction_sign=-1`` plants a mistake for the check that pins it."""
    f, pd, r1, r2 = (np.asarray(a, dtype=np.float64) for a in (f, pd, r1, r2))
    pd, r1, r2 = (np.broadcast_to(a, f.shape) for a in (pd, r1, r2))
    areas = np.zeros(len(taus)) if areas is None else areas
    n, m = quats.shape[:2]
    A = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    c = np.zeros((n, 3))
    for j in range(m):
        e1, e2 = np.exp(-0.5 * taus[j] * r1), np.exp(-0.5 * taus[j] * r2)
        E = np.zeros((n, 3, 3))
        E[:, 0, 0] = E[:, 1, 1] = e2
        E[:, 2, 2] = e1
        rec = np.zeros((n, 3))
        rec[:, 2] = pd * (1 - e1)
        half = rot_quat(0 * f, 0 * f, -0.5 * TWO_PI * (f * taus[j] + areas[j]))  # Rz(theta / 2)
        q = quats[:, j] / np.linalg.norm(quats[:, j], axis=-1, keepdims=True)
        R = quat_matrix(qmul(half, qmul(q, half)))
        A0, c0 = A, c
        for step in (E, R, E):
            A = step @ A
            c = np.einsum("nij,nj->ni", step, c)
            if step is E:
                c = c + rec
        if moments is not None:
            hx, hy = moments[0][:, j, 0], moments[0][:, j, 1]
            gxx, gxy, gyy = moments[1][:, j, 0], moments[1][:, j, 1], moments[1][:, j, 2]
            Sm = np.zeros((n, 3, 3))
            Sm[:, 0, 2] = Sm[:, 2, 0] = hy
            Sm[:, 1, 2] = Sm[:, 2, 1] = -hx
            Sm[:, 0, 0], Sm[:, 0, 1], Sm[:, 1, 0], Sm[:, 1, 1], Sm[:, 2, 2] = gyy, -gxy, -gxy, gxx, -(gxx + gyy)
            H = quat_matrix(half)
            left = H @ quat_matrix(q)
            delta = correction_sign * left @ ((r2 - r1)[:, None, None] * Sm) @ H
            A = A + delta @ A0
            hv = np.stack([hy, -hx, -0.5 * (gxx + gyy)], -1) * (r1 * pd)[:, None]
            c = c + np.einsum("nij,nj->ni", delta, c0) + correction_sign * np.einsum("nij,nj->ni", left, hv)
    return A - np.eye(3), c


# -- the network: d -> H -> H -> n_out, tanh, as flat FP64 parameters W1 (H, d), b1 (H), W2 (H, H), b2 (H),
#    W3 (n_out, H), b3 (n_out) (the order of the device's parameter vector)


def features(f, b, head, q_max: int) -> np.ndarray:
    """(N, 2 + 2 q_max) inputs of spins with offsets f [Hz] and B1 x peak b [Hz]: head = (c_f, h_f, c_b, h_b, S),
    columns (f - c_f) / h_f, (b - c_b) / h_b, then sin, cos of 2 pi f q S for q = 1 .. q_max."""
    f, b = np.asarray(f, dtype=np.float64), np.asarray(b, dtype=np.float64)
    cf, hf, cb, hb, base = head
    cols = [(f - cf) / hf, (b - cb) / hb]
    for q in range(1, q_max + 1):
        cols += [np.sin(TWO_PI * f * q * base), np.cos(TWO_PI * f * q * base)]
    return np.stack(cols, 1)

def mlp_unpack(p: np.ndarray, h: int, n_out: int, d: int):
    sizes = [(h, d), (h,), (h, h), (h,), (n_out, h), (n_out,)]
    out, o = [], 0
    for s in sizes:
        k = int(np.prod(s))
        out.append(p[o : o + k].reshape(s))
        o += k
    assert o == len(p)
    return out


def mlp_size(h: int, n_out: int, d: int) -> int:
    return h * d + h + h * h + h + n_out * h + n_out


def mlp_forward(p: np.ndarray, x: np.ndarray, h: int, n_out: int):
    """Outputs (N, n_out) and the activations (a1, a2) of inputs x (N, d)."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, h, n_out, x.shape[1])
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, (a1, a2)


def mlp_loss(p, x, y, h, n_out) -> float:
    """Mean over samples and outputs of the squared error."""
    out, _ = mlp_forward(p, x, h, n_out)
    return float(((out - y) ** 2).mean())


def mlp_grad(p: np.ndarray, x: np.ndarray, y: np.ndarray, h: int, n_out: int, tanh_derivative: bool = True):
    """Gradient of ``mlp_loss`` wrt the flat parameters (backpropagation). ``tanh_derivative=False`` plants a
    mistake (the activation derivative dropped) for the check that pins this function."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, h, n_out, x.shape[1])
    out, (a1, a2) = mlp_forward(p, x, h, n_out)
    d3 = 2.0 * (out - y) / out.size
    d2 = d3 @ W3
    if tanh_derivative:
        d2 = d2 * (1 - a2 * a2)
    d1 = d2 @ W2
    if tanh_derivative:
        d1 = d1 * (1 - a1 * a1)
    return np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ a1).ravel(), d2.sum(0), (d3.T @ a2).ravel(),
                           d3.sum(0)])


def adam(p, g, m, v, t: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, with bias correction), returning new (p, m, v); t counts from 1."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1**t)
    vh = v / (1 - beta2**t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
