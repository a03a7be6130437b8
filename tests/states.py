"""Seeded test states shared by the parity tests (host-side numpy; no physics kernels)."""
import numpy as np


def face_field_from_potential(desc, A1, A2, A3, dx):
    """Face B = curl A on a periodic global grid (edge-centred A), so div B = 0 to
    round-off.  Returns global face arrays b1[N3,N2,N1], b2, b3 (face i at the lower
    side of cell i)."""
    d1, d2, d3 = dx
    # A1 on x1-edges (k-1/2, j-1/2, i), A2 on x2-edges, A3 on x3-edges (periodic)
    b1 = (np.roll(A3, -1, axis=1) - A3) / d2 - (np.roll(A2, -1, axis=0) - A2) / d3
    b2 = (np.roll(A1, -1, axis=0) - A1) / d3 - (np.roll(A3, -1, axis=2) - A3) / d1
    b3 = (np.roll(A2, -1, axis=2) - A2) / d1 - (np.roll(A1, -1, axis=1) - A1) / d2
    return b1, b2, b3


def scatter_faces(desc, b1, b2, b3):
    """Global periodic face arrays -> block face arrays (interior faces only; ghosts 0)."""
    ng = desc.nghost
    nb1, nb2, nb3 = desc.nx1 // desc.bnx1, desc.nx2 // desc.bnx2, desc.nx3 // desc.bnx3
    bx1, bx2, bx3 = desc.bnx1, desc.bnx2, desc.bnx3
    n1, n2, n3 = bx1 + 2 * ng, bx2 + 2 * ng, bx3 + 2 * ng
    nb = nb1 * nb2 * nb3
    o1 = np.zeros((nb, n3, n2, n1 + 1))
    o2 = np.zeros((nb, n3, n2 + 1, n1))
    o3 = np.zeros((nb, n3 + 1, n2, n1))
    for g in range(nb):
        l1, l2, l3 = g % nb1, (g // nb1) % nb2, g // (nb1 * nb2)
        I = (l1 * bx1 + np.arange(bx1 + 1)) % desc.nx1
        J = (l2 * bx2 + np.arange(bx2 + 1)) % desc.nx2
        K = (l3 * bx3 + np.arange(bx3 + 1)) % desc.nx3
        o1[g, ng:ng + bx3, ng:ng + bx2, ng:ng + bx1 + 1] = b1[np.ix_(K[:-1], J[:-1], I)]
        o2[g, ng:ng + bx3, ng:ng + bx2 + 1, ng:ng + bx1] = b2[np.ix_(K[:-1], J, I[:-1])]
        o3[g, ng:ng + bx3 + 1, ng:ng + bx2, ng:ng + bx1] = b3[np.ix_(K, J[:-1], I[:-1])]
    return o1, o2, o3


def turbulent_state(desc, seed=0, mach=1.0, b0=(0.1, 0.0, 0.0), db=0.05, drho=0.2, gamma=5 / 3,
                    p0=None, vamp=None, iso=False):
    """Smooth random MHD state: a few random Fourier modes in v, rho, and a vector
    potential; returns block arrays (u, b1f, b2f, b3f)."""
    from tests.conftest import load_agx
    agx = load_agx()
    rng = np.random.default_rng(seed)
    N1, N2, N3 = desc.nx1, desc.nx2, desc.nx3
    L = (desc.x1max - desc.x1min, desc.x2max - desc.x2min, desc.x3max - desc.x3min)
    dx = (L[0] / N1, L[1] / N2, L[2] / N3)
    x = (np.arange(N1) + 0.5) / N1
    y = (np.arange(N2) + 0.5) / N2
    z = (np.arange(N3) + 0.5) / N3
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def field(amp):
        f = np.zeros_like(X)
        for _ in range(4):
            k = rng.integers(1, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            f += amp * rng.normal() * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
        return f

    rho = 1.0 + drho * np.tanh(field(1.0))
    p = (0.6 / gamma if p0 is None else p0) * np.ones_like(rho)
    cs = np.sqrt(gamma * p / rho)
    v = [(mach * cs if vamp is None else vamp) * field(0.5) for _ in range(3)]
    # edge-centred vector potential (sample at cell centres shifted; any smooth field works)
    A1, A2, A3 = (db * field(1.0) / (2 * np.pi) for _ in range(3))
    b1, b2, b3 = face_field_from_potential(desc, A1, A2, A3, dx)
    b1 += b0[0]
    b2 += b0[1]
    b3 += b0[2]
    bc1 = 0.5 * (b1 + np.roll(b1, -1, axis=2))
    bc2 = 0.5 * (b2 + np.roll(b2, -1, axis=1))
    bc3 = 0.5 * (b3 + np.roll(b3, -1, axis=0))
    U = np.zeros((5, N3, N2, N1))
    U[0] = rho
    U[1:4] = [rho * vv for vv in v]
    U[4] = p / (gamma - 1) + 0.5 * rho * sum(vv * vv for vv in v) + 0.5 * (bc1 ** 2 + bc2 ** 2 + bc3 ** 2)
    if iso:   # isothermal EOS: no energy variable (the oracle keeps the slot at 0)
        U[4] = 0.0
    u = agx.scatter_cells(desc, U)
    f1, f2, f3 = scatter_faces(desc, b1, b2, b3)
    return u, f1, f2, f3


def load_state(target, state):
    u, f1, f2, f3 = state
    target.set("u", u)
    target.set("b1f", f1)
    target.set("b2f", f2)
    target.set("b3f", f3)


def oracle_self_sensitivity(desc, state, advance, eps=6e-8, seeds=tuple(range(1, 9))):
    """How far the FP64 oracle moves from itself when the primitives its reconstruction
    reads are rounded to FP32 after every C2P, or multiplied there by (1 + eps u), u
    uniform in [-1, 1) (seeded; eps = 2^-24, the FP32 rounding bound): the FP64
    discretization's own response to perturbations of FP32-rounding size, which bounds
    what any form that stores or rounds the primitives in FP32 can be held to (PPM's
    limiter is discontinuous at its sign tests, so a perturbation that decides a near-tie
    the other way moves that cell by O(dq)).  advance(o) runs the workload.  Returns the
    max over the runs of the max-norm difference, relative to the unperturbed maximum, of
    rho, P and bcc (interior) and of each face field."""
    from oracle.orc import Oracle
    ng = desc.nghost
    sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    fs = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))

    def run(e, seed):
        o = Oracle(desc)
        load_state(o, state)
        o.initialize()
        o.set_noise(e, seed)
        advance(o)
        return o.get("w")[sl], [o.get(n)[fs] for n in ("b1f", "b2f", "b3f")], o.get("bcc")[sl]
    w0, b0, c0 = run(0.0, 0)
    out = {"rho": 0.0, "P": 0.0, "bcc": 0.0, "b1f": 0.0, "b2f": 0.0, "b3f": 0.0}
    for e, seed in [(-1.0, 0)] + [(eps, s) for s in seeds]:
        w1, b1, c1 = run(e, seed)
        out["bcc"] = max(out["bcc"], float(np.abs(c1 - c0).max() / np.abs(c0).max()))
        for k, n in (("rho", 0), ("P", 4)):
            out[k] = max(out[k], float(np.abs(w1[:, n] - w0[:, n]).max() / np.abs(w0[:, n]).max()))
        for k, a, b in zip(("b1f", "b2f", "b3f"), b1, b0):
            out[k] = max(out[k], float(np.abs(a - b).max() / np.abs(b).max()))
    return out
