"""Seeded synthetic initial conditions (host-side problem generators).

The reference builds its turbulent velocity field with an FFTW-based driver
(fft/turbulence.cpp:244-330) that is not available here; these generators produce the
benchmark inputs directly with numpy's PCG64 and FFT.
"""
import numpy as np


def gaussian_field(shape, seed, kmin=2, kmax=8, power=-4.0, box=(1.0, 1.0, 1.0)):
    """Real Gaussian random field on a periodic grid with |f_k|^2 ~ k^power on integer
    wavenumber shells kmin <= |k| <= kmax (inclusive), zero mean, unit rms.
    shape = (N3, N2, N1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N3, N2, N1 = shape
    k3 = np.fft.fftfreq(N3) * N3
    k2 = np.fft.fftfreq(N2) * N2
    k1 = np.fft.rfftfreq(N1) * N1
    K = np.sqrt(k3[:, None, None] ** 2 + k2[None, :, None] ** 2 + k1[None, None, :] ** 2)
    amp = np.zeros_like(K)
    sel = (K >= kmin) & (K <= kmax)
    amp[sel] = K[sel] ** (power / 2.0)
    fk = amp * (rng.normal(size=K.shape) + 1j * rng.normal(size=K.shape))
    f = np.fft.irfftn(fk, s=shape, axes=(0, 1, 2))
    f -= f.mean()
    f /= np.sqrt(np.mean(f * f))
    return f


def turbulence(desc, seed=2, mach=2.0, bx=0.1, p0=None):
    """Decaying-turbulence start: rho = 1, P = 0.6/gamma, uniform Bx, velocity from
    three independent k^-4 fields on 2 <= |k| <= 8, scaled to 3-D rms Mach `mach`.
    Returns global arrays (U[5,N3,N2,N1], b1, b2, b3) with face index i = lower face
    of cell i."""
    gamma = desc.gamma
    shape = (desc.nx3, desc.nx2, desc.nx1)
    p = 0.6 / gamma if p0 is None else p0
    rho = np.ones(shape)
    cs = np.sqrt(gamma * p / 1.0)
    v = [gaussian_field(shape, seed * 1000 + c) for c in range(3)]
    vrms = np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    scale = mach * cs / vrms
    v = [vv * scale for vv in v]
    U = np.empty((5,) + shape)
    U[0] = rho
    for c in range(3):
        U[1 + c] = rho * v[c]
    U[4] = p / (gamma - 1.0) + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) + 0.5 * bx * bx
    b1 = np.full(shape, bx)
    b2 = np.zeros(shape)
    b3 = np.zeros(shape)
    return U, b1, b2, b3


def turbulence_lognormal(desc, seed=5, mach=5.0, sigma_lnrho=1.0, beta=1.0, bx=None):
    """Large-run start: lognormal density rho = exp(sigma g) / mean (g a unit-rms k^-4
    field on 2 <= |k| <= 8, so <rho> = 1 and the rms of ln rho is sigma), isothermal-like
    P = 0.6 rho / gamma (uniform sound speed sqrt(0.6)), velocity as in turbulence() at
    rms Mach `mach`, uniform Bx with plasma beta 2 <P> / Bx^2 = `beta` (or Bx = `bx`
    when given)."""
    gamma = desc.gamma
    shape = (desc.nx3, desc.nx2, desc.nx1)
    rho = np.exp(sigma_lnrho * gaussian_field(shape, seed * 1000 + 7))
    rho /= rho.mean()
    pbar = 0.6 / gamma
    cs = np.sqrt(0.6)
    if bx is None:
        bx = np.sqrt(2.0 * pbar / beta)
    U = np.empty((5,) + shape)
    U[0] = rho
    U[4] = pbar * rho / (gamma - 1.0) + 0.5 * bx * bx
    v2 = np.zeros(shape)
    v = [gaussian_field(shape, seed * 1000 + c) for c in range(3)]
    vrms = np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    for c in range(3):
        vc = v[c] * (mach * cs / vrms)
        U[1 + c] = rho * vc
        v2 += vc * vc
    del v
    U[4] += 0.5 * rho * v2
    b1 = np.full(shape, bx)
    b2 = np.zeros(shape)
    b3 = np.zeros(shape)
    return U, b1, b2, b3


def blobs(desc, seed=3, nblobs=32):
    """Self-gravity benchmark density: 1 + sum of Gaussian blobs (widths 0.02-0.08,
    log-uniform amplitudes 1-100, periodic minimum-image distance) on the unit box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (desc.nx3, desc.nx2, desc.nx1)
    L = (desc.x1max - desc.x1min, desc.x2max - desc.x2min, desc.x3max - desc.x3min)
    x = [desc.x1min + (np.arange(n) + 0.5) * l / n for n, l in zip((desc.nx1, desc.nx2, desc.nx3), L)]
    rho = np.ones(shape)
    for _ in range(nblobs):
        c = [desc.x1min + rng.uniform() * L[0], desc.x2min + rng.uniform() * L[1],
             desc.x3min + rng.uniform() * L[2]]
        s = rng.uniform(0.02, 0.08)
        a = 10.0 ** rng.uniform(0.0, 2.0)
        d = []
        for ax in range(3):
            dd = np.abs(x[ax] - c[ax])
            d.append(np.minimum(dd, L[ax] - dd))
        g1 = np.exp(-d[0] ** 2 / (2 * s * s))
        g2 = np.exp(-d[1] ** 2 / (2 * s * s))
        g3 = np.exp(-d[2] ** 2 / (2 * s * s))
        rho += a * g3[:, None, None] * g2[None, :, None] * g1[None, None, :]
    return rho


def scatter_faces(desc, b1, b2, b3):
    """Global periodic face arrays -> block face arrays (interior faces; ghosts 0)."""
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


def divfree_blocks(faces, box, gamma=5.0 / 3.0, amp=0.2, b0=(0.6, 0.0, 0.0), seed=0):
    """Smooth periodic MHD state on blocks of any refinement level, given each block's face
    positions faces[g] = (x1f, x2f, x3f) (n + 1 each, ghosts included): density, pressure
    and velocity sinusoids at the cell centres; the face field B = curl A + b0 with
    A = (Ax(y, z), Ay(x, z), Az(x, y)) sinusoids and each face value the exact line
    integral of A around the face over its area, so the flux through a coarse face equals
    the sum over the fine faces that cover it and every cell's divergence vanishes up to
    round-off.  Returns (u (nb, 5, n, n, n), b1f, b2f, b3f) in the reference's layouts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = [b[1] - b[0] for b in box]
    k = [2 * np.pi / x for x in L]
    ph = rng.uniform(0, 2 * np.pi, 12)

    def Ax(y, z):
        return 0.1 * np.sin(k[1] * y + ph[0]) * np.cos(k[2] * z + ph[1])

    def Ay(x, z):
        return 0.1 * np.sin(k[2] * z + ph[2]) * np.cos(k[0] * x + ph[3])

    def Az(x, y):
        return 0.1 * np.sin(k[0] * x + ph[4]) * np.cos(k[1] * y + ph[5])
    nb, n = len(faces), len(faces[0][0]) - 1
    u = np.zeros((nb, 5, n, n, n))
    b1 = np.zeros((nb, n, n, n + 1))
    b2 = np.zeros((nb, n, n + 1, n))
    b3 = np.zeros((nb, n + 1, n, n))
    for g in range(nb):
        f1, f2, f3 = faces[g]
        c1, c2, c3 = 0.5 * (f1[1:] + f1[:-1]), 0.5 * (f2[1:] + f2[:-1]), 0.5 * (f3[1:] + f3[:-1])
        Z, Y, X = np.meshgrid(c3, c2, c1, indexing="ij")
        rho = 1.0 + amp * np.sin(k[0] * X + ph[6]) * np.cos(k[1] * Y + ph[7]) * np.sin(k[2] * Z + ph[8])
        vx = amp * np.sin(k[1] * Y + ph[9])
        vy = amp * np.sin(k[2] * Z + ph[10])
        vz = amp * np.sin(k[0] * X + ph[11])
        p = 1.0 + 0.5 * amp * np.cos(k[0] * X + ph[8]) * np.cos(k[2] * Z + ph[9])
        Zf, Yf, Xf = np.meshgrid(f3, f2, f1, indexing="ij")
        ax, ay, az = Ax(Yf, Zf), Ay(Xf, Zf), Az(Xf, Yf)
        dx, dy, dz = np.diff(f1), np.diff(f2), np.diff(f3)
        # x1 face (k, j, i): Bx = dAz/dy - dAy/dz as circulation over area
        b1[g] = ((ay[:-1, :-1, :] - ay[1:, :-1, :]) / dz[:, None, None] +
                 (az[:-1, 1:, :] - az[:-1, :-1, :]) / dy[None, :, None]) + b0[0]
        b2[g] = ((az[:-1, :, :-1] - az[:-1, :, 1:]) / dx[None, None, :] +
                 (ax[1:, :, :-1] - ax[:-1, :, :-1]) / dz[:, None, None]) + b0[1]
        b3[g] = ((ax[:, :-1, :-1] - ax[:, 1:, :-1]) / dy[None, :, None] +
                 (ay[:, :-1, 1:] - ay[:, :-1, :-1]) / dx[None, None, :]) + b0[2]
        bc = [0.5 * (b1[g][..., 1:] + b1[g][..., :-1]), 0.5 * (b2[g][:, 1:] + b2[g][:, :-1]),
              0.5 * (b3[g][1:] + b3[g][:-1])]
        u[g, 0] = rho
        u[g, 1], u[g, 2], u[g, 3] = rho * vx, rho * vy, rho * vz
        u[g, 4] = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * sum(x * x for x in bc)
    return u, b1, b2, b3
