"""Synthetic particle fields of the measured configurations.

Every generator is deterministic (numpy PCG64 with a fixed seed) and returns a
particle matrix in the reference's 46-double-per-particle layout.

  ring_cfg0()      one vortex ring, N = 8,000
  random_cloud()   uniform cloud, N = 65,536 (and N = 32,768 for the adjoint)
  leapfrog_cfg2()  two coaxial rings, N = 262,144
  jet_column()     round-jet-like column, N = 1,048,576
  ring_test()      the single ring of test/runtests_singlevortexring.jl:17-60
"""
import math

import numpy as np

from .fields import make_particles


def _ring(R, Gamma0, a, Nphi, layers, z0=0.0, sigma_factor=1.5):
    """A ring of radius R around the z axis at height z0.

    Station n sits at phi = 2 pi (n + 1/2) / Nphi.  Its cross section holds
    concentric rings k at r_c = a (k + 1/2) / len(layers) with layers[k]
    points each, at theta = 2 pi (m + 1/2) / layers[k].  Each particle carries
    Gamma = (Gamma0 / npts) (2 pi R / Nphi) t_hat (t_hat the azimuthal unit
    vector) and sigma = sigma_factor * 2 pi R / Nphi.
    """
    npts = int(sum(layers))
    phi = 2 * np.pi * (np.arange(Nphi) + 0.5) / Nphi
    rc, th = [], []
    nl = len(layers)
    for k, nk in enumerate(layers):
        rc.append(np.full(nk, a * (k + 0.5) / nl))
        th.append(2 * np.pi * (np.arange(nk) + 0.5) / nk)
    rc = np.concatenate(rc)
    th = np.concatenate(th)
    P = phi[:, None]
    rad = np.broadcast_to(R + rc[None, :] * np.cos(th[None, :]), (Nphi, npts))
    X = np.stack([rad * np.cos(P), rad * np.sin(P),
                  np.broadcast_to(z0 + rc[None, :] * np.sin(th[None, :]), rad.shape)], axis=-1)
    mag = Gamma0 / npts * (2 * np.pi * R / Nphi)
    tang = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], axis=-1)
    G = np.broadcast_to(mag * tang[:, None, :], X.shape)
    sigma = sigma_factor * 2 * np.pi * R / Nphi
    return X.reshape(-1, 3), G.reshape(-1, 3), np.full(Nphi * npts, sigma)


def ring_cfg0():
    """R=1, Gamma0=1, a=0.1, Nphi=200 x (4+8+12+16) = 8,000 particles."""
    X, G, S = _ring(1.0, 1.0, 0.1, 200, (4, 8, 12, 16))
    return make_particles(X, G, S)


def random_cloud(n=65536, seed=1):
    """X ~ U[0,1]^3, Gamma_c ~ N(0, 1e-3^2), sigma ~ U[0.02, 0.04]."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, (n, 3))
    G = rng.normal(0.0, 1e-3, (n, 3))
    S = rng.uniform(0.02, 0.04, n)
    return make_particles(X, G, S)


def adjoint_weights(n, seed=3):
    """w_i ~ N(0,1)^3 (weights of U) and W_i ~ N(0,1)^{3x3} (weights of J,
    flat in the J storage order)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 1.0, (n, 3)), rng.normal(0.0, 1.0, (n, 9))


def leapfrog_cfg2(Nphi=2048):
    """Rings R=1 at z=0 and R=1.1 at z=0.3, Gamma0=1, a=0.1; each cross
    section has 5 concentric rings of 4, 8, 12, 16, 24 points (64)."""
    layers = (4, 8, 12, 16, 24)
    X1, G1, S1 = _ring(1.0, 1.0, 0.1, Nphi, layers, 0.0)
    X2, G2, S2 = _ring(1.1, 1.0, 0.1, Nphi, layers, 0.3)
    return make_particles(np.vstack([X1, X2]), np.vstack([G1, G2]), np.concatenate([S1, S2]))


def jet_column(n=1 << 20, seed=4, radius=0.5, length=4.0, sigma=0.015, amp=1e-2):
    """Jittered-lattice column of exactly n particles.

    Lattice spacing h = (pi r^2 L / n)^(1/3); the lattice points with
    r <= radius are ranked by (r, lattice order) and the n innermost are kept,
    in lattice order (x slowest, z fastest); z is centred on 0.  Positions are
    jittered by U(-h/4, h/4) per coordinate.  Gamma = h^3 w(r) (z_hat +
    amp xi theta_hat), w(r) = 0.5 (1 - tanh((r - radius)/0.02)), xi ~ N(0,1).
    """
    rng = np.random.default_rng(seed)
    h = (math.pi * radius ** 2 * length / n) ** (1.0 / 3.0)
    while True:
        nx = int(math.floor(2 * radius / h)) + 1
        nz = int(math.floor(length / h)) + 1
        xs = (np.arange(nx) - (nx - 1) / 2) * h
        zs = (np.arange(nz) - (nz - 1) / 2) * h
        gx, gy = np.meshgrid(xs, xs, indexing="ij")
        r = np.hypot(gx, gy).ravel()
        inside = np.nonzero(r <= radius)[0]
        if inside.size * nz >= n:
            break
        h *= 0.999
    # rank columns by radius; keep whole z-columns, then trim the outermost
    order = inside[np.argsort(r[inside], kind="stable")]
    ncol_full = n // nz
    rem = n - ncol_full * nz
    keep_cols = np.sort(order[:ncol_full])
    X = np.empty((n, 3))
    gxf, gyf = gx.ravel(), gy.ravel()
    k = 0
    for c in keep_cols:
        X[k:k + nz, 0] = gxf[c]
        X[k:k + nz, 1] = gyf[c]
        X[k:k + nz, 2] = zs
        k += nz
    if rem:
        c = order[ncol_full]
        X[k:k + rem, 0] = gxf[c]
        X[k:k + rem, 1] = gyf[c]
        X[k:k + rem, 2] = zs[:rem]
    X += rng.uniform(-0.25 * h, 0.25 * h, X.shape)
    rr = np.hypot(X[:, 0], X[:, 1])
    w = 0.5 * (1.0 - np.tanh((rr - radius) / 0.02))
    th = np.arctan2(X[:, 1], X[:, 0])
    xi = rng.normal(0.0, 1.0, n)
    G = np.zeros((n, 3))
    G[:, 2] = h ** 3 * w
    G[:, 0] = -np.sin(th) * h ** 3 * w * amp * xi
    G[:, 1] = np.cos(th) * h ** 3 * w * amp * xi
    return make_particles(X, G, sigma)


def ring_test(Nphi=100, R=1.0, Rcross=0.15, Gamma0=1.0):
    """The nc=0 ring of test/runtests_singlevortexring.jl:17-60.

    With AR=1 and nc=0, addvortexring (examples/vortexrings/
    vortexrings_functions.jl:73-227) places one particle per cross section at
    R (sin phi_c, cos phi_c, 0), phi_c = 2 pi (n + 1/2)/Nphi, with
    Gamma = Gamma0 (2 pi R / Nphi) (-cos phi_c, sin phi_c, 0) (the core-area
    factor faux cancels) and sigma = Rcross.
    """
    phi = 2 * np.pi * (np.arange(Nphi) + 0.5) / Nphi
    X = np.stack([R * np.sin(phi), R * np.cos(phi), np.zeros(Nphi)], axis=1)
    T = np.stack([-np.cos(phi), np.sin(phi), np.zeros(Nphi)], axis=1)
    G = Gamma0 * (2 * np.pi * R / Nphi) * T
    return make_particles(X, G, Rcross)


def Uring(circulation, R, Rcross, beta):
    """Analytic ring speed (examples/vortexrings/vortexrings_functions.jl:47)."""
    return circulation / (4 * math.pi * R) * (math.log(8 * R / Rcross) - beta)


def ring_centroid(P):
    """|Gamma|-weighted centroid (vortexrings_functions.jl:275-316)."""
    w = np.linalg.norm(P[:, 3:6], axis=1)
    return (P[:, 0:3] * w[:, None]).sum(axis=0) / w.sum()


def morton_order(P):
    """Particles reordered along a 21-bit-per-axis Z-order curve (the order
    an octree code stores its leaves in)."""
    X = P[:, 0:3]
    lo, hi = X.min(0), X.max(0)
    ext = max(float((hi - lo).max()), 1e-300)
    u = ((X - lo) / ext * (2 ** 21 - 1)).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x1fffff)
        for s, m in ((32, 0x1f00000000ffff), (16, 0x1f0000ff0000ff), (8, 0x100f00f00f00f00f),
                     (4, 0x10c30c30c30c30c3), (2, 0x1249249249249249)):
            v = (v | (v << np.uint64(s))) & np.uint64(m)
        return v

    key = spread(u[:, 0]) | (spread(u[:, 1]) << np.uint64(1)) | (spread(u[:, 2]) << np.uint64(2))
    return P[np.argsort(key, kind="stable")].copy()


def leaf_direct_list(P, leaf=64, radius=3.0):
    """An FMM-style direct list over P (already in Morton order): leaves are
    runs of `leaf` consecutive particles; leaf j is a direct source of leaf i
    when their centres are within `radius` times the median leaf radius (i = j
    included).  Returns 0-based half-open (tb, te, sb, se), grouped by target
    leaf."""
    from scipy.spatial import cKDTree
    n = P.shape[0]
    starts = np.arange(0, n, leaf)
    ends = np.minimum(starts + leaf, n)
    C = np.add.reduceat(P[:, 0:3], starts, axis=0) / (ends - starts)[:, None]
    owner = np.repeat(np.arange(len(starts)), ends - starts)
    R = np.zeros(len(starts))
    np.maximum.at(R, owner, np.linalg.norm(P[:, 0:3] - C[owner], axis=1))
    pairs = cKDTree(C).query_pairs(radius * float(np.median(R)), output_type="ndarray")
    nl = len(starts)
    ii = np.concatenate([pairs[:, 0], pairs[:, 1], np.arange(nl)])
    jj = np.concatenate([pairs[:, 1], pairs[:, 0], np.arange(nl)])
    o = np.lexsort((jj, ii))
    ii, jj = ii[o], jj[o]
    return starts[ii], ends[ii], starts[jj], ends[jj]
