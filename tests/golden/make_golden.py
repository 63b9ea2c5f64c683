"""Writes the golden fixtures in tests/golden/*.json with mpmath at 50 digits.

Written from the closed-form mathematics of the regularized Biot-Savart pair
interaction (src/FLOWVPM_fmm.jl:113-161, src/FLOWVPM_kernel.jl:45-84,
src/FLOWVPM_subfilterscale_models.jl:16-41); it never imports or runs the
reference.  Inputs are exact doubles written with repr(), so the fixture
round-trips bit-for-bit.

    python tests/golden/make_golden.py        # rewrites the fixtures

Fixtures:
  erf.json       erf(x) on points covering the five fdlibm branches
  kernels.json   g, g', g'' and zeta for the four kernels on a q grid
  uj_<k>.json    U and J of small particle sets (random cloud + near pairs)
  check_fmm.json the 2-body case of scripts/check_fmm.jl:90-98 with its
                 analytic singular-kernel U and dU/dx (check_fmm.jl:39-71)
  estr.json      E_str SFS terms of a small set with prescribed J
  adjoint.json   dL/dX, dL/dGamma, dL/dsigma of a tiny set by mpmath
                 numerical differentiation of the exact loss
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))
PI = mp.pi
C4 = 1 / (4 * PI)
KERNELS = ["singular", "gaussian", "gaussianerf", "winckelmans"]


def g_closed(kernel, q):
    """g, g', g'' of each kernel (kernel.jl:46-47,61-62,52-53,75-76)."""
    q = mp.mpf(q)
    if kernel == "singular":
        return mp.mpf(1), mp.mpf(0), mp.mpf(0)
    if kernel == "gaussian":
        e = mp.exp(-q ** 3)
        return 1 - e, 3 * q ** 2 * e, (6 * q - 9 * q ** 4) * e
    if kernel == "gaussianerf":
        e = mp.exp(-q ** 2 / 2)
        c2 = mp.sqrt(2 / PI)
        return mp.erf(q / mp.sqrt(2)) - c2 * q * e, c2 * q ** 2 * e, c2 * (2 * q - q ** 3) * e
    if kernel == "winckelmans":
        t = q ** 2 + 1
        return (q ** 3 * (q ** 2 + mp.mpf(5) / 2) / t ** mp.mpf(2.5),
                mp.mpf(15) / 2 * q ** 2 / t ** mp.mpf(3.5),
                mp.mpf(15) / 2 * q * (2 - 5 * q ** 2) / t ** mp.mpf(4.5))
    raise ValueError(kernel)


def zeta_closed(kernel, q):
    q = mp.mpf(q)
    if kernel == "singular":
        return mp.mpf(1) if q == 0 else mp.mpf(0)
    if kernel == "gaussian":
        return 3 / (4 * PI) * mp.exp(-q ** 3)
    if kernel == "gaussianerf":
        return mp.exp(-q ** 2 / 2) / (2 * PI) ** mp.mpf(1.5)
    if kernel == "winckelmans":
        return mp.mpf(15) / (8 * PI) * (q ** 2 + 1) ** mp.mpf(-3.5)
    raise ValueError(kernel)


def pair_uj(kernel, xt, xs, G, sigma):
    """U (3) and J (3x3, J[a][b] = du_a/dx_b) induced at xt by one source."""
    d = [mp.mpf(xt[k]) - mp.mpf(xs[k]) for k in range(3)]
    r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
    if r2 == 0:
        return [mp.mpf(0)] * 3, [[mp.mpf(0)] * 3 for _ in range(3)]
    r = mp.sqrt(r2)
    q = r / mp.mpf(sigma)
    g, dg, _ = g_closed(kernel, q)
    G = [mp.mpf(x) for x in G]
    v = [d[1] * G[2] - d[2] * G[1], d[2] * G[0] - d[0] * G[2], d[0] * G[1] - d[1] * G[0]]
    F = -C4 * g / r ** 3
    H = -C4 * (q * dg - 3 * g) / r ** 5
    U = [F * v[a] for a in range(3)]
    eps = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}
    J = [[H * v[a] * d[b] + F * sum(eps.get((a, b, c), 0) * G[c] for c in range(3))
          for b in range(3)] for a in range(3)]
    return U, J


def field_uj(kernel, X, G, S):
    n = len(X)
    Us, Js = [], []
    for i in range(n):
        U = [mp.mpf(0)] * 3
        J = [[mp.mpf(0)] * 3 for _ in range(3)]
        for j in range(n):
            u, jj = pair_uj(kernel, X[i], X[j], G[j], S[j])
            for a in range(3):
                U[a] += u[a]
                for b in range(3):
                    J[a][b] += jj[a][b]
        Us.append([float(u) for u in U])
        # flat column-major, as the particle record stores it (fmm.jl:160)
        Js.append([float(J[a][b]) for b in range(3) for a in range(3)])
    return Us, Js


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as fh:
        json.dump(obj, fh, indent=0)
    print("wrote", name)


def make_erf():
    xs = [0.0, 1e-300, 1e-10, 1e-5, 0.1, 0.5, 0.84374, 0.84375, 0.9, 1.0, 1.2, 1.24999, 1.25,
          1.5, 2.0, 2.5, 2.857142857142857, 2.86, 3.5, 4.5, 5.5, 5.99999, 6.0, 7.0]
    xs = xs + [-x for x in xs if x != 0.0]
    rng = np.random.default_rng(11)
    xs += [float(x) for x in rng.uniform(-6.5, 6.5, 64)]
    write("erf.json", {"x": xs, "erf": [float(mp.erf(mp.mpf(x))) for x in xs]})


def make_kernels():
    qs = [0.0, 1e-6, 1e-4, 1e-3, 1e-2, 0.05, 0.1, 0.3, 0.5, 0.8, 1.0, 1.2, 1.5, 1.8, 2.0, 2.5,
          3.0, 3.5, 4.0, 5.0, 6.0, 6.5, 7.0, 8.0, 8.5, 9.0, 12.0, 20.0, 50.0]
    out = {"q": qs}
    for k in KERNELS:
        rows = [g_closed(k, q) for q in qs]
        out[k] = {"g": [float(r[0]) for r in rows], "dg": [float(r[1]) for r in rows],
                  "d2g": [float(r[2]) for r in rows],
                  "zeta": [float(zeta_closed(k, q)) for q in qs],
                  # reduced forms used by small-q code paths
                  "phi0": [float(r[0] / mp.mpf(q) ** 3) if q > 0 else None for r, q in zip(rows, qs)],
                  "phi1": [float((mp.mpf(q) * r[1] - 3 * r[0]) / mp.mpf(q) ** 5) if q > 0 else None
                           for r, q in zip(rows, qs)]}
    write("kernels.json", out)


def particle_set(seed, n, sig_lo, sig_hi):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, (n, 3))
    G = rng.normal(0.0, 1.0, (n, 3))
    S = rng.uniform(sig_lo, sig_hi, n)
    # near-coincident pairs at prescribed q: particle 2k+1 sits next to 2k
    qs = [1e-4, 1e-2, 0.5, 1.0, 3.0, 6.0, 8.5, 20.0]
    for k, q in enumerate(qs):
        i, j = 2 * k, 2 * k + 1
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        X[j] = X[i] + u * q * S[i]
    return X.tolist(), G.tolist(), S.tolist()


def make_uj():
    for k in KERNELS:
        X, G, S = particle_set(5 + KERNELS.index(k), 48, 0.05, 0.3)
        U, J = field_uj(k, X, G, S)
        write("uj_%s.json" % k, {"kernel": k, "X": X, "Gamma": G, "sigma": S, "U": U, "J": J})


def make_check_fmm():
    # scripts/check_fmm.jl:81-89: columns are bodies, rows x,y,z,sigma,Gx,Gy,Gz
    bodies = [[0.4, 0.1, -0.3, 1 / 8, 0.3, -0.1, 0.08],
              [0.1, -0.5, 0.2, 1 / 8, -0.4, -0.2, 0.5]]

    def u_an(xt, xs, G):            # check_fmm.jl:39-46
        d = [mp.mpf(xt[k]) - mp.mpf(xs[k]) for k in range(3)]
        n3 = mp.sqrt(sum(x * x for x in d)) ** 3
        return [1 / (4 * PI) / n3 * e for e in
                (-d[1] * G[2] + d[2] * G[1], -d[2] * G[0] + d[0] * G[2], -d[0] * G[1] + d[1] * G[0])]

    def du_an(xt, xs, G):           # check_fmm.jl:56-71, rows i, columns j
        x, y, z = [mp.mpf(xt[k]) - mp.mpf(xs[k]) for k in range(3)]
        gx, gy, gz = [mp.mpf(g) for g in G]
        n5 = mp.sqrt(x * x + y * y + z * z) ** 5
        m = [[3 * x * y * gz - 3 * x * z * gy, (2 * y * y - x * x - z * z) * gz - 3 * y * z * gy,
              3 * y * z * gz - (2 * z * z - x * x - y * y) * gy],
             [3 * x * z * gx - (2 * x * x - y * y - z * z) * gz, 3 * y * z * gx - 3 * x * y * gz,
              (2 * z * z - x * x - y * y) * gx - 3 * x * z * gz],
             [(2 * x * x - y * y - z * z) * gy - 3 * x * y * gx,
              3 * x * y * gy - (2 * y * y - x * x - z * z) * gx, 3 * x * z * gy - 3 * y * z * gx]]
        return [[1 / (4 * PI) * m[i][j] / n5 for j in range(3)] for i in range(3)]

    out = {"X": [b[0:3] for b in bodies], "sigma": [b[3] for b in bodies],
           "Gamma": [b[4:7] for b in bodies], "U": [], "J": []}
    for t, s in ((0, 1), (1, 0)):
        U = u_an(bodies[t][0:3], bodies[s][0:3], bodies[s][4:7])
        D = du_an(bodies[t][0:3], bodies[s][0:3], bodies[s][4:7])
        out["U"].append([float(u) for u in U])
        out["J"].append([float(D[a][b]) for b in range(3) for a in range(3)])
    write("check_fmm.json", out)


def make_estr():
    rng = np.random.default_rng(21)
    n = 32
    X = rng.uniform(0, 0.6, (n, 3)).tolist()
    G = rng.normal(0, 1, (n, 3)).tolist()
    S = rng.uniform(0.08, 0.2, n).tolist()
    J = rng.normal(0, 1, (n, 9)).tolist()
    static = [0.0] * n
    static[3] = 1.0
    static[10] = 1.0
    out = {"X": X, "Gamma": G, "sigma": S, "J": J, "static": static}
    for k in KERNELS:
        for transposed in (True, False):
            SFS = []
            for i in range(n):
                acc = [mp.mpf(0)] * 3
                if static[i] == 0:
                    for j in range(n):
                        if static[j] != 0:
                            continue
                        d = [mp.mpf(X[j][c]) - mp.mpf(X[i][c]) for c in range(3)]
                        r = mp.sqrt(sum(x * x for x in d))
                        sg = mp.mpf(S[j])
                        z = zeta_closed(k, r / sg) / sg ** 3
                        dJ = [mp.mpf(J[i][f]) - mp.mpf(J[j][f]) for f in range(9)]
                        for a in range(3):
                            if transposed:   # S_a = sum_c dJ[3a + c] Gamma_c
                                s = sum(dJ[3 * a + c] * mp.mpf(G[j][c]) for c in range(3))
                            else:            # S_a = sum_c dJ[3c + a] Gamma_c
                                s = sum(dJ[3 * c + a] * mp.mpf(G[j][c]) for c in range(3))
                            acc[a] += z * s
                SFS.append([float(a) for a in acc])
            out["%s_%s" % (k, "T" if transposed else "C")] = SFS
    write("estr.json", out)


def make_adjoint():
    rng = np.random.default_rng(31)
    n = 6
    X = rng.uniform(0, 0.5, (n, 3)).tolist()
    G = rng.normal(0, 1, (n, 3)).tolist()
    S = rng.uniform(0.1, 0.3, n).tolist()
    w = rng.normal(0, 1, (n, 3)).tolist()
    W = rng.normal(0, 1, (n, 9)).tolist()
    out = {"X": X, "Gamma": G, "sigma": S, "Ubar": w, "Jbar": W}
    mp.mp.dps = 40
    for k in ("gaussianerf", "winckelmans", "gaussian", "singular"):
        def loss(Xm, Gm, Sm):
            L = mp.mpf(0)
            for i in range(n):
                for j in range(n):
                    u, jj = pair_uj(k, Xm[i], Xm[j], Gm[j], Sm[j])
                    L += sum(mp.mpf(w[i][a]) * u[a] for a in range(3))
                    L += sum(mp.mpf(W[i][3 * b + a]) * jj[a][b] for a in range(3) for b in range(3))
            return L
        Xm = [[mp.mpf(x) for x in row] for row in X]
        Gm = [[mp.mpf(x) for x in row] for row in G]
        Sm = [mp.mpf(x) for x in S]
        dX, dG, dS = [], [], []
        for i in range(n):
            rx, rg = [], []
            for c in range(3):
                def fx(t, i=i, c=c):
                    Xp = [row[:] for row in Xm]
                    Xp[i][c] = t
                    return loss(Xp, Gm, Sm)

                def fg(t, i=i, c=c):
                    Gp = [row[:] for row in Gm]
                    Gp[i][c] = t
                    return loss(Xm, Gp, Sm)
                rx.append(float(mp.diff(fx, Xm[i][c])))
                rg.append(float(mp.diff(fg, Gm[i][c])))

            def fs(t, i=i):
                Sp = Sm[:]
                Sp[i] = t
                return loss(Xm, Gm, Sp)
            dX.append(rx)
            dG.append(rg)
            dS.append(float(mp.diff(fs, Sm[i])))
        out[k] = {"Xbar": dX, "Gbar": dG, "Sbar": dS}
    mp.mp.dps = 50
    write("adjoint.json", out)


if __name__ == "__main__":
    make_erf()
    make_kernels()
    make_uj()
    make_check_fmm()
    make_estr()
    make_adjoint()
