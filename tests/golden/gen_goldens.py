"""Writes tests/golden/goldens.npz -- small fixtures that pin the FP64 oracle.

Run in the build container (numpy 2.2 / scipy 1.15); the output is committed.
Nothing here imports the oracle: each golden comes from an independent route.

* dft_*: numpy.fft on seeded complex rows.  cuFFT's CUFFT_INVERSE is the
  unnormalised e^{+i} transform (n * ifft), CUFFT_FORWARD the e^{-i} one (fft).
* tri_*: dense numpy.linalg.solve of the per-mode Poisson rows written from
  the reference's coefficient rule (KaminoSolver.cu:117-163).
* proj_*: the whole projection at tiny sizes by dense linear algebra in the
  clean mathematical form (explicit DFT matrix, dense per-mode solves, n = 0
  mode removed, explicit synthesis, gradient subtraction); it shares no code
  path with the oracle's FFT / index shuffles / zero-mode cache.
* cubic_*: numpy.roots of G^2 x^3 + (1 + G v) x - u = 0 (the real root).
"""
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.npz")


def poisson_rows(n_theta, n_phi, n):
    h = 2 * np.pi / n_phi
    th = (np.arange(n_theta) + 0.5) * h
    a = 1 / h**2 - np.cos(th) / (2 * h * np.sin(th))
    b = -2 / h**2 - n * n / np.sin(th) ** 2
    c = 1 / h**2 + np.cos(th) / (2 * h * np.sin(th))
    if n != 0:
        b[0] += a[0]
        a[0] = 0
        b[-1] += c[-1]
        c[-1] = 0
    else:
        a[:] = 0
        b[:] = 1
        c[:] = 0
    return a, b, c


def dense(a, b, c):
    T = np.diag(b)
    T += np.diag(a[1:], -1)
    T += np.diag(c[:-1], 1)
    return T


def divergence(n_theta, n_phi, u_phi, u_theta):
    h = np.pi / n_theta
    th = (np.arange(n_theta) + 0.5) * h
    vN = np.zeros((n_theta, n_phi))
    vS = np.zeros((n_theta, n_phi))
    vN[1:] = u_theta
    vS[:-1] = u_theta
    uE = np.roll(u_phi, -1, axis=1)
    fac = 1 / (h * np.sin(th))[:, None]
    return fac * (vS * np.sin(th + h / 2)[:, None] - vN * np.sin(th - h / 2)[:, None]) + fac * (uE - u_phi)


def project_dense(n_theta, n_phi, u_phi, u_theta):
    h = np.pi / n_theta
    M = n_phi // 2
    d = divergence(n_theta, n_phi, u_phi, u_theta)
    i = np.arange(n_phi)
    modes = np.arange(-M, M)
    E = np.exp(-1j * np.outer(modes, i) * h)  # [n][i]
    dhat = (d @ E.T) / n_phi  # [theta][n]
    phat = np.zeros_like(dhat)
    for k, n in enumerate(modes):
        if n == 0:
            continue
        phat[:, k] = np.linalg.solve(dense(*poisson_rows(n_theta, n_phi, int(n))), dhat[:, k])
    p = np.real(phat @ np.conj(E))  # sum_n phat_n e^{+i n i h}
    th = (np.arange(n_theta) + 0.5) * h
    ut = u_theta - (p[1:] - p[:-1]) / h
    up = u_phi - (p - np.roll(p, 1, axis=1)) / (h * np.sin(th)[:, None])
    return up, ut, p, dhat


def main():
    rng = np.random.default_rng(20240607)
    g = {}
    lengths = [1, 2, 3, 5, 8, 12, 16, 30, 48, 64, 96, 97, 128]
    g["dft_lengths"] = np.array(lengths)
    for n in lengths:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        g[f"dft_in_{n}"] = x
        g[f"dft_fwd_{n}"] = np.fft.fft(x)
        g[f"dft_inv_{n}"] = np.fft.ifft(x) * n

    cases = [(4, 0), (4, 1), (8, 3), (8, 8), (16, 5), (16, 16)]
    g["tri_cases"] = np.array(cases)
    for nt, n in cases:
        a, b, c = poisson_rows(nt, 2 * nt, n)
        d = rng.standard_normal(nt)
        g[f"tri_a_{nt}_{n}"], g[f"tri_b_{nt}_{n}"], g[f"tri_c_{nt}_{n}"] = a, b, c
        g[f"tri_d_{nt}_{n}"] = d
        g[f"tri_x_{nt}_{n}"] = np.linalg.solve(dense(a, b, c), d)

    sizes = [4, 6, 8]
    g["proj_sizes"] = np.array(sizes)
    for nt in sizes:
        nph = 2 * nt
        up = rng.standard_normal((nt, nph))
        ut = rng.standard_normal((nt - 1, nph))
        oup, out, p, dhat = project_dense(nt, nph, up, ut)
        g[f"proj_uphi_in_{nt}"], g[f"proj_utheta_in_{nt}"] = up, ut
        g[f"proj_uphi_out_{nt}"], g[f"proj_utheta_out_{nt}"] = oup, out
        g[f"proj_p_{nt}"] = p
        g[f"proj_dhat_{nt}"] = dhat
        g[f"proj_div_{nt}"] = divergence(nt, nph, up, ut)

    G = np.array([1e-6, 1e-4, 3e-3, 0.05, 0.3, -0.2, -1e-5, 0.8])
    v = np.array([0.3, -0.7, 1.2, 0.5, -0.9, 0.4, 0.1, 0.6])
    u = np.array([0.2, -1.1, 0.05, 0.9, -0.3, 0.0, 1.5, -0.8])
    roots = []
    for Gi, vi, ui in zip(G, v, u):
        r = np.roots([Gi * Gi, 0.0, 1 + Gi * vi, -ui])
        real = r[np.abs(r.imag) < 1e-9 * max(1, np.abs(r).max())].real
        roots.append(real[np.argmin(np.abs(real - ui / (1 + Gi * vi)))])
    g["cubic_G"], g["cubic_v"], g["cubic_u"], g["cubic_x"] = G, v, u, np.array(roots)

    np.savez_compressed(OUT, **g)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
