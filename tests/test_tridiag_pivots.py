"""CPU model of the GPU tridiagonal solve's arithmetic (k_tridiag in
kaminogpu_amd/csrc/kamino_kernels.hip): L lanes per mode, Thomas over each
lane's K rows, then parallel cyclic reduction of the lanes' last unknowns,
replayed in numpy with the kernel's exact sequence of operations.

* In FP64 the model equals a dense solve of the oracle's Poisson rows
  (oracle.poisson_rows, after KaminoSolver.cu:117-163) to ~1e-13: the lane
  split and the reduction are an exact reformulation.
* In FP32 with the textbook pivots b_k - a_k ga_{k-1} the low modes lose up to
  ~1e-4 (their systems are nearly singular: diagonal excess n^2/sin^2 against
  off-diagonals ~ n_theta^2), which is why the kernel forms its FP32 pivots
  without subtraction (pivot D_k = c_k + sig_k, sig_k = m_k + a_k tau_{k-1};
  row excesses carried through the reduction).  With those, every mode stays
  within ~1e-6 of the dense FP64 solve.

The numbers come from the model, not from the GPU; the GPU parity tests bound
the kernel's own solve against dense FP64 solves."""
import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32


def lane_solve(a, b, c, f, L, stable, dt=F32):
    """x with T x = f for the rows (a, b, c) of one mode (FP64 inputs, rounded
    to dt as the kernel does), f of shape (2, nT): real and imaginary parts."""
    nT = len(a)
    K = -(-nT // L)
    pad = L * K - nT
    m = -b - a - c  # the kernel forms n^2 / sin^2 directly

    def lanes(v, fill):
        return np.concatenate([v, np.full(pad, fill)]).reshape(L, K).astype(dt)

    cnt = np.clip(nT - np.arange(L) * K, 0, K)
    A_, B_, C_, M_ = lanes(a, 0), lanes(b, 1), lanes(c, 0), lanes(m, 1)
    Fr = np.concatenate([f, np.zeros((2, pad))], 1).reshape(2, L, K).astype(dt)
    al = np.zeros((L, K), dt)
    ga = np.zeros((L, K), dt)
    dr = np.zeros((2, L, K), dt)
    one = dt(1)
    tau = np.ones(L, dt)
    ep_prev = np.zeros(L, dt)
    R = np.zeros(L, dt)
    prod = np.ones(L, dt)
    eE = np.zeros(L, dt)
    for k in range(K):
        act = k < cnt
        a_, b_, c_, m_ = A_[:, k], B_[:, k], C_[:, k], M_[:, k]
        al_prev = al[:, k - 1] if k else -np.ones(L, dt)
        dr_prev = dr[:, :, k - 1] if k else np.zeros((2, L), dt)
        if stable:
            sig = m_ + a_ * tau
            inv = one / (c_ + sig)
            ep = (m_ + a_ * ep_prev) * inv
            al_k = a_ * al_prev * inv
            ga_k = -c_ * inv
            dr_k = (a_ * dr_prev - Fr[:, :, k]) * inv
            tau = np.where(act, sig * inv, tau)
            ep_prev = np.where(act, ep, ep_prev)
            inner = k < cnt - 1
            R = np.where(act & inner, R + ep * prod, R)
            prod = np.where(act & inner, prod * (c_ * inv), prod)
            eE = np.where(act & ~inner, ep, eE)
        else:
            if k == 0:
                inv = one / b_
                al_k = a_ * inv
                dr_k = Fr[:, :, k] * inv
            else:
                inv = one / (b_ - a_ * ga[:, k - 1])
                al_k = -a_ * al_prev * inv
                dr_k = (Fr[:, :, k] - a_ * dr_prev) * inv
            ga_k = c_ * inv
        al[:, k] = np.where(act, al_k, 0)
        ga[:, k] = np.where(act, ga_k, 0)
        dr[:, :, k] = np.where(act, dr_k, 0)
    P = np.zeros((2, L), dt)
    Q = np.zeros(L, dt)
    S = np.ones(L, dt)
    for k in range(K - 2, -1, -1):
        act = k < cnt - 1
        P = np.where(act, dr[:, :, k] - ga[:, k] * P, P)
        Q = np.where(act, -al[:, k] - ga[:, k] * Q, Q)
        S = np.where(act, -ga[:, k] * S, S)
    idx = np.arange(L)
    e = np.maximum(cnt - 1, 0)
    gE, aE, dE = ga[idx, e], al[idx, e], dr[:, idx, e]
    has = idx + 1 < L
    nx = np.minimum(idx + 1, L - 1)
    live = cnt > 0
    A = np.where(live, aE, 0).astype(dt)
    C = np.where(live & has, gE * S[nx], 0).astype(dt)
    D = np.where(live, dE - np.where(has, gE * P[:, nx], 0), 0).astype(dt)
    if stable:
        E = np.where(live, eE + np.abs(gE) * np.where(has, R[nx], one), one).astype(dt)
        B = (np.abs(A) + np.abs(C) + E).astype(dt)
    else:
        B = np.where(live, one + np.where(has, gE * Q[nx], 0), one).astype(dt)
    st = 1
    while st < L:
        okm, okp = idx - st >= 0, idx + st < L
        sm, sp = np.where(okm, idx - st, idx), np.where(okp, idx + st, idx)
        Am, Bm, Cm = np.where(okm, A[sm], 0), np.where(okm, B[sm], 1), np.where(okm, C[sm], 0)
        Ap, Bp, Cp = np.where(okp, A[sp], 0), np.where(okp, B[sp], 1), np.where(okp, C[sp], 0)
        Dm, Dp = np.where(okm, D[:, sm], 0), np.where(okp, D[:, sp], 0)
        k1 = A * (one / Bm.astype(dt))
        k2 = C * (one / Bp.astype(dt))
        nA, nC = (-Am * k1).astype(dt), (-Cp * k2).astype(dt)
        D = (D - Dm * k1 - Dp * k2).astype(dt)
        if stable:
            Em, Ep = np.where(okm, E[sm], 1), np.where(okp, E[sp], 1)
            E = (E + np.abs(k1) * Em + np.abs(k2) * Ep).astype(dt)
            B = (np.abs(nA) + np.abs(nC) + E).astype(dt)
        else:
            B = (B - Cm * k1 - Ap * k2).astype(dt)
        A, C = nA, nC
        st *= 2
    X = D * (one / B)
    xm = np.concatenate([np.zeros((2, 1), dt), X[:, :-1]], 1)
    x = np.zeros((2, L, K), dt)
    v = X.copy()
    for k in range(K - 1, -1, -1):
        v = np.where(k < cnt - 1, dr[:, :, k] - al[:, k] * xm - ga[:, k] * v, v)
        x[:, :, k] = np.where(k < cnt, v, 0)
    return x.reshape(2, L * K)[:, :nT].astype(np.float64)


def dense(a, b, c, f):
    T = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    return np.linalg.solve(T, f.T).T


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


# (n_theta, lanes per mode) as the kernel's tile classes pick them
CASES = [(45, 8), (64, 16), (256, 32), (1024, 64), (2048, 64)]


def rhs(nt, n, kind, rng):
    th = (np.arange(nt) + 0.5) * np.pi / nt
    if kind == "random":
        f = rng.standard_normal((2, nt))
    else:
        s = np.sin(th) ** min(n, 8)
        f = np.stack([s * np.cos(3 * th), s * np.sin(5 * th)])
    return f.astype(F32).astype(np.float64)  # the staged spectrum is FP32


@pytest.mark.parametrize("nt,L", CASES)
def test_lane_split_is_exact_in_fp64(nt, L):
    rng = np.random.default_rng(nt)
    for n in (1, 2, 7, nt // 2, nt):
        a, b, c = O.poisson_rows(nt, 2 * nt, n)
        f = rhs(nt, n, "random", rng)
        for stable in (False, True):
            assert rel(lane_solve(a, b, c, f, L, stable, np.float64), dense(a, b, c, f)) < 1e-11


@pytest.mark.parametrize("nt,L", CASES)
def test_fp32_subtraction_free_pivots(nt, L):
    rng = np.random.default_rng(7 + nt)
    worst_plain, worst_sf = 0.0, 0.0
    for n in (1, 2, 4, 8, 16, 64, nt):
        if n > nt:
            continue
        a, b, c = O.poisson_rows(nt, 2 * nt, n)
        for kind in ("random", "smooth"):
            f = rhs(nt, n, kind, rng)
            x = dense(a, b, c, f)
            e_sf = rel(lane_solve(a, b, c, f, L, True), x)
            e_plain = rel(lane_solve(a, b, c, f, L, False), x)
            worst_sf, worst_plain = max(worst_sf, e_sf), max(worst_plain, e_plain)
            assert e_sf < 1.5e-6, (n, kind, e_sf)
    if nt >= 256:
        # the textbook pivots lose accuracy at the low modes; these do not
        assert worst_plain > 10 * worst_sf
