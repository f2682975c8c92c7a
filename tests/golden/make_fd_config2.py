"""Central finite differences of the config-2 action at h = 1e-6, in extended precision.

Config 2: honeycomb 6x6 torus (Nx=72), Nt=64, beta=6, U=3, mu=0, EXP hopping, particle-hole
basis, sigmaKappa=-1; phi ~ N(0, U beta/Nt) from RandomState(2000); 64 entries drawn from the
same generator right after phi (as tests/test_gpu_parity.py::test_config2_gradient_check does).

In FP64 a central difference at h = 1e-6 carries the action's rounding divided by h
(|S| ~ 350, so ~1e-4); here S(phi +- h e_i) is evaluated in numpy's extended precision
(x86 80-bit long double, eps ~1e-19), so the difference is accurate to ~1e-10.  The action is
S = -(log det M_p + log det M_h) with (hubbardFermiMatrixExp.cpp:574-612)
    log det M_p = log det(1 + F_p(Nt-1) ... F_p(0)),   F_p(t) = E_p diag(exp(i phi_{t-1})),
    log det M_h = const - i sum(phi) + log det(1 + G(0) ... G(Nt-1)),
                  G(t) = diag(exp(i phi_{t-1})) E_h^{-1},
(E_p = exp(kappa~), E_h^{-1} = exp(-sigmaKappa kappa~) = exp(kappa~) here), so phi at slice t
enters only slice t+1 of each product: with the prefix and suffix products precomputed,
each difference needs two products per species.  The FP64 exponentials are those of the
oracle (expm_sym); both signs of h use the same matrices, so their rounding cancels in the
difference.  Writes fd_config2_h1e-6.npz: idx (64 ints), fd (64 complex: -dS/dphi_i).
Run from the repository root: python tests/golden/make_fd_config2.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from oracle import hubbard_ref as ref  # noqa: E402

LD = np.clongdouble
H = 1e-6


def logdet_ld(a):
    """log det of a complex long-double matrix by LU with partial pivoting (branch of the
    principal log of each pivot; the sum is reduced modulo 2 pi by the caller)."""
    a = a.copy()
    n = a.shape[0]
    total = LD(0)
    for k in range(n):
        p = k + int(np.argmax(np.abs(a[k:, k])))
        if p != k:
            a[[k, p]] = a[[p, k]]
            total += LD(1j * np.pi)
        piv = a[k, k]
        total += np.log(piv)
        if k + 1 < n:
            mult = a[k + 1:, k] / piv
            a[k + 1:, k + 1:] -= np.outer(mult, a[k, k + 1:])
    return total


def main():
    nx, nt, beta, u = 72, 64, 6.0, 3.0
    kt = ref.honeycomb(6) * beta / nt
    ut = u * beta / nt
    rng = np.random.RandomState(2000)
    phi = rng.normal(0, np.sqrt(ut), nx * nt) + 0j
    idx = rng.choice(nx * nt, 64, replace=False)
    hfm = ref.HFMExp(kt, 0.0, -1)
    ep = hfm.exp_kappa[(ref.PARTICLE, False)].astype(LD)
    eh = hfm.exp_kappa[(ref.HOLE, True)].astype(LD)
    ph = np.exp(LD(1j) * phi.astype(LD)).reshape(nt, nx)  # exp(i phi_t)

    def fp(t, col):  # F_p(t) with exp(i phi_{t-1}) given as a vector
        return ep * col[None, :]

    def gh(t, col):  # G(t)
        return col[:, None] * eh

    prev = lambda t: (t - 1) % nt  # noqa: E731
    # particles: B = F(nt-1) ... F(0); L_t = F(t) ... F(0), R_t = F(nt-1) ... F(t+1)
    lp = [None] * nt
    acc = np.eye(nx, dtype=LD)
    for t in range(nt):
        acc = fp(t, ph[prev(t)]) @ acc
        lp[t] = acc
    rp = [None] * nt
    acc = np.eye(nx, dtype=LD)
    for t in range(nt - 1, -1, -1):
        rp[t] = acc
        acc = acc @ fp(t, ph[prev(t)])
    # holes: C = G(0) ... G(nt-1); lh_t = G(0) ... G(t-1), rh_t = G(t+1) ... G(nt-1)
    lh = [None] * nt
    acc = np.eye(nx, dtype=LD)
    for t in range(nt):
        lh[t] = acc
        acc = acc @ gh(t, ph[prev(t)])
    rh = [None] * nt
    acc = np.eye(nx, dtype=LD)
    for t in range(nt - 1, -1, -1):
        rh[t] = acc
        acc = gh(t, ph[prev(t)]) @ acc
    eye = np.eye(nx, dtype=LD)
    fd = np.empty(64, dtype=np.complex128)
    for q, i in enumerate(idx):
        ts, x = divmod(int(i), nx)
        tt = (ts + 1) % nt  # the slice whose phase holds phi_ts
        s = []
        for sign in (1, -1):
            col = ph[ts].copy()
            col[x] = np.exp(LD(1j) * (LD(phi[i]) + sign * LD(H)))
            if tt == 0:
                bp = rp[0] @ fp(0, col)
            else:
                bp = rp[tt] @ fp(tt, col) @ lp[tt - 1]
            bh = lh[tt] @ gh(tt, col) @ rh[tt]
            # S = -(ldp + ldh); the -i sum(phi) of ldh changes by -i (sign h)
            s.append(-(logdet_ld(eye + bp) + logdet_ld(eye + bh) - LD(1j) * sign * LD(H)))
        d = s[0] - s[1]
        d = complex(d.real, float(np.remainder(float(d.imag) + np.pi, 2 * np.pi) - np.pi))
        fd[q] = -d / (2 * H)
        print(q, int(i), fd[q], flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fd_config2_h1e-6.npz")
    np.savez(out, idx=idx, fd=fd)


if __name__ == "__main__":
    main()
