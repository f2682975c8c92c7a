r"""CPU oracle for the Hubbard fermion-action hot path.

THIS MODULE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or as the timed CPU baseline. The product package (``isle_amd``) never
imports it and has no CPU fallback.

It restates, operation for operation and in complex128 numpy/scipy (LAPACK
``zgetrf``/``zgetrs``/``zgetri`` exactly where the reference calls them), the
reference algorithms of evanberkowitz/isle:

===========================  =====================================================
function                     reference
===========================  =====================================================
``is_bipartite``             src/isle/cpp/lattice.cpp:16-84
``expm_sym``                 src/isle/cpp/math.hpp:242-253
``to_first_log_branch``      src/isle/cpp/math.hpp:168-171
``ilogdet``                  src/isle/cpp/math.hpp:265-293
``HFMExp``                   src/isle/cpp/hubbardFermiMatrixExp.cpp:24-239
``HFMDia``                   src/isle/cpp/hubbardFermiMatrixDia.cpp:25-244
``logdetM``                  hubbardFermiMatrixExp.cpp:572-625, hubbardFermiMatrixDia.cpp:571-597
``solveM``                   hubbardFermiMatrixExp.cpp:645-694, hubbardFermiMatrixDia.cpp:617-673
``force_direct_single_part`` src/isle/cpp/action/hubbardFermiAction.cpp:16-60
``force_direct_square``      src/isle/cpp/action/hubbardFermiAction.cpp:62-108
``qlu`` / ``solve_q``        src/isle/cpp/hubbardFermiMatrixExp.cpp:305-570 (QLU of Q, solveQ, logdetQ)
``logdet_q``                 (same; hubbardFermiMatrixDia.cpp:304-569 repeats it)
``hole_shortcut_possible``   src/isle/cpp/action/hubbardFermiAction.cpp:114-157
``HubbardFermiAction``       src/isle/cpp/action/hubbardFermiAction.cpp:159-298
``HubbardGaugeAction``       src/isle/cpp/action/hubbardGaugeAction.cpp:7-13
``SumAction``                src/isle/cpp/action/sumAction.cpp:46-58
``leapfrog``                 src/isle/cpp/integrator.cpp:9-36
``rk4_flow``                 src/isle/cpp/integrator.cpp:39-195 (k4 with the direction, 3/8-rule beta21 = 1/3)
``binary_select``            src/isle/evolver/selector.py:20-30
``const_step_leapfrog``      src/isle/evolver/leapfrog.py:38-66
===========================  =====================================================

Pinning (what checks each function independently of this restatement) is
listed per function in the tests under ``tests/test_oracle_*.py``: dense
``numpy.linalg.slogdet`` / ``solve`` of the assembled (NxNt)^2 matrix M,
closed forms at phi=0 and uniform phi, the DIRECT_SQUARE (dense Q^{-1})
algorithm, central finite differences, the particle/hole conjugation identity,
leapfrog reversibility, and the isBipartite truth table of the reference's own
test.  The stabilised functions (``logdetM_stable``, ``force_part_stable``)
are pinned by group-size independence and by agreement with the unstabilised
restatement where the latter is accurate.

Field layout everywhere: phi is a complex128 vector of length nx*nt in
time-major order, element (t, x) at index t*nx + x (src/isle/cpp/math.hpp:175-190).
"""

import math

import numpy as np
import scipy.linalg as sla

PARTICLE, HOLE = 0, 1
DIA, EXP = 0, 1
PARTICLE_HOLE, SPIN = 0, 1
DIRECT_SINGLE, DIRECT_SQUARE = 0, 1


# --------------------------------------------------------------------------- #
# lattice                                                                     #
# --------------------------------------------------------------------------- #

def is_bipartite(hopping):
    """2-colouring DFS plus the alternating-label requirement.

    Follows src/isle/cpp/lattice.cpp:16-84: neighbours are visited in increasing
    column order (blaze CompressedMatrix row iteration), recursion is skipped
    when the *current* site is the last row (lattice.cpp:26-28), and even sites
    must be colour A, odd sites colour B (lattice.cpp:44-61).
    """
    h = np.asarray(hopping)
    n = h.shape[0]
    if n == 0:
        return False
    unspecified, lat_a = 0, 1
    lat_b = ~lat_a
    colors = [unspecified] * n
    colors[0] = lat_a
    neigh = [np.nonzero(h[i])[0].tolist() for i in range(n)]

    def descend(cur):
        col = colors[cur]
        for nxt in neigh[cur]:
            if colors[nxt] == unspecified:
                colors[nxt] = ~col
                if cur < n - 1 and not descend(nxt):
                    return False
            elif colors[nxt] == col:
                return False
        return True

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10 * n + 100))
    try:
        bip = descend(0)
    finally:
        sys.setrecursionlimit(old)
    alternating = all((c == lat_a) if i % 2 == 0 else (c == lat_b)
                      for i, c in enumerate(colors))
    return bip and alternating


# --------------------------------------------------------------------------- #
# math helpers                                                                #
# --------------------------------------------------------------------------- #

def expm_sym(mat):
    """exp of a real symmetric matrix via its eigendecomposition (math.hpp:242-253)."""
    w, v = np.linalg.eigh(mat)
    return (v * np.exp(w)) @ v.T


def to_first_log_branch(z):
    """Project Im(z) with IEEE remainder (round-half-even) by 2*pi (math.hpp:168-171)."""
    z = complex(z)
    return complex(z.real, math.remainder(z.imag, 2 * math.pi))


def ilogdet(mat):
    """log det via zgetrf: sum log U_ii + i*pi*(row-swap parity), projected (math.hpp:265-293)."""
    lu, piv = sla.lu_factor(np.array(mat, dtype=np.complex128), check_finite=False)
    neg = (np.count_nonzero(piv != np.arange(len(piv))) % 2) == 1
    res = np.sum(np.log(np.diag(lu).astype(np.complex128)))
    if neg:
        res = res + 1j * math.pi
    return to_first_log_branch(res)


def spacevec(vec, t, nx):
    return vec[t * nx:(t + 1) * nx]


def get_nt(phi, nx):
    if len(phi) % nx != 0:
        raise RuntimeError("Vector dimension does not match, expected a spacetime vector.")
    return len(phi) // nx


# --------------------------------------------------------------------------- #
# fermion matrices                                                            #
# --------------------------------------------------------------------------- #

class HFMExp:
    """EXP discretisation (hubbardFermiMatrixExp.cpp:24-239)."""

    hopping = EXP

    def __init__(self, kappa_tilde, mu_tilde, sigma_kappa):
        k = np.asarray(kappa_tilde, dtype=np.float64)
        self.kappa = k
        self.mu = float(mu_tilde)
        self.sigma_kappa = int(sigma_kappa)
        ident = np.eye(k.shape[0])
        # computeExponential, hubbardFermiMatrixExp.cpp:24-45
        self.exp_kappa = {
            (PARTICLE, False): expm_sym(k - self.mu * ident),
            (PARTICLE, True): expm_sym(-k + self.mu * ident),
            (HOLE, False): expm_sym(self.sigma_kappa * k + self.mu * ident),
            (HOLE, True): expm_sym(-self.sigma_kappa * k - self.mu * ident),
        }
        # _logdetExpKappahInv, hubbardFermiMatrixExp.cpp:60
        self.logdet_exp_kappa_h_inv = ilogdet(self.exp_kappa[(HOLE, True)])

    def nx(self):
        return self.kappa.shape[0]

    def K(self, species):
        return np.eye(self.nx())

    def F(self, tp, phi, species, inv=False):
        """hubbardFermiMatrixExp.cpp:119-141."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        tm1 = nt - 1 if tp == 0 else tp - 1
        sign = 1j if ((species == PARTICLE and not inv) or (species == HOLE and inv)) else -1j
        ph = np.exp(sign * spacevec(phi, tm1, nx))
        e = self.exp_kappa[(species, inv)]
        if inv:
            return ph[:, None] * e
        return e * ph[None, :]

    def M(self, phi, species):
        """Dense (NxNt)^2 M: identity diagonal, -F(tp) at (tp, tp-1), +F(0) at (0, nt-1) (cpp:150-166)."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        m = np.eye(nx * nt, dtype=np.complex128)
        m[0:nx, (nt - 1) * nx:nt * nx] = self.F(0, phi, species)
        for tp in range(1, nt):
            m[tp * nx:(tp + 1) * nx, (tp - 1) * nx:tp * nx] = -self.F(tp, phi, species)
        return m

    def P(self):
        """cpp:180-186."""
        nx = self.nx()
        if self.sigma_kappa == -1:
            return 2 * np.eye(nx)
        return np.eye(nx) + self.exp_kappa[(PARTICLE, False)] @ self.exp_kappa[(HOLE, False)]

    def Tplus(self, tp, phi):
        """cpp:188-193."""
        anti = -1.0 if tp == 0 else 1.0
        return -anti * self.F(tp, phi, PARTICLE, False)

    def Tminus(self, tp, phi):
        """cpp:202-208."""
        nt = get_nt(phi, self.nx())
        anti = -1.0 if tp == nt - 1 else 1.0
        return -anti * self.F((tp + 1) % nt, phi, HOLE, False).T

    def Q(self, phi):
        return _assemble_q(self, phi)


class HFMDia:
    """DIA discretisation (hubbardFermiMatrixDia.cpp:25-244)."""

    hopping = DIA

    def __init__(self, kappa_tilde, mu_tilde, sigma_kappa):
        self.kappa = np.asarray(kappa_tilde, dtype=np.float64)
        self.mu = float(mu_tilde)
        self.sigma_kappa = int(sigma_kappa)

    def nx(self):
        return self.kappa.shape[0]

    def K(self, species):
        """cpp:96-103."""
        ident = np.eye(self.nx())
        if species == PARTICLE:
            return (1 + self.mu) * ident - self.kappa
        return (1 - self.mu) * ident - self.sigma_kappa * self.kappa

    def Kinv(self, species):
        """inverseK (cpp:25-42), guarded by isInvertible (math.hpp:229-234)."""
        k = self.K(species)
        if np.min(np.abs(np.linalg.eigvals(k))) <= 1e-15:
            raise RuntimeError("Matrix K is not invertible, did you forget to multiply kappa by delta?")
        return np.linalg.inv(k)

    def F(self, tp, phi, species, inv=False):
        """Diagonal e^{+-i phi_{tp-1}} (cpp:118-131), returned dense."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        tm1 = nt - 1 if tp == 0 else tp - 1
        if (inv and species == PARTICLE) or (species == HOLE and not inv):
            d = np.exp(-1j * spacevec(phi, tm1, nx))
        else:
            d = np.exp(1j * spacevec(phi, tm1, nx))
        return np.diag(d)

    def M(self, phi, species):
        """cpp:140-160."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        m = np.zeros((nx * nt, nx * nt), dtype=np.complex128)
        k = self.K(species)
        m[0:nx, 0:nx] = k
        m[0:nx, (nt - 1) * nx:nt * nx] = self.F(0, phi, species)
        for tp in range(1, nt):
            m[tp * nx:(tp + 1) * nx, (tp - 1) * nx:tp * nx] = -self.F(tp, phi, species)
            m[tp * nx:(tp + 1) * nx, tp * nx:(tp + 1) * nx] = k
        return m

    def P(self):
        """cpp:170-175."""
        nx, mu, sk, k = self.nx(), self.mu, self.sigma_kappa, self.kappa
        return (2 - mu * mu) * np.eye(nx) - (sk * (1 + mu) + 1 - mu) * k + sk * (k @ k)

    def Tplus(self, tp, phi):
        """cpp:183-195."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        tm1 = nt - 1 if tp == 0 else tp - 1
        anti = -1.0 if tp == 0 else 1.0
        t = (self.sigma_kappa * self.kappa - (1 - self.mu) * np.eye(nx)).astype(np.complex128)
        return (anti * np.exp(1j * spacevec(phi, tm1, nx)))[:, None] * t

    def Tminus(self, tp, phi):
        """cpp:203-214."""
        nx = self.nx()
        nt = get_nt(phi, nx)
        anti = -1.0 if tp == nt - 1 else 1.0
        t = (self.kappa - (1 + self.mu) * np.eye(nx)).astype(np.complex128)
        return t * (anti * np.exp(-1j * spacevec(phi, tp, nx)))[None, :]

    def Q(self, phi):
        return _assemble_q(self, phi)


def _assemble_q(hfm, phi):
    """Q = P on diagonal, T+ at (tp, tp-1), T- at (tp, tp+1), cyclic (Exp cpp:217-233, Dia cpp:217-233)."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    q = np.zeros((nx * nt, nx * nt), dtype=np.complex128)
    p = hfm.P()
    for tp in range(nt):
        q[tp * nx:(tp + 1) * nx, tp * nx:(tp + 1) * nx] = p
        c = (tp + nt - 1) % nt
        q[tp * nx:(tp + 1) * nx, c * nx:(c + 1) * nx] += hfm.Tplus(tp, phi)
        c = (tp + 1) % nt
        q[tp * nx:(tp + 1) * nx, c * nx:(c + 1) * nx] += hfm.Tminus(tp, phi)
    return q


# --------------------------------------------------------------------------- #
# logdetM / solveM                                                            #
# --------------------------------------------------------------------------- #

def _logdetM_exp(hfm, phi, species):
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    if species == PARTICLE:
        # hubbardFermiMatrixExp.cpp:574-591
        b = hfm.F(0, phi, PARTICLE, False)
        for t in range(1, nt):
            b = hfm.F(t, phi, PARTICLE, False) @ b
        b = b + np.eye(nx)
        return to_first_log_branch(ilogdet(b))
    # hubbardFermiMatrixExp.cpp:594-612
    aux = hfm.F(0, phi, HOLE, True)
    for t in range(1, nt):
        aux = aux @ hfm.F(t, phi, HOLE, True)
    aux = aux + np.eye(nx)
    return to_first_log_branch(-float(nt) * hfm.logdet_exp_kappa_h_inv
                               - 1j * np.sum(phi) + ilogdet(aux))


def _logdetM_dia(hfm, phi, species):
    """hubbardFermiMatrixDia.cpp:571-597 (aux = aux*f*k evaluated left to right)."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    k = hfm.K(species)
    aux = hfm.F(0, phi, species, True) @ k
    for t in range(1, nt):
        aux = (aux @ hfm.F(t, phi, species, True)) @ k
    aux = aux + np.eye(nx)
    sgn = 1j if species == PARTICLE else -1j
    return to_first_log_branch(sgn * np.sum(phi) + ilogdet(aux))


def logdetM(hfm, phi, species):
    phi = np.asarray(phi, dtype=np.complex128)
    if hfm.hopping == EXP:
        return _logdetM_exp(hfm, phi, species)
    return _logdetM_dia(hfm, phi, species)


def solveM(hfm, phi, species, rhss):
    """Solve M x = b for each row b of rhss (NRHS x NxNt); EXP cpp:645-694, DIA cpp:617-673."""
    phi = np.asarray(phi, dtype=np.complex128)
    rhss = np.asarray(rhss, dtype=np.complex128)
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    nrhs = rhss.shape[0]
    res = np.zeros((nrhs, nx * nt), dtype=np.complex128)
    k = hfm.K(species)
    partial = []
    if hfm.hopping == EXP:
        partial.append(hfm.F(0, phi, species, True))
        for t in range(1, nt):
            partial.append(partial[t - 1] @ hfm.F(t, phi, species, True))
        res[:, 0:nx] = rhss[:, 0:nx] @ partial[0].T
        for t in range(1, nt):
            res[:, t * nx:(t + 1) * nx] = rhss[:, t * nx:(t + 1) * nx] @ partial[t].T \
                + res[:, (t - 1) * nx:t * nx]
    else:
        finv = hfm.F(0, phi, species, True)
        res[:, 0:nx] = rhss[:, 0:nx] @ finv.T
        partial.append(finv @ k)
        for t in range(1, nt):
            finv = hfm.F(t, phi, species, True)
            p = partial[t - 1] @ finv
            res[:, t * nx:(t + 1) * nx] = rhss[:, t * nx:(t + 1) * nx] @ p.T \
                + res[:, (t - 1) * nx:t * nx]
            partial.append(p @ k)
    partial[nt - 1] = partial[nt - 1] + np.eye(nx)
    lus = [sla.lu_factor(p, check_finite=False) for p in partial]
    # getrs 'T' on row-major data == solve x A^T... i.e. rows x satisfy A x_row^T = b_row^T
    last = sla.lu_solve(lus[nt - 1], res[:, (nt - 1) * nx:nt * nx].T, trans=0, check_finite=False).T
    res[:, (nt - 1) * nx:nt * nx] = last
    for t in range(nt - 1):
        mat = res[:, t * nx:(t + 1) * nx] - last
        res[:, t * nx:(t + 1) * nx] = sla.lu_solve(lus[t], mat.T, trans=0, check_finite=False).T
    return res


# --------------------------------------------------------------------------- #
# forces                                                                      #
# --------------------------------------------------------------------------- #

def force_direct_single_part(hfm, phi, k, species):
    """hubbardFermiAction.cpp:16-60 (returns the force without the -i prefactor)."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    if nt < 2:
        raise ValueError("nt < 2 in HubbardFermiAction algorithm DIRECT_SINGLE not supported")
    lefts = [hfm.F(nt - 1, phi, species, True) @ k]
    for t in range(nt - 2, 0, -1):
        lefts.append(hfm.F(t, phi, species, True) @ k @ lefts[-1])
    ainv = hfm.F(0, phi, species, True) @ k @ lefts[-1]
    right = np.linalg.inv(np.eye(nx) + ainv)
    force = np.zeros(nx * nt, dtype=np.complex128)
    force[(nt - 1) * nx:] = np.einsum("ij,ji->i", ainv, right)
    for tau in range(nt - 1):
        right = right @ hfm.F(tau, phi, species, True) @ k
        force[tau * nx:(tau + 1) * nx] = np.einsum("ij,ji->i", lefts[nt - 1 - tau - 1], right)
    return force


def force_direct_square(hfm, phi):
    """Dense Q^{-1} force (hubbardFermiAction.cpp:62-108), small sizes only."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    qinv = np.linalg.inv(hfm.Q(phi))
    force = np.zeros(nx * nt, dtype=np.complex128)

    def blk(r, c):
        return qinv[r * nx:(r + 1) * nx, c * nx:(c + 1) * nx]

    for tau in range(nt):
        tp1 = (tau + 1) % nt
        tplus = hfm.Tplus(tp1, phi)
        tminus = hfm.Tminus(tau, phi)
        if hfm.hopping == DIA:
            f = 1j * np.einsum("ij,ji->i", tplus, blk(tau, tp1))
            f -= 1j * np.einsum("ij,ji->i", blk(tp1, tau), tminus)
        else:
            f = 1j * np.einsum("ij,ji->i", blk(tau, tp1), tplus)
            f -= 1j * np.einsum("ij,ji->i", tminus, blk(tp1, tau))
        force[tau * nx:(tau + 1) * nx] = f
    return force


def logdetQ_dense(hfm, phi):
    """-eval of DIRECT_SQUARE; the reference test compares logdetQ to dense logdet(Q)."""
    return ilogdet(hfm.Q(phi))


class QLUParts:
    """Blocks of the LU of the block-cyclic tridiagonal Q (hubbardFermiMatrixExp.hpp QLU):
    dinv[i] = d_i^{-1} (nt), u, l (nt-1), v, h (nt-2); lists of nx x nx complex arrays."""

    def __init__(self):
        self.dinv, self.u, self.l, self.v, self.h = [], [], [], [], []


def qlu(hfm, phi):
    """getQLU (hubbardFermiMatrixExp.cpp:482-492): nt == 1 (:368-383), nt == 2 (:386-416), nt > 2
    (:419-479); the Dia file repeats it (hubbardFermiMatrixDia.cpp:367-491)."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    p = hfm.P().astype(np.complex128)
    lu = QLUParts()
    if nt == 1:
        lu.dinv.append(np.linalg.inv(p + hfm.Tplus(0, phi) + hfm.Tminus(0, phi)))
        return lu
    if nt == 2:
        lu.dinv.append(np.linalg.inv(p))
        lu.l.append((hfm.Tplus(1, phi) + hfm.Tminus(1, phi)) @ lu.dinv[0])
        lu.u.append(hfm.Tplus(0, phi) + hfm.Tminus(0, phi))
        lu.dinv.append(np.linalg.inv(p - lu.l[0] @ lu.u[0]))
        return lu
    lu.dinv.append(np.linalg.inv(p))
    lu.u.append(hfm.Tminus(0, phi))
    lu.l.append(hfm.Tplus(1, phi) @ lu.dinv[0])
    lu.v.append(hfm.Tplus(0, phi))
    lu.h.append(hfm.Tminus(nt - 1, phi) @ lu.dinv[0])
    for i in range(1, nt - 2):
        lu.dinv.append(np.linalg.inv(p - lu.l[i - 1] @ lu.u[i - 1]))
        lu.l.append(hfm.Tplus(i + 1, phi) @ lu.dinv[i])
        lu.h.append(-lu.h[i - 1] @ lu.u[i - 1] @ lu.dinv[i])
        lu.v.append(-lu.l[i - 1] @ lu.v[i - 1])
        lu.u.append(hfm.Tminus(i, phi))
    lu.dinv.append(np.linalg.inv(p - lu.l[nt - 3] @ lu.u[nt - 3]))
    lu.u.append(hfm.Tminus(nt - 2, phi) - lu.l[nt - 3] @ lu.v[nt - 3])
    lu.l.append((hfm.Tplus(nt - 1, phi) - lu.h[nt - 3] @ lu.u[nt - 3]) @ lu.dinv[nt - 2])
    last = p - lu.l[nt - 2] @ lu.u[nt - 2]
    for i in range(nt - 2):
        last = last - lu.h[i] @ lu.v[i]
    lu.dinv.append(np.linalg.inv(last))
    return lu


def qlu_reconstruct(lu):
    """L U assembled back into Q (hubbardFermiMatrixExp.cpp:305-360), nt >= 3."""
    nt = len(lu.dinv)
    if nt < 3:
        raise ValueError("reconstruction needs nt >= 3")
    nx = lu.dinv[0].shape[0]
    q = np.zeros((nx * nt, nx * nt), dtype=np.complex128)

    def blk(r, c):
        return q[r * nx:(r + 1) * nx, c * nx:(c + 1) * nx]
    d = [np.linalg.inv(x) for x in lu.dinv]
    blk(0, 0)[:] = d[0]
    blk(0, 1)[:] = lu.u[0]
    blk(0, nt - 1)[:] = lu.v[0]
    for i in range(1, nt - 2):
        blk(i, i)[:] = d[i] + lu.l[i - 1] @ lu.u[i - 1]
        blk(i, i - 1)[:] = lu.l[i - 1] @ d[i - 1]
        blk(i, i + 1)[:] = lu.u[i]
        blk(i, nt - 1)[:] = lu.l[i - 1] @ lu.v[i - 1] + lu.v[i]
    blk(nt - 2, nt - 2)[:] = d[nt - 2] + lu.l[nt - 3] @ lu.u[nt - 3]
    blk(nt - 2, nt - 3)[:] = lu.l[nt - 3] @ d[nt - 3]
    blk(nt - 2, nt - 1)[:] = lu.l[nt - 3] @ lu.v[nt - 3] + lu.u[nt - 2]
    blk(nt - 1, 0)[:] = lu.h[0] @ d[0]
    for i in range(1, nt - 2):
        blk(nt - 1, i)[:] = lu.h[i - 1] @ lu.u[i - 1] + lu.h[i] @ d[i]
    blk(nt - 1, nt - 2)[:] = lu.h[nt - 3] @ lu.u[nt - 3] + lu.l[nt - 2] @ d[nt - 2]
    last = d[nt - 1] + lu.l[nt - 2] @ lu.u[nt - 2]
    for i in range(nt - 2):
        last = last + lu.h[i] @ lu.v[i]
    blk(nt - 1, nt - 1)[:] = last
    return q


def solve_q(lu, rhs):
    """Q x = rhs by forward then backward substitution with the QLU blocks
    (hubbardFermiMatrixExp.cpp:500-537)."""
    nt = len(lu.dinv)
    nx = lu.dinv[0].shape[0]
    b = np.asarray(rhs, dtype=np.complex128).reshape(nt, nx)
    y = np.empty_like(b)
    y[0] = b[0]
    for i in range(1, nt - 1):
        y[i] = b[i] - lu.l[i - 1] @ y[i - 1]
    if nt > 1:
        y[nt - 1] = b[nt - 1] - lu.l[nt - 2] @ y[nt - 2]
        for j in range(nt - 2):
            y[nt - 1] -= lu.h[j] @ y[j]
    x = np.empty_like(b)
    x[nt - 1] = lu.dinv[nt - 1] @ y[nt - 1]
    if nt > 1:
        x[nt - 2] = lu.dinv[nt - 2] @ (y[nt - 2] - lu.u[nt - 2] @ x[nt - 1])
        for i in range(nt - 3, -1, -1):
            x[i] = lu.dinv[i] @ (y[i] - lu.u[i] @ x[i + 1] - lu.v[i] @ x[nt - 1])
    return x.reshape(-1)


def logdet_q(lu):
    """log det Q = -sum_i log det d_i^{-1}, first branch (hubbardFermiMatrixExp.cpp:540-570)."""
    return to_first_log_branch(-sum(ilogdet(x) for x in lu.dinv))


# --------------------------------------------------------------------------- #
# actions                                                                     #
# --------------------------------------------------------------------------- #

def hole_shortcut_possible(basis, hopping, mu_tilde, sigma_kappa):
    """hubbardFermiAction.cpp:114-157."""
    if basis == SPIN:
        return False
    return is_bipartite(hopping) and mu_tilde == 0.0 and sigma_kappa == +1


class HubbardFermiAction:
    """hubbardFermiAction.hpp:96-142 with eval/force of hubbardFermiAction.cpp:159-298."""

    def __init__(self, kappa_tilde, mu_tilde, sigma_kappa, hopping=DIA,
                 basis=PARTICLE_HOLE, algorithm=DIRECT_SINGLE, allow_shortcut=False):
        cls = HFMExp if hopping == EXP else HFMDia
        self.hfm = cls(kappa_tilde, mu_tilde, sigma_kappa)
        self.hopping, self.basis, self.algorithm = hopping, basis, algorithm
        self.kp = self.hfm.K(PARTICLE)
        self.kh = self.hfm.K(HOLE)
        self.shortcut = bool(allow_shortcut) and hole_shortcut_possible(
            basis, kappa_tilde, mu_tilde, sigma_kappa)

    def eval(self, phi):
        phi = np.asarray(phi, dtype=np.complex128)
        if self.algorithm == DIRECT_SQUARE:
            if self.basis == PARTICLE_HOLE:
                return -to_first_log_branch(logdetQ_dense(self.hfm, phi))
            return -to_first_log_branch(logdetQ_dense(self.hfm, -1j * phi))
        if self.basis == PARTICLE_HOLE:
            ldp = logdetM(self.hfm, phi, PARTICLE)
            if self.shortcut:
                return -to_first_log_branch(ldp + np.conj(ldp))
            return -to_first_log_branch(ldp + logdetM(self.hfm, phi, HOLE))
        aux = -1j * phi
        return -to_first_log_branch(logdetM(self.hfm, aux, PARTICLE)
                                    + logdetM(self.hfm, aux, HOLE))

    def force(self, phi):
        phi = np.asarray(phi, dtype=np.complex128)
        if self.algorithm == DIRECT_SQUARE:
            if self.basis == PARTICLE_HOLE:
                return force_direct_square(self.hfm, phi)
            return -1j * force_direct_square(self.hfm, -1j * phi)
        if self.basis == PARTICLE_HOLE:
            fp = force_direct_single_part(self.hfm, phi, self.kp, PARTICLE)
            if self.shortcut:
                return -1j * (fp - np.conj(fp))
            return -1j * (fp - force_direct_single_part(self.hfm, phi, self.kh, HOLE))
        aux = -1j * phi
        return (force_direct_single_part(self.hfm, aux, self.kh, HOLE)
                - force_direct_single_part(self.hfm, aux, self.kp, PARTICLE))


class HubbardGaugeAction:
    """S = (phi, phi)/(2 Ut) with the bilinear product; F = -phi/Ut (hubbardGaugeAction.cpp:7-13)."""

    def __init__(self, utilde):
        self.utilde = float(utilde)

    def eval(self, phi):
        phi = np.asarray(phi, dtype=np.complex128)
        return complex(np.dot(phi, phi) / 2.0 / self.utilde)

    def force(self, phi):
        return -np.asarray(phi, dtype=np.complex128) / self.utilde


class SumAction:
    """sumAction.cpp:46-58."""

    def __init__(self, *actions):
        self.actions = list(actions)

    def eval(self, phi):
        return sum((a.eval(phi) for a in self.actions), 0j)

    def force(self, phi):
        res = np.zeros(len(phi), dtype=np.complex128)
        for a in self.actions:
            res = res + a.force(phi)
        return res


def leapfrog(phi, pi, action, length, nsteps, direction=1.0):
    """integrator.cpp:9-36: half kick with Re(force), nsteps drifts, final half kick, eval."""
    eps = direction * length / float(nsteps)
    pi_out = np.asarray(pi, dtype=np.complex128) + np.real(action.force(phi)) * (eps / 2)
    phi_out = np.asarray(phi, dtype=np.complex128) + pi_out * eps
    for _ in range(nsteps - 1):
        pi_out = pi_out + np.real(action.force(phi_out)) * eps
        phi_out = phi_out + pi_out * eps
    pi_out = pi_out + np.real(action.force(phi_out)) * (eps / 2)
    return phi_out, pi_out, action.eval(phi_out)


# Stage tables of integrator.cpp:39-70: classic RK4 (n = 0) and the 3/8 rule (n = 1):
# weights, then beta21, beta31, beta32, beta41, beta42, beta43.  The reference's 3/8 rule
# has beta21 = 1/30 (integrator.cpp:63); the textbook rule has 1/3.
_RK4 = {"reference": {0: ((1 / 6, 1 / 3, 1 / 3, 1 / 6), (0.5, 0.0, 0.5, 0.0, 0.0, 1.0)),
                      1: ((0.125, 0.375, 0.375, 0.125), (1 / 30, -1 / 3, 1.0, 1.0, -1.0, 1.0))},
        "textbook": {0: ((1 / 6, 1 / 3, 1 / 3, 1 / 6), (0.5, 0.0, 0.5, 0.0, 0.0, 1.0)),
                     1: ((0.125, 0.375, 0.375, 0.125), (1 / 3, -1 / 3, 1.0, 1.0, -1.0, 1.0))}}


def _rk4_step(phi, action, eps, direction, n, variant="reference"):
    """integrator.cpp:77-98: k_i = -eps direction conj(force(...)) for stages 1-3; stage 4
    of the reference uses -eps without the direction (integrator.cpp:96).  The "textbook"
    variant keeps the direction in stage 4 and uses the 3/8 rule's beta21 = 1/3."""
    w, b = _RK4[variant][n]
    e = eps * direction
    e4 = eps if variant == "reference" else e
    k1 = -e * np.conj(action.force(phi))
    k2 = -e * np.conj(action.force(phi + b[0] * k1))
    k3 = -e * np.conj(action.force(phi + b[1] * k1 + b[2] * k2))
    k4 = -e4 * np.conj(action.force(phi + b[3] * k1 + b[4] * k2 + b[5] * k3))
    out = phi + w[0] * k1 + w[1] * k2 + w[2] * k3 + w[3] * k4
    return out, action.eval(out)


def rk4_flow(phi, action, flow_time, step, act_val=None, n=0, direction=1.0, attenuation=0.9,
             threshold=1e-8, min_step=None, tolerance=1e-3, variant="reference"):
    """integrator.cpp:144-193: adaptive RK4 holomorphic flow; a step is accepted when
    |e^{i (Im S_old - Im S_new)} - 1| <= tolerance, the step size shrinks by
    attenuation (tol/err)^(1/5) (not below min_step) after a rejection and grows by
    min(max(attenuation (tol/err)^(1/4), 1), 2) after an error below threshold*tol.
    Returns (phi, action value, flow time reached).  ``variant`` selects the stage
    coefficients of _rk4_step."""
    phi = np.asarray(phi, dtype=np.complex128)
    if act_val is None:
        act_val = action.eval(phi)
    if min_step is None:
        min_step = max(step / 1000.0, 1e-12)
    t = 0.0
    while t < flow_time:
        if t + step > flow_time:
            step = flow_time - t
            if step < min_step:
                break
        new_phi, new_act = _rk4_step(phi, action, step, direction, n, variant)
        err = abs(np.exp(1j * (np.imag(act_val) - np.imag(new_act))) - 1.0)
        if err > tolerance:
            if step == min_step:
                break
            step = max(step * attenuation * (tolerance / err) ** 0.2, min_step)
        else:
            t += step
            phi, act_val = new_phi, new_act
            if err < threshold * tolerance:
                grow = 2.0 if err == 0 else min(max(attenuation * (tolerance / err) ** 0.25, 1.0), 2.0)
                step *= grow
    return phi, act_val, t


def binary_select(rng, energy0, energy1):
    """selector.py:20-30; the uniform is drawn only when deltaE >= 0 (short circuit)."""
    delta = np.real(energy1 - energy0)
    return 1 if delta < 0 or np.exp(-delta) > rng.uniform(0, 1) else 0


def const_step_leapfrog(rng, action, phi, act_val, length, nstep, transform=None, log_weights=None):
    """One ConstStepLeapfrog.evolve (leapfrog.py:38-66), with an optional transform
    (transform.py:73-101): the MD field is transform.backward(phi); log det J of the start
    is log_weights["logdetJ"] when given, else the one backward returns; after leapfrog,
    transform.forward gives the Monte Carlo field, its action and log det J.
    E0 = act_val + sum of the other log weights (+ log det J) + |pi|^2/2,
    E1 = act1 + log det J1 + |pi1|^2/2.

    Returns (phi_new, act_val_new, traj_point, energy0, energy1, log_weights_new)."""
    weights = dict(log_weights or {})
    if transform is None:
        phi_md = phi
    else:
        phi_md, logdet0 = transform.backward(phi)
        weights.setdefault("logdetJ", logdet0)
    pi = rng.normal(0, 1, len(phi)) + 0j
    phi1, pi1, act1 = leapfrog(phi_md, pi, action, length, nstep)
    logdet1 = 0
    if transform is not None:
        phi1, act1, logdet1 = transform.forward(phi1, act1)
    energy0 = act_val + sum(weights.values()) + np.linalg.norm(pi) ** 2 / 2
    energy1 = act1 + logdet1 + np.linalg.norm(pi1) ** 2 / 2
    tp = binary_select(rng, energy0, energy1)
    if tp == 1:
        return phi1, act1, 1, energy0, energy1, ({} if transform is None else {"logdetJ": logdet1})
    return np.asarray(phi), act_val, 0, energy0, energy1, weights


class ConstantShift:
    """constantShift.py:39-58: MC field = MD field + shift; the action value carried is
    the action of the MD field; log det J = 0."""

    def __init__(self, shift, action):
        self.shift = np.asarray(shift, dtype=np.complex128)
        self.action = action

    def forward(self, phi, act_val):
        return phi + self.shift, self.action.eval(phi), 0.0

    def backward(self, phi):
        return phi - self.shift, 0.0


def two_pi_jumps(rng, action, phi, act_val, n_batches, batch_size):
    """One TwoPiJumps.evolve (hubbard.py:58-80) with the action evaluated in full: per
    batch, sites = rng.choice(N, batch_size, replace=False), shifts =
    rng.choice((-2 pi, 2 pi), batch_size, replace=True), then binary_select on the
    action values.  Returns (phi, act_val, decisions); phi is the field after the last
    accepted batch (the reference's hubbard.py:75-77 returns the last proposal even when it
    was rejected; that defect is not restated)."""
    phi = np.asarray(phi, dtype=np.complex128).copy()
    decisions = []
    for _ in range(n_batches):
        sites = rng.choice(phi.size, batch_size, replace=False)
        shifts = rng.choice((-2 * np.pi, 2 * np.pi), batch_size, replace=True)
        new = phi.copy()
        new[sites] += shifts
        new_act = action.eval(new)
        if binary_select(rng, act_val, new_act):
            phi, act_val = new, new_act
            decisions.append(1)
        else:
            decisions.append(0)
    return phi, act_val, decisions


def uniform_jump(rng, action, phi, act_val, nt):
    """One UniformJump.evolve (hubbard.py:160-172): shift = rng.choice((-1, 1)) * 2 pi / Nt
    at every site, then binary_select on the action values (evaluated in full).
    Returns (phi, act_val, decision)."""
    new = np.asarray(phi, dtype=np.complex128) + rng.choice((-1, 1)) * 2 * np.pi / nt
    new_act = action.eval(new)
    if binary_select(rng, act_val, new_act):
        return new, new_act, 1
    return np.asarray(phi, dtype=np.complex128), act_val, 0


# --------------------------------------------------------------------------- #
# stabilised FP64 restatement (grouped block-cyclic dense LU)                  #
# --------------------------------------------------------------------------- #

def _finv_group_products(hfm, phi, species, g):
    """P_j = prod_{t=jg}^{jg+g-1} F^{-1}(t) K  (inverse form, left to right)."""
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    if nt % g != 0:
        raise ValueError("group size must divide nt")
    k = hfm.K(species)
    out = []
    for j in range(nt // g):
        p = hfm.F(j * g, phi, species, True) @ k
        for t in range(j * g + 1, (j + 1) * g):
            p = p @ hfm.F(t, phi, species, True) @ k
        out.append(p)
    return out


def _reduced_matrix(ps):
    """R with I on the diagonal, -P_j at (j, j+1) and +P_{n-1} at (n-1, 0); det R = det(1 + P_0...P_{n-1})."""
    n = len(ps)
    nx = ps[0].shape[0]
    r = np.eye(n * nx, dtype=np.complex128)
    if n == 1:
        return r + ps[0]
    for j in range(n - 1):
        r[j * nx:(j + 1) * nx, (j + 1) * nx:(j + 2) * nx] = -ps[j]
    r[(n - 1) * nx:, 0:nx] = ps[n - 1]
    return r


def logdet_one_plus_ainv_stable(hfm, phi, species, g):
    """log det(1 + A^{-1}) with A^{-1} = prod_t F^{-1}(t) K, by GEPP on the grouped block-cyclic matrix."""
    return ilogdet(_reduced_matrix(_finv_group_products(hfm, phi, species, g)))


def logdetM_stable(hfm, phi, species, g=8):
    """Same value as logdetM (mod 2 pi i) from the stabilised grouped factorisation.

    log det M = log det(1 + A^{-1}) + log det A, and log det A = -sum_t log det(F^{-1}(t) K):
    EXP particle: -Nt*logdet(e^{-k+mu}) + i*sum(phi); EXP hole: -Nt*logdet(e^{-s k-mu}) - i*sum(phi);
    DIA: +-i*sum(phi) (the K factors are part of the product, hubbardFermiMatrixDia.cpp:588-593).
    """
    phi = np.asarray(phi, dtype=np.complex128)
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    ld = logdet_one_plus_ainv_stable(hfm, phi, species, g)
    if hfm.hopping == EXP:
        if species == PARTICLE:
            extra = -nt * ilogdet(hfm.exp_kappa[(PARTICLE, True)]) + 1j * np.sum(phi)
        else:
            extra = -nt * hfm.logdet_exp_kappa_h_inv - 1j * np.sum(phi)
    else:
        extra = (1j if species == PARTICLE else -1j) * np.sum(phi)
    return to_first_log_branch(ld + extra)


def force_part_stable(hfm, phi, species, g=8):
    """forceDirectSinglePart from the stabilised grouped factorisation.

    force[tau] = diag(1 - Gt_tau), Gt_tau = (1 + A^{-1}_tau)^{-1}, with A^{-1}_tau the cyclic
    product of F^{-1}(t)K starting at t = tau+1.  Gt at tau = jg-1 is the j-th diagonal block of
    R^{-1}; inside a group Gt is wrapped with Gt_{tau+1} = X^{-1} Gt_tau X, X = F^{-1}(tau+1) K
    (at most g-1 wraps, condition number of a group).
    """
    phi = np.asarray(phi, dtype=np.complex128)
    nx = hfm.nx()
    nt = get_nt(phi, nx)
    k = hfm.K(species)
    ps = _finv_group_products(hfm, phi, species, g)
    n = len(ps)
    rinv = np.linalg.inv(_reduced_matrix(ps))
    force = np.zeros(nx * nt, dtype=np.complex128)
    for j in range(n):
        gt = rinv[j * nx:(j + 1) * nx, j * nx:(j + 1) * nx] if n > 1 else rinv
        tau = (j * g - 1) % nt
        force[tau * nx:(tau + 1) * nx] = 1 - np.diag(gt)
        for _ in range(g - 1):
            tau = (tau + 1) % nt
            x = hfm.F(tau, phi, species, True) @ k
            gt = np.linalg.solve(x, gt @ x)
            force[tau * nx:(tau + 1) * nx] = 1 - np.diag(gt)
    return force


# --------------------------------------------------------------------------- #
# synthetic lattices (configuration generators)                               #
# --------------------------------------------------------------------------- #

def honeycomb(l_i, l_j=None, periodic_j=True, kappa=1.0):
    """Honeycomb with sites A(i,j)=2(i*Lj+j), B=A+1; bonds A(i,j)-B(i,j), A(i,j)-B(i-1,j), A(i,j)-B(i,j-1).

    Periodic in i; periodic in j unless ``periodic_j`` is False (cylinder)."""
    l_j = l_i if l_j is None else l_j
    nx = 2 * l_i * l_j
    h = np.zeros((nx, nx))
    for i in range(l_i):
        for j in range(l_j):
            a = 2 * (i * l_j + j)
            for ii, jj in ((i, j), ((i - 1) % l_i, j), (i, j - 1)):
                if jj < 0:
                    if not periodic_j:
                        continue
                    jj %= l_j
                b = 2 * (ii * l_j + jj) + 1
                h[a, b] = h[b, a] = kappa
    return h
