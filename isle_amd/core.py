"""Fermion matrices, logdetM and solveM on the GPU.

Python mirror of the reference's pybind11 surface for this path
(src/isle/cpp/bind/bind_hubbardFermiMatrix.cpp:12-72, src/isle/cpp_wrappers.py:12-37):
``Species``, ``HubbardFermiMatrixExp``, ``HubbardFermiMatrixDia``, ``logdetM``,
``solveM`` and ``Vector``/``CDVector``.  All arithmetic runs in libisle_hfa.so.
"""
import ctypes
import enum
import math

import numpy as np
import torch

from . import _lib
from .lattice import Lattice, isBipartite


class Species(enum.IntEnum):
    """Species::PARTICLE / Species::HOLE (src/isle/cpp/species.hpp:10-13)."""

    PARTICLE = 0
    HOLE = 1

    @staticmethod
    def values():
        return iter((Species.PARTICLE, Species.HOLE))


def Vector(x, dtype=None):
    """Copy a 1-D array into a device tensor (complex128 for complex input, else float64).

    Like isle.Vector (cpp_wrappers.py:12-37) it copies; a buffer of another layout
    is accepted as long as it is one-dimensional."""
    t = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    if t.dim() != 1:
        raise RuntimeError("Incompatible buffer format: expected a one-dimensional array")
    if dtype is None:
        dtype = torch.complex128 if t.is_complex() else torch.float64
    return t.to(device=_lib.device(), dtype=dtype).clone()


def CDVector(x):
    return Vector(x, torch.complex128)


def as_field(phi, nx=None):
    """Contiguous complex128 device tensor for a field (no copy if already one)."""
    if isinstance(phi, torch.Tensor):
        t = phi
    else:
        t = torch.as_tensor(np.asarray(phi))
    t = t.to(device=_lib.device(), dtype=torch.complex128)
    if not t.is_contiguous():
        t = t.contiguous()
    if nx is not None and t.shape[-1] % nx != 0:
        raise RuntimeError("Vector dimension does not match, expected a spacetime vector.")
    return t


def to_first_log_branch(z):
    """Im(z) reduced by the IEEE remainder to [-pi, pi] (src/isle/cpp/math.hpp:168-171)."""
    z = complex(z)
    return complex(z.real, math.remainder(z.imag, 2 * math.pi))


class _HFMBase:
    """Owns a plan in libisle_hfa.so (device slice matrices, cached workspaces)."""

    HOPPING = None

    def __init__(self, *args):
        # (kappaTilde, muTilde, sigmaKappa) or (lat, beta, muTilde, sigmaKappa)
        # (bind_hubbardFermiMatrix.cpp:26-27)
        if len(args) == 4:
            lat, beta, mu_tilde, sigma_kappa = args
            kappa = np.asarray(lat.hopping(), dtype=np.float64) * beta / lat.nt()
        elif len(args) == 3:
            kappa, mu_tilde, sigma_kappa = args
            kappa = kappa.hopping() if isinstance(kappa, Lattice) else kappa
            kappa = np.ascontiguousarray(np.asarray(kappa, dtype=np.float64))
        else:
            raise TypeError("expected (kappaTilde, muTilde, sigmaKappa) or (lat, beta, muTilde, sigmaKappa)")
        if kappa.ndim != 2 or kappa.shape[0] != kappa.shape[1]:
            raise ValueError("Hopping matrix is not square.")
        self._kappa = np.ascontiguousarray(kappa)
        self._mu = float(mu_tilde)
        self._sigma_kappa = int(sigma_kappa)
        if self._sigma_kappa not in (-1, 1):
            import logging
            logging.getLogger(type(self).__name__).warning("sigmaKappa should be either -1 or +1.")
        if self._sigma_kappa == 1 and not isBipartite(self._kappa):
            import logging
            logging.getLogger(type(self).__name__).warning(
                "sigmaKappa should be -1 because the lattice is not bipartite.")
        _lib.device()
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.isle_hfm_create(ctypes.byref(h), self.HOPPING, kappa.shape[0], _lib.dptr(self._kappa),
                                       self._mu, self._sigma_kappa))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().isle_hfm_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        _lib.check(_lib.load().isle_hfm_set_stream(self._h, _lib.stream_ptr()))
        return self._h

    def nx(self):
        return self._kappa.shape[0]

    def pin_tile(self, kind, tile):
        """Launch shape `tile` for every step of launch kind `kind` (isle_hfm_set_tile; -1:
        the autotuner's choice again), for reproducible timings and for tests that cover
        every shape.  HIP graphs captured before keep their shapes."""
        _lib.check(_lib.load().isle_hfm_set_tile(self.handle, int(kind), int(tile)))

    def kappaTilde(self):
        return self._kappa.copy()

    def muTilde(self):
        return self._mu

    def sigmaKappa(self):
        return self._sigma_kappa

    def _slice(self, species, inv):
        out = np.empty((self.nx(), self.nx()))
        _lib.check(_lib.load().isle_hfm_slice_matrix(self.handle, int(species), int(bool(inv)), _lib.dptr(out)))
        return out

    def F(self, tp, phi, species, inv=False):
        """Dense device matrix F(tp) (hubbardFermiMatrixExp.cpp:119-141 / Dia.cpp:118-131)."""
        nx = self.nx()
        phi = as_field(phi, nx)
        nt = phi.shape[-1] // nx
        out = torch.empty((nx, nx), dtype=torch.complex128, device=phi.device)
        _lib.check(_lib.load().isle_hfm_F(self.handle, int(tp), _lib.ptr(phi), nt, int(species), int(bool(inv)),
                                          _lib.ptr(out)))
        return out

    def M(self, phi, species):
        """Dense (NxNt)^2 fermion matrix assembled on the device from F blocks.

        EXP: hubbardFermiMatrixExp.cpp:150-166; DIA: hubbardFermiMatrixDia.cpp:140-160.
        Intended for tests and small sizes."""
        nx = self.nx()
        phi = as_field(phi, nx)
        nt = phi.shape[-1] // nx
        dev = phi.device
        m = torch.zeros((nx * nt, nx * nt), dtype=torch.complex128, device=dev)
        k = torch.as_tensor(self.K(species), device=dev).to(torch.complex128)
        for tp in range(nt):
            m[tp * nx:(tp + 1) * nx, tp * nx:(tp + 1) * nx] = k
        m[0:nx, (nt - 1) * nx:nt * nx] += self.F(0, phi, species)
        for tp in range(1, nt):
            m[tp * nx:(tp + 1) * nx, (tp - 1) * nx:tp * nx] -= self.F(tp, phi, species)
        return m

    def Q(self, phi):
        """Dense (NxNt)^2 Q = M_p M_h^T from its blocks: P on the diagonal, Tplus(tp) at
        (tp, tp-1), Tminus(tp) at (tp, tp+1), cyclic (hubbardFermiMatrixExp.cpp:217-240,
        hubbardFermiMatrixDia.cpp:222-245).  Intended for tests and small sizes."""
        nx = self.nx()
        phi = as_field(phi, nx)
        nt = phi.shape[-1] // nx
        q = torch.zeros((nt, nx, nt, nx), dtype=torch.complex128, device=phi.device)
        p = torch.as_tensor(self.P(), device=phi.device).to(torch.complex128)
        for tp in range(nt):
            q[tp, :, tp] += p
            q[tp, :, (tp + nt - 1) % nt] += self.Tplus(tp, phi)
            q[tp, :, (tp + 1) % nt] += self.Tminus(tp, phi)
        return q.reshape(nx * nt, nx * nt)


class HubbardFermiMatrixExp(_HFMBase):
    """Exponential hopping discretisation (hubbardFermiMatrixExp.hpp:46-366)."""

    HOPPING = 1

    def expKappa(self, species, inv):
        """exp(+-(kappa-mu)) etc. (hubbardFermiMatrixExp.cpp:76-93), host float64."""
        return self._slice(species, inv)

    def K(self, species):
        return np.eye(self.nx())

    def Kinv(self, species):
        return np.eye(self.nx())

    def logdetKinv(self, species):
        return 0j

    def P(self):
        """Diagonal block of Q: 2 I for sigmaKappa = -1 (the exponentials are inverses of each
        other), else I + exp(kappa - mu) exp(sigmaKappa kappa + mu)
        (hubbardFermiMatrixExp.cpp:176-186); host float64."""
        nx = self.nx()
        if self.sigmaKappa() == -1:
            return 2 * np.eye(nx)
        return np.eye(nx) + self.expKappa(Species.PARTICLE, False) @ self.expKappa(Species.HOLE, False)

    def Tplus(self, tp, phi):
        """Block (tp, tp-1) of Q: -F_p(tp), +F_p(0) on the antiperiodic wrap
        (hubbardFermiMatrixExp.cpp:188-200); device complex128."""
        sign = 1.0 if tp == 0 else -1.0
        return sign * self.F(tp, phi, Species.PARTICLE, False)

    def Tminus(self, tp, phi):
        """Block (tp, tp+1) of Q: -F_h(tp+1)^T, + on the wrap tp = nt-1
        (hubbardFermiMatrixExp.cpp:202-215); device complex128."""
        phi = as_field(phi, self.nx())
        nt = phi.shape[-1] // self.nx()
        sign = 1.0 if tp == nt - 1 else -1.0
        return sign * self.F((tp + 1) % nt, phi, Species.HOLE, False).transpose(0, 1).contiguous()


class HubbardFermiMatrixDia(_HFMBase):
    """Diagonal hopping discretisation (hubbardFermiMatrixDia.hpp:46-344)."""

    HOPPING = 0

    def K(self, species):
        """K_p = (1+mu)I - kappa, K_h = (1-mu)I - sigmaKappa kappa (hubbardFermiMatrixDia.cpp:96-103)."""
        return self._slice(species, False)

    def Kinv(self, species):
        """K^{-1}, inverted on the device by LU (hubbardFermiMatrixDia.cpp:25-42)."""
        k = torch.as_tensor(self.K(species), device=_lib.device()).to(torch.complex128)
        return invert(k).real.cpu().numpy()

    def P(self):
        """Diagonal block of Q = K_p K_h^T + I: (2 - mu^2) I - (sigmaKappa (1 + mu) + 1 - mu) kappa
        + sigmaKappa kappa^2 (hubbardFermiMatrixDia.cpp:170-181); host float64."""
        k = np.asarray(self.kappaTilde(), dtype=np.float64)
        mu, sk = self.muTilde(), self.sigmaKappa()
        return (2 - mu * mu) * np.eye(self.nx()) - (sk * (1 + mu) + 1 - mu) * k + sk * (k @ k)

    def _phases(self, phi, t, sign):
        phi = as_field(phi, self.nx())
        return torch.exp(sign * 1j * phi[t * self.nx():(t + 1) * self.nx()])

    def Tplus(self, tp, phi):
        """Block (tp, tp-1) of Q: diag(e^{i phi_{tp-1}}) (sigmaKappa kappa - (1 - mu) I), negated
        on the antiperiodic wrap tp = 0 (hubbardFermiMatrixDia.cpp:183-201); device complex128."""
        phi = as_field(phi, self.nx())
        nt = phi.shape[-1] // self.nx()
        k = torch.as_tensor(self.kappaTilde(), device=phi.device).to(torch.complex128)
        t = self.sigmaKappa() * k - (1 - self.muTilde()) * torch.eye(self.nx(), dtype=k.dtype, device=k.device)
        sign = -1.0 if tp == 0 else 1.0
        return (sign * self._phases(phi, (tp + nt - 1) % nt, 1))[:, None] * t

    def Tminus(self, tp, phi):
        """Block (tp, tp+1) of Q: (kappa - (1 + mu) I) diag(e^{-i phi_tp}), negated on the wrap
        tp = nt-1 (hubbardFermiMatrixDia.cpp:203-220); device complex128."""
        phi = as_field(phi, self.nx())
        nt = phi.shape[-1] // self.nx()
        k = torch.as_tensor(self.kappaTilde(), device=phi.device).to(torch.complex128)
        t = k - (1 + self.muTilde()) * torch.eye(self.nx(), dtype=k.dtype, device=k.device)
        sign = -1.0 if tp == nt - 1 else 1.0
        return t * (sign * self._phases(phi, tp, -1))[None, :]


def logdet(matrix):
    """log det of a dense complex matrix by LU, Im in [-pi, pi] (math.hpp:265-334)."""
    a = torch.as_tensor(np.asarray(matrix) if not isinstance(matrix, torch.Tensor) else matrix)
    a = a.to(device=_lib.device(), dtype=torch.complex128).contiguous()
    if a.dim() != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("Invalid non-square matrix provided")
    out = np.empty(2)
    _lib.check(_lib.load().isle_logdet(_lib.ptr(a), a.shape[0], _lib.dptr(out), _lib.stream_ptr()))
    return complex(out[0], out[1])


def invert(matrix):
    """Dense inverse by LU with partial pivoting (math.hpp:222-226)."""
    a = torch.as_tensor(np.asarray(matrix) if not isinstance(matrix, torch.Tensor) else matrix)
    a = a.to(device=_lib.device(), dtype=torch.complex128).contiguous()
    if a.dim() != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("Invalid non-square matrix provided")
    out = torch.empty_like(a)
    _lib.check(_lib.load().isle_invert(_lib.ptr(a), a.shape[0], _lib.ptr(out), _lib.stream_ptr()))
    return out


def logdetM(hfm, phi, species):
    """log det M(phi) for one species, Im in [-pi, pi] (bind_hubbardFermiMatrix.cpp:51-52)."""
    nx = hfm.nx()
    phi = as_field(phi, nx)
    nt = phi.shape[-1] // nx
    out = np.empty(2)
    _lib.check(_lib.load().isle_logdetM(hfm.handle, _lib.ptr(phi), nt, int(species), _lib.dptr(out)))
    return complex(out[0], out[1])


def applyM(hfm, phi, species, x):
    """M(phi) applied to every row of x (NRHS x NxNt, time-major), in row form: with v a
    row split into time slices, (M v)_t = K v_t - F_t v_{t-1} for t >= 1 and
    (M v)_0 = K v_0 + F_0 v_{nt-1} (hubbardFermiMatrixExp.cpp:150-166, K = 1;
    hubbardFermiMatrixDia.cpp:140-160, F diagonal).  On the GPU (isle_applyM): one complex x
    real GEMM per time slice over all rows with the plan's slice matrix, between kernels that
    form the shifted, phase-scaled rows and interleave the result."""
    nx = hfm.nx()
    phi = as_field(phi, nx)
    nt = phi.shape[-1] // nx
    x = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    x = x.to(device=phi.device, dtype=torch.complex128)
    shape = x.shape
    rows = x.reshape(-1, nx * nt).contiguous()
    out = torch.empty_like(rows)
    _lib.check(_lib.load().isle_applyM(hfm.handle, _lib.ptr(phi.contiguous()), nt, int(species), _lib.ptr(rows),
                                       rows.shape[0], _lib.ptr(out)))
    return out.reshape(shape)


def solveM(hfm, phi, species, rhss, reuse_factor=False):
    """Solve M x = b for every row b of rhss (NRHS x NxNt); returns the same shape
    (bind_hubbardFermiMatrix.cpp:53-55, hubbardFermiMatrixExp.cpp:645-694).

    ``reuse_factor`` (no counterpart in the reference): phi is the field of the previous
    solveM on this matrix, unchanged, so its partial products and LUs are reused."""
    nx = hfm.nx()
    phi = as_field(phi, nx)
    nt = phi.shape[-1] // nx
    r = torch.as_tensor(np.asarray(rhss) if not isinstance(rhss, torch.Tensor) else rhss)
    r = r.to(device=phi.device, dtype=torch.complex128).contiguous()
    if r.dim() == 1:
        r = r.reshape(1, -1)
    if r.shape[1] != nx * nt:
        raise RuntimeError("Right hand side does not have correct size (spacetime vector)")
    out = torch.empty_like(r)
    _lib.check(_lib.load().isle_solveM_ex(hfm.handle, _lib.ptr(phi), nt, int(species), _lib.ptr(r), r.shape[0],
                                          _lib.ptr(out), int(bool(reuse_factor))))
    return out


def _invert_logdet(a, out):
    """out = a^{-1} and log det a from one HIP LU (a, out: contiguous device nx x nx)."""
    ld = np.empty(2)
    _lib.check(_lib.load().isle_invert_logdet(_lib.ptr(a), a.shape[0], _lib.ptr(out), _lib.dptr(ld),
                                              _lib.stream_ptr()))
    return complex(ld[0], ld[1])


class QLU:
    """Block LU of the block-cyclic tridiagonal Q = M_p M_h^T (HubbardFermiMatrixExp::QLU,
    hubbardFermiMatrixExp.hpp; hubbardFermiMatrixExp.cpp:282-360).

    Device complex128 stacks: ``dinv`` [nt, nx, nx] (d_i^{-1}), ``u``, ``l`` [nt-1, nx, nx],
    ``v``, ``h`` [nt-2, nx, nx]; ``logdet_d`` holds log det d_i from the LU that inverted d_i."""

    def __init__(self, dinv, u, l, v, h, logdet_d, hfm=None):
        self.dinv, self.u, self.l, self.v, self.h = dinv, u, l, v, h
        self.logdet_d = list(logdet_d)
        self.hfm = hfm  # the matrix it factorises (getQLU sets it): solveQ then runs natively

    def nt(self):
        return self.dinv.shape[0]

    def isConsistent(self):
        """Block counts fit nt (hubbardFermiMatrixExp.cpp:290-303)."""
        nt = self.dinv.shape[0]
        if nt == 0:
            return False
        if self.u.shape[0] != nt - 1 or self.l.shape[0] != nt - 1:
            return False
        return nt < 2 or (self.v.shape[0] == nt - 2 and self.h.shape[0] == nt - 2)

    def reconstruct(self):
        """L U multiplied out into the dense (NxNt)^2 Q (hubbardFermiMatrixExp.cpp:305-360).
        Needs nt >= 3: the reference's formula reads l[nt-3] and h[0]."""
        nt, nx = self.dinv.shape[0], self.dinv.shape[1]
        if nt < 3:
            raise ValueError("Reconstruction of LU needs nt >= 3")
        if not self.isConsistent():
            raise RuntimeError("Components of LU not initialized properly")
        d = torch.stack([invert(x) for x in self.dinv])
        l, u, v, h = self.l, self.u, self.v, self.h
        q = torch.zeros((nt, nx, nt, nx), dtype=torch.complex128, device=d.device)
        q[0, :, 0] = d[0]
        q[0, :, 1] = u[0]
        q[0, :, nt - 1] = v[0]
        for i in range(1, nt - 2):
            q[i, :, i] = d[i] + l[i - 1] @ u[i - 1]
            q[i, :, i - 1] = l[i - 1] @ d[i - 1]
            q[i, :, i + 1] = u[i]
            q[i, :, nt - 1] = l[i - 1] @ v[i - 1] + v[i]
        q[nt - 2, :, nt - 2] = d[nt - 2] + l[nt - 3] @ u[nt - 3]
        q[nt - 2, :, nt - 3] = l[nt - 3] @ d[nt - 3]
        q[nt - 2, :, nt - 1] = l[nt - 3] @ v[nt - 3] + u[nt - 2]
        q[nt - 1, :, 0] = h[0] @ d[0]
        for i in range(1, nt - 2):
            q[nt - 1, :, i] = h[i - 1] @ u[i - 1] + h[i] @ d[i]
        q[nt - 1, :, nt - 2] = h[nt - 3] @ u[nt - 3] + l[nt - 2] @ d[nt - 2]
        q[nt - 1, :, nt - 1] = d[nt - 1] + l[nt - 2] @ u[nt - 2] + _hcat(h) @ _vcat(v)
        return q.reshape(nx * nt, nx * nt)


def _hcat(h):
    """[h_0 h_1 ...] as one nx x (k nx) matrix."""
    k, nx = h.shape[0], h.shape[1]
    return h.permute(1, 0, 2).reshape(nx, k * nx)


def _vcat(v):
    """[v_0; v_1; ...] as one (k nx) x nx matrix."""
    return v.reshape(-1, v.shape[2])


def getQLU(hfm, phi):
    """LU of Q(phi) in blocks (getQLU, hubbardFermiMatrixExp.cpp:482-492; nt == 1 :368-383,
    nt == 2 :386-416, nt > 2 :419-479; the Dia file repeats it).

    The recursion is sequential in the time slice: each diagonal block
    d_i = P - l_{i-1} u_{i-1} is inverted by the HIP LU (which also gives log det d_i)
    before l_i and h_i can be formed.  For nt >= 3 the whole recursion is one native
    operation (isle_qlu): the off-diagonal blocks, the block products as complex GEMMs and
    one LU and solve per diagonal block, launched back to back with the log dets left on the
    device, and sum_i h_i v_i one GEMM over the stacked blocks; nt = 1, 2 are single blocks."""
    nx = hfm.nx()
    phi = as_field(phi, nx)
    if phi.dim() != 1:
        raise ValueError("getQLU takes one field")
    nt = phi.shape[0] // nx
    dev = phi.device
    c128 = torch.complex128
    dinv = torch.empty((nt, nx, nx), dtype=c128, device=dev)
    u = torch.empty((max(nt - 1, 0), nx, nx), dtype=c128, device=dev)
    l = torch.empty_like(u)
    v = torch.empty((max(nt - 2, 0), nx, nx), dtype=c128, device=dev)
    h = torch.empty_like(v)
    if nt >= 3:
        ld = np.empty(2 * nt)
        _lib.check(_lib.load().isle_qlu(hfm.handle, _lib.ptr(phi.contiguous()), nt, _lib.ptr(dinv), _lib.ptr(u),
                                        _lib.ptr(l), _lib.ptr(v), _lib.ptr(h), _lib.dptr(ld)))
        return QLU(dinv, u, l, v, h, [complex(ld[2 * i], ld[2 * i + 1]) for i in range(nt)], hfm)
    p = torch.as_tensor(hfm.P(), device=dev).to(c128)
    lds = []

    def inv(i, d):
        lds.append(_invert_logdet(d.contiguous(), dinv[i]))

    if nt == 1:
        inv(0, p + hfm.Tplus(0, phi) + hfm.Tminus(0, phi))
        return QLU(dinv, u, l, v, h, lds, hfm)
    inv(0, p)
    # nt == 2
    l[0] = (hfm.Tplus(1, phi) + hfm.Tminus(1, phi)) @ dinv[0]
    u[0] = hfm.Tplus(0, phi) + hfm.Tminus(0, phi)
    inv(1, p - l[0] @ u[0])
    return QLU(dinv, u, l, v, h, lds, hfm)


def logdetQ(hfm, phi=None):
    """log det Q(phi) = sum_i log det d_i, Im in [-pi, pi] (logdetQ,
    hubbardFermiMatrixExp.cpp:540-570, bind_hubbardFermiMatrix.cpp:44-46).
    ``logdetQ(qlu)`` reuses a factorisation from getQLU."""
    if not isinstance(hfm, QLU) and phi is None:
        raise TypeError("logdetQ takes (hfm, phi) or (qlu)")
    lu = hfm if isinstance(hfm, QLU) else getQLU(hfm, phi)
    if not lu.isConsistent():
        raise RuntimeError("Components of LU not initialized properly")
    return to_first_log_branch(sum(lu.logdet_d))


def solveQ(hfm, phi, rhs=None):
    """x with Q(phi) x = rhs (solveQ, hubbardFermiMatrixExp.cpp:494-537,
    bind_hubbardFermiMatrix.cpp:47-49): forward substitution with l and h, then backward
    with dinv, u and v.  ``solveQ(qlu, rhs)`` reuses a factorisation.  rhs: a spacetime
    vector (NxNt); a 2-D rhs solves every row.  Returns a device tensor of rhs's shape.
    With the QLU of getQLU (which records its matrix) the substitutions run as one native
    operation (isle_solveq, complex GEMMs of four real products); a QLU assembled by hand
    takes the same steps as torch products."""
    if isinstance(hfm, QLU):
        lu, rhs = hfm, phi
    if rhs is None:
        raise TypeError("solveQ takes (hfm, phi, rhs) or (qlu, rhs)")
    if not isinstance(hfm, QLU):
        lu = getQLU(hfm, phi)
    if not lu.isConsistent():
        raise RuntimeError("Components of LU not initialized properly")
    nt, nx = lu.dinv.shape[0], lu.dinv.shape[1]
    b = as_field(rhs)
    shape = b.shape
    if b.dim() not in (1, 2) or shape[-1] != nt * nx:
        raise RuntimeError("Right hand side does not have correct size (spacetime vector)")
    if lu.hfm is not None and lu.hfm.nx() == nx:
        # one native operation (isle_solveq): the substitutions as complex GEMMs on the device
        rows = b.reshape(-1, nt * nx).to(torch.complex128).contiguous()
        out = torch.empty_like(rows)
        parts = [t.to(device=rows.device, dtype=torch.complex128).contiguous()
                 for t in (lu.dinv, lu.u, lu.l, lu.v, lu.h)]
        _lib.check(_lib.load().isle_solveq(lu.hfm.handle, nt, rows.shape[0], *[_lib.ptr(t) for t in parts],
                                           _lib.ptr(rows), _lib.ptr(out)))
        return out.reshape(shape)
    # a QLU built by hand: the same substitutions as torch products; b[t] is nx x nrhs
    b = b.reshape(-1, nt, nx).permute(1, 2, 0).contiguous()
    y = torch.empty_like(b)
    y[0] = b[0]
    for i in range(1, nt):
        y[i] = b[i] - lu.l[i - 1] @ y[i - 1]
    if nt > 2:
        y[nt - 1] -= _hcat(lu.h) @ y[:nt - 2].reshape((nt - 2) * nx, -1)
    x = torch.empty_like(b)
    x[nt - 1] = lu.dinv[nt - 1] @ y[nt - 1]
    if nt > 1:
        x[nt - 2] = lu.dinv[nt - 2] @ (y[nt - 2] - lu.u[nt - 2] @ x[nt - 1])
    if nt > 2:
        vx = torch.matmul(lu.v, x[nt - 1])  # v_i x_{nt-1} for every i at once
        for i in range(nt - 3, -1, -1):
            x[i] = lu.dinv[i] @ (y[i] - lu.u[i] @ x[i + 1] - vx[i])
    return x.permute(2, 0, 1).reshape(shape)
