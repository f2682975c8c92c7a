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
    hubbardFermiMatrixDia.cpp:140-160, F diagonal).  One real GEMM of all slices' real and
    imaginary parts with the (real) slice matrix on the device; the slice matrices are the
    plan's own."""
    nx = hfm.nx()
    phi = as_field(phi, nx)
    nt = phi.shape[-1] // nx
    x = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    x = x.to(device=phi.device, dtype=torch.complex128)
    rows = x.reshape(-1, nt, nx)
    sign = 1j if Species(species) == Species.PARTICLE else -1j  # F(tp) = ... e^{sign i phi_{tp-1}}
    ph = torch.exp(sign * torch.roll(phi.reshape(nt, nx), 1, 0))
    y = torch.roll(rows, 1, 1) * ph
    y[:, 0] = -y[:, 0]  # + F_0 v_{nt-1}: the antiperiodic wrap
    s = torch.as_tensor(hfm._slice(species, False), device=phi.device).to(torch.float64).T.contiguous()

    def times_s(v):  # v @ s for complex v and real s: one real 2-D GEMM over the re/im parts
        shape = v.shape
        parts = torch.view_as_real(v).transpose(-1, -2).reshape(-1, nx)  # copied: rows (.., re/im)
        prod = torch.mm(parts, s).reshape(*shape[:-1], 2, nx)
        return torch.view_as_complex(prod.transpose(-1, -2).contiguous())
    out = rows - times_s(y) if hfm.HOPPING == 1 else times_s(rows) - y
    return out.reshape(x.shape)


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
