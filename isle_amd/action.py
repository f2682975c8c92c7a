"""Actions: the Action protocol, SumAction, HubbardGaugeAction and HubbardFermiAction.

Mirrors the reference's ``isle.action`` module (src/isle/cpp/bind/bind_action.cpp:100-324,
src/isle/cpp/action/*.cpp).  Fields are complex128 tensors on the GPU; ``eval`` returns a
Python complex and ``force`` a device tensor (the complex -dS/dphi).

Besides the reference's single-field ``eval``/``force``, HubbardFermiAction offers
``eval_batch``/``force_batch`` over a [batch, nx*nt] tensor of independent fields, and a
``precision`` switch:

* ``"fp64"``  - the reference DIRECT_SINGLE algorithm, every product in FP64;
* ``"mixed"`` - force from FP32 chains over groups of ``group`` time slices on MFMA,
  combined in FP64; eval (the Metropolis action) from FP64 group products, multiplied
  together (``logdet="grouped"``) or combined through the stabilising block-cyclic LU with
  groups of ``logdet_group`` slices (``logdet="stable"``, default 8).  With
  ``recompute=k`` (k dividing ``group``, Nx a multiple of 32) the force keeps the FP64
  suffix chains only every k slices and recomputes the others when they are needed:
  about a third of the force's memory for about half again its FP64 work.
* ``"stable"`` - for ill-conditioned products (large beta, spin basis, mu != 0, where the
  product of all time slices spans more decades than FP64 holds): eval and force from FP64
  group products combined through the structured pivoted elimination of their block-cyclic
  matrix, never through the product of all groups.  ``right_chains="fp32"`` runs the
  force's per-group right chains in FP32 on MFMA (with the force diagonal fused into them);
  the elimination, the suffix chains and the action stay FP64.
"""
import ctypes
import enum

import numpy as np
import torch

from . import _lib
from .core import (HubbardFermiMatrixDia, HubbardFermiMatrixExp, Lattice, as_field, invert, isBipartite, logdetQ,
                   to_first_log_branch)


# device workspace bound of one stable log det call; larger batches are split into chunks
_STABLE_LOGDET_BYTES = 24 << 30


class HFAHopping(enum.IntEnum):
    """hubbardFermiAction.hpp:16."""

    DIA = 0
    EXP = 1


class HFABasis(enum.IntEnum):
    """hubbardFermiAction.hpp:19."""

    PARTICLE_HOLE = 0
    SPIN = 1


class HFAAlgorithm(enum.IntEnum):
    """hubbardFermiAction.hpp:25."""

    DIRECT_SINGLE = 0
    DIRECT_SQUARE = 1


class Action:
    """Abstract action (src/isle/cpp/action/action.hpp:14-22).  Subclass and implement
    ``eval(phi) -> complex`` and ``force(phi) -> tensor``.  ``a + b`` builds a SumAction."""

    def eval(self, phi):
        raise NotImplementedError

    def force(self, phi):
        raise NotImplementedError

    def __add__(self, other):
        return SumAction(self, other)


class SumAction(Action):
    """Sum of actions (sumAction.hpp:21-59, sumAction.cpp:46-58).  Holds references to
    its terms (the reference keeps non-owning pointers alive through keep_alive)."""

    def __init__(self, *actions):
        self._actions = list(actions)

    def add(self, action):
        self._actions.append(action)
        return action

    def __getitem__(self, idx):
        if idx >= len(self._actions) or idx < -len(self._actions):
            raise IndexError("Index of action is out of range")
        return self._actions[idx]

    def __len__(self):
        return len(self._actions)

    def size(self):
        return len(self._actions)

    def clear(self):
        self._actions.clear()

    def eval(self, phi):
        res = 0j
        for a in self._actions:
            res += complex(a.eval(phi))
        return res

    def force(self, phi):
        phi = as_field(phi)
        res = torch.zeros_like(phi)
        for a in self._actions:
            res = res + as_field(a.force(phi))
        return res


class HubbardGaugeAction(Action):
    """S = (phi, phi)/(2 utilde), bilinear; force -phi/utilde (hubbardGaugeAction.cpp:7-13)."""

    def __init__(self, utilde):
        self._utilde = float(utilde)

    @property
    def utilde(self):
        return self._utilde

    def eval(self, phi):
        phi = as_field(phi)
        out = np.empty(4)
        n = phi.numel()
        _lib.check(_lib.load().isle_field_sums(_lib.ptr(phi), n, n, 1, _lib.dptr(out), _lib.stream_ptr()))
        return complex(out[2], out[3]) / 2.0 / self._utilde

    def force(self, phi):
        phi = as_field(phi)
        out = torch.empty_like(phi)
        _lib.check(_lib.load().isle_gauge_force(_lib.ptr(phi), _lib.ptr(out), phi.numel(), self._utilde, 0,
                                                _lib.stream_ptr()))
        return out


def _hole_shortcut_possible(basis, kappa, mu_tilde, sigma_kappa):
    """hubbardFermiAction.cpp:114-157."""
    import logging
    log = logging.getLogger("HubbardFermiAction")
    if basis == HFABasis.SPIN:
        log.info("Not using shortcut for hole determinant, spin basis is not supported.")
        return False
    if not isBipartite(kappa):
        log.info("Not using shortcut for hole determinant, lattice is not bipartite.")
        return False
    if mu_tilde != 0.0:
        log.info("Not using shortcut for hole determinant, chemical potential is not zero.")
        return False
    if sigma_kappa != +1:
        log.info("Not using shortcut for hole determinant, sigmaKappa is not +1")
        return False
    log.info("Using shortcut to calculate hole determinant from particle determinant.")
    return True


class HubbardFermiAction(Action):
    """Fermion action -log det M(phi) M(-phi) (hubbardFermiAction.hpp:96-142).

    DIRECT_SQUARE (hubbardFermiAction.cpp:62-108, 230-272) is the reference's small-size
    cross-check: eval = -logdetQ through the block LU of Q = M_p M_h^T (any size), and the
    force from the blocks of the dense (NxNt)^2 Q^-1 next to the off-diagonal blocks T+ and
    T- of Q, inverted by the HIP LU, so the force needs Nx*Nt <= 16383.  FP64 only; no
    per-species parts (logdets, force_parts).
    """

    def __init__(self, kappaTilde, muTilde, sigmaKappa, hopping=HFAHopping.DIA,
                 basis=HFABasis.PARTICLE_HOLE, algorithm=HFAAlgorithm.DIRECT_SINGLE,
                 allowShortcut=False, precision="fp64", group=32, recompute=0, right_chains="fp64",
                 logdet="grouped", logdet_group=8):
        if precision not in ("fp64", "mixed", "stable"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'stable'")
        if logdet not in ("grouped", "stable"):
            raise ValueError("logdet must be 'grouped' or 'stable'")
        if logdet == "stable" and precision != "mixed":
            raise ValueError("logdet='stable' selects the Metropolis log det of precision='mixed' "
                             "('fp64' is the reference algorithm, 'stable' is stable throughout)")
        if int(logdet_group) < 1:
            raise ValueError("logdet_group must be >= 1")
        if right_chains not in ("fp64", "fp32"):
            raise ValueError("right_chains must be 'fp64' or 'fp32'")
        if int(recompute) < 0:
            raise ValueError("recompute must be >= 0")
        if HFAAlgorithm(algorithm) == HFAAlgorithm.DIRECT_SQUARE and precision != "fp64":
            raise ValueError("HFAAlgorithm.DIRECT_SQUARE runs in FP64 only")
        kappa = kappaTilde.hopping() if isinstance(kappaTilde, Lattice) else np.asarray(kappaTilde, dtype=np.float64)
        self.hopping = HFAHopping(hopping)
        self.basis = HFABasis(basis)
        self.algorithm = HFAAlgorithm(algorithm)
        cls = HubbardFermiMatrixExp if self.hopping == HFAHopping.EXP else HubbardFermiMatrixDia
        self.hfm = cls(kappa, muTilde, sigmaKappa)
        self.precision = precision
        self.group = int(group)
        # memory schedule of the mixed force: keep the FP64 suffix chains every `recompute`
        # slices and recompute the rest (0: keep all; see isle_hfm_set_recompute)
        self.recompute = int(recompute)
        # the stable force's right chains: FP64 (default) or FP32 on MFMA with the force
        # diagonal fused into the chain products (as in the mixed form)
        self.right_chains = right_chains
        # the mixed form's Metropolis log det: FP64 products of `group` slices multiplied
        # together ("grouped"), or FP64 products of `logdet_group` slices combined through the
        # stabilising block-cyclic LU ("stable"), never through the product of all slices
        self.logdet_form = logdet
        self.logdet_group = int(logdet_group)
        if self.recompute:
            _lib.check(_lib.load().isle_hfm_set_recompute(self.hfm.handle, self.recompute))
        # log of the condition-number bound of one time slice: the FP32 right chains of
        # the mixed force span at most `group` slices and are kept within cond <= e^12
        r = float(np.max(np.sum(np.abs(kappa), axis=1))) if kappa.size else 0.0
        if self.hopping == HFAHopping.EXP:
            self._slice_spread = 2.0 * r
        else:
            mu = abs(float(muTilde))
            lo = 1.0 - mu - r
            self._slice_spread = float(np.log((1.0 + mu + r) / lo)) if lo > 0 else np.inf
        self.shortcut = bool(allowShortcut) and _hole_shortcut_possible(self.basis, kappa, float(muTilde),
                                                                         int(sigmaKappa))
        if self.algorithm == HFAAlgorithm.DIRECT_SQUARE:
            self.shortcut = False  # Q holds both species

    # ---- DIRECT_SQUARE (hubbardFermiAction.cpp:62-108): Q = M_p M_h^T
    SQUARE_MAX = 16383  # order of the dense LU of the force's Q^-1 (the HIP LU's limit)

    def _field(self, phi):
        return -1j * phi if self.basis == HFABasis.SPIN else phi

    def _q(self, field):
        n = field.numel()
        if n > self.SQUARE_MAX:
            raise ValueError(f"HFAAlgorithm.DIRECT_SQUARE force needs Nx*Nt <= {self.SQUARE_MAX} (dense LU), got {n}")
        return self.hfm.Q(field)

    def _t_pm(self, field, nt):
        """T+(tau + 1) and T-(tau) for tau = 0..nt-1, [nt, nx, nx] each (the blocks of Q at
        (tau+1, tau) and (tau, tau+1); hubbardFermiMatrixExp.cpp:188-215,
        hubbardFermiMatrixDia.cpp:183-220)."""
        tplus = [self.hfm.Tplus((tau + 1) % nt, field) for tau in range(nt)]
        tminus = [self.hfm.Tminus(tau, field) for tau in range(nt)]
        return torch.stack(tplus), torch.stack(tminus)

    def _eval_square(self, phi):
        # -logdetQ through the block LU of Q, as the reference evaluates it
        # (hubbardFermiAction.cpp:204-221, 275-292); no size limit
        return -logdetQ(self.hfm, self._field(phi))

    def _force_square(self, phi):
        field = self._field(phi)
        nx = self.hfm.nx()
        nt = field.numel() // nx
        blocks = invert(self._q(field)).reshape(nt, nx, nt, nx)
        tau = torch.arange(nt, device=field.device)
        up = blocks[tau, :, (tau + 1) % nt, :]    # Q^-1 block (tau, tau+1), [nt, nx, nx]
        down = blocks[(tau + 1) % nt, :, tau, :]  # Q^-1 block (tau+1, tau)
        tplus, tminus = self._t_pm(field, nt)

        def diag(a, b):  # diagonal of a b, per tau
            return (a * b.transpose(1, 2)).sum(dim=2)
        if self.hopping == HFAHopping.DIA:
            f = 1j * diag(tplus, up) - 1j * diag(down, tminus)
        else:
            f = 1j * diag(up, tplus) - 1j * diag(tminus, down)
        f = f.reshape(-1)
        return -1j * f if self.basis == HFABasis.SPIN else f

    def _mode(self, force=False):
        if force and self.precision == "stable" and self.right_chains == "fp32":
            return 3
        if not force and self.precision == "mixed" and self.logdet_form == "stable":
            return 2
        return {"fp64": 0, "mixed": 1, "stable": 2}[self.precision]

    def _logdet_group(self, nt):
        """Slices per group of the log det: the force's group, or for the mixed form's stable
        log det the largest divisor of nt not above ``logdet_group``."""
        if self.precision == "mixed" and self.logdet_form == "stable":
            g = min(self.logdet_group, nt)
            while nt % g:
                g -= 1
            return g
        return self._group(nt)

    def _group(self, nt):
        """Slices per group: the largest divisor of nt not above ``group`` that keeps the
        slice-conditioning bound g * spread <= 12 (cond <= e^12 over one group, for FP32
        right chains: the mixed force, "stable" with right_chains="fp32") or <= 24 (FP64
        chains of "stable")."""
        g = min(self.group, nt)
        bound = 24.0 if self.precision == "stable" and self.right_chains == "fp64" else 12.0
        if self._slice_spread > 0:
            g = max(1, min(g, int(bound / self._slice_spread)))
        while nt % g:
            g -= 1
        return g

    def _shape(self, phi):
        nx = self.hfm.nx()
        phi = as_field(phi, nx)
        if phi.dim() == 1:
            phi = phi.reshape(1, -1)
        return phi, phi.shape[0], phi.shape[1] // nx

    def _single_only(self):
        if self.algorithm == HFAAlgorithm.DIRECT_SQUARE:
            raise ValueError("per-species log dets and force parts belong to HFAAlgorithm.DIRECT_SINGLE")

    def logdets(self, phis):
        """[batch][nspecies] complex log det M of each field (aux = -i phi in the spin basis)."""
        self._single_only()
        phis, batch, nt = self._shape(phis)
        nsp = 1 if self.shortcut else 2
        out = np.empty((batch, nsp, 2))
        mode, g = self._mode(), self._logdet_group(nt)
        step = batch
        if mode == 2 and nt // g > 1:
            # the block-cyclic factorisation keeps every elimination step's 2Nx x 3Nx block
            # (12 Nx^2 doubles) and the group products (4 Nx^2 doubles) with LU work: about
            # 136 Nx^2 bytes per group, field and species
            nx = self.hfm.nx()
            per_field = 136 * nx * nx * (nt // g) * nsp
            step = max(1, min(batch, _STABLE_LOGDET_BYTES // per_field))
        for b0 in range(0, batch, step):
            b1 = min(batch, b0 + step)
            _lib.check(_lib.load().isle_logdetM_batch(
                self.hfm.handle, _lib.ptr(phis[b0:b1]), nt, b1 - b0, mode, g, nsp,
                int(self.basis == HFABasis.SPIN), _lib.dptr(out[b0:b1])))
        return out[..., 0] + 1j * out[..., 1]

    def eval_batch(self, phis):
        """Action of each field in a [batch, nx*nt] tensor (hubbardFermiAction.cpp:230-264)."""
        if self.algorithm == HFAAlgorithm.DIRECT_SQUARE:
            phis, batch, _ = self._shape(phis)
            return np.array([self._eval_square(phis[b]) for b in range(batch)])
        ld = self.logdets(phis)
        res = []
        for row in ld:
            if self.shortcut:
                res.append(-to_first_log_branch(row[0] + np.conj(row[0])))
            else:
                res.append(-to_first_log_branch(row[0] + row[1]))
        return np.array(res)

    def eval(self, phi):
        return complex(self.eval_batch(phi)[0])

    def force_parts(self, phis):
        """forceDirectSinglePart (hubbardFermiAction.cpp:16-60) of each species for each
        field: [batch, nspecies, nx*nt] (one species with the hole shortcut; evaluated at
        aux = -i phi in the spin basis)."""
        self._single_only()
        phis, batch, nt = self._shape(phis)
        nx = self.hfm.nx()
        nsp = 1 if self.shortcut else 2
        parts = torch.empty((batch, nsp, nx * nt), dtype=torch.complex128, device=phis.device)
        _lib.check(_lib.load().isle_force_parts(self.hfm.handle, _lib.ptr(phis), nt, batch, self._mode(force=True),
                                                self._group(nt), nsp, int(self.basis == HFABasis.SPIN),
                                                _lib.ptr(parts)))
        return parts

    def force_parts_multi(self, phis_list):
        """force_parts of several batches of fields of this lattice with different time
        extents, their operations merged into shared launches (isle_force_parts_multi: every
        chain step is one launch over all batches' descriptors).  ``phis_list``: tensors
        [batch_i, nx*nt_i] with distinct nt_i; returns [batch_i, nspecies, nx*nt_i] each."""
        self._single_only()
        shaped = [self._shape(p) for p in phis_list]
        nx = self.hfm.nx()
        nsp = 1 if self.shortcut else 2
        parts = [torch.empty((b, nsp, nx * nt), dtype=torch.complex128, device=p.device) for p, b, nt in shaped]
        k = len(shaped)
        ptrs = ctypes.c_void_p * k
        ints = ctypes.c_int * k
        phis_a = ptrs(*[p.data_ptr() for p, _, _ in shaped])
        outs_a = ptrs(*[t.data_ptr() for t in parts])
        nts = ints(*[nt for _, _, nt in shaped])
        batches = ints(*[b for _, b, _ in shaped])
        groups = ints(*[self._group(nt) for _, _, nt in shaped])
        as_vp = lambda a: ctypes.cast(a, ctypes.c_void_p)  # noqa: E731
        _lib.check(_lib.load().isle_force_parts_multi(
            self.hfm.handle, k, as_vp(phis_a), as_vp(nts), as_vp(batches), self._mode(force=True), as_vp(groups), nsp,
            int(self.basis == HFABasis.SPIN), as_vp(outs_a)))
        return parts

    def force_batch_multi(self, phis_list, outs):
        """force_batch of several batches of fields with different time extents (see
        force_parts_multi), written to ``outs`` (tensors shaped like the fields)."""
        parts = self.force_parts_multi(phis_list)
        nsp = 1 if self.shortcut else 2
        mode = 2 if self.basis == HFABasis.SPIN else (1 if self.shortcut else 0)
        lib = _lib.load()
        for pt, out in zip(parts, outs):
            batch, n = pt.shape[0], pt.shape[2]
            if out.stride(-1) != 1 or (batch > 1 and out.stride(0) < n):
                raise ValueError("out must hold the fields' forces as rows of contiguous elements")
            fh = pt[:, 1] if nsp == 2 else pt[:, 0]
            _lib.check(lib.isle_combine_force_batch(_lib.ptr(pt[:, 0]), _lib.ptr(fh), _lib.ptr(out), n, batch, nsp * n,
                                                    out.stride(0) if batch > 1 else 0, mode, 0, _lib.stream_ptr()))
        return outs

    def force_batch(self, phis, out=None, accumulate=False):
        """Force -dS/dphi of each field (hubbardFermiAction.cpp:243-272), [batch, nx*nt]."""
        if self.algorithm == HFAAlgorithm.DIRECT_SQUARE:
            phis, batch, _ = self._shape(phis)
            f = torch.stack([self._force_square(phis[b]) for b in range(batch)])
            if out is None:
                return f
            if accumulate:
                out += f
            else:
                out.copy_(f)
            return out
        phis, batch, nt = self._shape(phis)
        nx = self.hfm.nx()
        nsp = 1 if self.shortcut else 2
        parts = self.force_parts(phis)
        lib = _lib.load()
        if out is None:
            out = torch.empty((batch, nx * nt), dtype=torch.complex128, device=phis.device)
            accumulate = False
        mode = 2 if self.basis == HFABasis.SPIN else (1 if self.shortcut else 0)
        if out.stride(-1) != 1 or (batch > 1 and out.stride(0) < nx * nt):
            raise ValueError("out must hold the fields' forces as rows of contiguous elements")
        fh = parts[:, 1] if nsp == 2 else parts[:, 0]  # one launch for the whole batch
        _lib.check(lib.isle_combine_force_batch(_lib.ptr(parts[:, 0]), _lib.ptr(fh), _lib.ptr(out), nx * nt, batch,
                                                nsp * nx * nt, out.stride(0) if batch > 1 else 0, mode,
                                                int(bool(accumulate)), _lib.stream_ptr()))
        return out

    def force(self, phi):
        return self.force_batch(phi)[0]


def makeHubbardFermiAction(lat_or_kappa, *args, **kwargs):
    """Run-time dispatch of HubbardFermiAction (bind_action.cpp:114-222).

    makeHubbardFermiAction(kappaTilde, muTilde, sigmaKappa, hopping=DIA, basis=PARTICLE_HOLE,
                           algorithm=DIRECT_SINGLE, allowShortcut=False)
    makeHubbardFermiAction(lat, beta, muTilde, sigmaKappa, hopping=DIA, ...)  with
    kappaTilde = lat.hopping() * beta / lat.nt() (bind_action.cpp:206-221).
    Extra keyword arguments: precision ('fp64' | 'mixed' | 'stable'), group (slices per group).
    """
    if isinstance(lat_or_kappa, Lattice):
        beta, mu, sk, *rest = args
        kappa = lat_or_kappa.hopping() * beta / lat_or_kappa.nt()
    else:
        mu, sk, *rest = args
        kappa = np.asarray(lat_or_kappa, dtype=np.float64)
    names = ("hopping", "basis", "algorithm", "allowShortcut")
    defaults = {"hopping": HFAHopping.DIA, "basis": HFABasis.PARTICLE_HOLE,
                "algorithm": HFAAlgorithm.DIRECT_SINGLE, "allowShortcut": False}
    for name, val in zip(names, rest):
        defaults[name] = val
    defaults.update(kwargs)
    return HubbardFermiAction(kappa, mu, sk, **defaults)
