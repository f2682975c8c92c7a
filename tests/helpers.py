"""Shared numeric helpers for the tests (no oracle logic lives here)."""
import cmath
import math

import numpy as np


def mod2pi_diff(a, b):
    """|a - b| with the imaginary parts compared modulo 2*pi."""
    d = complex(a) - complex(b)
    return abs(complex(d.real, math.remainder(d.imag, 2 * math.pi)))


def rel_mod2pi(a, b):
    return mod2pi_diff(a, b) / max(abs(complex(b)), 1e-300)


def random_phi(rng, n, real=True, imag=True, std=0.2):
    r = rng.normal(0, std, n) if real else np.zeros(n)
    i = rng.normal(0, std, n) if imag else np.zeros(n)
    return r + 1j * i


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
