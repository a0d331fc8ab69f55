"""Analytic laser fields as descriptors (sfparticles/fields.py).

The reference JIT-compiles a user Python function ``(x, y, z, t) -> (Ex, Ey,
Ez, Bx, By, Bz)`` and superposes fields by building a new closure
(fields.py:11-61).  Here a field is a list of *terms*, each the user-level
arguments of one builder, evaluated on the GPU by a fixed device function per
kind; ``+`` concatenates the term lists and the terms are summed left to right
as the reference's ``__add__`` does.  Arbitrary Python field functions cannot
run inside a HIP kernel: ``Fields(fn)`` keeps the function for reference but
any GPU operation on it raises ``NotImplementedError``.
"""
from typing import Callable, List, Sequence, Tuple

from .constants import c, e, m_e, pi  # noqa: F401  (re-exported like the reference)

STATIC, SIMPLE, GAUSSIAN, FOCUSED = 0, 1, 2, 3
MAX_TERMS = 8


class Fields(object):
    """A superposition of analytic field terms (fields.py:11-61)."""

    def __init__(self, field_func: Callable = None, terms: Sequence[Tuple[int, Sequence[float]]] = None):
        if terms is None and field_func is None:
            raise ValueError("Fields needs a field function or a list of terms")
        self.field_func_input = field_func
        self.terms: List[Tuple[int, Tuple[float, ...]]] = (
            [(int(k), tuple(float(v) for v in p)) for k, p in terms] if terms is not None else None)
        if self.terms is not None and len(self.terms) > MAX_TERMS:
            raise ValueError(f"at most {MAX_TERMS} field terms are supported")

    @property
    def is_analytic(self) -> bool:
        return self.terms is not None

    def require_terms(self):
        if self.terms is None:
            raise NotImplementedError(
                "Fields built from an arbitrary Python function cannot be evaluated by the HIP "
                "kernels; build the field from simple_laser_pulse / gaussian_laser_pulse / "
                "focused_gaussian_laser_pulse / static_field and combine them with '+'")
        return self.terms

    def __add__(self, other: "Fields") -> "Fields":
        if self.terms is None or other.terms is None:
            return Fields(field_func=_sum_functions(self, other))
        return Fields(terms=self.terms + other.terms)

    def __call__(self, x, y, z, t):
        """Evaluate on the host for a single point is not provided: the field
        is evaluated by the GPU kernels (``Particles._eval_field``)."""
        if self.field_func_input is not None:
            return self.field_func_input(x, y, z, t)
        raise NotImplementedError("analytic fields are evaluated on the GPU")


def _sum_functions(a, b):
    fa = a.field_func_input
    fb = b.field_func_input

    def out(x, y, z, t):
        if fa is None or fb is None:
            raise NotImplementedError("analytic fields are evaluated on the GPU")
        A = fa(x, y, z, t)
        B = fb(x, y, z, t)
        return tuple(p + q for p, q in zip(A, B))
    return out


def simple_laser_pulse(a0, w0, ctau, t_order=2, direction=1, x0=0, wavelength=0.8e-6,
                       pol_angle=0, cep=0):
    """Plane-phase linearly polarised pulse (fields.py:69-118)."""
    assert direction == 1 or direction == -1, "direction must be +1 (+x) or -1 (-x)"
    assert int(t_order) == t_order and t_order >= 0, "t_order must be a non-negative integer"
    return Fields(terms=[(SIMPLE, (a0, w0, ctau, t_order, direction, x0, wavelength, pol_angle,
                                   cep))])


def gaussian_laser_pulse(a0, w0, ctau, direction=1, x0=0.0, l0=0.8e-6, pol_angle=0.0, cep=0.0):
    """Paraxial Gaussian pulse focused at x = 0 (fields.py:120-170)."""
    assert direction == 1 or direction == -1, "direction must be +1 (+x) or -1 (-x)"
    return Fields(terms=[(GAUSSIAN, (a0, w0, ctau, direction, x0, l0, pol_angle, cep))])


def focused_gaussian_laser_pulse(a0, w0, ctau, direction=1, xfoc=0.0, x0=0.0, l0=0.8e-6,
                                 pol_angle=0.0, cep=0.0):
    """Tightly focused Gaussian pulse to fifth order in w0/zR (fields.py:173-265)."""
    assert direction == 1 or direction == -1, "direction must be +1 (+x) or -1 (-x)"
    return Fields(terms=[(FOCUSED, (a0, w0, ctau, direction, xfoc, x0, l0, pol_angle, cep))])


def static_field(Ex=0, Ey=0, Ez=0, Bx=0, By=0, Bz=0):
    """Uniform static field (fields.py:268-271)."""
    return Fields(terms=[(STATIC, (Ex, Ey, Ez, Bx, By, Bz))])


def terms_to_c(terms):
    """ctypes array of sfp_field for the library."""
    from ._lib import Field
    arr = (Field * max(1, len(terms)))()
    for i, (kind, params) in enumerate(terms):
        arr[i].kind = int(kind)
        for j, v in enumerate(params):
            arr[i].p[j] = float(v)
    return arr
