"""Smoother specifications and per-level setup (mirror of amjax/relaxation/smoothing.py).

A specification is a name (``'jacobi'``), a ``(name, kwargs)`` tuple, ``None``
(no smoothing), or a list of those, one per level with the last entry repeated
(smoothing.py:36-56).  Jacobi takes ``iterations`` (default 1), ``omega``
(default 1.0) and ``withrho`` (divide omega by the power-iteration estimate of
rho(D^-1 A), computed once per level and shared by pre and post smoother,
smoothing.py:85-99).  The resolved (kind, iterations, omega) of every level is
what the native cycle runs; the Python callables exist for API parity and run
the same GPU kernel.

Beyond the reference (which lists them as not implemented, smoothing.py:120-127),
the polynomial smoothers of PyAMG's roadmap are available as fused sweeps:
``'chebyshev'`` (lower_bound 1/30, upper_bound 1.1, degree 3, iterations 1) and
``'richardson'`` (omega 1, iterations 1), with PyAMG's definitions
(pyamg.relaxation.smoothing setup_chebyshev / setup_richardson and
relaxation.polynomial).  rho(A) is estimated with the power iteration of
``approximate_spectral_radius`` on A itself (D^-1 = I) instead of PyAMG's
Lanczos/Arnoldi estimate.
"""

import numpy as np
import torch

from . import relaxation

DEFAULT_NITER = 1
SYMMETRIC_RELAXATION = ["jacobi", "chebyshev", "richardson", None]

_NOT_IMPLEMENTED_SMOOTHERS = {
    "gauss_seidel", "sor", "block_jacobi", "block_gauss_seidel",
    "schwarz", "strength_based_schwarz",
    "jacobi_ne", "gauss_seidel_ne", "gauss_seidel_nr",
    "gauss_seidel_indexed", "cf_jacobi", "fc_jacobi",
    "cf_block_jacobi", "fc_block_jacobi", "gmres", "cg", "cgne", "cgnr",
    "bsr_gauss_seidel", "bsr_jacobi", "bsr_jacobi_indexed",
}


def _unpack_arg(v):
    if isinstance(v, tuple):
        return v[0], v[1]
    return v, {}


def _as_list(v):
    if isinstance(v, list):
        return v
    if isinstance(v, (str, tuple)) or v is None:
        return [v]
    raise ValueError(f"Unrecognized smoother specification: {v!r}")


def _expand_to_n(specs, n):
    if not specs:
        raise ValueError("Empty smoother specification.")
    return (list(specs) + [specs[-1]] * n)[:n]


def _bake_kwargs(kw, smoother):
    """Replace ``withrho=True`` by the resolved omega (smoothing.py:59-82); polynomial
    smoothers keep their resolved coefficients."""
    if getattr(smoother, "_kind", None) == "poly":
        return {"iterations": smoother._iterations, "coefficients": list(smoother._coefficients)}
    if not kw.get("withrho", False):
        return dict(kw)
    return {**kw, "omega": smoother._omega, "withrho": False}


def setup_jacobi(lvl, iterations=DEFAULT_NITER, omega=1.0, withrho=False):
    """Jacobi smoother of one level (smoothing.py:85-99)."""
    if withrho:
        if not hasattr(lvl, "_spectral_radius"):
            lvl._spectral_radius = relaxation.approximate_spectral_radius(lvl.A, lvl.Dinv)
        omega = omega / lvl._spectral_radius
    Dinv = lvl.Dinv

    def smoother(A, x, b):
        return relaxation.jacobi(A, x, b, Dinv, iterations=iterations, omega=omega)

    smoother.__name__ = "jacobi"
    smoother._omega = omega
    smoother._iterations = int(iterations)
    smoother._kind = "jacobi"
    return smoother


def _rho_A(lvl):
    if not hasattr(lvl, "_spectral_radius_A"):
        d = lvl.Dinv
        one = torch.ones_like(d) if torch.is_tensor(d) else np.ones_like(np.asarray(d))
        lvl._spectral_radius_A = relaxation.approximate_spectral_radius(lvl.A, one)
    return lvl._spectral_radius_A


def chebyshev_polynomial_coefficients(a, b, degree):
    """Coefficients (highest power first) of the monic polynomial whose roots are the
    ``degree`` Chebyshev points of [a, b], scaled to C(0) = 1 (PyAMG's
    relaxation.chebyshev.chebyshev_polynomial_coefficients)."""
    if a >= b or a <= 0:
        raise ValueError(f"invalid interval [{a},{b}]")
    std_roots = np.cos(np.pi * (np.arange(degree) + 0.5) / degree)
    scaled_roots = 0.5 * (b - a) * (1 + std_roots) + a
    scaled_poly = np.poly(scaled_roots)
    scaled_poly /= np.polyval(scaled_poly, 0)
    return scaled_poly


def _poly_smoother(name, coefficients, iterations):
    coefficients = [float(c) for c in coefficients]
    if not 1 <= len(coefficients) <= 8:
        raise ValueError("polynomial smoothers take 1 to 8 coefficients")

    def smoother(A, x, b):
        return relaxation.polynomial(A, x, b, coefficients, iterations=iterations)

    smoother.__name__ = name
    smoother._omega = coefficients[0]
    smoother._coefficients = coefficients
    smoother._iterations = int(iterations)
    smoother._kind = "poly"
    return smoother


def setup_chebyshev(lvl, lower_bound=1.0 / 30.0, upper_bound=1.1, degree=3, iterations=DEFAULT_NITER,
                    coefficients=None):
    """Chebyshev polynomial smoother: x += p(A) (b - A x) with 1 - l p(l) the degree-``degree``
    Chebyshev polynomial on [rho lower_bound, rho upper_bound] scaled to 1 at 0.
    ``coefficients`` (p's, highest power first) skip the setup."""
    if coefficients is None:
        rho = _rho_A(lvl)
        coefficients = -chebyshev_polynomial_coefficients(rho * lower_bound, rho * upper_bound, degree)[:-1]
    return _poly_smoother("chebyshev", coefficients, iterations)


def setup_richardson(lvl, iterations=DEFAULT_NITER, omega=1.0, coefficients=None):
    """Richardson smoother: x += (omega / rho(A)) (b - A x)."""
    if coefficients is None:
        coefficients = [omega / _rho_A(lvl)]
    return _poly_smoother("richardson", coefficients, iterations)


def setup_none(lvl, **kwargs):
    """Identity smoother (smoothing.py:102-117)."""
    del lvl, kwargs

    def smoother(A, x, b):
        del A, b
        return x

    smoother.__name__ = "none"
    smoother._omega = 1.0
    smoother._iterations = 0
    smoother._kind = "none"
    return smoother


_REGISTRY = {"jacobi": setup_jacobi, "chebyshev": setup_chebyshev, "richardson": setup_richardson,
             "none": setup_none}


def _setup_call(fn):
    """Return the setup function of smoother ``fn`` (smoothing.py:130-158)."""
    if fn is None:
        fn = "none"
    if not isinstance(fn, str):
        raise ValueError(f"Smoother name must be a string or None, got {fn!r}")
    if fn in _NOT_IMPLEMENTED_SMOOTHERS:
        raise NotImplementedError(f"'{fn}' smoother is not yet implemented in amjax_amd. "
                                  f"Available: {sorted(_REGISTRY)}")
    if fn not in _REGISTRY:
        raise ValueError(f"Unknown smoother {fn!r}. Available: {sorted(_REGISTRY)}")
    return _REGISTRY[fn]


def change_smoothers(ml, presmoother, postsmoother):
    """Set pre/post smoothers on every non-coarsest level (smoothing.py:161-228).

    Sets ``ml.symmetric_smoothing`` (same method and iteration count, method in
    SYMMETRIC_RELAXATION).  Invalidates the solver's native plan.
    """
    n = len(ml.levels) - 1
    if n <= 0:
        return
    pre_specs = _expand_to_n(_as_list(presmoother), n)
    post_specs = _expand_to_n(_as_list(postsmoother), n)
    # validate every spec before touching the levels
    for s in pre_specs + post_specs:
        _setup_call(_unpack_arg(s)[0])
    ml.symmetric_smoothing = True
    for i in range(n):
        pre_fn, pre_kw = _unpack_arg(pre_specs[i])
        post_fn, post_kw = _unpack_arg(post_specs[i])
        lvl = ml.levels[i]
        lvl.presmoother = _setup_call(pre_fn)(lvl, **pre_kw)
        lvl.postsmoother = _setup_call(post_fn)(lvl, **post_kw)
        lvl._presmoother_spec = (pre_fn, _bake_kwargs(pre_kw, lvl.presmoother))
        lvl._postsmoother_spec = (post_fn, _bake_kwargs(post_kw, lvl.postsmoother))
        it1 = pre_kw.get("iterations", DEFAULT_NITER)
        it2 = post_kw.get("iterations", DEFAULT_NITER)
        if it1 != it2 or pre_fn != post_fn or pre_fn not in SYMMETRIC_RELAXATION:
            ml.symmetric_smoothing = False
    if hasattr(ml, "_invalidate"):
        ml._invalidate()


def rebuild_smoother(lvl):
    """Rebuild a level's smoothers from their stored specs (smoothing.py:231-254)."""
    if getattr(lvl, "_presmoother_spec", None) is None:
        raise AttributeError("Missing _presmoother_spec on level.")
    if getattr(lvl, "_postsmoother_spec", None) is None:
        raise AttributeError("Missing _postsmoother_spec on level.")
    fn1, kw1 = lvl._presmoother_spec
    fn2, kw2 = lvl._postsmoother_spec
    lvl.presmoother = _setup_call(fn1)(lvl, **kw1)
    lvl.postsmoother = _setup_call(fn2)(lvl, **kw2)
