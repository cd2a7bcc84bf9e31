from .relaxation import (
    matrix_diagonal,
    inverse_diagonal,
    jacobi,
    polynomial,
    approximate_spectral_radius,
    none,
)
