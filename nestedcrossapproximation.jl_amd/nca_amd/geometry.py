"""Point sets of the benchmark configurations."""
import numpy as np


def fibonacci_sphere(n, radius=1.0):
    """Fibonacci lattice on a sphere: z = 1 - 2(i + 1/2)/n, rho = sqrt(1 - z^2),
    phi = i * pi * (3 - sqrt(5)) (the lattice formula of
    src/admissibility/directionalsubdivision/boundingballdirections.jl:19-28)."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - 2.0 * (i + 0.5) / n
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    pts = np.empty((n, 3), dtype=np.float64)
    pts[:, 0] = radius * rho * np.cos(phi)
    pts[:, 1] = radius * rho * np.sin(phi)
    pts[:, 2] = radius * z
    return pts


def grid_square(nside=32):
    """[SVector(x, y, 0) for x in range(0, 1; length=n), y in ...][:] — x varies
    fastest (README.md:110-113)."""
    g = np.linspace(0.0, 1.0, nside)
    pts = np.zeros((nside * nside, 3), dtype=np.float64)
    pts[:, 0] = np.tile(g, nside)
    pts[:, 1] = np.repeat(g, nside)
    return pts


def ellipsoid(n, axes):
    """Fibonacci lattice mapped by diag(a, b, c)."""
    return fibonacci_sphere(n) * np.asarray(axes, dtype=np.float64)[None, :]


def complex_gaussian(rng, shape):
    """i.i.d. complex Gaussian entries with re, im ~ N(0, 1/2)."""
    s = np.sqrt(0.5)
    return (rng.standard_normal(shape) * s + 1j * rng.standard_normal(shape) * s).astype(np.complex128)
