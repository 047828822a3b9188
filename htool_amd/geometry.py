"""Seeded synthetic point clouds and kernel parameters for the H-matrix workloads.

All generators use numpy's PCG64 (``numpy.random.default_rng(seed)``) so the
same seed gives the same points on every machine. Points are returned as a
C-contiguous ``(N, 3)`` float64 array, which is the interleaved ``xt[3*i+d]``
layout the reference's ``HMatrix::build(A, xt, xs)`` takes
(include/htool/types/hmatrix.hpp:285).
"""
from __future__ import annotations

import math

import numpy as np

# Kernel families, shared with include/htool_amd.h (HTG_KERNEL_*).
KERNEL_LAPLACE = 0          # 1/(4 pi r), value `diag` on the diagonal of a square same-point matrix
KERNEL_INV_DELTA_R = 1      # 1/(delta + r)
KERNEL_INV_DELTA_4PIR = 2   # 1/(delta + 4 pi r)   (GeneratorTestDoubleSymmetric)
KERNEL_SMOOTH = 3           # 1/(4 pi sqrt(delta + r^2))   (examples/smallest_example.cpp)
KERNEL_HELMHOLTZ = 4        # e^{ikr}/(4 pi r), complex; `diag` on the diagonal of a square same-point matrix
KERNEL_COMPLEX_TEST = 5     # (1 + i)/(4 pi r), complex   (GeneratorTestComplex)
KERNEL_ZERO = 6             # every entry 0   (GeneratorTestZero, test_hmat_zero.hpp)
KERNEL_HERMITIAN_TEST = 7   # (1 + sign(x_i - x_j) i)/(1e-5 + 4 pi r), complex   (GeneratorTestComplexHermitian)
COMPLEX_KERNELS = (KERNEL_HELMHOLTZ, KERNEL_COMPLEX_TEST, KERNEL_HERMITIAN_TEST)


def cylinder(n: int, seed: int = 0, radius: float = 1.0, height: float = 1.0) -> np.ndarray:
    """Uniform points on the lateral surface of a cylinder (theta~U[0,2pi), z~U[0,h))."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * math.pi, n)
    z = rng.uniform(0.0, height, n)
    return np.ascontiguousarray(np.stack([radius * np.cos(theta), radius * np.sin(theta), z], axis=1))


def sphere(n: int, seed: int = 0, radius: float = 1.0) -> np.ndarray:
    """Uniform points on a sphere: normalised Gaussian triples."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray(radius * g)


def ellipsoid(n: int, axes=(1.0, 1.0, 4.0), seed: int = 0) -> np.ndarray:
    """Area-uniform points on an ellipsoid with semi-axes ``axes``.

    Points of the unit sphere are mapped by diag(axes) and kept with probability
    proportional to the local area stretch (rejection sampling), which makes the
    accepted points uniform with respect to surface area.
    """
    a, b, c = (float(v) for v in axes)
    rng = np.random.default_rng(seed)
    out = []
    have = 0
    # Largest area stretch of the map x -> diag(a,b,c) x on the unit sphere.
    gmax = max(a * b, b * c, a * c)
    while have < n:
        m = max(2 * (n - have), 1024)
        g = rng.standard_normal((m, 3))
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        stretch = np.sqrt((b * c * g[:, 0]) ** 2 + (a * c * g[:, 1]) ** 2 + (a * b * g[:, 2]) ** 2)
        keep = rng.uniform(0.0, gmax, m) < stretch
        p = g[keep] * np.array([a, b, c])
        out.append(p)
        have += p.shape[0]
    return np.ascontiguousarray(np.concatenate(out, axis=0)[:n])


def prolate_spheroid_area(a: float, c: float) -> float:
    """Surface area of the spheroid with semi-axes (a, a, c), c > a:
    2 pi a^2 (1 + c / (a e) asin e), e = sqrt(1 - a^2 / c^2)."""
    e = np.sqrt(1.0 - (a / c) ** 2)
    return float(2 * np.pi * a * a * (1 + c / (a * e) * np.arcsin(e)))


def disk(n: int, z: float, seed: int = 0) -> np.ndarray:
    """Uniform points on the unit disk in the plane ``z`` (the reference's test geometry,
    include/htool/testing/geometry.hpp:10-22, drawn with PCG64 instead of C rand())."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 1.0, n)
    theta = rng.uniform(0.0, 1.0, n)
    return np.ascontiguousarray(
        np.stack([np.sqrt(rho) * np.cos(2 * math.pi * theta), np.sqrt(rho) * np.sin(2 * math.pi * theta),
                  np.full(n, z)], axis=1))


def sphere_spacing(n: int, radius: float = 1.0) -> float:
    """Mean point spacing h = sqrt(area / N) on a sphere."""
    return math.sqrt(4.0 * math.pi * radius * radius / n)


def laplace_diag(h: float) -> float:
    """Diagonal value 1/(4 pi h) used for the singular Laplace kernel."""
    return 1.0 / (4.0 * math.pi * h)
