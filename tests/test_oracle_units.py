"""Oracle functions against what the mathematics fixes, independent of the
reference's outputs (CPU only, through oracle/_build/liboracle.so).

* minmod: the slope of least magnitude when all three agree in sign, else 0;
* the multigrid operator: on a flat column thickness H and a periodic grid,
  cos(2 pi k i / N) is an eigenfunction of L with eigenvalue
  H (2 cos(2 pi k / N) - 2) / dx^2 (a dropped term, a wrong sign or a
  transposed index fails it); constants are its null space, also with walls,
  where the wall faces carry no flux;
* restriction (2:1 full weighting and the general area-weighted form) keeps
  constants and the area-weighted mean; prolongation keeps constants and,
  for 2:1, reproduces linear fields away from the periodic seam;
* V-cycles drive the residual of a manufactured problem to the tolerance.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ORACLE_LIB, REPO

dbl = ctypes.c_double
P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True)
    L = ctypes.CDLL(ORACLE_LIB)
    L.or_minmod.restype = dbl
    L.or_minmod.argtypes = [dbl, dbl, dbl]
    L.or_mg_init.restype = ctypes.c_bool
    L.or_mg_init.argtypes = [ctypes.c_uint, ctypes.c_uint, dbl, dbl, P, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool,
                             dbl, ctypes.c_uint, dbl]
    L.or_mg_nlevels.restype = ctypes.c_int
    L.or_mg_max_residual.restype = dbl
    return L


def ptr(a):
    return a.ctypes.data_as(P)


def init(lib, nx, ny, H, walls=(False, False), L=(1.0, 1.0)):
    Hc = np.full((nx, ny), H) if np.isscalar(H) else np.ascontiguousarray(H, dtype=np.float64)
    assert lib.or_mg_init(nx, ny, L[0], L[1], ptr(Hc), walls[0], walls[1], False, 1e-12, 1000, 1.0)
    return lib.or_mg_nlevels() - 1


def level_size(lib, m):
    a, b = ctypes.c_uint(), ctypes.c_uint()
    lib.or_mg_level_size(m, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def test_minmod_is_least_magnitude_of_agreeing_slopes(lib):
    rng = np.random.default_rng(5)
    for a, b, c in rng.normal(size=(2000, 3)) * rng.choice([1e-3, 1.0, 1e3], size=(2000, 1)):
        s = np.sign([a, b, c])
        want = s[0] * min(abs(a), abs(b), abs(c)) if (s[0] == s[1] == s[2] and s[0] != 0) else 0.0
        assert lib.or_minmod(a, b, c) == want
    assert lib.or_minmod(0.0, 1.0, 2.0) == 0.0 and lib.or_minmod(-1.0, -3.0, -2.0) == -1.0


@pytest.mark.parametrize("kx,ky", [(1, 0), (0, 2), (3, 1), (8, 8)])
def test_operator_eigenfunction_periodic(lib, kx, ky):
    nx, ny, H, Lx, Ly = 32, 16, 700.0, 3.2e6, 1.6e6
    m = init(lib, nx, ny, H, L=(Lx, Ly))
    dx, dy = Lx / nx, Ly / ny
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pi = np.cos(2 * np.pi * kx * i / nx) * np.cos(2 * np.pi * ky * j / ny)
    lam = H * (2 * np.cos(2 * np.pi * kx / nx) - 2) / dx ** 2 + H * (2 * np.cos(2 * np.pi * ky / ny) - 2) / dy ** 2
    res = np.zeros_like(pi)
    lib.or_mg_residual(m, ptr(pi), ptr(np.zeros_like(pi)), ptr(res))  # L(pi) - 0
    assert np.allclose(res, lam * pi, rtol=0, atol=1e-12 * H / dx ** 2)


@pytest.mark.parametrize("walls", [(True, True), (False, True), (True, False), (False, False)])
def test_operator_null_space_and_wall_faces(lib, walls):
    nx, ny = 24, 20
    rng = np.random.default_rng(1)
    m = init(lib, nx, ny, 500.0 + 100 * rng.random((nx, ny)), walls=walls)
    Ow, Os, Osum = (np.zeros((nx, ny)) for _ in range(3))
    lib.or_mg_operators(m, ptr(Ow), ptr(Os), ptr(Osum))
    if walls[0]:
        assert np.all(Ow[0] == 0) and np.all(Ow[1:] > 0)
    if walls[1]:
        assert np.all(Os[:, 0] == 0) and np.all(Os[:, 1:] > 0)
    res = np.zeros((nx, ny))
    lib.or_mg_residual(m, ptr(np.full((nx, ny), 3.0)), ptr(np.zeros((nx, ny))), ptr(res))
    assert np.max(np.abs(res)) <= 1e-13 * np.max(Osum)
    # the operator is symmetric: <L a, b> = <a, L b>
    a, b = rng.normal(size=(2, nx, ny))
    La, Lb = np.zeros((nx, ny)), np.zeros((nx, ny))
    z = np.zeros((nx, ny))
    lib.or_mg_residual(m, ptr(a), ptr(z), ptr(La))
    lib.or_mg_residual(m, ptr(b), ptr(z), ptr(Lb))
    assert abs(np.sum(La * b) - np.sum(a * Lb)) <= 1e-12 * np.sum(np.abs(La * b))


@pytest.mark.parametrize("nx,ny", [(32, 32), (24, 20), (15, 17)])
def test_transfer_operators(lib, nx, ny):
    m = init(lib, nx, ny, 1000.0)
    cx, cy = level_size(lib, m - 1)
    fine = np.full((nx, ny), 2.5)
    coarse = np.zeros((cx, cy))
    lib.or_mg_restrict(m, ptr(fine), ptr(coarse))
    assert np.allclose(coarse, 2.5, rtol=1e-15, atol=0)
    back = np.zeros((nx, ny))
    lib.or_mg_interpolate(m, ptr(np.full((cx, cy), -1.25)), ptr(back))
    assert np.allclose(back, -1.25, rtol=1e-15, atol=0)
    # area-weighted mean is kept: each coarse cell covers (nx/cx) x (ny/cy) fine cells
    rng = np.random.default_rng(2)
    fine = rng.normal(size=(nx, ny))
    lib.or_mg_restrict(m, ptr(fine), ptr(coarse))
    assert abs(coarse.mean() - fine.mean()) <= 1e-13
    if nx == 2 * cx and ny == 2 * cy:
        # bilinear 3/4-1/4 prolongation reproduces a linear field at interior points
        ic, jc = np.meshgrid(np.arange(cx), np.arange(cy), indexing="ij")
        lin_c = 0.3 * (2 * ic + 0.5) - 0.7 * (2 * jc + 0.5)
        lib.or_mg_interpolate(m, ptr(np.ascontiguousarray(lin_c)), ptr(back))
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        lin_f = 0.3 * i - 0.7 * j
        sl = (slice(1, nx - 1), slice(1, ny - 1))
        assert np.allclose(back[sl], lin_f[sl], rtol=0, atol=1e-12)


def test_vcycles_solve_manufactured_problem(lib):
    nx, ny = 64, 64
    rng = np.random.default_rng(3)
    m = init(lib, nx, ny, 800.0 + 200 * rng.random((nx, ny)), walls=(True, True), L=(1e6, 1e6))
    exact = rng.normal(size=(nx, ny))
    rhs = np.zeros((nx, ny))
    lib.or_mg_residual(m, ptr(exact), ptr(np.zeros((nx, ny))), ptr(rhs))  # rhs = L(exact)
    pi = np.zeros((nx, ny))
    r0 = lib.or_mg_max_residual(m, ptr(pi), ptr(rhs))
    for _ in range(30):
        lib.or_mg_vcycle(ptr(pi), ptr(rhs))
    assert lib.or_mg_max_residual(m, ptr(pi), ptr(rhs)) <= 1e-9 * r0
    d = (pi - pi.mean()) - (exact - exact.mean())
    assert np.max(np.abs(d)) <= 1e-6 * np.max(np.abs(exact))
