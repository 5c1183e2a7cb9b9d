"""Pinning the oracle's GP noise psi (gpnoise.cpp:51-87, kernel.h:39-78) with checks
independent of its code: the impulse dump against the integer restatement, the
27-cell truncation rule, central differences, and the closed-form variance."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import kernels, random_points
from tests.test_rng import py_make_seed, py_pcg


def f32fma(a, b, c):
    """Correctly rounded float32 fma via extended precision (exact for these magnitudes)."""
    r = np.longdouble(np.float32(a)) * np.longdouble(np.float32(b)) + np.longdouble(np.float32(c))
    return np.float32(r)


def py_cell_impulses(kernel, cell, seed):
    """gpnoise.cpp:61-74 restated: state = make_seed(c, seed) + 1 (u32), draws d0..d3 per
    impulse, offset (d2, d1, d0), weight d3 < 0.5 ? -1 : 1, position fma(off, kr, n*kr)."""
    kr = np.float32(kernel["length_scale"])
    cs = py_make_seed(*cell, seed)
    state = (cs + 1) & 0xFFFFFFFF
    pos, w = [], []
    for _ in range(kernel["points_per_cell"]):
        d, state = py_pcg(state, 4)
        u = [np.array([(x >> 9) | 0x3F800000], np.uint32).view(np.float32)[0] - np.float32(1)
             for x in d]
        off = (u[2], u[1], u[0])
        pos.append([f32fma(off[a], kr, np.float32(cell[a]) * kr) for a in range(3)])
        w.append(-1 if d[3] < 2**31 else 1)
    return cs, np.array(pos, np.float32), np.array(w, np.int8)


@pytest.mark.parametrize("ki", range(len(kernels())))
def test_cell_impulses_match_restatement(ki):
    k = kernels()[ki]
    rng = np.random.default_rng(ki)
    cells = rng.integers(-500, 500, (40, 3)).astype(np.int32)
    cells[0] = [0, 0, 0]
    cells[1] = [-1, -1, -1]
    seeds = rng.integers(0, 2**32, 40, dtype=np.uint64).astype(np.uint32)
    cs, pos, w = O.cell_impulses(k, cells, seeds)
    for i in range(40):
        ecs, epos, ew = py_cell_impulses(k, [int(c) for c in cells[i]], int(seeds[i]))
        assert cs[i] == ecs
        assert np.array_equal(pos[i].view(np.uint32), epos.view(np.uint32))
        assert np.array_equal(w[i], ew)
        kr = np.float32(k["length_scale"])
        lo = cells[i].astype(np.float64) * kr
        assert np.all(pos[i] >= lo - 1e-6 * np.abs(lo)) and np.all(pos[i] <= lo + kr * (1 + 1e-6))


def psi_from_dump(kernel, p, seed, floor=False):
    """sigma/sqrt(Var) * sum_{27 cells} sum_i w_i exp(-|p-s_i|^2/L^2) from the impulse dump."""
    kr = np.float32(kernel["length_scale"])
    q = np.float32(p) / kr
    c = np.floor(q).astype(np.int64) if floor else np.trunc(q).astype(np.int64)
    nb = np.array([[dx, dy, dz] for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)])
    cells = (c[None, :] + nb).astype(np.int32)
    _, pos, w = O.cell_impulses(kernel, cells, np.full(27, seed, np.uint32))
    L2 = float(np.sum(np.square(np.float32(kernel["aniso"])), dtype=np.float32))
    d = p[None, None, :].astype(np.float64) - pos.astype(np.float64)
    h = np.exp(-np.sum(d * d, axis=2) / L2)
    ppc = kernel["points_per_cell"]
    var = np.float32(ppc / float(kernel["length_scale"]) ** 3 * np.pi ** 1.5)
    return float(np.float32(kernel["sigma"])) / np.sqrt(np.float64(var)) * np.sum(w * h)


@pytest.mark.parametrize("ki", [0, 1, 2, 3])
def test_psi_is_the_truncated_27_cell_sum(ki):
    """Cell = trunc(p / kr) (toward zero), 27 neighbours, sigma / sqrt(Var) normalisation."""
    k = kernels()[ki]
    rng = np.random.default_rng(50 + ki)
    kr = k["length_scale"]
    p = rng.uniform(-3 * kr, 3 * kr, (60, 3)).astype(np.float32)
    p[:10, 0] = -0.5 * kr  # trunc and floor disagree here
    seeds = rng.integers(0, 2**32, 60, dtype=np.uint64).astype(np.uint32)
    got, _ = O.psi_eval(k, p, seeds, fp64=True)
    differ = 0
    for i in range(60):
        want = psi_from_dump(k, p[i], int(seeds[i]))
        assert abs(got[i, 0] - want) <= 1e-9 * (1 + abs(want))
        alt = psi_from_dump(k, p[i], int(seeds[i]), floor=True)
        differ += abs(alt - want) > 1e-6
    assert differ >= 10  # the floor rule would give other values at negative coordinates


@pytest.mark.parametrize("ki", range(len(kernels())))
def test_gradient_is_derivative(ki):
    """grad psi against FP64 central differences, away from cell boundaries."""
    k = kernels()[ki]
    rng = np.random.default_rng(70 + ki)
    kr = k["length_scale"]
    p = rng.uniform(2 * kr, 40 * kr, (200, 3))
    frac = p / kr - np.floor(p / kr)
    keep = np.all((frac > 0.05) & (frac < 0.95), axis=1)
    p = p[keep].astype(np.float32)
    seeds = rng.integers(0, 2**32, len(p), dtype=np.uint64).astype(np.uint32)
    f0, l1 = O.psi_eval(k, p, seeds, fp64=True)
    h = np.float32(1e-3 * kr)
    for a in range(3):
        e = np.zeros(3, np.float32)
        e[a] = h
        fp, _ = O.psi_eval(k, p + e, seeds, fp64=True)
        fm, _ = O.psi_eval(k, p - e, seeds, fp64=True)
        step = ((p + e)[:, a].astype(np.float64) - (p - e)[:, a].astype(np.float64))
        fd = (fp[:, 0] - fm[:, 0]) / step
        assert np.all(np.abs(fd - f0[:, 1 + a]) <= 2e-4 * (1 + l1 / kr)), a


def test_variance_closed_form():
    """E[psi^2] = sigma^2 (L^2/2)^(3/2) when the support is inside the 27 cells (L << kr)."""
    k = {"sigma": 1.7, "length_scale": 4.0, "aniso": [0.5, 0.5, 0.0], "points_per_cell": 5}
    L2 = 0.5
    rng = np.random.default_rng(3)
    n = 40000
    p = rng.uniform(40.0, 400.0, (n, 3)).astype(np.float32)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    f, _ = O.psi_eval(k, p, seeds, fp64=True)
    want = k["sigma"] ** 2 * (L2 / 2) ** 1.5
    got = np.mean(f[:, 0] ** 2)
    assert abs(got / want - 1) < 0.05, (got, want)
    assert abs(np.mean(f[:, 0])) < 0.02 * np.sqrt(want)


@pytest.mark.parametrize("ki", range(len(kernels())))
def test_fp32_faithful_within_1e5_of_fp64(ki):
    k = kernels()[ki]
    rng = np.random.default_rng(90 + ki)
    p = random_points(rng, 3000, 500.0 * k["length_scale"])
    seeds = rng.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    a, l1 = O.psi_eval(k, p, seeds, fp64=True)
    b, _ = O.psi_eval(k, p, seeds, fp64=False)
    assert np.all(np.abs(a[:, 0] - b[:, 0]) <= 1e-5 * l1)


def test_var_matches_definition():
    for k in kernels():
        want = np.float32(k["points_per_cell"] / float(np.float32(k["length_scale"])) ** 3 * np.pi ** 1.5)
        assert np.float32(O.psi_var(k)) == want
