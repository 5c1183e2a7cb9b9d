"""Pinning the oracle's SDF and grid media (sdfmedium.cpp, externalgpmedium.cpp,
GridMeanFunction in meanfunction.cpp:41-86): hash11 against an integer
restatement, the trilinear grid against fields it must reproduce, sphere-traced
marches against analytic ray-sphere distances, and the medium priority of the
scene lowering (controller.cpp:230-234)."""
import numpy as np
import pytest

from mupsi_amd import _abi as A
from mupsi_amd import scenes
from mupsi_amd.scene import build_desc
from oracle import oracle as O

M32 = 0xFFFFFFFF


def _hash11_int(x, y, z, seed):
    """externalgpmedium.cpp:8-17 in Python integers."""
    h = int(seed)
    for v in (x, y, z):
        b = int(np.float32(v).view(np.uint32))
        h = (h * 0x9E3779B1 + b) & M32
    h = ((h ^ (h >> 16)) * 0x85EBCA6B) & M32
    h = ((h ^ (h >> 13)) * 0xC2B2AE35) & M32
    h = h ^ (h >> 16)
    return np.float32(np.float32(h & 0xFFFFFF) / np.float32(0xFFFFFF) * np.float32(2.0) - np.float32(1.0))


def test_hash11_matches_integer_restatement():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 300, (2000, 3)).astype(np.float32)
    p[:5] = [[0, 0, 0], [-0.0, 0, 0], [1, 2, 3], [np.float32(1e-30), 7, -7], [65504, -65504, 3.5]]
    seeds = rng.integers(0, 2 ** 32, 2000, dtype=np.uint64).astype(np.uint32)
    seeds[0] = 0
    got = O.hash11(p, seeds)
    ref = np.array([_hash11_int(*p[i], seeds[i]) for i in range(len(p))], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert got[0] == -1.0                       # all-zero input hashes to 0 -> -1
    assert np.all(np.abs(got) <= 1.0) and got.std() > 0.5


def _grid_scene(mean, res, bmin, bmax, variance=None, noise_scale=1.0):
    sc = scenes.medium_scene("grid", width=16, height=16)
    sc["external_gp_medium"].update(resolution=res, bounds_min=bmin, bounds_max=bmax,
                                    mean=mean, variance=variance, noise_scale=noise_scale)
    return sc


def _cell_centres(res, bmin, bmax):
    return [bmin[a] + (np.arange(res) + 0.5) * (bmax[a] - bmin[a]) / res for a in range(3)]


def test_grid_reproduces_linear_fields_and_clamps_outside():
    """Trilinear interpolation over cell centres is exact for affine fields inside
    the centre lattice; outside it the clamped weights hold the boundary value
    along that axis; the central-difference gradient is the field's slope."""
    res, bmin, bmax = 12, [-1.0, -2.0, 0.5], [1.0, 2.0, 2.5]
    a, b = np.array([0.7, -0.3, 0.45]), 0.2
    c = _cell_centres(res, bmin, bmax)
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    mean = (a[0] * X + a[1] * Y + a[2] * Z + b).astype(np.float32).reshape(-1)
    osc = O.OracleScene(_grid_scene(mean, res, bmin, bmax))
    rng = np.random.default_rng(2)
    lo = np.array([c[k][0] for k in range(3)])
    hi = np.array([c[k][-1] for k in range(3)])
    p = (lo + rng.uniform(0, 1, (500, 3)) * (hi - lo)).astype(np.float32)
    f = osc.field_eval(p, np.zeros(500, np.uint32))
    np.testing.assert_allclose(f[:, 0], p.astype(np.float64) @ a + b, atol=2e-6)
    inner = np.all((p - lo > 2e-4) & (hi - p > 2e-4), axis=1)   # stencil inside the lattice
    np.testing.assert_allclose(f[inner, 1:], np.tile(a, (inner.sum(), 1)), atol=3e-3)
    # beyond the last centre along x the value is held at x = c_x[-1]
    q = p.copy()
    q[:, 0] = bmax[0] + rng.uniform(0.0, 3.0, 500)
    fq = osc.field_eval(q, np.zeros(500, np.uint32))
    held = q.astype(np.float64).copy()
    held[:, 0] = c[0][-1]
    np.testing.assert_allclose(fq[:, 0], held @ a + b, atol=2e-6)


def test_grid_noise_is_scaled_hash():
    """value = mean + noise_scale * sqrt(var) * hash11(p, seed) with constant grids."""
    res, bmin, bmax = 4, [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    mean = np.full(res ** 3, 0.25, np.float32)
    var = np.full(res ** 3, 4.0, np.float32)
    osc = O.OracleScene(_grid_scene(mean, res, bmin, bmax, variance=var, noise_scale=0.5))
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.5, 0.5, (300, 3)).astype(np.float32)
    seeds = rng.integers(0, 2 ** 32, 300, dtype=np.uint64).astype(np.uint32)
    f = osc.field_eval(p, seeds)
    h = O.hash11(p, seeds).astype(np.float64)
    np.testing.assert_allclose(f[:, 0], 0.25 + 0.5 * 2.0 * h, atol=1e-6)
    # zero variance: no noise
    osc0 = O.OracleScene(_grid_scene(mean, res, bmin, bmax, variance=np.zeros_like(var)))
    np.testing.assert_allclose(osc0.field_eval(p, seeds)[:, 0], 0.25, atol=1e-7)


def _rays_to(center, radius, n, rng, origin_r):
    """Rays from points at distance origin_r from center, aimed near the centre."""
    d0 = rng.normal(size=(n, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    o = center + d0 * origin_r
    tgt = center + rng.normal(0, radius * 0.6, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1.5e-3
    rays[:, 7] = np.finfo(np.float32).max
    return rays


def _sphere_t(rays, c, r, far=False):
    o = rays[:, :3].astype(np.float64) - c
    d = rays[:, 3:6].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = np.sum(o * d, 1)
    disc = b * b - (np.sum(o * o, 1) - r * r)
    t = -b + (1 if far else -1) * np.sqrt(np.maximum(disc, 0))
    return np.where(disc > 0, t, np.inf)


def test_sdf_sphere_march_hits_the_analytic_sphere():
    """Sphere tracing toward a sphere mean approaches it in steps of |f| until
    the 1e-3 minimum step crosses it; bisection then lands within float precision
    of the analytic distance, the normal is radial, and a miss runs out at ft or
    after 512 steps.  The reference detects a crossing by f_prev * f < 0, so a
    step that lands where f rounds to exactly 0 hides the entry and the march
    reports the exit on the far side (sdfmedium.cpp:36); that happens here for
    a few rays and is kept."""
    sc = scenes.medium_scene("sdf", width=16, height=16)
    osc = O.OracleScene(sc)
    c = np.array([0.0, 0.0, -500.0])
    rng = np.random.default_rng(4)
    rays = _rays_to(c, 70.0, 400, rng, 120.0)
    m = osc.march(rays, np.zeros(400, np.uint32))
    ta = _sphere_t(rays, c, 70.0)
    te = _sphere_t(rays, c, 70.0, far=True)
    hit = m["hit"].astype(bool)
    assert np.array_equal(hit, np.isfinite(ta)) and hit.sum() > 100
    entry = np.abs(m["t"] - ta) <= 2e-3
    exit_ = np.abs(m["t"] - te) <= 2e-3
    assert np.all(entry[hit] | exit_[hit])
    assert entry[hit].mean() > 0.9
    radial = m["p"][hit] - c
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    np.testing.assert_allclose(m["n"][hit], radial, atol=1e-4)
    assert np.all(m["steps"][~hit] <= 512)


def test_grid_march_without_noise_finds_the_sampled_sphere():
    """The trilinear zero set of a sampled sphere distance lies within the grid's
    interpolation error (here < 0.3 for 5.4-unit cells on a radius-70 sphere) of
    the sphere; the hit lies in the last step's bisection bracket."""
    sc = scenes.medium_scene("grid", width=16, height=16, noise_variance=0.0)
    osc = O.OracleScene(sc)
    c = np.array([0.0, 0.0, -500.0])
    rng = np.random.default_rng(6)
    rays = _rays_to(c, 70.0, 300, rng, 110.0)
    m = osc.march(rays, np.zeros(300, np.uint32))
    ta = _sphere_t(rays, c, 70.0)
    clear = np.isfinite(ta) & (np.abs(np.linalg.norm(np.cross(rays[:, 3:6], rays[:, :3] - c), axis=1) - 70.0) > 3.0)
    hit = m["hit"].astype(bool)
    assert hit[clear].all() and clear.sum() > 80
    np.testing.assert_allclose(m["t"][clear], ta[clear], atol=0.3)
    miss = ~np.isfinite(ta)
    assert not hit[miss].any()


def test_medium_priority_of_the_scene_lowering():
    """int_medium takes gp_medium, else sdf_medium, else external_gp_medium."""
    gp = scenes.preset_scene("sphere_renewal_256")
    sdf = scenes.medium_scene("sdf")
    grid = scenes.medium_scene("grid")
    both = dict(gp)
    both["sdf_medium"] = sdf["sdf_medium"]
    d, _ = build_desc(both)
    assert d.medium_kind == A.MEDIUM_GP and d.has_gp_medium == 1
    s2 = dict(sdf)
    s2["external_gp_medium"] = grid["external_gp_medium"]
    d, _ = build_desc(s2)
    assert d.medium_kind == A.MEDIUM_SDF and d.mean_radius == 70.0
    d, _ = build_desc(grid)
    assert d.medium_kind == A.MEDIUM_GRID and d.grid.resolution == 48


def test_grid_file_reads_zeros_where_absent(tmp_path):
    """GridMeanFunction zero-fills before reading: a short or missing file."""
    from mupsi_amd.scene import read_grid
    np.arange(10, dtype="<f4").tofile(tmp_path / "short.raw")
    g = read_grid(tmp_path / "short.raw", 3)
    assert g.shape == (27,) and np.array_equal(g[:10], np.arange(10)) and not g[10:].any()
    assert not read_grid(tmp_path / "missing.raw", 2).any()


@pytest.mark.parametrize("kind", ["sdf", "grid"])
def test_media_integrator_sky_and_unbiased_light(kind):
    """Paths through the SDF / grid media end on the sky or the Lambertian
    surface; images are finite and non-negative and the medium region darkens
    the sky behind it."""
    sc = scenes.medium_scene(kind, width=32, height=32)
    osc = O.OracleScene(sc)
    img, paths, _ = osc.render(0, 2)
    assert paths == 2 * 32 * 32
    assert np.isfinite(img).all() and (img >= 0).all()
    centre = img[12:20, 12:20].mean()
    corner = img[:4, :4].mean()
    assert centre != corner
