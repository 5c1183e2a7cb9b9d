"""Pins of the oracle's light sampling and next-event estimation by mathematics
(no reference binary is available).

Sphere::sampleDirect (src/primitives/sphere.cpp:91-133): a uniform direction in
the cone of half-angle theta_max (cos theta_max = sqrt(d^2 - 1)/d, d the centre
distance in radii) towards the sphere, the distance to the first intersection,
pdf 1/(2 pi (1 - cos theta_max)) and weight Le / pdf.  Mesh::sampleDirect
(src/primitives/mesh.cpp:252-297): a face drawn with probability proportional to
its area, a uniform point on it, pdf 1/area (area measure), weight Le * area.
Scene::chooseLight (src/geometry/scene.cpp:50-61): a uniform emitter index
nextI() % count, and no draw at all without emitters.  The next-event term of
handleSurface (src/rendering/trace.cpp:43-58) adds T * weight(bsdf, wi = d) *
Le / pdf * |d.n|; with a Lambertian surface weight() is the albedo for d above
the surface and 0 below, so under a sphere light fully above the horizon its
expectation is albedo * Le * pi * sin^2(theta_max) * cos(alpha) (alpha between
the normal and the light's centre), and under a mesh light albedo * Le * area *
E[max(d.n, 0)].
"""
import numpy as np
import pytest
from scipy import stats

from mupsi_amd import scenes
from oracle import oracle as O

MASK64 = (1 << 64) - 1


def _pcg_next(state: int):
    """Random::nextI (PCG-XSH-RR, increment 1): (output, new state)."""
    old = state
    state = (old * 6364136223846793005 + 1) & MASK64
    xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
    rot = old >> 59
    return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF, state


def _scene(prims, *, extensions=False, camera=None, width=8, height=8, max_bounce=1):
    sc = {"width": width, "height": height, "camera": camera or dict(scenes.CAMERA),
          "max_bounce": max_bounce, "primitives": prims}
    if extensions:
        sc["extensions"] = {"mesh_emission": True}
    return sc


def _states(rng, n):
    return rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)


SPHERES = [([0.0, 0.0, 0.0], 1.0), ([3.0, -2.0, 5.0], 2.5)]


@pytest.mark.parametrize("center,radius", SPHERES)
@pytest.mark.parametrize("dist_in_radii", [1.5, 4.0, 20.0])
def test_sphere_light_cone_distance_pdf_weight(center, radius, dist_in_radii):
    Le = np.array([3.0, 5.0, 7.0])
    sc = _scene([{"type": "sphere", "center": center, "radius": radius, "bsdf": "lambertian",
                  "emission": Le.tolist()}])
    osc = O.OracleScene(sc)
    rng = np.random.default_rng(int(dist_in_radii * 10) + int(radius))
    n = 20000
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    c = np.asarray(center, np.float64)
    p = (c + axis * dist_in_radii * radius).astype(np.float32)
    P = np.repeat(p[None, :], n, axis=0)
    r = osc.choose_light(P, _states(rng, n))
    assert np.all(r["light"] == 0)
    pc = c - p.astype(np.float64)
    D = np.linalg.norm(pc)
    cos_max = np.sqrt(D * D - radius * radius) / D
    a = pc / D
    d = r["d"].astype(np.float64)
    cosang = d @ a
    # inside the cone, and the cosine uniform on [cos_max, 1] (uniform cap)
    assert np.all(cosang >= cos_max - 2e-5)
    ks = stats.kstest((cosang - cos_max) / (1 - cos_max), "uniform")
    assert ks.pvalue > 1e-3, ks
    # E[d] = a (1 + cos_max) / 2 for the uniform cap: the azimuth is uniform too
    m = d.mean(axis=0)
    se = d.std(axis=0) / np.sqrt(n)
    assert np.all(np.abs(m - a * (1 + cos_max) / 2) < 5 * se + 1e-6)
    # the distance is the first intersection of the ray with the sphere
    q = p.astype(np.float64) + d * r["dist"][:, None]
    assert np.all(np.abs(np.linalg.norm(q - c, axis=1) - radius) < 2e-5 * D)
    # ... and the nearer one: at most b = d.(c - p), the midpoint of the two roots;
    # away from grazing directions (where the root is ill-conditioned) the closed form
    b = d @ pc
    det = np.sqrt(np.maximum(b * b - (D * D - radius * radius), 0))
    assert np.all(r["dist"] <= b + 2e-5 * D)
    ok = det > 0.2 * radius
    assert ok.mean() > 0.5
    assert np.all(np.abs(r["dist"] - (b - det))[ok] < 2e-5 * D)
    # solid-angle pdf and weight Le / pdf
    pdf = 1.0 / (2 * np.pi * (1 - cos_max))
    assert np.all(np.abs(r["pdf"] / pdf - 1) < 1e-4 / (1 - cos_max) * 1e-2 + 2e-5)
    np.testing.assert_allclose(r["w"], (Le[None, :] / r["pdf"][:, None]).astype(np.float32),
                               rtol=1e-6)


def test_sphere_light_from_inside_draws_but_fails():
    sc = _scene([{"type": "sphere", "center": [0, 0, 0], "radius": 2.0, "bsdf": "lambertian",
                  "emission": [1, 1, 1]}])
    osc = O.OracleScene(sc)
    st = np.array([12345], np.uint64)
    r = osc.choose_light(np.array([[0.5, 0.2, -0.3]], np.float32), st)
    assert r["light"][0] == -2
    _, s1 = _pcg_next(12345)
    assert int(r["state"][0]) == s1      # the index draw happened, nothing more


def _tri_mesh():
    """Three separated triangles of areas 0.5, 1 and 1.5 in the plane y = 4,
    normals -y (facing a point below)."""
    v, f = [], []
    for i, s in enumerate([1.0, np.sqrt(2.0), np.sqrt(3.0)]):
        x0 = 5.0 * i
        base = len(v)
        v += [[x0, 4.0, 0.0], [x0, 4.0, s], [x0 + s, 4.0, 0.0]]
        f.append([base, base + 1, base + 2])
    return {"vertices": np.array(v, np.float32), "faces": np.array(f, np.int32),
            "normals": None, "texcoords": None, "face_tex": None}


@pytest.mark.parametrize("scale", [1.0, 2.0])
def test_mesh_light_faces_by_area_uniform_points(scale):
    Le = np.array([2.0, 4.0, 8.0])
    m = _tri_mesh()
    prim = {"type": "mesh", "mesh": m, "bsdf": "lambertian", "emission": Le.tolist()}
    if scale != 1.0:
        prim["transform"] = {"scale": scale, "position": [1.0, -2.0, 0.5]}
    osc = O.OracleScene(_scene([prim], extensions=True))
    rng = np.random.default_rng(3)
    n = 60000
    p = np.array([2.0, -1.0, 0.3], np.float32)
    r = osc.choose_light(np.repeat(p[None, :], n, axis=0), _states(rng, n))
    assert np.all(r["light"] == 0)
    q = p.astype(np.float64) + r["d"].astype(np.float64) * r["dist"][:, None]
    # back to local space
    pos = np.array([1.0, -2.0, 0.5]) if scale != 1.0 else np.zeros(3)
    ql = (q - pos) / scale
    assert np.all(np.abs(ql[:, 1] - 4.0) < 2e-4)
    face = np.floor(ql[:, 0] / 5.0 + 0.5).astype(int)
    face = np.clip(face, 0, 2)
    areas = np.array([0.5, 1.0, 1.5])
    counts = np.bincount(face, minlength=3)
    chi = stats.chisquare(counts, areas / areas.sum() * n)
    assert chi.pvalue > 1e-3, (counts, chi)
    V = m["vertices"].astype(np.float64)
    for i in range(3):
        sel = ql[face == i]
        cen = V[3 * i:3 * i + 3].mean(axis=0)
        se = sel.std(axis=0) / np.sqrt(len(sel))
        assert np.all(np.abs(sel.mean(axis=0) - cen) < 5 * se + 1e-6), i
        # every point inside its triangle (barycentrics in [0, 1])
        a, b, c = V[3 * i], V[3 * i + 1], V[3 * i + 2]
        M = np.stack([b - a, c - a], axis=1)[[0, 2]]
        uv = np.linalg.solve(M, (sel - a)[:, [0, 2]].T).T
        assert np.all(uv > -1e-5) and np.all(uv.sum(axis=1) < 1 + 1e-5)
    # area measure in the mesh's local space, no cos / dist^2 conversion
    A = areas.sum()
    np.testing.assert_allclose(r["pdf"], np.float32(1.0 / A), rtol=1e-6)
    np.testing.assert_allclose(r["w"], (Le * A)[None, :].repeat(n, 0), rtol=1e-5)


def test_choose_light_uniform_index_and_no_draw_without_lights():
    prims = [{"type": "sphere", "center": [0, 10, 0], "radius": 1.0, "bsdf": "lambertian",
              "emission": [1, 1, 1]},
             {"type": "sphere", "center": [0, -10, 0], "radius": 1.0, "bsdf": "null"},
             {"type": "sphere", "center": [10, 0, 0], "radius": 1.0, "bsdf": "lambertian",
              "emission": [2, 2, 2]},
             {"type": "mesh", "mesh": _tri_mesh(), "bsdf": "lambertian", "emission": [3, 3, 3]}]
    osc = O.OracleScene(_scene(prims, extensions=True))
    rng = np.random.default_rng(9)
    n = 30000
    st = _states(rng, n)
    r = osc.choose_light(np.zeros((n, 3), np.float32), st)
    lights = np.array([0, 2, 3])
    # the index is nextI() % 3 of the incoming state, then that light's sampleDirect
    exp = np.array([lights[_pcg_next(int(s))[0] % 3] for s in st[:2000]])
    assert np.array_equal(r["light"][:2000], exp)
    counts = np.bincount(r["light"], minlength=4)[lights]
    assert stats.chisquare(counts).pvalue > 1e-3, counts
    # no emitter: false without touching the sampler
    osc2 = O.OracleScene(_scene([prims[1]]))
    st2 = _states(rng, 100)
    r2 = osc2.choose_light(np.zeros((100, 3), np.float32), st2.copy())
    assert np.all(r2["light"] == -1) and np.array_equal(r2["state"], st2)


def _plane(y, half=30.0):
    v = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2]], np.int32)  # normals +y
    return {"vertices": v, "faces": f, "normals": None, "texcoords": None, "face_tex": None}


DOWN_CAMERA = {"pos": [0.0, 6.0, 0.0], "dir": [0.0, -1.0, 0.0], "up": [0.0, 0.0, -1.0], "fov": 40.0}


def _nee_pixels(osc, W, H):
    """Pixels whose camera ray hits the plane (prim 0) at y = 0: a ray through
    the quad's diagonal can pass between its two triangles (the Moller-Trumbore
    bounds of both reject it, as in the reference) and is left out."""
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    x, y = xs.ravel().astype(np.int32), ys.ravel().astype(np.int32)
    d = osc.camera_rays(x, y).astype(np.float64)
    o = np.array(DOWN_CAMERA["pos"], np.float64)
    rays = np.zeros((x.size, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 7] = np.finfo(np.float32).max
    hit = osc.mesh_intersect(0, rays)[0] >= 0
    assert hit.mean() > 0.9
    t = (0.0 - o[1]) / d[:, 1]          # hit on the plane y = 0
    return x[hit], y[hit], (o + d * t[:, None])[hit], d[hit]


def test_nee_expectation_under_sphere_light():
    """max_bounce 1: a camera path's radiance is exactly the next-event term at
    the first (Lambertian, albedo 0.5) surface hit; its mean over passes is
    albedo Le pi sin^2(theta_max) cos(alpha) where the light is fully visible."""
    Le, alb, c, R = 4.0, 0.5, np.array([0.8, 2.0, 0.3]), 0.4
    W = H = 24
    sc = _scene([{"type": "mesh", "mesh": _plane(0.0), "bsdf": "lambertian",
                  "texture": np.full((1, 1, 3), alb, np.float32)},
                 {"type": "sphere", "center": c.tolist(), "radius": R, "bsdf": "lambertian",
                  "emission": [Le, Le, Le]}],
                camera=DOWN_CAMERA, width=W, height=H, max_bounce=1)
    osc = O.OracleScene(sc)
    x, y, P, d = _nee_pixels(osc, W, H)
    o = np.array(DOWN_CAMERA["pos"])
    # pixels whose camera ray passes the light, or whose cone dips below the horizon, are left out
    oc = c - o
    tca = d @ oc
    miss = np.linalg.norm(oc[None, :] - d * tca[:, None], axis=1) > R * 1.2
    v = c[None, :] - P
    D = np.linalg.norm(v, axis=1)
    cos_a = v[:, 1] / D
    sin_max = R / D
    above = np.arccos(np.clip(cos_a, -1, 1)) + np.arcsin(sin_max) < np.pi / 2 - 0.02
    sel = miss & above
    assert sel.sum() > 200
    K = 48
    xx, yy = np.repeat(x[sel], K), np.repeat(y[sel], K)
    kk = np.tile(np.arange(K, dtype=np.int32), sel.sum())
    L = osc.trace_paths(xx, yy, kk)[:, 0].astype(np.float64).reshape(-1, K)
    expect = alb * Le * np.pi * sin_max[sel] ** 2 * cos_a[sel]
    ratio = L / expect[:, None]
    z = (ratio.mean() - 1.0) / (ratio.std() / np.sqrt(ratio.size))
    assert abs(z) < 5, (ratio.mean(), z)
    # and every pixel's mean within its own 5 sigma
    se = L.std(axis=1) / np.sqrt(K) + 1e-9
    assert np.all(np.abs(L.mean(axis=1) - expect) < 5 * se + 1e-4 * expect)


def test_nee_expectation_under_mesh_light():
    """The mesh light's term has no cos/dist^2 factor (area-measure pdf used as
    is): mean albedo Le A E[max(d.n, 0)] over uniform points of the quad."""
    Le, alb = 5.0, 0.6
    quad = scenes.emissive_quad([0.4, 2.5, -0.2], 0.3)
    W = H = 16
    sc = _scene([{"type": "mesh", "mesh": _plane(0.0), "bsdf": "lambertian",
                  "texture": np.full((1, 1, 3), alb, np.float32)},
                 {"type": "mesh", "mesh": quad, "bsdf": "lambertian", "emission": [Le, Le, Le]}],
                extensions=True, camera=DOWN_CAMERA, width=W, height=H, max_bounce=1)
    osc = O.OracleScene(sc)
    x, y, P, d = _nee_pixels(osc, W, H)
    # camera rays through the quad see the emitter instead of the plane
    o = np.array(DOWN_CAMERA["pos"])
    tq = (2.5 - o[1]) / d[:, 1]
    hq = o + d * tq[:, None]
    sel = (np.abs(hq[:, 0] - 0.4) > 0.2) | (np.abs(hq[:, 2] + 0.2) > 0.2)
    g = (np.arange(64) + 0.5) / 64 * 0.3 - 0.15
    Q = np.stack(np.meshgrid(0.4 + g, [2.5], -0.2 + g, indexing="ij"), -1).reshape(-1, 3)
    A = 0.09
    v = Q[None, :, :] - P[sel][:, None, :]
    cosn = np.maximum(v[..., 1] / np.linalg.norm(v, axis=2), 0).mean(axis=1)
    expect = alb * Le * A * cosn
    K = 48
    xx, yy = np.repeat(x[sel], K), np.repeat(y[sel], K)
    kk = np.tile(np.arange(K, dtype=np.int32), sel.sum())
    L = osc.trace_paths(xx, yy, kk)[:, 0].astype(np.float64).reshape(-1, K)
    se = L.std(axis=1) / np.sqrt(K) + 1e-9
    assert np.all(np.abs(L.mean(axis=1) - expect) < 5 * se + 2e-4 * expect)
    ratio = (L / expect[:, None]).ravel()
    z = (ratio.mean() - 1) / (ratio.std() / np.sqrt(ratio.size))
    assert abs(z) < 5, z
