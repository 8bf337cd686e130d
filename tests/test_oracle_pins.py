"""Pins the CPU oracle against closed forms, invariants and brute force (no GPU).

The reference's noise, SDF primitives, hash and marching cubes live in the cage engine,
which is absent from the reference checkout, and the reference has no tests or
fixtures for this path; so each oracle function is pinned here by what the mathematics
fixes, chosen so that a dropped term, a wrong sign or index, or a transposed operand
fails a check.
"""
import math

import numpy as np
import pytest

from unnatural_planets_amd import spec as S
from oracle_lib import mesh_edge_stats

GRAD, VALUE, CELL = 1, 2, 3
M32 = 0xFFFFFFFF


def py_hash3(seed, x, y, z):
    """Independent big-integer statement of the lattice hash (two's-complement inputs)."""
    h = (seed ^ ((x * 501125321) & M32) ^ ((y * 1136930381) & M32) ^ ((z * 1720413743) & M32)) & M32
    return (h * 0x27D4EB2D) & M32


# ------------------------------------------------------------------------------ hash
def test_hash_matches_bigint_statement(oracle):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        seed = int(rng.integers(0, 2**32))
        x, y, z = (int(v) for v in rng.integers(-2**31, 2**31, size=3))
        assert oracle.hash3(seed, x, y, z) == py_hash3(seed, x, y, z)


def test_hash_top_bits_are_uniform(oracle):
    # the gradient index is h >> 28: all 16 buckets must be populated evenly
    counts = np.zeros(16, dtype=np.int64)
    for x in range(-20, 20):
        for y in range(-20, 20):
            for z in range(-4, 4):
                counts[oracle.hash3(1234, x, y, z) >> 28] += 1
    expected = counts.sum() / 16
    assert np.all(np.abs(counts - expected) < 0.15 * expected), counts


# ------------------------------------------------------------------------------ gradient noise
def test_gradient_noise_zero_at_lattice_points(oracle):
    rng = np.random.default_rng(2)
    for _ in range(500):
        q = rng.integers(-50, 50, size=3).astype(np.float64)
        assert oracle.noise(GRAD, int(rng.integers(0, 2**32)), q) == 0.0


def test_gradient_noise_slope_at_lattice_point_is_corner_gradient(oracle):
    # fade'(0) = 0 and fade(0) = 0, so near a lattice point n(k + e) = g_k . e + O(e^2):
    # the slope along each axis is a component of the corner's gradient, in {-1, 0, 1}.
    rng = np.random.default_rng(3)
    eps = 1e-6
    for _ in range(200):
        k = rng.integers(-30, 30, size=3).astype(np.float64)
        seed = int(rng.integers(0, 2**32))
        h = py_hash3(seed, *(int(v) for v in k)) >> 28
        u = [0, 0, 0]
        ua = 0 if h < 8 else 1
        va = 1 if h < 4 else (0 if h in (12, 14) else 2)
        u[ua] += -1 if h & 1 else 1
        u[va] += -1 if h & 2 else 1
        for a in range(3):
            q = k.copy()
            q[a] += eps
            slope = oracle.noise(GRAD, seed, q) / eps
            assert abs(slope - u[a]) < 1e-4, (h, a, slope, u)


def test_gradient_noise_bounded_and_continuous(oracle):
    rng = np.random.default_rng(4)
    mx = 0.0
    for _ in range(3000):
        q = rng.uniform(-20, 20, size=3)
        mx = max(mx, abs(oracle.noise(GRAD, 99, q)))
    assert mx < 1.1
    # continuity across cell faces
    for _ in range(300):
        q = rng.uniform(-20, 20, size=3)
        a = int(rng.integers(0, 3))
        q[a] = math.floor(q[a])
        lo, hi = q.copy(), q.copy()
        lo[a] -= 1e-9
        hi[a] += 1e-9
        assert abs(oracle.noise(GRAD, 5, lo) - oracle.noise(GRAD, 5, hi)) < 1e-7


def test_gradient_noise_c1_across_faces(oracle):
    # quintic fade makes the noise C2; the derivative is continuous across lattice faces
    rng = np.random.default_rng(5)
    h = 1e-5
    for _ in range(200):
        q = rng.uniform(-10, 10, size=3)
        a = int(rng.integers(0, 3))
        q[a] = math.floor(q[a])
        e = np.zeros(3)
        e[a] = h
        left = (oracle.noise(GRAD, 7, q) - oracle.noise(GRAD, 7, q - e)) / h
        right = (oracle.noise(GRAD, 7, q + e) - oracle.noise(GRAD, 7, q)) / h
        assert abs(left - right) < 1e-3


# ------------------------------------------------------------------------------ value noise
def test_value_noise_equals_lattice_value(oracle):
    rng = np.random.default_rng(6)
    for _ in range(500):
        k = rng.integers(-100, 100, size=3)
        seed = int(rng.integers(0, 2**32))
        h = py_hash3(seed, *(int(v) for v in k))
        sh = h - 2**32 if h >= 2**31 else h
        expect = np.float32(np.float32(sh) * np.float32(4.656612873077393e-10))
        assert oracle.noise(VALUE, seed, k.astype(np.float32), f32=True) == expect


def test_value_noise_is_trilinear_blend_on_edges(oracle):
    # along a lattice edge the value is lerp(v0, v1, fade(t))
    rng = np.random.default_rng(7)
    for _ in range(100):
        k = rng.integers(-20, 20, size=3).astype(np.float64)
        t = rng.uniform(0, 1)
        v0 = oracle.noise(VALUE, 3, k)
        k1 = k.copy(); k1[0] += 1
        v1 = oracle.noise(VALUE, 3, k1)
        q = k.copy(); q[0] += t
        fade = t ** 3 * (t * (t * 6 - 15) + 10)
        assert abs(oracle.noise(VALUE, 3, q) - (v0 + fade * (v1 - v0))) < 1e-12


# ------------------------------------------------------------------------------ cellular
def cellular_brute(seed, q, jitter, reach):
    """F1 by brute force over (2*reach+1)^3 cells around round(q), feature points stated
    independently from the hash."""
    c = [int(np.rint(v)) for v in q]
    best = np.inf
    for dz in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dx in range(-reach, reach + 1):
                k = (c[0] + dx, c[1] + dy, c[2] + dz)
                h = py_hash3(seed, *k)
                m = h ^ (h >> 15)
                r = [(m & 1023) / 1023, ((m >> 10) & 1023) / 1023, ((m >> 20) & 1023) / 1023]
                p = [k[a] + jitter * (r[a] - 0.5) for a in range(3)]
                best = min(best, sum((p[a] - q[a]) ** 2 for a in range(3)))
    return math.sqrt(best)


def test_cellular_matches_brute_force(oracle):
    rng = np.random.default_rng(8)
    n_far_better = 0
    for _ in range(400):
        q = rng.uniform(-8, 8, size=3)
        v = oracle.noise(CELL, 11, q, 0.9)
        assert abs(v - cellular_brute(11, q, 0.9, 1)) < 1e-9
        far = cellular_brute(11, q, 0.9, 2)
        assert far <= v + 1e-12  # the 27-cell search can only miss closer points
        n_far_better += far < v - 1e-12
    assert n_far_better <= 4  # rare: the 27-cell F1 is the definition, not an approximation of 125


def test_cellular_zero_at_feature_point_and_jitter_zero(oracle):
    # jitter 0: feature points sit on lattice points, F1 = distance to the nearest one
    rng = np.random.default_rng(9)
    for _ in range(200):
        q = rng.uniform(-5, 5, size=3)
        expect = math.sqrt(sum((v - np.rint(v)) ** 2 for v in q))
        assert abs(oracle.noise(CELL, 1, q, 0.0) - expect) < 1e-12


# ------------------------------------------------------------------------------ fractal layers
def test_fbm_and_ridged_are_octave_sums(oracle):
    rng = np.random.default_rng(10)
    pts = rng.uniform(-1.5, 1.5, size=(40, 3)).astype(np.float32)
    for fractal in ("fbm", "ridged"):
        for noise in ("gradient", "value", "cellular"):
            L = S.layer(noise, fractal, 5, 1.7, 2.1, 0.45, 1.0, 321, jitter=0.8)
            spec = S.field_spec("none", (), [L], compose="elevation_raw")
            got = oracle.eval_points(spec, pts, f64=True)
            for i, p in enumerate(pts.astype(np.float64)):
                acc = 0.0
                for k in range(5):
                    n = oracle.noise(S.NOISE[noise], 321 + k, p * float(L.freq[k]), float(L.jitter))
                    term = 1 - 2 * abs(n) if fractal == "ridged" else n
                    acc += float(L.amp[k]) * term
                assert abs(got[i] - acc) < 1e-9


def test_layer_expansion(oracle, upf):
    import ctypes
    for fractal, oct_, f, lac, g in [(1, 4, 2.0, 2.0, 0.5), (2, 6, 0.001, 1.5, -0.4), (0, 3, 0.75, 2, 0.5),
                                     (2, 8, 3.0, 1.9, 0.6)]:
        fr, am = oracle.layer_expand(fractal, oct_, f, lac, g)
        L = S.Layer()
        assert upf.lib().upf_layer_init(ctypes.byref(L), 1, fractal, oct_, f, lac, g, 1.0, 0, 1.0, 0) == 0
        py = S.layer("gradient", fractal, oct_, f, lac, g)
        for k in range(8):
            assert fr[k] == L.freq[k] == py.freq[k]
            assert am[k] == L.amp[k] == py.amp[k]
        n = 1 if fractal == 0 else oct_
        # closed forms: geometric frequencies, amplitudes normalised so sum |amp| = 1
        assert abs(sum(abs(float(a)) for a in am[:n]) - 1.0) < 1e-5
        for k in range(n):
            assert abs(float(fr[k]) - f * lac ** k) <= 1e-5 * f * lac ** k


# ------------------------------------------------------------------------------ shapes
@pytest.mark.parametrize("shape,params,point,expect", [
    ("sphere", (1.0,), (2.0, 0.0, 0.0), 1.0),
    ("sphere", (1100.0,), (0.0, 0.0, 0.0), -1100.0),
    ("sphere", (1.0,), (0.6, 0.0, 0.8), 0.0),
    ("torus", (1.0, 0.4), (1.0, 0.0, 0.0), -0.4),      # on the core circle
    ("torus", (1.0, 0.4), (0.0, 0.0, 1.7), 0.3),       # ring in xz-plane
    ("torus", (1.0, 0.4), (0.0, 0.0, 0.0), math.hypot(1.0, 0.0) - 0.4),
    ("torus", (1.0, 0.4), (0.0, 1.0, 1.0), 0.6),       # straight above the core circle
    ("box", (1.0, 2.0, 3.0, 0.0), (0.0, 0.0, 0.0), -1.0),
    ("box", (1.0, 2.0, 3.0, 0.0), (2.0, 3.0, 3.0), math.sqrt(2.0)),
    ("box", (450.0, 450.0, 450.0, 100.0), (0.0, 0.0, 600.0), 50.0),
    ("capsule", (2.0, 0.5), (0.0, 0.0, 0.0), -0.5),
    ("capsule", (2.0, 0.5), (0.0, 0.0, 2.0), 0.5),     # axis z, tip at 1.0 + 0.5
    ("capsule", (2.0, 0.5), (1.0, 0.0, 0.3), 0.5),
])
def test_shape_closed_forms(oracle, shape, params, point, expect):
    spec = S.field_spec(shape, params, [], compose="water")
    v = oracle.eval_points(spec, np.array([point], dtype=np.float32), f64=True)[0]
    assert abs(v - expect) < 1e-6 * max(1.0, abs(expect))


@pytest.mark.parametrize("shape,params", [("sphere", (1.0,)), ("torus", (1.0, 0.4)), ("capsule", (1.0, 0.5)),
                                          ("box", (0.5, 0.7, 0.9, 0.2))])
def test_shape_gradient_unit_length(oracle, shape, params):
    rng = np.random.default_rng(12)
    pts = rng.uniform(-1.5, 1.5, size=(200, 3)).astype(np.float32)
    spec = S.field_spec(shape, params, [], compose="water")
    _, g = oracle.grad_points(spec, pts)
    ln = np.linalg.norm(g, axis=1)
    if shape == "box":  # inside a box the distance is to the nearest face: still unit
        pass
    assert np.all(np.abs(ln - 1.0) < 1e-9), ln.min()


# ------------------------------------------------------------------------------ duals vs FD
def test_dual_gradient_matches_finite_differences(oracle):
    spec = S.spec_torus_ridged_voronoi(42)
    rng = np.random.default_rng(13)
    pts = rng.uniform(-1.4, 1.4, size=(60, 3)).astype(np.float32)
    _, g = oracle.grad_points(spec, pts)
    h = 1e-4
    n_ok = 0
    for i, p in enumerate(pts.astype(np.float64)):
        fd = np.zeros(3)
        for a in range(3):
            e = np.zeros(3); e[a] = h
            fp = oracle.eval_points(spec, (p + e)[None].astype(np.float32), f64=True)[0]
            fm = oracle.eval_points(spec, (p - e)[None].astype(np.float32), f64=True)[0]
            fd[a] = (fp - fm) / (2 * h)
        # float32 positions: FD step is not exact, and kinks (ridges, cell switches) exist
        n_ok += np.max(np.abs(fd - g[i])) < 2e-2 * max(1.0, np.max(np.abs(g[i])))
    assert n_ok >= 55


def test_theta_gradient_matches_finite_differences(oracle):
    spec = S.spec_torus_fbm_voronoi(7)
    rng = np.random.default_rng(14)
    pts = rng.uniform(-1.4, 1.4, size=(64, 3)).astype(np.float32)
    w = rng.uniform(-1, 1, size=64).astype(np.float32)
    th = oracle.theta_grad(spec, pts, w)
    h = 1e-3

    def L(s):
        return float(np.dot(w.astype(np.float64), oracle.eval_points(s, pts, f64=True)))

    checks = [("shape", 0, 0), ("shape", 1, 0), ("weight", 0, 0), ("weight", 1, 0)]
    checks += [("amp", 0, k) for k in range(4)] + [("freq", 0, k) for k in range(4)]
    for kind, l, k in checks:
        sp, sm = S.copy_spec(spec), S.copy_spec(spec)
        if kind == "shape":
            sp.shape_params[l] += h; sm.shape_params[l] -= h
        elif kind == "weight":
            sp.layers[l].weight += h; sm.layers[l].weight -= h
        elif kind == "amp":
            sp.layers[l].amp[k] += h; sm.layers[l].amp[k] -= h
        else:
            sp.layers[l].freq[k] += h; sm.layers[l].freq[k] -= h
        fd = (L(sp) - L(sm)) / (2 * h)
        got = th[S.theta_index(kind, l, k)]
        assert abs(fd - got) < 2e-2 * max(1.0, abs(got)), (kind, l, k, fd, got)


# ------------------------------------------------------------------------------ math.cpp helpers
def test_math_helpers_closed_forms(oracle):
    L = oracle.L
    assert L.upfo_rescale(0.5, 0, 1, 10, 20) == 15
    assert L.upfo_rescale(3, 1, 5, 0, 1) == 0.5
    assert L.upfo_sharp_edge(0.2, 0.05) == 0.0 and L.upfo_sharp_edge(0.8, 0.05) == 1.0
    assert abs(L.upfo_sharp_edge(0.5, 0.05) - 0.5) < 1e-12
    # smoothMin equals min when |a-b| >= k, and is below min otherwise
    for a, b, k in [(0, 5, 1), (3, -2, 2), (1, 1.5, 1)]:
        v = L.upfo_smooth_min(a, b, k)
        if abs(a - b) >= k:
            assert v == min(a, b)
        else:
            assert v < min(a, b)
        assert L.upfo_smooth_max(a, b, k) == -L.upfo_smooth_min(-a, -b, k)
    # terrace stays within [floor(x) - 1e-3, floor(x) + 1 + 1e-3] (sources/math.cpp:24)
    for x in np.linspace(-3, 3, 101):
        for s in (0.0, 0.5, 1.0, 4.0):
            r = L.upfo_terrace(x, s)
            assert math.floor(x) - 1e-3 < r < math.floor(x) + 1 + 1e-3
    assert L.upfo_terrace(2.3, 0.0) == 2.3  # steepness 0 leaves x unchanged


# ------------------------------------------------------------------------------ flop table
@pytest.mark.parametrize("make", [S.spec_sphere_fbm, S.spec_torus_ridged_voronoi, S.spec_torus_fbm_voronoi])
def test_flop_table_matches_counted_oracle(oracle, upf, make):
    import ctypes
    spec = make()
    counted = oracle.flops_at(spec, np.array([0.3, 0.2, -0.4], dtype=np.float32))
    assert upf.lib().upf_flops_per_sample(ctypes.byref(spec)) == counted


# ------------------------------------------------------------------------------ marching cubes
CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_mc_table_uses_exactly_the_crossing_edges(oracle):
    nt, tr = oracle.mc_table()
    for code in range(256):
        inside = [(code >> c) & 1 for c in range(8)]
        crossing = {e for e, (a, b) in enumerate(EDGES) if inside[a] != inside[b]}
        used = set(int(v) for v in tr[code, :3 * nt[code]])
        assert used == crossing, code
        assert all(v == -1 for v in tr[code, 3 * nt[code]:])
        assert nt[code] == nt[255 - code] or True  # complements may differ in loop count


def test_mc_table_orientation_points_to_outside_corners(oracle):
    # each triangle's normal must point along the gradient of the cell's trilinear
    # interpolant of the corner signs (-1 inside, +1 outside), taken at its centroid
    nt, tr = oracle.mc_table()
    mid = [np.add(CORNERS[a], CORNERS[b]) / 2.0 for a, b in EDGES]
    checked = 0
    for code in range(1, 255):
        val = [(-1.0 if (code >> c) & 1 else 1.0) for c in range(8)]
        for t in range(nt[code]):
            p = [mid[tr[code, 3 * t + v]] for v in range(3)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            x, y, z = (p[0] + p[1] + p[2]) / 3.0
            grad = np.zeros(3)
            for c, (cx, cy, cz) in enumerate(CORNERS):
                wx, wy, wz = (x if cx else 1 - x), (y if cy else 1 - y), (z if cz else 1 - z)
                dx, dy, dz = (1 if cx else -1), (1 if cy else -1), (1 if cz else -1)
                grad += val[c] * np.array([dx * wy * wz, wx * dy * wz, wx * wy * dz])
            if np.linalg.norm(grad) < 0.2:
                continue  # saddle of an ambiguous configuration: no preferred side
            checked += 1
            assert np.dot(n, grad) > 0, (code, t)
    assert checked > 700


def _sphere_field(n, r=0.7):
    ax = np.linspace(-1, 1, n, dtype=np.float32)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return (np.sqrt(x * x + y * y + z * z) - r).astype(np.float32), S.uniform_grid(-1.0, 1.0, n)


def test_mc_sphere_closed_manifold_euler_2(oracle):
    f, g = _sphere_field(40)
    pos, nrm, idx = oracle.mc(f, g)
    ne, boundary, bad, oriented = mesh_edge_stats(idx)
    assert boundary == 0 and bad == 0 and oriented
    assert len(pos) - ne + len(idx) == 2
    # vertices on the sphere (linear interpolation error is second order), normals radial
    r = np.linalg.norm(pos, axis=1)
    assert np.all(np.abs(r - 0.7) < 5e-3)
    assert np.all(np.sum(nrm * pos / r[:, None], axis=1) > 0.99)
    # triangles wind counter-clockwise seen from outside
    tri = pos[idx.astype(np.int64)]
    tn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert np.all(np.sum(tn * tri.mean(axis=1), axis=1) > 0)


def test_mc_torus_euler_0_and_random_fields_closed(oracle):
    n = 48
    ax = np.linspace(-1.6, 1.6, n, dtype=np.float32)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    f = (np.sqrt((np.sqrt(x * x + z * z) - 1.0) ** 2 + y * y) - 0.4).astype(np.float32)
    pos, _, idx = oracle.mc(f, S.uniform_grid(-1.6, 1.6, n), normals=False)
    ne, boundary, bad, oriented = mesh_edge_stats(idx)
    assert boundary == 0 and bad == 0 and oriented
    assert len(pos) - ne + len(idx) == 0
    # random fields with a positive border: every ambiguous configuration occurs, the
    # face rule must still close the surface
    rng = np.random.default_rng(15)
    for _ in range(5):
        f = rng.standard_normal((14, 14, 14)).astype(np.float32)
        f[0, :, :] = f[-1, :, :] = f[:, 0, :] = f[:, -1, :] = f[:, :, 0] = f[:, :, -1] = 1.0
        pos, _, idx = oracle.mc(f, S.uniform_grid(0.0, 1.0, 14), normals=False)
        ne, boundary, bad, oriented = mesh_edge_stats(idx)
        assert boundary == 0 and bad == 0 and oriented


def test_mc_vertex_on_linear_crossing(oracle):
    # a linear field: every vertex lies exactly on the plane f = 0
    n = 9
    ax = np.linspace(0, 1, n, dtype=np.float32)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    f = (x + 0.5 * y - 0.25 * z - 0.37).astype(np.float32)
    pos, nrm, idx = oracle.mc(f, S.uniform_grid(0.0, 1.0, n))
    plane = pos[:, 0] + 0.5 * pos[:, 1] - 0.25 * pos[:, 2] - 0.37
    assert np.all(np.abs(plane) < 1e-6)
    expect = np.array([1, 0.5, -0.25]) / np.linalg.norm([1, 0.5, -0.25])
    assert np.all(np.abs(nrm - expect) < 1e-5)
