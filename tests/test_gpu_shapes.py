"""Extended base shapes on the GPU against the oracle: values on a grid (within 1e-5 of
max(|f|, the grid's half extent) -- the scale FP32 rounds the metre-scale coordinates
at, as 1e-5 of max(|f|, 1) is for the unit-scale workloads), position gradients against
the FP64 duals, and a closed marching-cubes surface per shape.

The Mandelbulb is a 20-step chaotic iteration of sin / cos / acos / atan2 / pow, so the
last-ulp differences between the device's and the C library's functions grow along it:
its bars are 1e-4 for values and 98% of points for gradients."""
import numpy as np
import pytest

import unnatural_planets_amd as upfm
from oracle_lib import mesh_edge_stats
from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu

CASES = [("cylinder", (2000.0, 800.0), 1500), ("plane", (0.0, 0.6, -0.8, 100.0), 1500),
         ("octahedron", (800.0, 100.0), 1200), ("double_torus", (900.0, 400.0, 200.0), 1600),
         ("torus_cross", (750.0, 400.0, 200.0), 2400), ("h2o", (), 1500), ("bowl", (), 3000),
         ("inside_cube", (), 1700), ("wormhole", (), 2400), ("pipe", (), 1800), ("hemispheres", (), 3300),
         ("h3o", (), 1800), ("h4o", (), 1800), ("gear", (), 2000), ("mandelbulb", (), 3200), ("mobius", (), 1900),
         ("fibers", S.SHAPE_REFERENCE_PARAMS["fibers"], 2200), ("asteroid", S.SHAPE_REFERENCE_PARAMS["asteroid"], 2200),
         ("twisted_plane", (), 1500), ("belt", (), 2500), ("triangular_prism", (), 1100),
         ("hexagonal_prism", (), 700), ("twisted_hex_prism", (), 1200), ("tetrahedron", (), 800),
         ("knot", (), 2100)]


def spec_of(name, params):
    # the shape plus a mild gradient-noise layer, so the path through the layers is exercised too
    return S.field_spec(name, params, [S.layer("gradient", "fbm", 3, 0.002, 2.0, 0.5, 1.0, 17)],
                        elevation_scale=30.0)


@pytest.mark.parametrize("name,params,half", CASES)
def test_shape_grid_matches_oracle(upf, oracle, name, params, half):
    s = spec_of(name, params)
    g = S.uniform_grid(-half, half, 65)
    for variant in (upfm.EVAL_ROWS, upfm.EVAL_PER_LANE):
        f = upfm.field_eval_grid(s, g, variant=variant).cpu().numpy().astype(np.float64)
        ref = oracle.eval_grid(s, g, f64=True)
        err = np.abs(f - ref) / np.maximum(np.abs(ref), half)
        assert err.max() < (1e-4 if name == "mandelbulb" else 1e-5), (name, variant, err.max())


@pytest.mark.parametrize("name,params,half", CASES)
def test_shape_gradients_match_duals(upf, oracle, name, params, half):
    s = spec_of(name, params)
    pts = np.random.default_rng(8).uniform(-half, half, size=(4096, 3)).astype(np.float32)
    import torch
    f, g = upfm.field_eval_points(s, torch.from_numpy(pts).cuda(), with_grad=True)
    _, rg = oracle.grad_points(s, pts)
    err = np.linalg.norm(g.cpu().numpy() - rg, axis=1) / np.maximum(np.linalg.norm(rg, axis=1), 1e-3)
    assert np.mean(err < 1e-4) > (0.98 if name == "mandelbulb" else 0.99), (name, np.quantile(err, [0.5, 0.99]))


@pytest.mark.parametrize("name,params,half", [c for c in CASES if c[0] not in ("plane", "bowl", "inside_cube",
                                                                                 "wormhole", "hemispheres", "pipe",
                                                                                 "twisted_plane", "belt")])
def test_shape_meshes_are_closed(upf, name, params, half):
    m = upfm.mesh_generate_generic(spec_of(name, params), S.uniform_grid(-half, half, 96))
    E, boundary, bad, oriented = mesh_edge_stats(m.indices.cpu().numpy().astype(np.int64))
    assert m.n_triangles > 1000 and boundary == 0 and bad == 0 and oriented
