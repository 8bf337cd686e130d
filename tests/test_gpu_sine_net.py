"""GPU parity of UPF_SHAPE_SINE_NET (the architecture of sdfBunny / sdfMonkeyHead,
sources/sdf.cpp:252-303) with the oracle, through the C-ABI: weights in device memory
for the library, the same weights in host memory for the oracle.

Bars: values within 1e-5 of max(|f|, 2000) (metre scale; the network's output is
divided by scale = 0.0005); gradients within 1e-3 (unit-scale slopes of a smooth
network); the grid evaluation equals the oracle's at the nodes (the grid runs the
per-lane kernels of the further shapes) to the same bar, and marching cubes of the grid
yields a closed, consistently oriented mesh on the sphere a ball-shaped network has as
its zero set.
"""
import numpy as np
import pytest

from unnatural_planets_amd import spec as S
from oracle_lib import mesh_edge_stats

pytestmark = pytest.mark.gpu

N = S.SINE_NET_WEIGHTS
SCALE = 0.0005


def spec_with(ptr, scale=SCALE, d1=1.0, d2=1.4, compose="water"):
    s = S.field_spec("sine_net", (scale, d1, d2), [], compose=compose)
    s.shape_weights = ptr
    return s


def random_weights(seed, sd=0.6):
    return (sd * np.random.default_rng(seed).standard_normal(N)).astype(np.float32)


def points(n, seed):
    """Inside and outside the unit ball of q, away from its sphere (where the two branches
    meet, a last-bit difference in |q| could pick the other branch)."""
    rng = np.random.default_rng(seed)
    q = rng.uniform(-1.2, 1.2, size=(2 * n, 3))
    q = q[np.abs(np.linalg.norm(q, axis=1) - 1.0) > 1e-3][:n]
    return (q / SCALE).astype(np.float32)


@pytest.mark.parametrize("seed", [7, 21])
def test_points_and_gradients_match_oracle(upf, oracle, seed):
    import torch
    w = random_weights(seed)
    wd = torch.from_numpy(w).cuda()
    pts = points(20000, seed + 1)
    print(f"case: sine net points seed {seed}", flush=True)
    f, g = upf.field_eval_points(spec_with(wd.data_ptr()), torch.from_numpy(pts).cuda(), with_grad=True)
    f, g = f.cpu().numpy().astype(np.float64), g.cpu().numpy()
    rf = oracle.eval_points(spec_with(w.ctypes.data), pts, f64=True)
    _, rg = oracle.grad_points(spec_with(w.ctypes.data), pts)
    err = np.abs(f - rf) / np.maximum(np.abs(rf), 2000.0)
    print(f"value err max {err.max():.2e}, gradient err max {np.abs(g - rg).max():.2e}")
    assert err.max() < 1e-5
    np.testing.assert_allclose(g, rg, atol=1e-3)


def test_grid_matches_oracle_and_meshes_closed(upf, oracle):
    import torch
    # inside the unit ball of q the network is -0.3 + 0.2 sin(q.y) < 0, outside the
    # reference's (|q| - 0.8) / scale > 0.1 / scale: the zero set is the sphere |q| = 1
    # (radius 2000 m), inside the grid box
    w = np.zeros(N, dtype=np.float32)
    w[624] = -0.3
    w[0] = 1.0      # A of unit 0: q.y
    w[608] = 0.2    # U of unit 0
    wd = torch.from_numpy(w).cuda()
    g = S.uniform_grid(-2400.0, 2400.0, 72)
    print("case: sine net grid 72^3", flush=True)
    field = upf.field_eval_grid(spec_with(wd.data_ptr()), g).cpu().numpy().astype(np.float64)
    ref = oracle.eval_grid(spec_with(w.ctypes.data), g, f64=True).reshape(field.shape)
    err = np.abs(field - ref) / np.maximum(np.abs(ref), 2000.0)
    assert err.max() < 1e-5, err.max()
    m = upf.mc_extract(torch.from_numpy(field.astype(np.float32)).cuda(), g, 0.0)
    assert m.n_triangles > 1000
    _, boundary, bad, oriented = mesh_edge_stats(m.indices.cpu().numpy().astype(np.int64))
    assert boundary == 0 and bad == 0 and oriented
    r = np.linalg.norm(m.positions.cpu().numpy(), axis=1)
    assert np.all(np.abs(r - 2000.0) < 4800.0 / 71)  # within a cell of the sphere
