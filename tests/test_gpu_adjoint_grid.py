"""The grid adjoint (wave-cooperative rows, upf_field_adjoint_grid) against the
per-point adjoint at the same nodes (upf_field_adjoint, itself pinned to the oracle's
FP64 dual numbers) and against the oracle directly: f, df/dx and the accumulated
dL/dtheta.  The row version regroups the interpolation sums, so values agree to a few
ulps and gradients to 1e-5 relative; theta sums to 1e-4 of their scale."""
import numpy as np
import pytest

import unnatural_planets_amd as upfm
from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu


def value_ridged():
    return S.field_spec("sphere", (1.0,), [S.layer("value", "ridged", 5, 2.5, 2.0, 0.5, 0.05, 77)])


CASES = {
    "torus_ridged_voronoi": (S.spec_torus_ridged_voronoi(42), S.uniform_grid(-1.6, 1.6, 70)),
    "sphere_fbm_all_add": (S.spec_sphere_fbm(1234), S.uniform_grid(-1.3, 1.3, 66)),
    "value_ridged": (value_ridged(), S.uniform_grid(-1.3, 1.3, 64)),
    "navigation_torus": (upfm._with_compose(S.spec_torus_fbm_voronoi(7), "navigation"), S.uniform_grid(-1.6, 1.6, 65)),
    "simplex_in_stack": (S.field_spec("torus", (1.0, 0.4), [S.layer("simplex", "fbm", 3, 2.0, 2.0, 0.5, 0.05, 9),
                                                            S.layer("gradient", "ridged", 3, 1.5, 2.0, 0.5, 0.04, 10)]),
                         S.uniform_grid(-1.6, 1.6, 64)),
    "octahedron_metres": (S.field_spec("octahedron", (800.0, 100.0),
                                       [S.layer("gradient", "fbm", 3, 0.002, 2.0, 0.5, 1.0, 17)], elevation_scale=30.0),
                          S.uniform_grid(-1200.0, 1200.0, 67)),
}


def grid_points(g):
    nx, ny, nz = S.grid_dims(g)
    i = np.arange(nx, dtype=np.float32) * np.float32(g.spacing[0])
    j = np.arange(ny, dtype=np.float32) * np.float32(g.spacing[1])
    k = np.arange(nz, dtype=np.float32) * np.float32(g.spacing[2])
    x = np.float32(g.origin[0]) + i
    y = np.float32(g.origin[1]) + j
    z = np.float32(g.origin[2]) + k
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([xx, yy, zz], -1).reshape(-1, 3)


@pytest.mark.parametrize("case", list(CASES))
def test_grid_adjoint_matches_points_and_oracle(upf, oracle, case):
    import torch
    s, g = CASES[case]
    nx, ny, nz = S.grid_dims(g)
    rng = np.random.default_rng(9)
    w = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    f, dfdx, th = upfm.field_adjoint_grid(s, g, torch.from_numpy(w).cuda())
    pts = grid_points(g)
    pf, pg, pth = upfm.field_adjoint(s, torch.from_numpy(pts).cuda(), torch.from_numpy(w.reshape(-1)).cuda())
    f, dfdx, th = f.cpu().numpy().reshape(-1), dfdx.cpu().numpy().reshape(-1, 3), th.cpu().numpy()
    pf, pg, pth = pf.cpu().numpy(), pg.cpu().numpy(), pth.cpu().numpy()
    scale = max(1.0, float(np.abs(pts).max()))
    assert np.max(np.abs(f - pf)) <= 2e-6 * max(scale, float(np.abs(pf).max()))
    gn = np.maximum(np.linalg.norm(pg, axis=1), 1e-3)
    gerr = np.linalg.norm(dfdx - pg, axis=1) / gn
    assert np.mean(gerr < 1e-5) > 0.998, np.quantile(gerr, [0.5, 0.998, 1.0])
    ts = max(float(np.abs(pth).max()), 1e-6)
    assert np.max(np.abs(th - pth)) <= 1e-4 * ts, (np.abs(th - pth).max(), ts)
    # and the oracle's FP64 dual numbers for the theta sum
    ref = oracle.theta_grad(s, pts, w.reshape(-1))
    assert np.max(np.abs(th - ref)) <= 2e-3 * max(float(np.abs(ref).max()), 1e-6)


def test_config4_grid_256_forward_adjoint(upf, oracle):
    """The differentiation workload at size: torus R=1, r=0.4 + 4-octave gradient fBm +
    cellular F1 (the 11-parameter theta) on a 256^3 grid over [-1.6, 1.6]^3.  dL/df is
    nonzero at 20,000 seeded nodes, so the grid's dL/dtheta is the oracle's FP64 dual
    sum over those nodes; f and df/dx are checked there against the FP64 oracle, and the
    whole grid's f against the forward row kernel."""
    import torch
    s, g = S.spec_torus_fbm_voronoi(7), S.uniform_grid(-1.6, 1.6, 256)
    n = 256
    rng = np.random.default_rng(7)
    sel = rng.choice(n ** 3, size=20000, replace=False)
    w = np.zeros(n ** 3, dtype=np.float32)
    w[sel] = rng.standard_normal(len(sel)).astype(np.float32)
    print("case: config 4 grid adjoint 256^3", flush=True)
    f, dfdx, th = upfm.field_adjoint_grid(s, g, torch.from_numpy(w.reshape(n, n, n)).cuda())
    fwd = upfm.field_eval_grid(s, g)
    assert float((f - fwd).abs().max()) <= 2e-6 * max(1.0, float(fwd.abs().max()))
    f = f.cpu().numpy().reshape(-1)[sel]
    dfdx = dfdx.cpu().numpy().reshape(-1, 3)[sel]
    th = th.cpu().numpy()
    k, j, i = np.unravel_index(sel, (n, n, n))
    sp = np.float32(g.spacing[0])
    pts = np.stack([np.float32(g.origin[0]) + i.astype(np.float32) * sp, np.float32(g.origin[1]) + j.astype(np.float32) * sp,
                    np.float32(g.origin[2]) + k.astype(np.float32) * sp], 1).astype(np.float32)
    ref_f = oracle.eval_points(s, pts, f64=True)
    assert np.max(np.abs(f - ref_f) / np.maximum(1.0, np.abs(ref_f))) <= 1e-5
    _, ref_g = oracle.grad_points(s, pts)
    gerr = np.linalg.norm(dfdx - ref_g, axis=1) / np.maximum(np.linalg.norm(ref_g, axis=1), 1e-3)
    assert np.mean(gerr <= 1e-4) >= 0.998, np.quantile(gerr, [0.5, 0.998, 1.0])
    ref_th = oracle.theta_grad(s, pts, w[sel])
    idx = [S.theta_index("amp", 0, o) for o in range(4)] + [S.theta_index("freq", 0, o) for o in range(4)] + \
        [S.theta_index("weight", 1), S.theta_index("shape", 0), S.theta_index("shape", 1)]
    scale = float(np.abs(ref_th[idx]).max())
    assert np.max(np.abs(th[idx] - ref_th[idx])) <= 2e-3 * scale, (th[idx], ref_th[idx])
