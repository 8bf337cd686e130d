"""GPU parity of the full material stack (upf_material_tiles, upf_bake_material) with
the oracle's restatement, through the C-ABI.

The stack is a chain of thresholds (early returns at 1e-7, biome and type boundaries,
the boulder and flower tests), so a float a few ulps away could switch a branch; the
kernels therefore round every operation as the restatement does, and both use the
project's pow and atan (upf_math.hpp).  Bars: biome and type equal on every tile; every
float output of every tile and texel equal bit for bit; NaN exactly where the oracle
has NaN.
"""
import numpy as np
import pytest

import unnatural_planets_amd as upfm
from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu


def _planet_points(n, seed, r0=1080, r1=1140):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * rng.uniform(r0, r1, size=(n, 1))).astype(np.float32), d.astype(np.float32)


CASES = [("lakes", "default", "land"), ("islands", "default", "water"), ("lakes", "default", "navigation"),
         ("craters", "barren", "land"), ("legacy", "barren", "water"), ("simple", "debug", "land")]


@pytest.mark.parametrize("mode,coloring,purpose", CASES)
def test_material_tiles_match_oracle(upf, oracle, mode, coloring, purpose):
    import torch
    m = S.material_spec(S.spec_elevation_mode(mode, 13), coloring, purpose, seed=21)
    pos, nrm = _planet_points(20000, 3)
    print(f"case: tiles {mode} {coloring} {purpose} 20000 points", flush=True)
    got = upf.material_tiles(m, torch.from_numpy(pos).cuda(), torch.from_numpy(nrm).cuda()).cpu().numpy()
    ref = oracle.material_tiles(m, pos, nrm)
    G, R = S.tiles_fields(got), S.tiles_fields(ref)
    for k in ("biome", "type"):
        np.testing.assert_array_equal(G[k], R[k], err_msg=k)
    print("biomes:", np.bincount(G["biome"], minlength=12).tolist())
    np.testing.assert_array_equal(got[:, :11].view(np.int32), ref[:, :11].view(np.int32))


@pytest.mark.parametrize("purpose", ["land", "water"])
def test_bake_material_matches_oracle(upf, oracle, purpose):
    f = S.spec_elevation_mode("islands", 5)
    g = S.grid_reference(72)
    mesh_spec = f if purpose == "land" else upfm._with_compose(f, "water")
    mesh = upfm.mesh_generate_generic(mesh_spec, g)
    m = S.material_spec(f, "default", purpose, seed=8)
    print(f"case: bake {purpose} 512x384 V={mesh.n_vertices}", flush=True)
    uv = upfm.atlas_per_triangle(mesh.n_triangles, 512, 384)
    tri = upfm.bake_rasterize(uv, 512, 384)
    a, s, h = upfm.bake_material(m, mesh, uv, tri)
    ra, rs, rh = oracle.bake_material(m, mesh.positions.cpu().numpy(), mesh.normals.cpu().numpy(),
                                      mesh.indices.cpu().numpy(), uv.cpu().numpy(), tri.cpu().numpy())
    a, s, h = a.cpu().numpy(), s.cpu().numpy(), h.cpu().numpy()
    assert a.shape == ra.shape and a.shape[2] == (4 if purpose == "water" else 3)
    for x, y in ((a, ra), (s, rs), (h, rh)):
        nan = np.isnan(y)
        np.testing.assert_array_equal(np.isnan(x), nan)
        np.testing.assert_array_equal(x[~nan].view(np.int32), y[~nan].view(np.int32))
    print(f"covered {int((~np.isnan(rh)).sum())} texels, bit-exact")


def test_generate_textures_pipeline(upf):
    f = S.spec_elevation_mode("lakes", 6)
    mesh = upfm.mesh_generate_generic(f, S.grid_reference(64))
    m = S.material_spec(f, seed=2)
    a, s, h, uv = upfm.generate_textures(m, mesh, 256, 256)
    assert a.shape == (256, 256, 3) and s.shape == (256, 256, 2) and h.shape == (256, 256)
    import torch
    assert a.dtype == torch.uint8 and int(a.max()) > 0
    alb, pbr, nrm = upfm.export_maps(a, s, h)
    assert pbr.shape == (256, 256, 3) and nrm.shape == (256, 256, 3)
