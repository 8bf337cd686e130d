"""Narrow-band marching cubes against dense marching cubes of the same fine grid.

With a band that holds the surface (band >= L x half the block diagonal: every point of
a block lies within half its diagonal of one of its corners; the oracle's FP64 gradient
gives max |grad f| = 1.31 for the sphere workload and 1.70 for the torus one, and the
tests use L = 2), the narrow-band mesh is the dense mesh of the row-kernel field up to
order: the bricks are evaluated in rows with the row kernels' lattice column folds and
reads (`brick_field`), which give the row kernel's values bit for bit.  Sorting its vertices by key (edge id = node * 3 + axis) gives the dense
order, so:
  * vs the oracle's marching cubes of the GPU row-kernel field (129^3, 257^3 sphere fBm):
    same vertex count, positions bit-identical after the key sort, same triangles;
  * vs the GPU dense path at 513^3 (torus, ridged + Voronoi + mask): the same;
  * normals are the normalized analytic gradient: vs the oracle's FP64 dual gradient at
    the vertex positions, within 1e-4 on >= 99.9% of vertices;
  * band 0 on the ridged torus leaves surface outside the active blocks: UPF_ERR_BAND;
  * a workspace too small reports the active count; dims off the block size: BAD_GRID.
"""
import ctypes

import numpy as np
import pytest

import unnatural_planets_amd as upfm
from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu

L_BOUND = 2.0


def band_for(g: S.Grid) -> float:
    diag = S.NB_BLOCK * float(np.linalg.norm([g.spacing[0], g.spacing[1], g.spacing[2]]))
    return L_BOUND * diag / 2


def canonical(keys, pos, nrm, idx):
    """Vertices sorted by key; triangles renumbered, each rotated to start at its
    smallest id (orientation kept), rows sorted."""
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    assert np.all(np.diff(sk) > 0), "duplicate vertex keys"
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    return sk, pos[order], (nrm[order] if nrm is not None else None), canon_tris(inv[idx])


def canon_tris(t):
    t = np.asarray(t, dtype=np.int64)
    r = np.argmin(t, axis=1)
    rows = np.arange(len(t))
    t = np.stack([t[rows, r], t[rows, (r + 1) % 3], t[rows, (r + 2) % 3]], 1)
    return t[np.lexsort((t[:, 2], t[:, 1], t[:, 0]))]


def nb_numpy(s, g, band):
    m, keys, st = upfm.nb_extract(s, g, band, keys=True)
    assert st.unresolved == 0
    return (keys.cpu().numpy(), m.positions.cpu().numpy(), m.normals.cpu().numpy(),
            m.indices.cpu().numpy().astype(np.int64), st)


@pytest.mark.parametrize("n", [129, 257])
def test_nb_equals_oracle_dense_mc(upf, oracle, n):
    s = S.spec_sphere_fbm(1234)
    g = S.uniform_grid(-1.3, 1.3, n)
    keys, pos, nrm, idx, st = nb_numpy(s, g, band_for(g))
    f = upfm.field_eval_grid(s, g, variant=upfm.EVAL_ROWS).cpu().numpy()
    rpos, _, ridx = oracle.mc(f, g, normals=False)
    sk, spos, _, stri = canonical(keys, pos, None, idx)
    print(f"{n}^3: {st.active_blocks}/{st.blocks} blocks active, V={len(pos)} T={len(idx)}")
    assert len(sk) == len(rpos) and len(stri) == len(ridx)
    np.testing.assert_array_equal(spos, rpos)
    np.testing.assert_array_equal(stri, canon_tris(ridx.astype(np.int64)))
    # every key names an edge whose end nodes straddle the iso value
    node, ax = sk // 3, sk % 3
    step = np.array([1, n, n * n])[ax]
    ff = f.reshape(-1)
    assert np.all((ff[node] < 0) != (ff[node + step] < 0))


def test_nb_normals_are_analytic_gradients(upf, oracle):
    s = S.spec_torus_ridged_voronoi(42)
    g = S.uniform_grid(-1.6, 1.6, 257)
    keys, pos, nrm, idx, st = nb_numpy(s, g, band_for(g))
    _, grad = oracle.grad_points(s, pos)
    ref = grad / np.maximum(np.linalg.norm(grad, axis=1, keepdims=True), 1e-30)
    err = np.linalg.norm(nrm - ref, axis=1)
    print(f"normal error: median {np.median(err):.2e}, p99.9 {np.quantile(err, 0.999):.2e}, max {err.max():.2e}")
    assert np.mean(err <= 1e-4) >= 0.999
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)


def test_nb_equals_dense_gpu_torus_513(upf):
    import torch
    s = S.spec_torus_ridged_voronoi(42)
    g = S.uniform_grid(-1.6, 1.6, 513)
    keys, pos, nrm, idx, st = nb_numpy(s, g, band_for(g))
    f = upfm.field_eval_grid(s, g, variant=upfm.EVAL_ROWS)
    dense = upfm.mc_extract(f, g, normals=False)
    del f
    torch.cuda.empty_cache()
    sk, spos, _, stri = canonical(keys, pos, None, idx)
    print(f"513^3: {st.active_blocks}/{st.blocks} blocks active ({st.active_blocks / st.blocks:.3%}), "
          f"V={len(pos)} T={len(idx)}")
    assert len(sk) == dense.n_vertices and len(stri) == dense.n_triangles
    np.testing.assert_array_equal(spos, dense.positions.cpu().numpy())
    np.testing.assert_array_equal(stri, canon_tris(dense.indices.cpu().numpy().astype(np.int64)))


def test_nb_fallback_layer_equals_dense_rows(upf):
    """A layer too fine for both row forms (over 512 lattice columns per 64-node dense row,
    over 512 / 21 per 9-node brick row) runs per lane in both, beside the other layers'
    row forms, so the meshes stay equal bit for bit."""
    s = S.spec_sphere_fbm(1234)
    s.layers[s.n_layers] = S.layer("gradient", "none", 1, 900.0, 2.0, 0.5, 1e-4, 77)
    s.n_layers += 1
    g = S.uniform_grid(-1.3, 1.3, 257)
    keys, pos, nrm, idx, st = nb_numpy(s, g, band_for(g))
    f = upfm.field_eval_grid(s, g, variant=upfm.EVAL_ROWS)
    dense = upfm.mc_extract(f, g, normals=False)
    sk, spos, _, stri = canonical(keys, pos, None, idx)
    print(f"fallback layer 257^3: V={len(pos)} T={len(idx)}")
    assert len(sk) == dense.n_vertices and len(stri) == dense.n_triangles
    np.testing.assert_array_equal(spos, dense.positions.cpu().numpy())
    np.testing.assert_array_equal(stri, canon_tris(dense.indices.cpu().numpy().astype(np.int64)))


def test_nb_composition_shape_equals_dense_rows(upf):
    """A composition shape (cylinder) under the fBm layer: the bricks run the
    composition-capable instantiation and still give the dense row field's mesh."""
    s = S.field_spec("cylinder", (1.6, 0.8), [S.layer("gradient", "fbm", 4, 2.0, 2.0, 0.5, 0.05, 31)])
    g = S.uniform_grid(-1.3, 1.3, 129)
    keys, pos, nrm, idx, st = nb_numpy(s, g, band_for(g))
    f = upfm.field_eval_grid(s, g, variant=upfm.EVAL_ROWS)
    dense = upfm.mc_extract(f, g, normals=False)
    sk, spos, _, stri = canonical(keys, pos, None, idx)
    print(f"cylinder 129^3: V={len(pos)} T={len(idx)}")
    assert len(stri) > 1000
    assert len(sk) == dense.n_vertices and len(stri) == dense.n_triangles
    np.testing.assert_array_equal(spos, dense.positions.cpu().numpy())
    np.testing.assert_array_equal(stri, canon_tris(dense.indices.cpu().numpy().astype(np.int64)))


def test_nb_errors(upf):
    s = S.spec_torus_ridged_voronoi(42)
    g = S.uniform_grid(-1.6, 1.6, 257)
    with pytest.raises(upfm.UpfError) as e:
        upfm.nb_extract(s, g, 0.0)
    assert e.value.status == S.ERR_BAND
    L = upfm.lib()
    ws = upfm.NbWorkspace(g, 16)
    st = S.NbStats()
    mb = S.MeshBuffers()
    r = L.upf_nb_extract(ctypes.byref(s), ctypes.byref(g), 0.0, band_for(g), ctypes.byref(mb), None,
                         ctypes.c_void_p(ws.buf.data_ptr()), ws.bytes, ctypes.byref(st), None)
    assert r == S.ERR_WORKSPACE and st.active_blocks > 16
    bad = S.uniform_grid(-1.6, 1.6, 100)
    assert L.upf_nb_workspace_bytes(ctypes.byref(bad), 16) == 0
    with pytest.raises(upfm.UpfError) as e:
        upfm.nb_extract(s, bad, 0.1, workspace=ws)
    assert e.value.status == S.ERR_BAD_GRID
    # a band that finds no surface: an empty mesh, not an error
    far = S.field_spec("sphere", (10.0,), [])
    m, _, st = upfm.nb_extract(far, g, 0.01)
    assert m.n_triangles == 0 and st.active_blocks == 0


def _edge_stats_gpu(idx):
    """(unique edges, boundary edges, edges in > 2 triangles, consistently oriented) of a
    triangle list on the GPU (10^7 triangles sort in milliseconds)."""
    import torch
    t = idx.to(torch.int64)
    e = torch.cat([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]], 0)
    lo, hi = torch.minimum(e[:, 0], e[:, 1]), torch.maximum(e[:, 0], e[:, 1])
    key = lo * (1 << 32) + hi
    fwd = (e[:, 0] < e[:, 1]).to(torch.int64)
    key, order = torch.sort(key)
    fwd = fwd[order]
    uniq, counts = torch.unique_consecutive(key, return_counts=True)
    start = torch.cumsum(counts, 0) - counts
    dsum = torch.zeros_like(counts).index_add_(0, torch.repeat_interleave(torch.arange(len(counts), device=key.device),
                                                                          counts), fwd)
    oriented = bool(torch.all(dsum[counts == 2] == 1))
    return len(uniq), int((counts == 1).sum()), int((counts > 2).sum()), oriented


def test_nb_config5_planet0_1025_equals_dense_gpu(upf):
    """The combined workload's fine grid (1025^3 over [-1.6, 1.6]^3) for its first planet:
    the narrow-band mesh equals the dense GPU marching cubes of the row field bit for bit."""
    import torch
    name, s = S.nb_planets()[0]
    g = S.uniform_grid(-1.6, 1.6, 1025)
    print(f"case: config 5 {name} 1025^3 vs dense", flush=True)
    keys, pos, nrm, idx, st = nb_numpy(s, g, S.nb_band(g))
    f = upfm.field_eval_grid(s, g, variant=upfm.EVAL_ROWS)
    dense = upfm.mc_extract(f, g, normals=False)
    del f
    torch.cuda.empty_cache()
    sk, spos, _, stri = canonical(keys, pos, None, idx)
    print(f"1025^3: {st.active_blocks}/{st.blocks} blocks active, V={len(pos)} T={len(idx)}")
    assert len(stri) > 1_000_000
    assert len(sk) == dense.n_vertices and len(stri) == dense.n_triangles
    np.testing.assert_array_equal(spos, dense.positions.cpu().numpy())
    np.testing.assert_array_equal(stri, canon_tris(dense.indices.cpu().numpy().astype(np.int64)))


def test_nb_config5_all_planets_closed_with_unit_normals(upf):
    """All eight planets of the combined workload at 1025^3: every mesh closed (each edge
    in exactly two triangles, opposite directions), unit normals, no unresolved edges."""
    import torch
    g = S.uniform_grid(-1.6, 1.6, 1025)
    ws = None
    for name, s in S.nb_planets():
        print(f"case: config 5 {name}", flush=True)
        m, _, st = upfm.nb_extract(s, g, S.nb_band(g), workspace=ws)
        ws = upfm.NbWorkspace(g, int(st.active_blocks)) if ws is None or st.active_blocks > ws.max_active else ws
        assert st.unresolved == 0 and m.n_triangles > 1_000_000, name
        ne, boundary, bad, oriented = _edge_stats_gpu(m.indices)
        assert boundary == 0 and bad == 0 and oriented, (name, boundary, bad)
        nl = torch.linalg.norm(m.normals, dim=1)
        assert float((nl - 1).abs().max()) <= 1e-5, name
        print(f"  V={m.n_vertices} T={m.n_triangles} Euler={m.n_vertices - ne + m.n_triangles}")
        del m
