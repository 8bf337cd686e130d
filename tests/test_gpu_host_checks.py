"""Host-side argument checks of the Python mirror, and calls on a side stream.

Every wrong shape below raises ValueError in Python before any kernel is launched (the
kernels read 3 values per row of xyz / positions and write cap rows of each output).
Calls given `stream=` allocate their outputs and temporaries on that stream; their
results equal the same calls on the current stream."""
import numpy as np
import pytest
import torch

from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu


def _spec():
    return S.spec_sphere_fbm(1234)


@pytest.mark.parametrize("shape", [(30,), (10, 2), (10, 4), (5, 3, 1)])
def test_points_must_be_rows_of_three(upf, shape):
    bad = torch.zeros(shape, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        upf.field_eval_points(_spec(), bad)
    with pytest.raises(ValueError):
        upf.field_adjoint(_spec(), bad)
    m = S.material_spec(S.spec_elevation_mode("lakes", 3))
    with pytest.raises(ValueError):
        upf.material_tiles(m, bad, torch.zeros((10, 3), dtype=torch.float32, device="cuda"))


def test_mesh_async_checks_every_output_buffer(upf):
    g = S.grid_sphere_fbm(32)
    field = torch.empty((32, 32, 32), dtype=torch.float32, device="cuda")
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    good_p = torch.empty((100, 3), dtype=torch.float32, device="cuda")
    good_i = torch.empty((200, 3), dtype=torch.int32, device="cuda")
    bad = [
        upf.Mesh(good_p, torch.empty((99, 3), dtype=torch.float32, device="cuda"), good_i),  # normals short
        upf.Mesh(good_p, None, torch.empty((200, 3), dtype=torch.int64, device="cuda")),     # indices dtype
        upf.Mesh(good_p, None, torch.empty((200, 2), dtype=torch.int32, device="cuda")),     # indices shape
        upf.Mesh(torch.empty((300,), dtype=torch.float32, device="cuda"), None, good_i),     # positions shape
    ]
    for m in bad:
        with pytest.raises(ValueError):
            upf.field_mesh_async(_spec(), g, field, m, counts)


def test_mc_extract_async_checks_buffers_and_counts(upf):
    g = S.grid_sphere_fbm(32)
    field = torch.zeros((32, 32, 32), dtype=torch.float32, device="cuda")
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    good_p = torch.empty((100, 3), dtype=torch.float32, device="cuda")
    good_i = torch.empty((200, 3), dtype=torch.int32, device="cuda")
    for m in (upf.Mesh(good_p, torch.empty((99, 3), dtype=torch.float32, device="cuda"), good_i),
              upf.Mesh(good_p, None, torch.empty((200, 3), dtype=torch.int64, device="cuda")),
              upf.Mesh(torch.empty((300,), dtype=torch.float32, device="cuda"), None, good_i)):
        with pytest.raises(ValueError):
            upf.mc_extract_async(field, g, m, counts)
    ok = upf.Mesh(good_p, None, good_i)
    for c in (torch.zeros(2, dtype=torch.int32, device="cuda"), torch.zeros(1, dtype=torch.int64, device="cuda"),
              torch.zeros(2, dtype=torch.int64)):
        with pytest.raises(ValueError):
            upf.mc_extract_async(field, g, ok, c)
    with pytest.raises(ValueError):
        upf.mc_extract_async(field[:, :, :31], g, ok, counts)


def test_bake_rasterize_checks_uv_and_method(upf):
    uv = upf.atlas_per_triangle(10, 32, 32)
    for bad in (uv.reshape(10, 6), uv.reshape(10, 2, 3), uv.double()):
        with pytest.raises(ValueError):
            upf.bake_rasterize(bad, 32, 32)
    with pytest.raises(ValueError):
        upf.bake_rasterize(uv, 32, 32, method="bogus")
    with pytest.raises(upf.UpfError) as e:
        upf.bake_rasterize(uv, (1 << 24) + 1, 1)
    assert e.value.status == S.ERR_BAD_GRID


def test_bake_checks_indices(upf):
    s = _spec()
    g = S.grid_sphere_fbm(24)
    m = upf.mc_extract(upf.field_eval_grid(s, g), g)
    uv = upf.atlas_per_triangle(m.n_triangles, 256, 256)
    tri = upf.bake_rasterize(uv, 256, 256)
    wrong = upf.Mesh(m.positions, m.normals, m.indices.reshape(-1))
    with pytest.raises(ValueError):
        upf.bake(S.bake_spec(s), wrong, uv, tri)


def test_side_stream_calls_match_current_stream(upf):
    s = _spec()
    g = S.grid_sphere_fbm(64)
    ref_f = upf.field_eval_grid(s, g)
    ref_m = upf.mc_extract(ref_f, g)
    ref_post = upf.mesh_remove_disconnected(ref_m)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    f = upf.field_eval_grid(s, g, stream=side)
    m = upf.mc_extract(f, g, stream=side)          # workspace allocated on the side stream
    post = upf.mesh_remove_disconnected(m, stream=side)
    # churn the allocator on the current stream while the side stream may still run
    junk = [torch.full((1 << 20,), 7.0, device="cuda") for _ in range(8)]
    side.synchronize()
    del junk
    assert torch.equal(f, ref_f)
    assert torch.equal(m.indices, ref_m.indices) and torch.equal(m.positions, ref_m.positions)
    assert torch.equal(post.indices, ref_post.indices)
    img = torch.full((64, 64, 3), float("nan"), device="cuda")
    img[20:40, 20:40] = 0.5
    ref_img = upf.image_dilate(img.clone(), 7)
    src = img.clone()  # made on the current stream: the side stream waits for it
    side.wait_stream(torch.cuda.current_stream())
    got = upf.image_dilate(src, 7, stream=side)
    side.synchronize()
    assert torch.equal(torch.nan_to_num(got, -1.0), torch.nan_to_num(ref_img, -1.0))
    pts = torch.from_numpy(np.random.default_rng(1).uniform(-1.2, 1.2, (5000, 3)).astype(np.float32)).cuda()
    f_ref = upf.field_eval_points(s, pts)
    f_side = upf.field_eval_points(s, pts, stream=side)
    side.synchronize()
    assert torch.equal(f_ref, f_side)
