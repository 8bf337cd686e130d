"""The allocation-free entry point upf_field_mesh_async_static, called through the C-ABI,
against upf_field_mesh_async with a caller-allocated workspace: field, mesh and counts
equal bit for bit, on two calls in a row whose second grid is smaller (the shared
scratch keeps the first call's contents, which the second must not read), and a grid
whose workspace exceeds the scratch is refused before any launch."""
import ctypes

import pytest

import unnatural_planets_amd as upfm
from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu


def _run(L, s, g, static: bool, cap_v: int, cap_t: int):
    import torch
    nx, ny, nz = S.grid_dims(g)
    field = torch.full((nz, ny, nx), float("nan"), dtype=torch.float32, device="cuda")
    pos = torch.zeros((cap_v, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros((cap_v, 3), dtype=torch.float32, device="cuda")
    idx = torch.full((cap_t, 3), -1, dtype=torch.int32, device="cuda")
    counts = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    mb = S.MeshBuffers()
    mb.positions, mb.normals, mb.indices = pos.data_ptr(), nrm.data_ptr(), idx.data_ptr()
    mb.cap_vertices, mb.cap_triangles = cap_v, cap_t
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if static:
        st = L.upf_field_mesh_async_static(ctypes.byref(s), ctypes.byref(g), 0.0, vp(field), None, ctypes.byref(mb),
                                           vp(counts), stream)
    else:
        ws = upfm.McWorkspace(g)
        st = L.upf_field_mesh_async(ctypes.byref(s), ctypes.byref(g), 0.0, vp(field), None, ctypes.byref(mb),
                                    vp(counts), vp(ws.buf), ws.bytes, stream)
    assert st == S.OK, upfm.status_string(st)
    torch.cuda.synchronize()
    nv, nt = (int(v) for v in counts.tolist())
    assert 0 < nv <= cap_v and 0 < nt <= cap_t
    return field, pos[:nv], nrm[:nv], idx[:nt], (nv, nt)


def test_static_scratch_equals_workspace_call_twice(upf):
    import torch
    L = upfm.lib()
    assert L.upf_static_scratch_bytes() == 64 << 20
    cases = [(S.spec_torus_ridged_voronoi(42), S.grid_torus(256), 400000, 800000),
             (S.spec_sphere_fbm(1234), S.grid_sphere_fbm(96), 100000, 200000)]  # second call: smaller
    for s, g, cap_v, cap_t in cases:
        print(f"case: static scratch {S.grid_dims(g)}", flush=True)
        a = _run(L, s, g, True, cap_v, cap_t)
        b = _run(L, s, g, False, cap_v, cap_t)
        assert a[4] == b[4]
        for x, y in zip(a[:4], b[:4]):
            assert torch.equal(x.contiguous().view(torch.int32), y.contiguous().view(torch.int32))


def test_static_scratch_refuses_a_grid_too_large(upf):
    L = upfm.lib()
    g = S.uniform_grid(-1.0, 1.0, 640)  # its workspace is over 64 MiB
    assert L.upf_mc_workspace_bytes(ctypes.byref(g)) > L.upf_static_scratch_bytes()
    mb = S.MeshBuffers()
    # no buffers at all: the size check comes first, and a null field would be refused
    # before any launch as well
    st = L.upf_field_mesh_async_static(ctypes.byref(S.spec_sphere_fbm()), ctypes.byref(g), 0.0, None, None,
                                       ctypes.byref(mb), None, None)
    assert st == S.ERR_WORKSPACE
