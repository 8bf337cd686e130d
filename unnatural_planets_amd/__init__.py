"""unnatural_planets_amd -- MI355X (gfx950) field evaluation and meshing for unnatural-planets.

Host-side mirror of the reference's interface for this path, over the C-ABI of
libupf.so.  torch supplies device memory and streams only; every computation runs in
the library's HIP kernels, and every call raises when libupf.so is missing.

Reference functions mirrored (paths relative to the reference checkout):
  terrain_sdf_land / water / navigation / elevation / elevation_raw
      sources/modesConfiguration.cpp:230-279 (point-wise field)
  mesh_generate_generic          sources/meshGeneration.cpp:32-45
  mesh_generate_base_land        sources/meshGeneration.cpp:47-55
  mesh_generate_base_water       sources/meshGeneration.cpp:57-98
  mesh_remove_disconnected       sources/meshGeneration.cpp:41
  mesh_water_mask                sources/meshGeneration.cpp:62-95
  mesh_generate_base_navigation  sources/meshGeneration.cpp:100-107
  generate_textures_land         sources/textures.cpp:45-71, 86-90 (rasterize + per-texel material)
  texture_post                   sources/textures.cpp:73-81 (dilation by 7, U8 conversion)

Streams: a call given `stream=` runs with that stream current (torch.cuda.stream), so
its outputs and temporaries are allocated on it and the caching allocator reuses a
temporary only after the call's kernels on that stream.  The outputs belong to that
stream: a caller reading them on another stream synchronises first, as with any torch
tensor made on a side stream.
"""
from __future__ import annotations

import ctypes
import functools
import inspect
from dataclasses import dataclass

from . import spec as spec_mod
from ._lib import UpfError, check, lib, status_string  # noqa: F401
from .spec import (BakeSpec, FieldSpec, Grid, Layer, MaterialSpec, MeshBuffers, bake_spec, copy_spec,  # noqa: F401
                   field_spec, grid_dims, layer, material_spec, uniform_grid)

__all__ = [
    "FieldSpec", "Grid", "Layer", "Mesh", "UpfError", "field_spec", "layer", "uniform_grid",
    "field_eval_grid", "field_eval_points", "field_adjoint", "field_adjoint_grid", "mc_extract",
    "mc_case_codes", "terrain_sdf_land", "terrain_sdf_water", "terrain_sdf_navigation",
    "terrain_sdf_elevation", "terrain_sdf_elevation_raw", "mesh_generate_generic",
    "mesh_generate_base_land", "mesh_generate_base_water", "mesh_generate_base_navigation",
    "reference_grid", "flops_per_sample", "adjoint_flops_per_sample", "BakeSpec", "bake_spec",
    "Textures", "mc_extract_async", "atlas_per_triangle", "bake_rasterize", "bake", "generate_textures_land",
    "NbWorkspace", "nb_extract", "field_mesh_count", "mc_emit_counted", "image_dilate", "image_to_u8",
    "image_height_to_normal", "image_special_to_gltf_pbr", "export_maps",
    "texture_post", "mesh_remove_disconnected", "mesh_water_mask",
    "MaterialSpec", "material_spec", "material_tiles", "bake_material", "generate_textures", "texture_post_images",
]


def _torch():
    import torch
    return torch


def _stream(stream=None):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev(device):
    torch = _torch()
    return torch.device(device if device is not None else "cuda")


def _on_stream(fn):
    """Run fn with its `stream` argument (when given) as torch's current stream, so that
    everything fn allocates is allocated on the stream its kernels run on."""
    sig = inspect.signature(fn)

    @functools.wraps(fn)
    def wrapper(*args, **kw):
        stream = sig.bind(*args, **kw).arguments.get("stream")
        if stream is None:
            return fn(*args, **kw)
        with _torch().cuda.stream(stream):
            return fn(*args, **kw)
    return wrapper


def _check_f32(t, name, shape=None):
    torch = _torch()
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _check_rows3(t, name, dtype="float32"):
    """t must be a contiguous CUDA tensor of shape [n, 3] (the kernels read 3 values per row)."""
    torch = _torch()
    want = torch.float32 if dtype == "float32" else torch.int32
    if t.dtype != want or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor")
    if t.dim() != 2 or int(t.shape[1]) != 3:
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected [n, 3]")


def flops_per_sample(s: FieldSpec) -> float:
    return float(lib().upf_flops_per_sample(ctypes.byref(s)))


def adjoint_flops_per_sample(s: FieldSpec) -> float:
    return float(lib().upf_adjoint_flops_per_sample(ctypes.byref(s)))


def reference_grid(debug: bool = False) -> Grid:
    """The reference's sweep box [-2000, 2000]^3 at 500^3 nodes (110^3 in debug builds),
    sources/meshGeneration.cpp:19-28,35-36."""
    return uniform_grid(-2000.0, 2000.0, 110 if debug else 500)


# ---------------------------------------------------------------------------- field
EVAL_AUTO, EVAL_ROWS, EVAL_PER_LANE = 0, 1, 2


class EvalOpts(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


@_on_stream
def field_eval_grid(s: FieldSpec, g: Grid, out=None, stream=None, variant: int = EVAL_AUTO):
    """f at every node of g -> float32 tensor [nz, ny, nx] (upf_field_eval_grid_ex)."""
    torch = _torch()
    nx, ny, nz = grid_dims(g)
    if out is None:
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    _check_f32(out, "out", (nz, ny, nx))
    opts = EvalOpts()
    opts.variant = variant
    check(lib().upf_field_eval_grid_ex(ctypes.byref(s), ctypes.byref(g), _ptr(out), ctypes.byref(opts),
                                       _stream(stream)), "upf_field_eval_grid_ex")
    return out


@_on_stream
def field_eval_points(s: FieldSpec, xyz, with_grad: bool = False, stream=None):
    """f (and df/dx when with_grad) at points xyz [n, 3] (upf_field_eval_points)."""
    torch = _torch()
    _check_rows3(xyz, "xyz")
    n = xyz.shape[0]
    f = torch.empty((n,), dtype=torch.float32, device=xyz.device)
    g = torch.empty((n, 3), dtype=torch.float32, device=xyz.device) if with_grad else None
    check(lib().upf_field_eval_points(ctypes.byref(s), _ptr(xyz), n, _ptr(f), _ptr(g), _stream(stream)),
          "upf_field_eval_points")
    return (f, g) if with_grad else f


@_on_stream
def field_adjoint(s: FieldSpec, xyz, dLdf=None, want_f=True, want_dfdx=True, dtheta=None, stream=None):
    """Fused forward + adjoint at points (upf_field_adjoint).  Returns (f, dfdx, dtheta);
    dtheta is a float32 [76] tensor in upf_theta_grad order (sum over points of
    dLdf * df/dθ), None when dLdf is None."""
    torch = _torch()
    _check_rows3(xyz, "xyz")
    n = xyz.shape[0]
    f = torch.empty((n,), dtype=torch.float32, device=xyz.device) if want_f else None
    g = torch.empty((n, 3), dtype=torch.float32, device=xyz.device) if want_dfdx else None
    if dLdf is not None:
        _check_f32(dLdf, "dLdf", (n,))
        if dtheta is None:
            dtheta = torch.zeros((spec_mod.THETA_COUNT,), dtype=torch.float32, device=xyz.device)
    else:
        dtheta = None
    check(lib().upf_field_adjoint(ctypes.byref(s), _ptr(xyz), n, _ptr(dLdf), _ptr(f), _ptr(g), _ptr(dtheta),
                                  _stream(stream)), "upf_field_adjoint")
    return f, g, dtheta


@_on_stream
def field_adjoint_grid(s: FieldSpec, gr: Grid, dLdf=None, want_f=True, want_dfdx=True, stream=None):
    """Fused forward + adjoint over grid nodes (upf_field_adjoint_grid)."""
    torch = _torch()
    nx, ny, nz = grid_dims(gr)
    f = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda") if want_f else None
    g = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device="cuda") if want_dfdx else None
    dtheta = None
    if dLdf is not None:
        _check_f32(dLdf, "dLdf", (nz, ny, nx))
        dtheta = torch.zeros((spec_mod.THETA_COUNT,), dtype=torch.float32, device="cuda")
    check(lib().upf_field_adjoint_grid(ctypes.byref(s), ctypes.byref(gr), _ptr(dLdf), _ptr(f), _ptr(g),
                                       _ptr(dtheta), _stream(stream)), "upf_field_adjoint_grid")
    return f, g, dtheta


def _with_compose(s: FieldSpec, compose: str) -> FieldSpec:
    r = copy_spec(s)
    r.compose = spec_mod.COMPOSE[compose]
    return r


def terrain_sdf_land(s, xyz):
    """terrainSdfLand at points (sources/modesConfiguration.cpp:248-258)."""
    return field_eval_points(_with_compose(s, "land"), xyz)


def terrain_sdf_water(s, xyz):
    """terrainSdfWater (sources/modesConfiguration.cpp:260-267)."""
    return field_eval_points(_with_compose(s, "water"), xyz)


def terrain_sdf_navigation(s, xyz):
    """terrainSdfNavigation (sources/modesConfiguration.cpp:269-279)."""
    return field_eval_points(_with_compose(s, "navigation"), xyz)


def terrain_sdf_elevation(s, xyz):
    """terrainSdfElevation = shape / elevation_scale (sources/modesConfiguration.cpp:230-237)."""
    return field_eval_points(_with_compose(s, "elevation"), xyz)


def terrain_sdf_elevation_raw(s, xyz):
    """terrainSdfElevationRaw (sources/modesConfiguration.cpp:239-246)."""
    return field_eval_points(_with_compose(s, "elevation_raw"), xyz)


# ---------------------------------------------------------------------------- meshing
@dataclass
class Mesh:
    positions: object  # float32 [V, 3]
    normals: object    # float32 [V, 3] or None
    indices: object    # int32 [T, 3] (uint32 values)

    @property
    def n_vertices(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_triangles(self) -> int:
        return int(self.indices.shape[0])


class McWorkspace:
    """Reusable marching-cubes workspace for one grid shape."""

    def __init__(self, g: Grid, device=None):
        torch = _torch()
        self.bytes = int(lib().upf_mc_workspace_bytes(ctypes.byref(g)))
        if self.bytes == 0:
            raise UpfError(spec_mod.ERR_BAD_GRID, "upf_mc_workspace_bytes")
        self.buf = torch.empty((self.bytes,), dtype=torch.uint8, device=_dev(device))


@_on_stream
def mc_extract(field, g: Grid, iso: float = 0.0, normals: bool = True, workspace: McWorkspace | None = None,
               stream=None) -> Mesh:
    """Marching cubes over field [nz, ny, nx] (upf_mc_count + upf_mc_emit)."""
    torch = _torch()
    nx, ny, nz = grid_dims(g)
    _check_f32(field, "field", (nz, ny, nx))
    ws = workspace or McWorkspace(g, field.device)
    nv, nt = ctypes.c_int64(0), ctypes.c_int64(0)
    st = _stream(stream)
    check(lib().upf_mc_count(_ptr(field), ctypes.byref(g), iso, _ptr(ws.buf), ws.bytes, ctypes.byref(nv),
                             ctypes.byref(nt), st), "upf_mc_count")
    pos = torch.empty((nv.value, 3), dtype=torch.float32, device=field.device)
    nrm = torch.empty((nv.value, 3), dtype=torch.float32, device=field.device) if normals else None
    idx = torch.empty((nt.value, 3), dtype=torch.int32, device=field.device)
    mb = MeshBuffers()
    mb.positions = pos.data_ptr() if nv.value else None
    mb.normals = nrm.data_ptr() if (normals and nv.value) else None
    mb.indices = idx.data_ptr() if nt.value else None
    mb.cap_vertices, mb.cap_triangles = nv.value, nt.value
    mb.n_vertices, mb.n_triangles = nv.value, nt.value
    check(lib().upf_mc_emit(_ptr(field), ctypes.byref(g), iso, ctypes.byref(mb), _ptr(ws.buf), ws.bytes, st),
          "upf_mc_emit")
    return Mesh(pos, nrm, idx)


@_on_stream
def mc_case_codes(field, g: Grid, iso: float = 0.0, stream=None):
    torch = _torch()
    nx, ny, nz = grid_dims(g)
    _check_f32(field, "field", (nz, ny, nx))
    out = torch.empty((nz - 1, ny - 1, nx - 1), dtype=torch.uint8, device=field.device)
    check(lib().upf_mc_case_codes(_ptr(field), ctypes.byref(g), iso, _ptr(out), _stream(stream)),
          "upf_mc_case_codes")
    return out


@_on_stream
def field_mesh_count(s: FieldSpec, g: Grid, iso: float = 0.0, field=None, workspace: McWorkspace | None = None,
                     variant: int = EVAL_AUTO, stream=None):
    """Field evaluation fused with marching-cubes classification (upf_field_mesh_count).
    Returns (field [nz, ny, nx], workspace, n_vertices, n_triangles); mc_emit_counted
    then writes the mesh."""
    torch = _torch()
    nx, ny, nz = grid_dims(g)
    if field is None:
        field = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    _check_f32(field, "field", (nz, ny, nx))
    ws = workspace or McWorkspace(g, field.device)
    opts = EvalOpts()
    opts.variant = variant
    nv, nt = ctypes.c_int64(0), ctypes.c_int64(0)
    check(lib().upf_field_mesh_count(ctypes.byref(s), ctypes.byref(g), iso, _ptr(field), ctypes.byref(opts),
                                     _ptr(ws.buf), ws.bytes, ctypes.byref(nv), ctypes.byref(nt), _stream(stream)),
          "upf_field_mesh_count")
    return field, ws, int(nv.value), int(nt.value)


@_on_stream
def mc_emit_counted(field, g: Grid, iso: float, ws: McWorkspace, n_vertices: int, n_triangles: int,
                    normals: bool = True, stream=None) -> Mesh:
    """upf_mc_emit after upf_mc_count / upf_field_mesh_count left their state in ws."""
    torch = _torch()
    pos = torch.empty((n_vertices, 3), dtype=torch.float32, device=field.device)
    nrm = torch.empty((n_vertices, 3), dtype=torch.float32, device=field.device) if normals else None
    idx = torch.empty((n_triangles, 3), dtype=torch.int32, device=field.device)
    mb = MeshBuffers()
    mb.positions = pos.data_ptr() if n_vertices else None
    mb.normals = nrm.data_ptr() if (normals and n_vertices) else None
    mb.indices = idx.data_ptr() if n_triangles else None
    mb.cap_vertices, mb.cap_triangles = n_vertices, n_triangles
    mb.n_vertices, mb.n_triangles = n_vertices, n_triangles
    check(lib().upf_mc_emit(_ptr(field), ctypes.byref(g), iso, ctypes.byref(mb), _ptr(ws.buf), ws.bytes,
                            _stream(stream)), "upf_mc_emit")
    return Mesh(pos, nrm, idx)


@_on_stream
def field_mesh_async(s: FieldSpec, g: Grid, field, mesh: Mesh, counts, iso: float = 0.0,
                     workspace: McWorkspace | None = None, variant: int = EVAL_AUTO, stream=None) -> McWorkspace:
    """Field + marching cubes in one stream-ordered call (upf_field_mesh_async): the mesh
    goes into the preallocated `mesh` tensors (their row counts are the capacities) and
    the required counts into `counts` (int64 [2] on the device).  Nothing synchronises;
    the mesh is complete when counts <= the capacities.  Returns the workspace."""
    nx, ny, nz = grid_dims(g)
    _check_f32(field, "field", (nz, ny, nx))
    if counts.dtype != _torch().int64 or counts.numel() < 2 or not counts.is_contiguous():
        raise ValueError("counts must be a contiguous int64 tensor of 2 elements")
    ws = workspace or McWorkspace(g, field.device)
    opts = EvalOpts()
    opts.variant = variant
    _check_rows3(mesh.positions, "mesh.positions")
    _check_rows3(mesh.indices, "mesh.indices", "int32")
    cap_v, cap_t = int(mesh.positions.shape[0]), int(mesh.indices.shape[0])
    if mesh.normals is not None:
        _check_f32(mesh.normals, "mesh.normals", (cap_v, 3))
    mb = MeshBuffers()
    mb.positions = mesh.positions.data_ptr() if cap_v else None
    mb.normals = mesh.normals.data_ptr() if (mesh.normals is not None and cap_v) else None
    mb.indices = mesh.indices.data_ptr() if cap_t else None
    mb.cap_vertices, mb.cap_triangles = cap_v, cap_t
    check(lib().upf_field_mesh_async(ctypes.byref(s), ctypes.byref(g), iso, _ptr(field), ctypes.byref(opts),
                                      ctypes.byref(mb), _ptr(counts), _ptr(ws.buf), ws.bytes, _stream(stream)),
          "upf_field_mesh_async")
    return ws


def _cap_buffers(mesh: Mesh):
    _check_rows3(mesh.positions, "mesh.positions")
    _check_rows3(mesh.indices, "mesh.indices", "int32")
    cap_v, cap_t = int(mesh.positions.shape[0]), int(mesh.indices.shape[0])
    if mesh.normals is not None:
        _check_f32(mesh.normals, "mesh.normals", (cap_v, 3))
    mb = MeshBuffers()
    mb.positions = mesh.positions.data_ptr() if cap_v else None
    mb.normals = mesh.normals.data_ptr() if (mesh.normals is not None and cap_v) else None
    mb.indices = mesh.indices.data_ptr() if cap_t else None
    mb.cap_vertices, mb.cap_triangles = cap_v, cap_t
    return mb


def _check_counts(counts):
    if counts.dtype != _torch().int64 or counts.numel() < 2 or not counts.is_contiguous() or not counts.is_cuda:
        raise ValueError("counts must be a contiguous int64 CUDA tensor of 2 elements")


@_on_stream
def mc_extract_async(field, g: Grid, mesh: Mesh, counts, iso: float = 0.0, workspace: McWorkspace | None = None,
                     stream=None) -> McWorkspace:
    """Marching cubes over a stored field in one stream-ordered call (upf_mc_extract_async):
    the mesh goes into the preallocated `mesh` tensors (their row counts are the
    capacities), the required counts into `counts` (int64 [2] on the device).  Nothing
    synchronises; the mesh is complete when counts <= the capacities."""
    nx, ny, nz = grid_dims(g)
    _check_f32(field, "field", (nz, ny, nx))
    _check_counts(counts)
    ws = workspace or McWorkspace(g, field.device)
    mb = _cap_buffers(mesh)
    check(lib().upf_mc_extract_async(_ptr(field), ctypes.byref(g), iso, ctypes.byref(mb), _ptr(counts), _ptr(ws.buf),
                                      ws.bytes, _stream(stream)), "upf_mc_extract_async")
    return ws


def _mesh_buffers(m: Mesh):
    mb = MeshBuffers()
    mb.positions = m.positions.data_ptr() if m.n_vertices else None
    mb.normals = m.normals.data_ptr() if (m.normals is not None and m.n_vertices) else None
    mb.indices = m.indices.data_ptr() if m.n_triangles else None
    mb.cap_vertices = mb.n_vertices = m.n_vertices
    mb.cap_triangles = mb.n_triangles = m.n_triangles
    return mb


@_on_stream
def _mesh_post(fn, m: Mesh, *lead, stream=None) -> Mesh:
    torch = _torch()
    _check_rows3(m.positions, "positions")
    _check_rows3(m.indices, "indices", "int32")
    if m.normals is not None:
        _check_f32(m.normals, "normals", tuple(m.positions.shape))
    dev = m.positions.device
    bytes_ = int(lib().upf_mesh_post_workspace_bytes(m.n_vertices, m.n_triangles))
    if bytes_ == 0:
        raise UpfError(spec_mod.ERR_BAD_GRID, "upf_mesh_post_workspace_bytes")
    ws = torch.empty((bytes_,), dtype=torch.uint8, device=dev)
    pos = torch.empty_like(m.positions)
    nrm = torch.empty_like(m.normals) if m.normals is not None else None
    idx = torch.empty_like(m.indices)
    out = _mesh_buffers(Mesh(pos, nrm, idx))
    check(fn(*lead, ctypes.byref(_mesh_buffers(m)), ctypes.byref(out), _ptr(ws), bytes_, _stream(stream)),
          fn.__name__)
    nv, nt = int(out.n_vertices), int(out.n_triangles)
    return Mesh(pos[:nv], nrm[:nv] if nrm is not None else None, idx[:nt])


@_on_stream
def mesh_remove_disconnected(m: Mesh, stream=None) -> Mesh:
    """meshRemoveDisconnected (sources/meshGeneration.cpp:41): the connected component with
    the most triangles, order preserved (upf_mesh_remove_disconnected)."""
    return _mesh_post(lib().upf_mesh_remove_disconnected, m, stream=stream)


@_on_stream
def mesh_water_mask(s: FieldSpec, m: Mesh, threshold: float = 0.1, stream=None) -> Mesh:
    """The water-mesh vertex mask of sources/meshGeneration.cpp:62-95: vertices with raw
    elevation below threshold, grown by two triangle rings (upf_mesh_water_mask)."""
    return _mesh_post(lib().upf_mesh_water_mask, m, ctypes.byref(s), ctypes.c_float(threshold), stream=stream)


def mesh_generate_generic(s: FieldSpec, g: Grid, iso: float = 0.0, normals: bool = True,
                          remove_disconnected: bool = True) -> Mesh:
    """Grid sweep + marching cubes (sources/meshGeneration.cpp:32-45), as one fused
    field + classification pass and the emission, then meshRemoveDisconnected (:41).
    meshFlipNormals (:42) is not needed: normals and winding already point outward
    (towards increasing f)."""
    f, ws, nv, nt = field_mesh_count(s, g, iso)
    m = mc_emit_counted(f, g, iso, ws, nv, nt, normals)
    return mesh_remove_disconnected(m) if remove_disconnected else m


def mesh_generate_base_land(s: FieldSpec, g: Grid | None = None) -> Mesh:
    """sources/meshGeneration.cpp:47-55: land mesh, raising on an empty result."""
    m = mesh_generate_generic(_with_compose(s, "land"), g or reference_grid())
    if m.n_triangles == 0:
        raise UpfError(spec_mod.ERR_EMPTY_MESH, "generated empty base land mesh")
    return m


def mesh_generate_base_water(s: FieldSpec, g: Grid | None = None, threshold: float = 0.1) -> Mesh:
    """sources/meshGeneration.cpp:57-98: water mesh, then the vertex mask (raw elevation
    below threshold, grown by two triangle rings) of :62-95."""
    m = mesh_generate_generic(_with_compose(s, "water"), g or reference_grid())
    return mesh_water_mask(s, m, threshold)


def mesh_generate_base_navigation(s: FieldSpec, g: Grid | None = None) -> Mesh:
    """sources/meshGeneration.cpp:100-107: navigation mesh, raising on an empty result."""
    m = mesh_generate_generic(_with_compose(s, "navigation"), g or reference_grid())
    if m.n_triangles == 0:
        raise UpfError(spec_mod.ERR_EMPTY_MESH, "generated empty base navigation mesh")
    return m


# ---------------------------------------------------------------------------- narrow band
class NbWorkspace:
    """Narrow-band workspace for one fine grid shape, holding up to `max_active` blocks."""

    def __init__(self, g: Grid, max_active: int, device=None):
        torch = _torch()
        self.max_active = int(max_active)
        self.bytes = int(lib().upf_nb_workspace_bytes(ctypes.byref(g), self.max_active))
        if self.bytes == 0:
            raise UpfError(spec_mod.ERR_BAD_GRID, "upf_nb_workspace_bytes")
        self.buf = torch.empty((self.bytes,), dtype=torch.uint8, device=_dev(device))


@_on_stream
def nb_extract(s: FieldSpec, g: Grid, band: float, iso: float = 0.0, normals: bool = True, keys: bool = False,
               workspace: NbWorkspace | None = None, device=None, stream=None):
    """Narrow-band marching cubes of the fine grid g (upf_nb_extract).  Grows the
    workspace and the mesh buffers as the library reports (each retry reruns the
    extraction).  Returns (Mesh, keys or None, NbStats)."""
    torch = _torch()
    dev = _dev(device)
    ws = workspace or NbWorkspace(g, 1024, dev)
    stats = spec_mod.NbStats()
    cap_v = cap_t = 0
    st = _stream(stream)
    for _ in range(4):
        pos = torch.empty((cap_v, 3), dtype=torch.float32, device=dev)
        nrm = torch.empty((cap_v, 3), dtype=torch.float32, device=dev) if normals else None
        idx = torch.empty((cap_t, 3), dtype=torch.int32, device=dev)
        kk = torch.empty((cap_v,), dtype=torch.int64, device=dev) if keys else None
        mb = MeshBuffers()
        mb.positions = pos.data_ptr() if cap_v else None
        mb.normals = nrm.data_ptr() if (normals and cap_v) else None
        mb.indices = idx.data_ptr() if cap_t else None
        mb.cap_vertices, mb.cap_triangles = cap_v, cap_t
        r = lib().upf_nb_extract(ctypes.byref(s), ctypes.byref(g), iso, band, ctypes.byref(mb),
                                 _ptr(kk) if (keys and cap_v) else None, _ptr(ws.buf), ws.bytes,
                                 ctypes.byref(stats), st)
        if r == spec_mod.ERR_WORKSPACE and stats.active_blocks > ws.max_active:
            ws = NbWorkspace(g, stats.active_blocks, dev)
            continue
        if r == spec_mod.ERR_CAPACITY:
            cap_v, cap_t = int(stats.n_vertices), int(stats.n_triangles)
            continue
        check(r, "upf_nb_extract")
        n_v, n_t = int(stats.n_vertices), int(stats.n_triangles)
        return Mesh(pos[:n_v], nrm[:n_v] if normals else None, idx[:n_t]), (kk[:n_v] if keys else None), stats
    raise UpfError(spec_mod.ERR_CAPACITY, "upf_nb_extract did not converge")


# ---------------------------------------------------------------------------- texel bake
@dataclass
class Textures:
    albedo: object     # float32 [h, w, 3]; NaN where no triangle covers the texel
    bump: object       # float32 [h, w]
    roughness: object  # float32 [h, w]
    uv: object         # float32 [T, 3, 2]
    tri: object        # int32 [h, w], covering triangle or -1


@_on_stream
def atlas_per_triangle(n_triangles: int, width: int, height: int, device=None, stream=None):
    """Per-triangle UV atlas (upf_atlas_per_triangle) -> float32 [T, 3, 2]."""
    torch = _torch()
    uv = torch.empty((n_triangles, 3, 2), dtype=torch.float32, device=_dev(device))
    check(lib().upf_atlas_per_triangle(n_triangles, width, height, _ptr(uv) if n_triangles else None,
                                       _stream(stream)), "upf_atlas_per_triangle")
    return uv


@_on_stream
def bake_rasterize(uv, width: int, height: int, stream=None, method: str = "tiles"):
    """Covering triangle per texel -> int32 [h, w], -1 where uncovered.

    method "tiles": upf_bake_rasterize_ws (tile-binned, the map stored once; sides up to
    2^24 texels); "atomic": upf_bake_rasterize (one thread per triangle, atomicMax into a
    -1 filled map).  Both give the same map."""
    torch = _torch()
    _check_f32(uv, "uv")
    if uv.dim() != 3 or tuple(uv.shape[1:]) != (3, 2):
        raise ValueError("uv must have shape [T, 3, 2]")
    tri = torch.empty((height, width), dtype=torch.int32, device=uv.device)
    n = int(uv.shape[0])
    L = lib()
    if method == "atomic":
        check(L.upf_bake_rasterize(_ptr(uv) if n else None, n, width, height, _ptr(tri), _stream(stream)),
              "upf_bake_rasterize")
    elif method == "tiles":
        nbytes = L.upf_bake_rasterize_workspace_bytes(n, width, height)
        if nbytes == 0:
            check(5, "upf_bake_rasterize_workspace_bytes")  # UPF_ERR_BAD_GRID
        ws = torch.empty(nbytes, dtype=torch.uint8, device=uv.device)
        check(L.upf_bake_rasterize_ws(_ptr(uv) if n else None, n, width, height, _ptr(tri), _ptr(ws), nbytes,
                                      _stream(stream)), "upf_bake_rasterize_ws")
    else:
        raise ValueError(f"unknown method {method!r}")
    return tri


@_on_stream
def bake(b: BakeSpec, mesh: Mesh, uv, tri, stream=None):
    """Per-texel material (upf_bake) -> (albedo [h, w, 3], bump [h, w], roughness [h, w])."""
    torch = _torch()
    h, w = (int(x) for x in tri.shape)
    n = mesh.n_triangles
    _check_rows3(mesh.positions, "positions")
    if mesh.normals is None:
        raise ValueError("the bake needs vertex normals")
    _check_f32(mesh.normals, "normals", tuple(mesh.positions.shape))
    _check_rows3(mesh.indices, "indices", "int32")
    _check_f32(uv, "uv", (n, 3, 2))
    if tri.dtype != torch.int32 or not tri.is_contiguous() or not tri.is_cuda:
        raise ValueError("tri must be a contiguous int32 CUDA tensor")
    dev = tri.device
    albedo = torch.empty((h, w, 3), dtype=torch.float32, device=dev)
    bump = torch.empty((h, w), dtype=torch.float32, device=dev)
    rough = torch.empty((h, w), dtype=torch.float32, device=dev)
    nz = n > 0
    check(lib().upf_bake(ctypes.byref(b), _ptr(mesh.positions) if nz else None, _ptr(mesh.normals) if nz else None,
                         mesh.n_vertices, _ptr(mesh.indices) if nz else None, n, _ptr(uv) if nz else None, _ptr(tri), w, h,
                         _ptr(albedo), _ptr(bump), _ptr(rough), _stream(stream)), "upf_bake")
    return albedo, bump, rough


@_on_stream
def generate_textures_land(b: BakeSpec, mesh: Mesh, width: int, height: int, stream=None) -> Textures:
    """Atlas + rasterize + per-texel material for a land mesh (sources/textures.cpp:45-71,
    86-90): the reference's Generator<false>::generate up to its dilation and U8 steps."""
    uv = atlas_per_triangle(mesh.n_triangles, width, height, mesh.positions.device, stream)
    tri = bake_rasterize(uv, width, height, stream)
    albedo, bump, rough = bake(b, mesh, uv, tri, stream)
    return Textures(albedo, bump, rough, uv, tri)


# ---------------------------------------------------------------------------- texture post
@_on_stream
def image_dilate(img, rounds: int = 7, stream=None):
    """In-place dilation of an [h, w] or [h, w, c] float32 image (upf_image_dilate)."""
    torch = _torch()
    _check_f32(img, "img")
    if img.dim() not in (2, 3):
        raise ValueError("img must be [h, w] or [h, w, c]")
    h, w = int(img.shape[0]), int(img.shape[1])
    c = 1 if img.dim() == 2 else int(img.shape[2])
    scratch = torch.empty_like(img)
    check(lib().upf_image_dilate(_ptr(img), _ptr(scratch), w, h, c, rounds, _stream(stream)), "upf_image_dilate")
    return img


@_on_stream
def image_to_u8(img, stream=None):
    """NaN -> 0, else round(saturate(v) * 255) (upf_image_to_u8)."""
    torch = _torch()
    _check_f32(img, "img")
    out = torch.empty(img.shape, dtype=torch.uint8, device=img.device)
    check(lib().upf_image_to_u8(_ptr(img), img.numel(), _ptr(out), _stream(stream)), "upf_image_to_u8")
    return out


@_on_stream
def texture_post(tex: Textures, rounds: int = 7, stream=None):
    """The reference's images after the bake (sources/textures.cpp:73-81): albedo [h, w, 3],
    special [h, w, 2] = (roughness, metallic = 0) and height [h, w], each dilated by
    `rounds` texels and converted to U8.  Returns (albedo_u8, special_u8, height_u8)."""
    torch = _torch()
    albedo = tex.albedo.clone()
    special = torch.stack([tex.roughness, torch.where(torch.isnan(tex.roughness), tex.roughness,
                                                      torch.zeros_like(tex.roughness))], dim=2).contiguous()
    height = tex.bump.clone()
    for im in (albedo, special, height):
        image_dilate(im, rounds, stream)
    return image_to_u8(albedo, stream), image_to_u8(special, stream), image_to_u8(height, stream)


@_on_stream
def image_height_to_normal(height_u8, strength: float = 1.0, stream=None):
    """U8 height map [h, w] -> U8 normal map [h, w, 3] (upf_image_height_to_normal; the
    reference's imageConvertHeigthToNormal(heightMap, 1), sources/generator.cpp:138)."""
    torch = _torch()
    if height_u8.dtype != torch.uint8 or height_u8.dim() != 2 or not height_u8.is_contiguous():
        raise ValueError("height_u8 must be a contiguous [h, w] uint8 tensor")
    h, w = int(height_u8.shape[0]), int(height_u8.shape[1])
    out = torch.empty((h, w, 3), dtype=torch.uint8, device=height_u8.device)
    check(lib().upf_image_height_to_normal(_ptr(height_u8), w, h, float(strength), _ptr(out), _stream(stream)),
          "upf_image_height_to_normal")
    return out


@_on_stream
def image_special_to_gltf_pbr(special_u8, stream=None):
    """U8 (roughness, metallic) [..., 2] -> glTF metallicRoughness [..., 3] = (255, roughness,
    metallic) (upf_image_special_to_gltf_pbr; imageConvertSpecialToGltfPbr,
    sources/generator.cpp:136)."""
    torch = _torch()
    if special_u8.dtype != torch.uint8 or special_u8.shape[-1] != 2 or not special_u8.is_contiguous():
        raise ValueError("special_u8 must be a contiguous [..., 2] uint8 tensor")
    out = torch.empty(tuple(special_u8.shape[:-1]) + (3,), dtype=torch.uint8, device=special_u8.device)
    check(lib().upf_image_special_to_gltf_pbr(_ptr(special_u8), special_u8.numel() // 2, _ptr(out), _stream(stream)),
          "upf_image_special_to_gltf_pbr")
    return out


@_on_stream
def export_maps(albedo_u8, special_u8, height_u8, stream=None):
    """The conversions LandProcessor::chunkEntry applies before writing a chunk's images
    (sources/generator.cpp:134-139): albedo unchanged, special -> glTF PBR, height ->
    normal map with strength 1.  Returns (albedo, pbr, normal)."""
    return albedo_u8, image_special_to_gltf_pbr(special_u8, stream), image_height_to_normal(height_u8, 1.0, stream)


# ---------------------------------------------------------------------------- material stack
@_on_stream
def material_tiles(m: MaterialSpec, positions, normals, stream=None):
    """terrainTile at points with unit normals (upf_material_tiles; the navmesh loop of
    sources/tileProperties.cpp:258-273).  Returns a float32 [n, 16] tensor of upf_tile
    rows (spec.tiles_fields names the columns)."""
    torch = _torch()
    _check_rows3(positions, "positions")
    n = int(positions.shape[0])
    _check_f32(normals, "normals", (n, 3))
    out = torch.empty((n, 16), dtype=torch.float32, device=positions.device)
    check(lib().upf_material_tiles(ctypes.byref(m), _ptr(positions), _ptr(normals), n, _ptr(out), _stream(stream)),
          "upf_material_tiles")
    return out


@_on_stream
def bake_material(m: MaterialSpec, mesh: Mesh, uv, tri, stream=None):
    """Per-texel full material (upf_bake_material; Generator::pixel, sources/textures.cpp:
    24-43) -> (albedo [h, w, 3] or [h, w, 4] for water, special [h, w, 2], height [h, w])."""
    torch = _torch()
    h, w = (int(x) for x in tri.shape)
    n = mesh.n_triangles
    _check_rows3(mesh.positions, "positions")
    if mesh.normals is None:
        raise ValueError("the bake needs vertex normals")
    _check_f32(mesh.normals, "normals", tuple(mesh.positions.shape))
    _check_rows3(mesh.indices, "indices", "int32")
    _check_f32(uv, "uv", (n, 3, 2))
    if tri.dtype != torch.int32 or not tri.is_contiguous() or not tri.is_cuda:
        raise ValueError("tri must be a contiguous int32 CUDA tensor")
    dev = tri.device
    ac = 4 if m.purpose == spec_mod.PURPOSE["water"] else 3
    albedo = torch.empty((h, w, ac), dtype=torch.float32, device=dev)
    special = torch.empty((h, w, 2), dtype=torch.float32, device=dev)
    height = torch.empty((h, w), dtype=torch.float32, device=dev)
    nz = n > 0
    check(lib().upf_bake_material(ctypes.byref(m), _ptr(mesh.positions) if nz else None,
                                  _ptr(mesh.normals) if nz else None, mesh.n_vertices,
                                  _ptr(mesh.indices) if nz else None, n, _ptr(uv) if nz else None, _ptr(tri), w, h,
                                  _ptr(albedo), _ptr(special), _ptr(height), _stream(stream)), "upf_bake_material")
    return albedo, special, height


@_on_stream
def generate_textures(m: MaterialSpec, mesh: Mesh, width: int, height: int, stream=None):
    """generateTexturesLand / generateTexturesWater (sources/textures.cpp:86-96) with the
    full material: atlas, rasterization and per-texel tiles, then the dilation by 7 and
    the U8 conversion (:73-81).  Returns (albedo_u8, special_u8, height_u8, uv)."""
    uv = atlas_per_triangle(mesh.n_triangles, width, height, mesh.positions.device, stream)
    tri = bake_rasterize(uv, width, height, stream)
    albedo, special, hmap = bake_material(m, mesh, uv, tri, stream)
    a, s, h = texture_post_images(albedo, special, hmap, 7, stream)
    return a, s, h, uv


@_on_stream
def texture_post_images(albedo, special, height, rounds: int = 7, stream=None):
    """Dilation and U8 conversion of the three images (sources/textures.cpp:73-81)."""
    ims = [albedo.clone(), special.clone(), height.clone()]
    for im in ims:
        image_dilate(im, rounds, stream)
    return tuple(image_to_u8(im, stream) for im in ims)

