"""Batched chunks over the upf_batch_* C-ABI: many uneven grids evaluated and meshed by
one set of persistent launches, optionally captured into a HIP graph."""
from __future__ import annotations

import ctypes

from . import Mesh, _ptr, _stream, check, lib
from .spec import BatchOpts, Chunk, ChunkMesh

class Batch:
    """Plan + buffers for one batch of chunks (ctypes Chunk array)."""

    def __init__(self, chunks, opts: BatchOpts | None = None, normals: bool = True):
        import torch
        self.L = lib()
        self.n = len(chunks)
        self.chunks = chunks
        self.h = ctypes.c_void_p()
        check(self.L.upf_batch_create(chunks, self.n, ctypes.byref(opts) if opts else None, ctypes.byref(self.h)),
              "upf_batch_create")
        self.field = torch.empty((int(self.L.upf_batch_field_floats(self.h)),), dtype=torch.float32, device="cuda")
        self.ws_bytes = int(self.L.upf_batch_workspace_bytes(self.h))
        self.ws = torch.empty((self.ws_bytes,), dtype=torch.uint8, device="cuda")
        self.meshes_dev = torch.zeros((self.n, 4), dtype=torch.int64, device="cuda")
        self.normals = normals
        self.cap_v = self.cap_t = 0
        self.pos = self.nrm = self.idx = None

    def _args(self):
        return (self.h, _ptr(self.field), _ptr(self.ws), self.ws_bytes, _ptr(self.pos), _ptr(self.nrm),
                _ptr(self.idx), self.cap_v, self.cap_t, _ptr(self.meshes_dev))

    def allocate(self, cap_v: int, cap_t: int):
        import torch
        self.cap_v, self.cap_t = int(cap_v), int(cap_t)
        self.pos = torch.empty((max(1, self.cap_v), 3), dtype=torch.float32, device="cuda")
        self.nrm = torch.empty((max(1, self.cap_v), 3), dtype=torch.float32, device="cuda") if self.normals else None
        self.idx = torch.empty((max(1, self.cap_t), 3), dtype=torch.int32, device="cuda")

    def run(self, stream=None):
        check(self.L.upf_batch_run(*self._args(), _stream(stream)), "upf_batch_run")

    def totals(self, stream=None):
        nv, nt = ctypes.c_int64(), ctypes.c_int64()
        check(self.L.upf_batch_totals(self.h, _ptr(self.ws), ctypes.byref(nv), ctypes.byref(nt), _stream(stream)),
              "upf_batch_totals")
        return nv.value, nt.value

    def size_outputs(self, slack: float = 1.0):
        """Run once without outputs to learn the totals, then allocate them."""
        self.allocate(0, 0)
        self.run()
        nv, nt = self.totals()
        self.allocate(int(nv * slack) + 1, int(nt * slack) + 1)
        return nv, nt

    def capture(self, stream):
        check(self.L.upf_batch_capture(*self._args(), _stream(stream)), "upf_batch_capture")

    def replay(self, stream=None):
        check(self.L.upf_batch_replay(self.h, _stream(stream)), "upf_batch_replay")

    def autotune(self, stream=None) -> BatchOpts:
        best = BatchOpts()
        check(self.L.upf_batch_autotune(*self._args(), _stream(stream), ctypes.byref(best)), "upf_batch_autotune")
        return best

    def set_opts(self, opts: BatchOpts):
        check(self.L.upf_batch_set_opts(self.h, ctypes.byref(opts)), "upf_batch_set_opts")

    def get_opts(self) -> BatchOpts:
        o = BatchOpts()
        check(self.L.upf_batch_get_opts(self.h, ctypes.byref(o)), "upf_batch_get_opts")
        return o

    def field_of(self, c: int):
        g = self.chunks[c].grid
        off = int(self.L.upf_batch_field_offset(self.h, c))
        n = int(g.dims[0]) * int(g.dims[1]) * int(g.dims[2])
        return self.field[off:off + n].view(int(g.dims[2]), int(g.dims[1]), int(g.dims[0]))

    def mesh_table(self):
        return [ChunkMesh(*[int(v) for v in row]) for row in self.meshes_dev.cpu().tolist()]

    def mesh_of(self, c: int, table=None) -> Mesh:
        m = (table or self.mesh_table())[c]
        v0, v1 = m.vertex_offset, m.vertex_offset + m.n_vertices
        t0, t1 = m.triangle_offset, m.triangle_offset + m.n_triangles
        return Mesh(self.pos[v0:v1], self.nrm[v0:v1] if self.nrm is not None else None, self.idx[t0:t1])

    def __del__(self):
        try:
            if self.h:
                self.L.upf_batch_destroy(self.h)
                self.h = ctypes.c_void_p()
        except Exception:
            pass
