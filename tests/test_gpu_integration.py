"""The integration guide's C++ binding, compiled against include/upf.h and linked with
libupf.so (tests/integration/mesh_generate_generic.cpp, built by build()), run as a
separate process: its meshGenerateGeneric mesh equals the Python mirror's
mesh_generate_generic on the same spec and grid bit for bit, with the workspace call and
with the allocation-free one."""
import os
import subprocess

import numpy as np
import pytest

from unnatural_planets_amd import spec as S

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "integration", "mesh_generate_generic")


def _read(path):
    raw = open(path, "rb").read()
    nv, nt = np.frombuffer(raw[:16], dtype=np.int64)
    o = 16
    pos = np.frombuffer(raw[o:o + 12 * nv], dtype=np.float32).reshape(nv, 3)
    o += 12 * nv
    nrm = np.frombuffer(raw[o:o + 12 * nv], dtype=np.float32).reshape(nv, 3)
    o += 12 * nv
    idx = np.frombuffer(raw[o:o + 12 * nt], dtype=np.uint32).reshape(nt, 3)
    return pos, nrm, idx


@pytest.mark.parametrize("mode", ["workspace", "static"])
def test_cpp_caller_equals_python_mirror(upf, tmp_path, mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run build()")
    n = 128
    out = tmp_path / "mesh.bin"
    args = [BIN, str(out), str(n)] + (["static"] if mode == "static" else [])
    print("case: C++ caller", mode, n, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    pos, nrm, idx = _read(out)
    # the same spec as the caller's makeSpec: sphere 1100 m, ratio 10, 6-octave gradient
    # fBm (frequency 0.0005, lacunarity 2, gain 0.4, weight 1000, seed 1234)
    s = S.field_spec("sphere", (1100.0,), [S.layer("gradient", "fbm", 6, 0.0005, 2.0, 0.4, 1000.0, 1234)],
                     elevation_scale=1.0 / 10)
    m = upf.mesh_generate_generic(s, S.uniform_grid(-2000.0, 2000.0, n))
    assert len(idx) == m.n_triangles > 1000 and len(pos) == m.n_vertices
    np.testing.assert_array_equal(idx, m.indices.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(pos.view(np.int32), m.positions.cpu().numpy().view(np.int32))
    np.testing.assert_array_equal(nrm.view(np.int32), m.normals.cpu().numpy().view(np.int32))
