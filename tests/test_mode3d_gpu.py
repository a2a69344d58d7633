"""GPU parity of the Mode3D mesh field and the mesh-field tracer against the
CPU oracle: the row-stencil field (cMode3DMeshFieldEval::GetB_T,
CutoffRigidityMode3D.cpp:702-800) bit for bit on AMR meshes at points
crowded around block faces and coarse/fine interfaces, and TraceAllowed3D
trajectories (termination, steps, time, path, x, p) bit for bit with the
Mode3D step selector."""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import cutoff as CU
from amps_amd import mesh as M
from oracle import oracle as O
from tests import tracer_setup as TS
from tests.test_trace_gpu import _oparams, _random_tasks, _same

pytestmark = pytest.mark.gpu


def _points(mesh, rng, n, planes):
    x = rng.uniform(mesh.xmin, mesh.xmax, size=(n, 3))
    k = n // 2
    a = rng.integers(0, 3, k)
    x[np.arange(k), a] = rng.choice(planes, k) + rng.uniform(-0.04, 0.04, k)
    return np.clip(x, mesh.xmin, np.nextafter(mesh.xmax, -np.inf) - 1e-12)


@pytest.mark.parametrize("which", ["octant", "three_level"])
def test_mode3d_field_bitexact(gpu_ctx, which):
    mesh = M.central_octant_mesh() if which == "octant" else M.three_level_mesh()
    rng = np.random.default_rng(61)
    Bm = rng.normal(size=(mesh.nblocks, 512, 3))
    print(f"mode3d field case {which} blocks={mesh.nblocks}", flush=True)
    f = CU.Mode3DField(gpu_ctx, mesh, Bm)
    planes = [-0.5, -0.25, 0.0, 0.25, 0.5] if which == "octant" else [-0.3, -0.1875, -0.0625, 0.0, 0.0625, 0.125]
    x = _points(mesh, rng, 60000, planes)
    x[:5] = [[2.0, 0, 0], [0, -1.5, 0], [0, 0, 1.0], [-1.0, -1.0, -1.0], [0.999, 0.999, 0.999]]
    g = f.at(x)
    o = O.Mesh(mesh).mode3d_field(Bm, x)
    bad = np.flatnonzero(np.any(g != o, axis=1))
    assert len(bad) == 0, (len(bad), x[bad[:3]], g[bad[:3]], o[bad[:3]])
    assert np.all(g[:2] == 0.0)


@pytest.mark.parametrize("mover", ["BORIS", "HC4", "RK4"])
def test_mode3d_trace_bitexact(gpu_ctx, mover):
    mesh = TS.geospace_mesh()
    P = CU.mode3d_params(mover=mover)
    P.max_time_s = 30.0 if mover == "BORIS" else 10.0
    Bm = TS.dipole_cells(mesh, P)
    print(f"mode3d trace case {mover} blocks={mesh.nblocks}", flush=True)
    CU.Mode3DField(gpu_ctx, mesh, Bm)
    om = O.Mesh(mesh)
    om.set_trace_field(Bm)
    x, d, R = _random_tasks(600, 17 + CU.MOVERS[mover], rhi=3.0)
    g, steps = CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP64)
    o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    assert steps == int(o["steps"].sum())
    assert {0, 1} <= set(np.unique(o["termination"]).tolist())


def test_mode3d_trace_argument_checks(gpu_ctx):
    mesh = M.uniform_mesh(1, xmin=(-1e8,) * 3, xmax=(1e8,) * 3)
    CU.Mode3DField(gpu_ctx, mesh, np.zeros((mesh.nblocks, 512, 3)))
    P = CU.mode3d_params()
    x, d, R = _random_tasks(4, 1)
    with pytest.raises(RuntimeError, match="FP32"):
        CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP32)
    P.field_model = 2
    with pytest.raises(RuntimeError, match="field_model"):
        CU.trace_batch(gpu_ctx, P, x, d, R)
