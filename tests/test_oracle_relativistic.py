"""Pins the oracle's relativistic Boris mover (PIC::Mover::Relativistic::Boris,
pic_mover_relativistic_boris.cpp:16-583) with closed forms:

* non-relativistic limit: equals the classic Boris mover;
* uniform B: |u| conserved, rotation by 2 atan(|t|) per sub-step with
  t = q dt B / (2 m gamma) and sub-steps of at most one gyro period
  (the gyro-frequency limit of :113-124);
* uniform E: momentum grows by q E dt exactly (u = gamma v).
"""
import numpy as np

from amps_amd import mesh as M
from amps_amd.pic import Species
from oracle import oracle as O

C_LIGHT = 299792458.0
Q_P = 1.602176565e-19
M_P = 1.67262192e-27


def _uniform_field_mesh(E0, B0):
    mesh = M.uniform_mesh(2, xmin=(-100.0,) * 3, xmax=(100.0,) * 3)
    E = np.zeros((mesh.nblocks, 729, 3))
    B = np.zeros((mesh.nblocks, 729, 3))
    E[...] = E0
    B[...] = B0
    return mesh, E, B


def _one(mesh, x, v):
    x = np.asarray(x, np.float64).reshape(3, 1)
    v = np.asarray(v, np.float64).reshape(3, 1)
    blk = mesh.locate(x.T).astype(np.int32)
    cell = M.cell_of(mesh, x, blk)
    return M.Particles(x, v, np.zeros(1, np.int32), blk, cell)


def test_nonrelativistic_limit_equals_boris():
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    parts = M.uniform_particles(mesh, 20000, 0.01, seed_x=11, seed_v=12)
    dt = 0.1 / np.max(np.linalg.norm(B, axis=-1))  # one sub-step everywhere
    sp = Species([1.0], [1.0], [dt])
    om = O.Mesh(mesh)
    r, rc = O.push(om, E, B, sp, parts, mover=2, bc_mode=1, r_sphere=0.2)
    b, bc = O.push(om, E, B, sp, parts, mover=1, bc_mode=1, r_sphere=0.2)
    assert np.array_equal(rc, bc)
    ok = (rc == 3)
    dv = np.abs(r.v[:, ok] - b.v[:, ok]).max() / np.abs(b.v[:, ok]).max()
    dx = np.abs(r.x[:, ok] - b.x[:, ok]).max()
    assert dv < 1e-13 and dx < 1e-15, (dv, dx)
    assert np.array_equal(r.blk[ok], b.blk[ok]) and np.array_equal(r.cell[ok], b.cell[ok])


def _rotation_angle(q, m, B0, dt, gamma):
    return 2.0 * np.arctan(q * dt * B0 / (2.0 * m * gamma))


def test_uniform_B_rotation_and_subcycling():
    B0 = 2.0
    mesh, E, B = _uniform_field_mesh((0.0, 0.0, 0.0), (0.0, 0.0, B0))
    om = O.Mesh(mesh)
    speed = 0.8 * C_LIGHT
    gamma = 1.0 / np.sqrt(1.0 - 0.8 ** 2)
    period = 2.0 * np.pi * M_P * gamma / (Q_P * B0)  # 1 / GetGyroFrequency
    radius = gamma * M_P * speed / (Q_P * B0)
    assert radius < 3.0
    for frac, subs in ((0.3, [0.3]), (2.5, [1.0, 1.0, 0.5])):
        dt = frac * period
        sp = Species([Q_P], [M_P], [dt])
        p = _one(mesh, (10.0, 10.0, 10.0), (speed, 0.0, 0.0))
        o, code = O.push(om, E, B, sp, p, mover=2, bc_mode=1)
        assert code[0] == 3
        v = o.v[:, 0]
        assert abs(np.linalg.norm(v) / speed - 1.0) < 1e-14
        # positive charge, B along +z: clockwise gyration seen from +z
        ang = sum(_rotation_angle(Q_P, M_P, B0, s * period, gamma) for s in subs)
        want = speed * np.array([np.cos(ang), -np.sin(ang), 0.0])
        assert np.abs(v - want).max() / speed < 1e-12, (frac, v, want)
        # position: chords of the sub-steps with the post-rotation velocity
        x = np.array([10.0, 10.0, 10.0])
        a = 0.0
        for s in subs:
            a += _rotation_angle(Q_P, M_P, B0, s * period, gamma)
            x = x + speed * np.array([np.cos(a), -np.sin(a), 0.0]) * (s * period)
        assert np.abs(o.x[:, 0] - x).max() < 1e-12 * 100.0


def test_uniform_E_momentum_gain():
    E0 = 1.0e6
    mesh, E, B = _uniform_field_mesh((E0, 0.0, 0.0), (0.0, 0.0, 0.0))
    om = O.Mesh(mesh)
    v0 = 0.5 * C_LIGHT
    g0 = 1.0 / np.sqrt(1.0 - 0.25)
    dt = 1e-7
    sp = Species([Q_P], [M_P], [dt])
    p = _one(mesh, (-50.0, 0.0, 0.0), (v0, 0.0, 0.0))
    o, code = O.push(om, E, B, sp, p, mover=2, bc_mode=1)
    assert code[0] == 3
    u = g0 * v0 + Q_P * E0 * dt / M_P  # u = gamma v = p / m
    v = u / np.sqrt(1.0 + (u / C_LIGHT) ** 2)
    assert abs(o.v[0, 0] / v - 1.0) < 1e-14
    assert abs(o.x[0, 0] - (-50.0 + v * dt)) < 1e-12 * 50.0
    assert o.v[1, 0] == 0.0 and o.v[2, 0] == 0.0
