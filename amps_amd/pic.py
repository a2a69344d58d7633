"""Host mirror of the reference's PIC mover / ECSIM deposit interface.

``PicState`` owns one library context holding the mesh, corner fields,
species table and the resident SoA particle population.  Its methods carry
the reference's names:

* ``MoveParticles``  — PIC::Mover::MoveParticles (pic_mover.cpp:580-1088)
  through the per-particle Lapenta2017 (pic_mover_boris.cpp:876-1380) or
  Boris (:126-553) mover.
* ``UpdateJMassMatrix`` — ECSIM::UpdateJMassMatrix
  (pic_field_solver_ecsim.cpp:3244-3995) through ProcessCell (:1882-2438).

All work runs in the HIP library; there is no CPU path here.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import Particles


class Species:
    def __init__(self, charge, mass, dt, weight=None):
        self.charge = np.asarray(charge, np.float64)
        self.mass = np.asarray(mass, np.float64)
        self.dt = np.broadcast_to(np.asarray(dt, np.float64), self.charge.shape).copy()
        self.weight = (np.ones_like(self.charge) if weight is None
                       else np.broadcast_to(np.asarray(weight, np.float64), self.charge.shape).copy())

    @property
    def n(self):
        return len(self.charge)


class PicState:
    def __init__(self, mesh, E_corner, B_corner, species, ctx=None):
        self.ctx = ctx if ctx is not None else L.Context(0)
        self.lib = self.ctx.lib
        self.mesh = mesh
        self.species = species
        d = L.MeshDesc()
        for i in range(3):
            d.xmin[i] = mesh.xmin[i]
            d.xmax[i] = mesh.xmax[i]
        d.lmax = mesh.lmax
        d.nblocks = mesh.nblocks
        lv, ijk = mesh.desc_arrays()
        L.check(self.lib.amps_gpu_set_mesh(self.ctx.h, C.byref(d), L._ptr(lv, C.c_int32),
                                           L._ptr(ijk, C.c_int64)), "set_mesh")
        self.set_fields(E_corner, B_corner)
        s = species
        L.check(self.lib.amps_gpu_set_species(self.ctx.h, s.n, L._ptr(s.charge, C.c_double),
                                              L._ptr(s.mass, C.c_double), L._ptr(s.dt, C.c_double),
                                              L._ptr(s.weight, C.c_double)), "set_species")
        self.nnodes = int(self.lib.amps_gpu_node_count(self.ctx.h))

    def set_fields(self, E_corner, B_corner):
        E = np.ascontiguousarray(E_corner, np.float64)
        B = np.ascontiguousarray(B_corner, np.float64)
        assert E.shape == (self.mesh.nblocks, 729, 3) and B.shape == E.shape
        L.check(self.lib.amps_gpu_set_fields(self.ctx.h, L._ptr(E, C.c_double),
                                             L._ptr(B, C.c_double)), "set_fields")

    def node_map(self):
        noc = np.zeros(self.mesh.nblocks * 729, np.int32)
        keys = np.zeros(self.nnodes, np.int64)
        L.check(self.lib.amps_gpu_get_node_map(self.ctx.h, L._ptr(noc, C.c_int32),
                                               L._ptr(keys, C.c_int64)), "get_node_map")
        return noc, keys

    def set_particles(self, p: Particles, precision=L.FP64):
        L.check(self.lib.amps_gpu_set_particles(
            self.ctx.h, p.n, L._ptr(p.x, C.c_double), L._ptr(p.v, C.c_double),
            L._ptr(p.spec, C.c_int32), L._ptr(p.blk, C.c_int32), L._ptr(p.cell, C.c_int32),
            L._ptr(p.id, C.c_int64), precision), "set_particles")

    def generate_particles(self, n_per_species, vth, drift=(0.0, 0.0, 0.0), seed=1234, precision=L.FP64):
        """Device-side synthetic population (uniform in every cell, Maxwellian)."""
        vth = np.ascontiguousarray(vth, np.float64)
        drift = np.ascontiguousarray(drift, np.float64)
        L.check(self.lib.amps_gpu_generate_particles(self.ctx.h, int(n_per_species), len(vth),
                                                     L._ptr(vth, C.c_double), L._ptr(drift, C.c_double),
                                                     seed, precision), "generate_particles")

    def run_steps(self, nsteps, deposit_dt, mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.0,
                  use_graph=True, tile=0):
        """nsteps of MoveParticles + migration + UpdateJMassMatrix."""
        o = L.PushOpts(mover, bc_mode, 1, tile, r_sphere)
        L.check(self.lib.amps_gpu_run_steps(self.ctx.h, C.byref(o), deposit_dt, int(nsteps),
                                            1 if use_graph else 0), "run_steps")

    def set_center_fields(self, B_center):
        """Cell-centred B with one ghost layer per block, [nb][1000][3]."""
        Bc = np.ascontiguousarray(B_center, np.float64)
        if Bc.shape != (self.mesh.nblocks, 1000, 3):
            raise ValueError("B_center must be [nblocks][1000][3]")
        L.check(self.lib.amps_gpu_set_center_fields(self.ctx.h, L._ptr(Bc, C.c_double)), "set_center_fields")

    def interp_center(self, x, blk):
        """B at points x[3][n] of blocks blk from the cell-centred field."""
        x = np.ascontiguousarray(x, np.float64)
        blk = np.ascontiguousarray(blk, np.int32)
        n = x.shape[1]
        B = np.zeros((3, n))
        L.check(self.lib.amps_gpu_interp_center(self.ctx.h, n, L._ptr(x, C.c_double), L._ptr(blk, C.c_int32),
                                                L._ptr(B, C.c_double)), "interp_center")
        return B

    def particle_count(self):
        n = int(self.lib.amps_gpu_particle_count(self.ctx.h))
        if n < 0:
            L.check(L.AMPS_OK + 3, "particle_count")
        return n

    def get_particles(self):
        n = self.particle_count()
        x = np.zeros((3, n))
        v = np.zeros((3, n))
        spec = np.zeros(n, np.int32)
        blk = np.zeros(n, np.int32)
        cell = np.zeros(n, np.int32)
        pid = np.zeros(n, np.int64)
        code = np.zeros(n, np.int32)
        L.check(self.lib.amps_gpu_get_particles(
            self.ctx.h, L._ptr(x, C.c_double), L._ptr(v, C.c_double), L._ptr(spec, C.c_int32),
            L._ptr(blk, C.c_int32), L._ptr(cell, C.c_int32), L._ptr(pid, C.c_int64),
            L._ptr(code, C.c_int32)), "get_particles")
        return Particles(x, v, spec, blk, cell, pid), code

    def get_restart(self, layout):
        """The particle part of a restart file (SaveParticleDataBlock,
        pic_restart.cpp:312-370): per-block ParticleNumberTable [nb, 512] in
        its [i][j][k] order and the records, bytes [n, record_bytes]."""
        nb = self.mesh.nblocks
        counts = np.zeros((nb, 512), np.int32)
        n = C.c_int64()
        L.check(self.lib.amps_gpu_get_particles_restart(self.ctx.h, C.byref(layout), L._ptr(counts, C.c_int32),
                                                        None, 0, C.byref(n)), "get_particles_restart")
        rec = np.zeros((n.value, layout.record_bytes), np.uint8)
        L.check(self.lib.amps_gpu_get_particles_restart(self.ctx.h, C.byref(layout), L._ptr(counts, C.c_int32),
                                                        rec.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
                "get_particles_restart")
        return counts, rec

    def set_restart(self, layout, counts, records, precision=L.FP64):
        """The inverse (ReadParticleDataBlock, pic_restart.cpp:430-512)."""
        counts = np.ascontiguousarray(counts, np.int32)
        records = np.ascontiguousarray(records, np.uint8)
        L.check(self.lib.amps_gpu_set_particles_restart(self.ctx.h, C.byref(layout), L._ptr(counts, C.c_int32),
                                                        len(records), records.ctypes.data_as(C.c_void_p),
                                                        precision), "set_particles_restart")

    def MoveParticles(self, mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.0,
                      compact=True, tile=0):
        o = L.PushOpts(mover, bc_mode, 1 if compact else 0, tile, r_sphere)
        L.check(self.lib.amps_gpu_move_particles(self.ctx.h, C.byref(o)), "move_particles")

    def get_exits(self, clear=True):
        """The exit list of the USER_FUNCTION boundary mode: the arguments of
        PIC::Mover::ProcessOutsideDomainParticles (pic.h:6043-6044) for every
        particle that left through an external face since the last clear, as
        a dict of arrays: id, x [n, 3] and v [n, 3] on the face, spec, face,
        blk (-1: no block holds the face state)."""
        n = C.c_int64()
        L.check(self.lib.amps_gpu_get_exits(self.ctx.h, 0, C.byref(n), None, None, None, None, None, None, 0),
                "get_exits")
        m = n.value
        out = {"id": np.zeros(m, np.int64), "x": np.zeros((m, 3)), "v": np.zeros((m, 3)),
               "spec": np.zeros(m, np.int32), "face": np.zeros(m, np.int32), "blk": np.zeros(m, np.int32)}
        L.check(self.lib.amps_gpu_get_exits(self.ctx.h, m, C.byref(n), L._ptr(out["id"], C.c_int64),
                                            L._ptr(out["x"], C.c_double), L._ptr(out["v"], C.c_double),
                                            L._ptr(out["spec"], C.c_int32), L._ptr(out["face"], C.c_int32),
                                            L._ptr(out["blk"], C.c_int32), 1 if clear else 0), "get_exits")
        return out

    def compact(self):
        L.check(self.lib.amps_gpu_compact(self.ctx.h), "compact")

    def UpdateJMassMatrix(self, dt, copy_out=True):
        """copy_out: True (J and M), False (neither) or "J" (J only)."""
        if not copy_out:
            L.check(self.lib.amps_gpu_update_j_mass_matrix(self.ctx.h, dt, None, None), "deposit")
            return None, None
        J = np.zeros((self.nnodes, 3))
        if copy_out == "J":
            L.check(self.lib.amps_gpu_update_j_mass_matrix(self.ctx.h, dt, L._ptr(J, C.c_double), None), "deposit")
            return J, None
        M = np.zeros((self.nnodes, 243))
        L.check(self.lib.amps_gpu_update_j_mass_matrix(self.ctx.h, dt, L._ptr(J, C.c_double),
                                                       L._ptr(M, C.c_double)), "deposit")
        return J, M

    def ParticleEnergyCfl(self, dt):
        """Total particle energy and per-species max CFL of UpdateJMassMatrix
        (pic_field_solver_ecsim.cpp:2229-2238, 2355-2358, 3925-3929,
        3972-3987) over the resident particles."""
        e = C.c_double()
        cfl = np.zeros(self.species.n)
        L.check(self.lib.amps_gpu_particle_energy_cfl(self.ctx.h, float(dt), C.byref(e), L._ptr(cfl, C.c_double)),
                "particle_energy_cfl")
        return e.value, cfl

    # ---- the ECSIM field equation over the resident J and M ---------------
    @staticmethod
    def ecsim_params(dt, theta=0.5, cDt=None, E_conv=1.0, B_conv=1.0, length_conv=1.0):
        """amps_ecsim_params; cDt defaults to dt (c = 1, normalized units)."""
        return L.EcsimParams(dt, theta, dt if cDt is None else cDt, E_conv, B_conv, length_conv)

    def EcsimOperator(self, params, x, reps=1):
        """A x of GetStencil_Legacy + UpdateMatrixElement
        (pic_field_solver_ecsim.cpp:740-1228, 6004-6016), [nnodes][3]."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros((self.nnodes, 3))
        L.check(self.lib.amps_gpu_ecsim_operator(self.ctx.h, C.byref(params), L._ptr(x, C.c_double),
                                                 L._ptr(y, C.c_double), int(reps)), "ecsim_operator")
        return y

    def EcsimRhs(self, params, E_n, B_c):
        """UpdateRhs_Legacy (pic_field_solver_ecsim.cpp:6040-6137), [nnodes][3]."""
        E_n = np.ascontiguousarray(E_n, np.float64)
        B_c = np.ascontiguousarray(B_c, np.float64)
        out = np.zeros((self.nnodes, 3))
        L.check(self.lib.amps_gpu_ecsim_rhs(self.ctx.h, C.byref(params), L._ptr(E_n, C.c_double),
                                            L._ptr(B_c, C.c_double), L._ptr(out, C.c_double)), "ecsim_rhs")
        return out

    def EcsimSolve(self, params, rhs, tol=1e-10, max_iter=500, restart=40):
        """GMRES on the device: (dE, iterations, relative residual)."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        dE = np.zeros((self.nnodes, 3))
        it = C.c_int32()
        rr = C.c_double()
        L.check(self.lib.amps_gpu_ecsim_solve(self.ctx.h, C.byref(params), L._ptr(rhs, C.c_double), float(tol),
                                              int(max_iter), int(restart), L._ptr(dE, C.c_double), C.byref(it),
                                              C.byref(rr)), "ecsim_solve")
        return dE, it.value, rr.value

    def EcsimUpdateB(self, params, E_theta, B_c):
        """UpdateB (pic_field_solver_ecsim.cpp:5160-5270), [nblocks * 512][3]."""
        E_theta = np.ascontiguousarray(E_theta, np.float64)
        B_c = np.ascontiguousarray(B_c, np.float64)
        out = np.zeros_like(B_c)
        L.check(self.lib.amps_gpu_ecsim_update_b(self.ctx.h, C.byref(params), L._ptr(E_theta, C.c_double),
                                                 L._ptr(B_c, C.c_double), L._ptr(out, C.c_double)), "ecsim_update_b")
        return out

    def EcsimTimeStep(self, params, E, B_c, tol=1e-10, max_iter=500, restart=40):
        """One field step (ECSIM::TimeStep): returns (E^{n+1}, E^{n+theta},
        B^{n+1}, iterations, relative residual)."""
        E = np.array(E, np.float64, copy=True, order="C")
        B = np.array(B_c, np.float64, copy=True, order="C")
        Eth = np.zeros_like(E)
        it = C.c_int32()
        rr = C.c_double()
        L.check(self.lib.amps_gpu_ecsim_time_step(self.ctx.h, C.byref(params), L._ptr(E, C.c_double),
                                                  L._ptr(Eth, C.c_double), L._ptr(B, C.c_double), float(tol),
                                                  int(max_iter), int(restart), C.byref(it), C.byref(rr)),
                "ecsim_time_step")
        return E, Eth, B, it.value, rr.value

    def MassMatrixApply(self, E_nodes, c1, c2=1.0, reps=1, seed=None):
        """The ECSIM mass-matrix term over the resident M of the last
        UpdateJMassMatrix (UpdateRhs_Legacy, pic_field_solver_ecsim.cpp:
        6116-6124): out[n, c] = sum over the 27 neighbours of
        (E . M[n, slot, c, :]) * c1 * c2, with c1 = -4 pi dt theta and
        c2 = E_conv in the reference.  E_nodes is [nnodes, 3] in node-key
        order; seed, when given, is the row sum the terms are added onto
        (the curl and current parts of UpdateRhs_Legacy).  Single-level
        meshes only, as the reference's solver."""
        E = np.ascontiguousarray(E_nodes, np.float64)
        if E.shape != (self.nnodes, 3):
            raise ValueError("E_nodes must be [nnodes, 3]")
        if seed is None:
            out = np.zeros((self.nnodes, 3))
        else:
            out = np.array(seed, np.float64, copy=True, order="C")
            if out.shape != (self.nnodes, 3):
                raise ValueError("seed must be [nnodes, 3]")
        L.check(self.lib.amps_gpu_ecsim_mass_apply(self.ctx.h, L._ptr(E, C.c_double), float(c1), float(c2),
                                                   L._ptr(out, C.c_double), 0 if seed is None else 1, int(reps)),
                "ecsim_mass_apply")
        return out
