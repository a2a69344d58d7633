"""Benchmark of the particle hot path on one MI355X.

Default workload: the full particle step at scale — 2^27 particles (two
species, electrons and ions, mass ratio 100, Maxwellian + drift) on a
3-level AMR mesh (16^3 level-4 root blocks of 8^3 cells refined twice around
the r<0.3 shell and the z=0 slab), each step = Lapenta2017 push (FP64) +
migration (compaction by block/cell) + ECSIM J/mass-matrix deposit (FP64),
captured in a hipGraph.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
PUSH_BYTES = 104          # per particle-step: x,v read (48) + write (48) + ids (8)
DEPOSIT_BYTES_P = 48      # per particle: x, v read
NODE_BYTES = 246 * 8      # per unique node: J (3) + mass matrix (243) written once


def cpu_baseline(target_s=12.0):
    """The CPU oracle (FP64 C restatement, OpenMP) timed on this host's cores
    on a bounded sample of the same per-cell workload: the 2-level AMR mesh
    (120 blocks) with 16+16 particles per cell, push + deposit per step."""
    import numpy as np

    from amps_amd import mesh as M
    from amps_amd.pic import Species
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = M.central_octant_mesh()
    E, B = M.corner_fields(m)
    parts = M.per_cell_particles(m, [16, 16], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    bmax = float(np.max(np.linalg.norm(B, axis=-1)))
    dt = 0.1 / (bmax / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    om = O.Mesh(m)
    noc, keys = om.build_nodes()
    steps, t_push, t_dep = 0, 0.0, 0.0
    p = parts
    while t_push + t_dep < target_s or steps < 2:
        t0 = time.perf_counter()
        p, code = O.push(om, E, B, sp, p, mover=0, bc_mode=1, r_sphere=0.0, nthreads=threads)
        t1 = time.perf_counter()
        O.deposit(om, B, sp, dt, p, noc, len(keys), nthreads=threads)
        t2 = time.perf_counter()
        t_push += t1 - t0
        t_dep += t2 - t1
        steps += 1
    n = parts.n
    return {"value": n * steps / (t_push + t_dep), "unit": "particle-pushes/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} particles (2-level AMR, 120 blocks, 16+16 per cell), {steps} push+deposit "
                      f"steps, {t_push + t_dep:.1f} s (push {t_push:.1f} s, deposit {t_dep:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles-log2", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "gpurun_out", "tuning.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from amps_amd import _lib as L
    from amps_amd import autotune
    from amps_amd import mesh as M
    from amps_amd.pic import PicState, Species

    ctx = L.Context(0)
    arch, ncu, hbm = ctx.device_info()
    mesh = M.three_level_mesh()
    E, B = M.corner_fields(mesh)
    bmax = float(np.max(np.linalg.norm(B, axis=-1)))
    dt = 0.1 / (bmax / 0.01)  # max omega_c dt = 0.1 over corners and species
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    st = PicState(mesh, E, B, sp, ctx=ctx)
    nps = (1 << args.particles_log2) // 2
    st.generate_particles(nps, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=4)
    n = st.particle_count()
    opts = dict(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.0)
    tune = None
    if not args.no_tune:
        os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
        sig = f"{arch}|pic|nb{mesh.nblocks}|n{n}"
        tune, cached = autotune.tune_pic(ctx, st, opts, dt, sig, cache=args.tune_cache)
        tune = {"tile": tune["tile"], "deposit_chunks": tune["deposit_chunks"], "cached": cached}
    # warm-up, then the timed region (graph-captured steps)
    if args.warmup > 0:
        st.run_steps(args.warmup, dt, use_graph=True, **opts)
    ctx.sync()
    t0 = time.perf_counter()
    st.run_steps(args.steps, dt, use_graph=True, **opts)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    n_end = st.particle_count()
    # per-kernel HIP-event durations on the library stream (separate eager steps)
    ctx.set_timing(True)
    ctx.reset_timing()
    st.run_steps(4, dt, use_graph=False, **opts)
    ctx.set_timing(False)
    ms, cnt = ctx.timing()
    avg = {k: ms[i] / max(cnt[i], 1) for k, i in [("push", L.K_PUSH), ("compact", L.K_COMPACT),
                                                  ("deposit_cell", L.K_DEPOSIT_CELL),
                                                  ("deposit_node", L.K_DEPOSIT_NODE)]}
    nb, nnodes = mesh.nblocks, st.nnodes
    push_bytes = n * PUSH_BYTES + nb * 729 * 6 * 8
    dep_bytes = n * DEPOSIT_BYTES_P + nnodes * NODE_BYTES + nb * 729 * 3 * 8
    push_gbps = push_bytes / (avg["push"] * 1e-3) / 1e9
    dep_ms = avg["deposit_cell"] + avg["deposit_node"]
    dep_gbps = dep_bytes / (dep_ms * 1e-3) / 1e9
    pd_gbps = (push_bytes + dep_bytes) / ((avg["push"] + dep_ms) * 1e-3) / 1e9
    copy_bytes = 1 << 30
    copy_ms = ctx.copy_probe(copy_bytes, reps=10)
    copy_gbps = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    if avg["push"] >= dep_ms:
        roof = {"kernel": "k_push_f64<Lapenta2017>", "bound": "hbm", "achieved": round(push_gbps, 1),
                "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(push_gbps / HBM_PEAK_GBPS, 4),
                "traffic": None}
    else:
        roof = {"kernel": "ECSIM deposit (k_deposit_cell + k_deposit_node)", "bound": "hbm",
                "achieved": round(dep_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(dep_gbps / HBM_PEAK_GBPS, 4), "traffic": None}
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "particle-pushes/s (full ECSIM particle step: Lapenta2017 push + migration + J/mass-matrix deposit, FP64)",
        "value": n * args.steps / elapsed,
        "unit": "particle-pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated Maxwellian population, analytic dipole + uniform fields)",
        "config": {"workload": "full particle step at scale: 2^27 particles, 2 species, 3-level AMR "
                               f"({nb} blocks of 8^3 cells), push + migration + deposit, hipGraph",
                   "particles": int(n), "particles_end": int(n_end), "blocks": int(nb),
                   "nodes": int(nnodes), "device": arch, "cus": int(ncu)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_ms": {k: round(v, 4) for k, v in avg.items()},
        "push_GBps": round(push_gbps, 1),
        "deposit_GBps": round(dep_gbps, 1),
        "push_deposit_GBps": round(pd_gbps, 1),
        "push_deposit_frac_of_hbm_peak": round(pd_gbps / HBM_PEAK_GBPS, 4),
        "push_deposit_frac_of_copy": round(pd_gbps / copy_gbps, 4),
        "measured_copy_GBps": round(copy_gbps, 1),
        "autotune": tune,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
