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
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
PUSH_BYTES = 104          # per particle-step: x,v read (48) + write (48) + ids (8)
PUSH_BYTES_F32 = 80       # FP32 velocities: x read+write (48) + v read+write (24) + ids (8)
DEPOSIT_BYTES_P = 48      # per particle: x, v read
# a step's push also makes the previous migration's record move: per particle
# it reads the source index (4), x, v (48), species (1) and id (8), and writes
# x, v (48), species (1), id (8), key (4) and code (1)
PUSH_MOVE_BYTES = 61 + 62
PUSH_MOVE_BYTES_F32 = 49 + 50   # the same with FP32 velocities
NODE_BYTES = 246 * 8      # per unique node: J (3) + mass matrix (243) written once
ECSIM_NODE_BYTES = 243 * 8 + 27 * 4 + 24 + 24  # mass-term: M row, neighbour ids, own E once, result


COPY_PROBE_BYTES = 1 << 30
# kernel identifiers of the counter passes -> report keys
PMC_KERNELS = (("k_push_f64", "push"), ("k_push_f32", "push"), ("k_deposit_cell", "deposit_cell"),
               ("k_deposit_cell_rec", "deposit_cell"), ("k_deposit_cell_mfma", "deposit_cell"),
               ("k_deposit_node", "deposit_node"), ("k_deposit_node_staged", "deposit_node"),
               ("k_scatter", "scatter"), ("k_mig_scatter", "scatter"), ("k_mig_dest", "mig_dest"),
               ("k_mig_move", "mig_move"),
               ("k_mig_classify", "mig_classify"),
               ("k_mig_order", "mig_order"), ("k_mig_order_slab", "mig_order"), ("k_copy_probe", "copy"),
               ("k_ecsim_mass", "ecsim_mass"), ("k_ecsim_op", "ecsim_op"),
               ("k_deposit_cell_lanes", "deposit_cell"))
# kernels of the particle step: counted only after the copy probe, which the
# counter child launches between the ECSIM set-up (a deposit of its own) and
# the step
STEP_KEYS = ("push", "deposit_cell", "deposit_node", "scatter", "mig_dest", "mig_move", "mig_classify",
             "mig_order")


def _pmc_pass(args, tune, counter, outdir):
    """One rocprofv3 counter pass over two eager steps of the same workload
    (a child process; the counter in KiB per dispatch: the second step's
    dispatch for the step's kernels, the average for the others)."""
    import glob
    import subprocess

    cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc", counter, "--output-format", "csv",
           "-d", outdir, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--pmc-child",
           "--workload", "pic", "--particles-log2", str(args.particles_log2), "--tile", str(tune["tile"]),
           "--deposit-kernel", str(tune["deposit_kernel"]), "--deposit-chunks", str(tune["deposit_chunks"]),
           "--node-batch", str(tune.get("node_batch", args.node_batch)),
           "--push-waves", str(tune.get("push_waves", args.push_waves)),
           "--node-staged", str(args.node_staged), "--deposit-paired", str(args.deposit_paired),
           "--deterministic", str(args.deterministic)]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 {counter} pass: exit {r.returncode}, no counter file; "
                           + r.stdout.decode(errors="replace")[-400:])
    return _pmc_values(files, counter)


def _pmc_values(files, counter):
    """The counter per report key (PMC_KERNELS) from rocprofv3 counter CSV
    files, in KiB per dispatch: for the step's kernels (STEP_KEYS) their last
    dispatch after the last copy-probe dispatch (the counter child's second
    step), for the others the average over their dispatches."""
    import csv

    rows = []
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                for ident, key in PMC_KERNELS:
                    if re.search(r"\b" + ident + r"\b", name):
                        rows.append((int(row["Start_Timestamp"]), key, float(row["Counter_Value"])))
                        break
    marker = max((t for t, k, _ in rows if k == "copy"), default=0)
    sums, last = {}, {}
    for t, key, val in sorted(rows):
        if key in STEP_KEYS:
            # the step kernels of the second step (its push moves the records
            # the first step's migration left)
            if t >= marker:
                last[key] = val
            continue
        tot, cnt = sums.get(key, (0.0, 0))
        sums[key] = (tot + val, cnt + 1)
    out = {k: tot / cnt for k, (tot, cnt) in sums.items()}
    out.update(last)
    return out


def ecsim_mass_setup():
    """A second context holding the single-level mesh of the ECSIM field
    solve (16^3 level-4 blocks, 2.1 M cells, 16 + 16 particles per cell) and
    the M of one deposit on it."""
    import numpy as np

    from amps_amd import _lib as L
    from amps_amd import mesh as M
    from amps_amd.pic import PicState, Species

    mesh = M.uniform_mesh(4)
    E, B = M.corner_fields(mesh)
    dt = 0.1 / (float(np.max(np.linalg.norm(B, axis=-1))) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    c2 = L.Context(0)
    st = PicState(mesh, E, B, sp, ctx=c2)
    st.generate_particles(mesh.ncells * 16, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=5)
    st.UpdateJMassMatrix(dt, copy_out=False)
    En = np.random.default_rng(7).normal(size=(st.nnodes, 3))
    return c2, st, En, dt


ECSIM_OP_NODE_BYTES = 243 * 8 + 27 * 4 + 27 + 24 + 24  # M row, neighbours, their boundary flags, own x, result


def ecsim_field_bench(ctx2, st2, En, dt):
    """The ECSIM field equation on the same single-level mesh and resident
    J/M: the operator's time per launch (HIP events) against its algorithmic
    bytes, and one field step (right-hand side, GMRES to 1e-8, E and B
    updates) timed whole, with c dt theta / dx = 1."""
    import numpy as np

    from amps_amd import _lib as L

    dx = float(st2.mesh.cell_size()[0][0])
    p = st2.ecsim_params(dt, theta=0.5, cDt=2 * dx)
    x = np.random.default_rng(3).normal(size=(st2.nnodes, 3))
    st2.EcsimOperator(p, x)  # tables and work space
    ctx2.set_timing(True)
    ctx2.reset_timing()
    st2.EcsimOperator(p, x, reps=20)
    ctx2.set_timing(False)
    ms, cnt = ctx2.timing()
    op_ms = ms[L.K_ECSIM] / max(cnt[L.K_ECSIM], 1)
    op_bytes = st2.nnodes * ECSIM_OP_NODE_BYTES
    Bc = np.random.default_rng(4).normal(scale=1e-3, size=(st2.mesh.nblocks * 512, 3))
    st2.EcsimTimeStep(p, En, Bc, tol=1e-8, max_iter=400, restart=40)
    ctx2.sync()
    t0 = time.perf_counter()
    _, _, _, it, rr = st2.EcsimTimeStep(p, En, Bc, tol=1e-8, max_iter=400, restart=40)
    t_step = time.perf_counter() - t0
    return {"kernel": "k_ecsim_op", "operator_ms": round(op_ms, 4), "operator_GBps": round(op_bytes / (op_ms * 1e-3) / 1e9, 1),
            "operator_frac_of_hbm_peak": round(op_bytes / (op_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "operator_algorithmic_bytes": int(op_bytes), "step_s": round(t_step, 4), "gmres_iterations": int(it),
            "relative_residual": rr, "c_dt_theta_over_dx": 1.0,
            "note": "field step = host arrays in and out (PCIe included), right-hand side, GMRES(40) to 1e-8, "
                    "E and B updates"}


def pmc_traffic(args, tune):
    """HBM bytes per launch of the step's kernels from FETCH_SIZE and
    WRITE_SIZE, each scaled by the factor that makes the copy kernel of known
    size read and write exactly its bytes."""
    import shutil
    import tempfile

    try:
        per = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = tempfile.mkdtemp(prefix="amps_pmc_", dir="/tmp")
            try:
                per[counter] = _pmc_pass(args, tune, counter, d)
            finally:
                shutil.rmtree(d, ignore_errors=True)
        return _traffic_bytes(per)
    except Exception as e:  # the line stays valid without the counter passes
        return {"error": repr(e)[:300]}


def _traffic_bytes(per):
    """Bytes per launch from the two counter passes ({"FETCH_SIZE": {key:
    KiB}, "WRITE_SIZE": {...}}): each counter scaled by the factor that makes
    the copy probe (COPY_PROBE_BYTES read, the same written) read exactly
    its bytes; total and fetch bytes per key present in both passes."""
    kib = COPY_PROBE_BYTES / 1024.0
    f_fetch = kib / per["FETCH_SIZE"]["copy"]
    f_write = kib / per["WRITE_SIZE"]["copy"]
    out = {"calibration": {"fetch": round(f_fetch, 4), "write": round(f_write, 4)}}
    for key in per["FETCH_SIZE"]:
        if key == "copy" or key not in per["WRITE_SIZE"]:
            continue
        out[key] = int(1024 * (per["FETCH_SIZE"][key] * f_fetch + per["WRITE_SIZE"][key] * f_write))
        out[key + "_fetch"] = int(1024 * per["FETCH_SIZE"][key] * f_fetch)
    return out


def cpu_baseline(target_s=10.0, nblk=256):
    """The CPU oracle (FP64 C restatement built -O3 with AVX2, OpenMP) timed on
    this host on a bounded sample of the headline step: the same 3-level AMR
    mesh (7456 blocks), fields, species, time step and particle density (18 +
    17 per cell, as 2^27 particles give) in nblk blocks spread over the mesh.
    A step is the Lapenta push with reflection (particles in parallel) and the
    deposit in the reference's parity-set arrangement (the cells of a set in
    parallel, each added straight into its corners).  Timed on all host cores
    and on one core, about target_s each; the migration is the deposit's
    counting sort of the particles by cell, inside the deposit time."""
    import numpy as np

    from amps_amd import mesh as M
    from amps_amd.pic import Species
    from oracle import oracle as O

    m = M.three_level_mesh()
    E, B = M.corner_fields(m)
    bmax = float(np.max(np.linalg.norm(B, axis=-1)))
    dt = 0.1 / (bmax / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    om = O.Mesh(m)
    noc, keys = om.build_nodes()
    blocks = np.unique(np.linspace(0, m.nblocks - 1, nblk).astype(np.int32))
    parts0 = M.per_cell_particles(m, [18, 17], [0.05, 0.005], drift=(0.01, 0.0, 0.0), blocks=blocks)
    out = (np.empty((len(keys), 3)), np.empty((len(keys), 243)))
    code = np.zeros(parts0.n, np.int32)
    host_cores = os.cpu_count() or 1
    all_threads = int(os.environ.get("OMP_NUM_THREADS", host_cores))

    def run(threads, target):
        p = parts0.copy()
        steps, t_push, t_dep = 0, 0.0, 0.0
        while t_push + t_dep < target or steps < 2:
            t0 = time.perf_counter()
            O.push(om, E, B, sp, p, mover=0, bc_mode=1, r_sphere=0.0, nthreads=threads, copy=False, code=code)
            t1 = time.perf_counter()
            O.deposit_colored(om, B, sp, dt, p, noc, len(keys), nthreads=threads, out=out)
            t2 = time.perf_counter()
            t_push += t1 - t0
            t_dep += t2 - t1
            steps += 1
        return p.n * steps / (t_push + t_dep), steps, t_push, t_dep

    v_all, s_all, tp_all, td_all = run(all_threads, target_s)
    v_one, s_one, tp_one, td_one = run(1, target_s)
    n = parts0.n
    return {"value": v_all, "unit": "particle-pushes/s", "cores": all_threads, "kind": "port",
            "value_1core": v_one, "host_cores": host_cores,
            "sample": f"{n} particles (35 per cell in {len(blocks)} of the 7456 blocks of the 3-level mesh), "
                      f"push + parity-set deposit: {all_threads} threads {s_all} steps in "
                      f"{tp_all + td_all:.1f} s (push {tp_all:.1f} s, deposit {td_all:.1f} s); 1 thread "
                      f"{s_one} steps in {tp_one + td_one:.1f} s (push {tp_one:.1f} s, deposit {td_one:.1f} s)"}


MLP_FWD_FLOP = 2 * (6 * 128 + 128 * 128 + 128)  # per sample, 6 -> 128 -> 128 -> 1
BF16_PEAK_TFLOPS = 2500.0                          # MI355X dense BF16 matrix peak


def mlp_bench(ctx, arch, args):
    """The surrogate network alone: Adam training steps (batch 4096) and
    batched inference on 2^20 synthetic samples, timed with the device
    synchronised on both sides.  FLOPs count the network's multiply-adds
    (forward 34.6 kFLOP per sample, training 3x that: forward, input and
    weight gradients)."""
    import numpy as np

    from amps_amd import surrogate as SG

    rng = np.random.default_rng(3)
    n = 1 << args.train_samples_log2
    lat, lon = rng.uniform(-90, 90, n), rng.uniform(-180, 180, n)
    zen, az = rng.uniform(0, 60, n), rng.uniform(0, 360, n)
    X = SG.features(lat, lon, zen, az)
    T = (np.cos(np.radians(lat)) ** 4 * (1.0 + 0.1 * np.sin(np.radians(lon))) + 0.05).astype(np.float32)
    sur = SG.Surrogate(ctx, seed=1)
    sur.set_data(X, T)
    batch = 4096
    sur.train(max(args.warmup, 1) * 16, batch=batch)
    ctx.sync()
    nsteps = 256 * args.steps
    t0 = time.perf_counter()
    sur.train(nsteps, batch=batch)
    ctx.sync()
    t_train = time.perf_counter() - t0
    sur.predict(X[:4096])
    ctx.sync()
    t0 = time.perf_counter()
    reps = 4
    for _ in range(reps):
        sur.predict(X)
    ctx.sync()
    t_inf = (time.perf_counter() - t0) / reps
    train_tflops = 3 * MLP_FWD_FLOP * batch * nsteps / t_train / 1e12
    inf_tflops = MLP_FWD_FLOP * n / t_inf / 1e12
    return {"metric": "surrogate MLP training samples/s (BF16 MFMA, FP32 accumulate, Adam)",
            "value": batch * nsteps / t_train, "unit": "samples/s", "n_gpus": 1, "steps": nsteps,
            "warmup": max(args.warmup, 1) * 16, "ms_per_step": 1e3 * t_train / nsteps, "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "bf16", "data": "synthetic features and targets",
            "config": {"workload": "MLP 6-128-128-1 GELU, batch 4096, 2^%d samples" % args.train_samples_log2,
                       "device": arch},
            "roofline": {"kernel": "k_mlp_train + k_mlp_adam (training step)", "bound": "mfma",
                         "achieved": round(train_tflops, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(train_tflops / BF16_PEAK_TFLOPS, 6), "traffic": None},
            "inference": {"samples": n, "s": t_inf, "samples_per_s": n / t_inf, "TFLOPs": round(inf_tflops, 3),
                          "frac_of_bf16_peak": round(inf_tflops / BF16_PEAK_TFLOPS, 6)},
            "train_step_us": 1e6 * t_train / nsteps}


def scan_bench(ctx, arch, args):
    """The plain full-grid cutoff map (64x32 points x 8 directions x 48
    rigidities, every bin traced) once in FP64 and once in FP32, after one
    small warm-up scan: the tracer kernel's own rate, for profiles."""
    import numpy as np

    from amps_amd import _lib as L
    from amps_amd import cutoff as CU

    P = CU.config2_params()
    pts, lat, lon = CU.config2_points()
    zen, az = CU.config2_directions(8)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    CU.cutoff_scan(ctx, P, x0[::256], dirs[::256], grid, mode=CU.SCAN_FULL, precision=L.FP32)
    out = {}
    for prec, tag in ((L.FP64, "fp64"), (L.FP32, "fp32")):
        ctx.sync()
        t0 = time.perf_counter()
        _, _, st = CU.cutoff_scan(ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=prec)
        t = time.perf_counter() - t0
        out[tag] = {"s": t, "steps": int(st), "steps_per_s": st / t}
    return {"metric": "cutoff-map Boris steps/s, plain full scan (FP64)", "value": out["fp64"]["steps_per_s"],
            "unit": "steps/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * out["fp64"]["s"],
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "64x32 lat/lon x 8 directions x 48 rigidities, every bin traced, 1e5-step cap",
                       "device": arch},
            "scan": out}


def cutoff_map_bench(ctx, n_cu, n_train=1 << 20, epochs=20, cpu_check_points=16, tune_cache=None, n_batch=16):
    """Batched backward cutoff map (64x32 points x 8 directions x 48 bins)
    and the surrogate-bracketed map.  Returns a dict of timings and errors."""
    import numpy as np

    from amps_amd import _lib as L
    from amps_amd import autotune
    from amps_amd import cutoff as CU
    from amps_amd import surrogate as SG

    out = {}
    P = CU.config2_params()
    pts, lat, lon = CU.config2_points()
    zen, az = CU.config2_directions(8)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    flat_lat, flat_lon = np.repeat(lat, len(zen)), np.repeat(lon, len(zen))
    flat_zen, flat_az = np.tile(zen, len(pts)), np.tile(az, len(pts))
    grid = CU.build_search_grid(0.1, 20.0, 48)
    nloc = len(x0)
    sub = slice(0, nloc, 64)

    def run_sub():
        CU.cutoff_scan(ctx, P, x0[sub], dirs[sub], grid, mode=CU.SCAN_FULL, precision=L.FP32)

    tune, cached = autotune.tune_tracer(ctx, run_sub, n_cu, f"trace|{n_cu}", cache=tune_cache)
    out["tracer_tune"] = {"workgroups": tune["workgroups"], "chunk": tune["chunk"], "cached": cached}
    # FP64 map (bit-exact with the CPU oracle): the reference map of this run
    t0 = time.perf_counter()
    s64, b64, st64 = CU.cutoff_scan(ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    out["plain_full_fp64_s"] = time.perf_counter() - t0
    out["plain_full_fp64_steps"] = int(st64)
    out["plain_full_fp64_steps_per_s"] = st64 / out["plain_full_fp64_s"]
    # the same FP64 map from start points moved by 1e-12 (relative): how far
    # the map itself moves under a perturbation of rounding size
    _, bp, _ = CU.cutoff_scan(ctx, P, x0 * (1.0 + 1e-12), dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    out["fp64_perturbed_1e-12_bins_off_gt1"] = float(np.mean(np.abs(bp.astype(int) - b64.astype(int)) > 1))
    # plain FP32 map, every bin
    t0 = time.perf_counter()
    s32, b32, st32 = CU.cutoff_scan(ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP32)
    out["plain_full_fp32_s"] = time.perf_counter() - t0
    out["plain_full_fp32_steps"] = int(st32)
    out["plain_full_fp32_steps_per_s"] = st32 / out["plain_full_fp32_s"]
    out["trajectories_full"] = int(nloc * len(grid))
    out["fp32_states_equal_fp64"] = float(np.mean(s32 == s64))
    out["fp32_bins_off_gt1"] = float(np.mean(np.abs(b32.astype(int) - b64.astype(int)) > 1))
    # CPU oracle on a subsample of map points (every bin)
    if cpu_check_points > 0:
        from oracle import oracle as O
        Q = O.TraceParams()
        for name, _t in L.TraceParams._fields_:
            v = getattr(P, name)
            if hasattr(v, "__len__"):
                for i in range(len(v)):
                    getattr(Q, name)[i] = v[i]
            else:
                setattr(Q, name, v)
        idx = np.linspace(0, nloc - 1, cpu_check_points).astype(int)
        cpu_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.perf_counter()
        r = O.trace_batch(Q, np.repeat(x0[idx], 48, axis=0), np.repeat(dirs[idx], 48, axis=0),
                          np.tile(grid, len(idx)), nthreads=cpu_threads)
        t_cpu = time.perf_counter() - t0
        ost = (r["termination"] == 0).astype(np.int8).reshape(-1, 48)
        cpu_steps = int(r["steps"].sum())
        # the same FP64 scan on the device runs every bin of every point: its
        # per-step rate against the CPU oracle's on this subsample
        out["cpu_oracle_check"] = {"points": int(len(idx)), "states_bitexact": bool(np.array_equal(ost, s64[idx])),
                                   "threads": cpu_threads, "s": t_cpu, "trajectories": int(len(r)),
                                   "steps": cpu_steps, "steps_per_s": cpu_steps / t_cpu,
                                   "gpu_fp64_steps_per_s_ratio": out["plain_full_fp64_steps_per_s"]
                                   / (cpu_steps / t_cpu)}
    # surrogate: training set from the tracer itself (UPPER_SCAN, FP32)
    t0 = time.perf_counter()
    Xtr, Ttr, tb, tsteps = SG.training_set(ctx, P, n_train, grid, seed=6, precision=L.FP32)
    out["train_data_s"] = time.perf_counter() - t0
    out["train_samples"] = int(n_train)
    sur = SG.Surrogate(ctx, seed=0)
    sur.set_data(Xtr, Ttr)
    nsteps = max(1, epochs * n_train // 4096)
    t0 = time.perf_counter()
    losses = sur.train(nsteps, batch=4096, lr=1e-3, record_loss=False)
    ctx.sync()
    out["train_s"] = time.perf_counter() - t0
    out["train_steps"] = int(nsteps)
    g, tl0 = sur.grad(4096)
    out["train_loss_end"] = float(tl0)
    # surrogate-bracketed maps (predict + bracket scan + full-scan fallback)
    feats = SG.features(flat_lat, flat_lon, flat_zen, flat_az)
    for prec, tag, plain in ((L.FP64, "fp64", "plain_full_fp64_s"), (L.FP32, "fp32", "plain_full_fp32_s")):
        t0 = time.perf_counter()
        bs, fb, pb, sst = SG.bracketed_map(ctx, sur, P, x0, dirs, feats, grid, half=3, precision=prec)
        dt_s = time.perf_counter() - t0
        # the same map with the fallback as a second launch after the bracket scan
        t0 = time.perf_counter()
        bs_2, _, _, sst_2 = SG.bracketed_map(ctx, sur, P, x0, dirs, feats, grid, half=3, precision=prec,
                                             one_launch=False)
        dt_2 = time.perf_counter() - t0
        out[f"surrogate_{tag}"] = {
            "map_s": dt_s, "steps": int(sst), "speedup_vs_plain_full": out[plain] / dt_s,
            "two_launch_map_s": dt_2, "two_launch_steps": int(sst_2),
            "two_launch_bins_equal": bool(np.array_equal(bs, bs_2)),
            "fallback_frac": float(np.mean(fb)),
            "bins_off_gt1_vs_fp64": float(np.mean(np.abs(bs.astype(int) - b64.astype(int)) > 1))}
    out["surrogate_pred_bin_mae"] = float(np.mean(np.abs(pb.astype(int) - np.where(b64 < 0, 47, b64))))
    # many maps per launch: the same grid with n_batch direction sets (seeds
    # 5, 6, ...), e.g. the maps of a time series; the scans are then
    # throughput-bound instead of bound by the longest single trajectory
    if n_batch > 1:
        xs, ds, fs = [], [], []
        for i in range(n_batch):
            z_i, a_i = CU.config2_directions(8, seed=5 + i)
            ds.append(CU.arrival_dirs(pts, z_i, a_i).reshape(-1, 3))
            xs.append(np.repeat(pts, len(z_i), axis=0))
            fs.append(SG.features(np.repeat(lat, len(z_i)), np.repeat(lon, len(z_i)),
                                  np.tile(z_i, len(pts)), np.tile(a_i, len(pts))))
        xb, db, fb_ = np.concatenate(xs), np.concatenate(ds), np.concatenate(fs)
        bt = {"maps": n_batch, "pairs": int(len(xb))}
        ref_bins = None
        for prec, tag in ((L.FP64, "fp64"), (L.FP32, "fp32")):
            t0 = time.perf_counter()
            _, bpl, _ = CU.cutoff_scan(ctx, P, xb, db, grid, mode=CU.SCAN_FULL, precision=prec)
            t_plain = time.perf_counter() - t0
            if ref_bins is None:
                ref_bins = bpl
            t0 = time.perf_counter()
            bs2, fb2, _, _ = SG.bracketed_map(ctx, sur, P, xb, db, fb_, grid, half=3, precision=prec)
            t_sur = time.perf_counter() - t0
            bt[tag] = {"plain_full_s": t_plain, "surrogate_s": t_sur, "speedup": t_plain / t_sur,
                       "fallback_frac": float(np.mean(fb2)),
                       "bins_off_gt1_vs_fp64": float(np.mean(np.abs(bs2.astype(int) - ref_bins.astype(int)) > 1))}
        out["batch"] = bt
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles-log2", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of tuning results to reuse (default: tune on every run)")
    ap.add_argument("--workload", choices=["all", "pic", "map", "mlp", "scan"], default="all",
                    help="scan: the plain cutoff-map scans alone (FP64, FP32), for tracer profiles")
    ap.add_argument("--train-samples-log2", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--tile", type=int, default=0, help="push tile (skips the tuner with the two below)")
    ap.add_argument("--deposit-kernel", type=int, default=-1)
    ap.add_argument("--deposit-chunks", type=int, default=0)
    ap.add_argument("--node-batch", type=int, default=0, help="nodes per wavefront in the node gather")
    ap.add_argument("--push-waves", type=int, default=0, help="push register target: 1 or 4 waves per SIMD")
    ap.add_argument("--node-staged", type=int, default=-1, help="1: staged node gather, 0: per-value gather")
    ap.add_argument("--deposit-paired", type=int, default=-1, help="matrix-core cell pass: 1 paired phases, 0 not")
    ap.add_argument("--deterministic", type=int, default=1,
                    help="1: migration layout independent of atomic order (default), 0: atomic order")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
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
    if args.tune_cache:
        os.makedirs(os.path.dirname(os.path.abspath(args.tune_cache)), exist_ok=True)
    if args.workload == "map":
        t0 = time.perf_counter()
        mp = cutoff_map_bench(ctx, ncu, n_train=1 << args.train_samples_log2, epochs=args.epochs,
                              tune_cache=args.tune_cache)
        t_map = mp["surrogate_fp64"]["map_s"]
        line = {"metric": "cutoff-rigidity map time, surrogate-bracketed FP64 scan (vs plain full GPU scan in cutoff_map)",
                "value": t_map, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
                "ms_per_step": 1e3 * t_map, "higher_is_better": False, "scaling": "strong",
                "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                "config": {"workload": "64x32 lat/lon x 8 directions x 48 rigidities 0.1-20 GV, tilted "
                                       "dipole + Bz -5 nT, 0.01 gyro-period Boris steps, 1e5-step cap",
                           "device": arch},
                "cutoff_map": mp, "wall_s": time.perf_counter() - t0}
        print(json.dumps(line), flush=True)
        ctx.close()
        return
    if args.workload == "mlp":
        print(json.dumps(mlp_bench(ctx, arch, args)), flush=True)
        ctx.close()
        return
    if args.workload == "scan":
        print(json.dumps(scan_bench(ctx, arch, args)), flush=True)
        ctx.close()
        return
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
    fixed = args.tile > 0 and args.deposit_kernel >= 0 and args.deposit_chunks > 0
    if args.node_batch > 0:
        ctx.set_param("node_batch", args.node_batch)
    ctx.set_param("deterministic", args.deterministic)
    if args.push_waves > 0:
        ctx.set_param("push_waves", args.push_waves)
    if args.node_staged >= 0:
        ctx.set_param("node_staged", args.node_staged)
    if args.deposit_paired >= 0:
        ctx.set_param("deposit_paired", args.deposit_paired)
    if fixed:
        ctx.set_param("tile", args.tile)
        autotune.set_deposit_variant(ctx, args.deposit_kernel, args.deposit_chunks)
        tune = {"tile": args.tile, "deposit_kernel": args.deposit_kernel,
                "deposit_chunks": args.deposit_chunks, "cached": "given"}
    if args.pmc_child:
        # the ECSIM kernels on their own mesh, the copy kernel of known size
        # (calibration, and the marker after which the step's kernels count),
        # then two eager steps, for the counter passes
        c2, st2, En2, dt2 = ecsim_mass_setup()
        st2.MassMatrixApply(En2, -4 * np.pi * dt2 * 0.5)
        st2.EcsimOperator(st2.ecsim_params(dt2, cDt=2 * float(st2.mesh.cell_size()[0][0])), En2)
        c2.sync()
        c2.close()
        ctx.copy_probe(COPY_PROBE_BYTES, reps=1)
        ctx.sync()
        st.run_steps(2, dt, use_graph=False, **opts)
        ctx.sync()
        ctx.close()
        return
    if not args.no_tune and not fixed:
        sig = f"{arch}|pic|nb{mesh.nblocks}|n{n}"
        tune, cached = autotune.tune_pic(ctx, st, opts, dt, sig, cache=args.tune_cache)
        tune = {k: tune[k] for k in ("tile", "push_waves", "deposit_kernel", "deposit_chunks", "node_batch",
                                     "push_ms", "push_waves_ms", "deposit_ms", "step_ms", "node_batch_ms")
                if k in tune}
        tune["cached"] = cached
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
    # with the deterministic migration and the single-phase matrix-core or
    # the lane cell pass the push moves the migrated records (the cell pass
    # reads them where they were): 3 of the 4 timed pushes, all of the
    # hipGraph steps'
    dep_kernel_used = tune["deposit_kernel"] if tune else (args.deposit_kernel if args.deposit_kernel >= 0 else 2)
    fused = args.deterministic == 1 and (dep_kernel_used == 3 or (dep_kernel_used == 2 and args.deposit_paired != 1))
    push_bytes = n * (PUSH_MOVE_BYTES if fused else PUSH_BYTES) + nb * 729 * 6 * 8
    dep_bytes = n * DEPOSIT_BYTES_P + nnodes * NODE_BYTES + nb * 729 * 3 * 8
    push_gbps = push_bytes / (avg["push"] * 1e-3) / 1e9
    dep_ms = avg["deposit_cell"] + avg["deposit_node"]
    dep_gbps = dep_bytes / (dep_ms * 1e-3) / 1e9
    pd_gbps = (push_bytes + dep_bytes) / ((avg["push"] + dep_ms) * 1e-3) / 1e9
    copy_bytes = COPY_PROBE_BYTES
    copy_ms = ctx.copy_probe(copy_bytes, reps=10)
    copy_gbps = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    traffic = None
    if not args.no_traffic:
        if tune is None:
            tune_used = {"tile": 8192, "deposit_kernel": 1, "deposit_chunks": 12}
        else:
            tune_used = tune
        traffic = pmc_traffic(args, tune_used)
    if avg["push"] >= dep_ms:
        roof = {"kernel": ("k_push_f64<Lapenta2017> with the migration's record move" if fused else
                           "k_push_f64<Lapenta2017>"), "bound": "hbm", "achieved": round(push_gbps, 1),
                "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(push_gbps / HBM_PEAK_GBPS, 4),
                "traffic": traffic["push"] if traffic and "push" in traffic else None,
                "algorithmic_bytes": int(push_bytes)}
    else:
        tr = None
        if traffic and "deposit_cell" in traffic and "deposit_node" in traffic:
            tr = traffic["deposit_cell"] + traffic["deposit_node"]
        roof = {"kernel": "ECSIM deposit (deposit cell pass + node gather, one launch each)", "bound": "hbm",
                "achieved": round(dep_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(dep_gbps / HBM_PEAK_GBPS, 4), "traffic": tr,
                "algorithmic_bytes": int(dep_bytes)}
    # the ECSIM mass-matrix term over the resident M of a deposit on the
    # single-level mesh the reference's field solver runs on (27-neighbour
    # stencil, 243 entries per node): HIP-event time per launch
    ecsim = None
    try:
        c2, st2, En, dt2 = ecsim_mass_setup()
        st2.MassMatrixApply(En, -4 * np.pi * dt2 * 0.5)  # builds the neighbour table once
        c2.set_timing(True)
        c2.reset_timing()
        st2.MassMatrixApply(En, -4 * np.pi * dt2 * 0.5, reps=20)
        c2.set_timing(False)
        mse, cnte = c2.timing()
        e_ms = mse[L.K_ECSIM] / max(cnte[L.K_ECSIM], 1)
        e_nodes = st2.nnodes
        e_bytes = e_nodes * ECSIM_NODE_BYTES
        ecsim = {"kernel": "k_ecsim_mass", "mesh": "uniform level 4 (16^3 blocks of 8^3 cells), 16+16 per cell",
                 "ms": round(e_ms, 4), "nodes": int(e_nodes),
                 "traffic": traffic.get("ecsim_mass") if traffic else None,
                 "algorithmic_bytes": int(e_bytes), "GBps": round(e_bytes / (e_ms * 1e-3) / 1e9, 1),
                 "frac_of_hbm_peak": round(e_bytes / (e_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
        ecsim["field_solve"] = ecsim_field_bench(c2, st2, En, dt2)
        c2.close()
    except Exception as e:  # the particle-step line stays valid
        ecsim = {"error": repr(e)}
    # the same step with the FP32 push (FP32 stencil, fields and velocities,
    # FP64 positions and FP64 deposit): same population, regenerated
    st.generate_particles(nps, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=4, precision=L.FP32)
    if args.warmup > 0:
        st.run_steps(args.warmup, dt, use_graph=True, **opts)
    ctx.sync()
    t0 = time.perf_counter()
    st.run_steps(args.steps, dt, use_graph=True, **opts)
    ctx.sync()
    el32 = time.perf_counter() - t0
    ctx.set_timing(True)
    ctx.reset_timing()
    st.run_steps(4, dt, use_graph=False, **opts)
    ctx.set_timing(False)
    ms32, cnt32 = ctx.timing()
    avg32 = {k: ms32[i] / max(cnt32[i], 1) for k, i in [("push", L.K_PUSH), ("compact", L.K_COMPACT),
                                                        ("deposit_cell", L.K_DEPOSIT_CELL),
                                                        ("deposit_node", L.K_DEPOSIT_NODE)]}
    push32_bytes = n * (PUSH_MOVE_BYTES_F32 if fused else PUSH_BYTES_F32) + nb * 729 * 6 * 8
    fp32_variant = {"value": n * args.steps / el32, "unit": "particle-pushes/s", "ms_per_step": 1e3 * el32 / args.steps,
                    "kernel_ms": {k: round(v, 4) for k, v in avg32.items()},
                    "push_GBps": round(push32_bytes / (avg32["push"] * 1e-3) / 1e9, 1)}
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
        "traffic_bytes_per_launch": traffic,
        "fp32_push_variant": fp32_variant,
        "ecsim_mass_term": ecsim,
    }
    if args.workload == "all":
        try:
            line["cutoff_map"] = cutoff_map_bench(ctx, ncu, n_train=1 << args.train_samples_log2,
                                                  epochs=args.epochs, tune_cache=args.tune_cache)
        except Exception as e:  # the PIC line stays valid if the map part fails
            line["cutoff_map"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
