"""Host logic of the benchmark, no GPU: how the counter passes attribute
rocprofv3's per-dispatch rows to the report's kernel keys, and how the copy
probe calibrates them into bytes per launch."""
import csv
import os
import subprocess
import sys

import bench

FIELDS = ["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp"]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))
    return str(path)


def test_counter_rows_attributed_to_report_keys(tmp_path):
    """Kernel names as rocprofv3 writes them (template arguments, argument
    lists, anonymous namespaces); a step kernel counts by its last dispatch
    after the copy probe (the set-up deposit before it does not leak in),
    any other by its average; names that merely start like a listed kernel,
    other counters and unlisted kernels are left out; rows may come from
    several files."""
    fetch = "FETCH_SIZE"
    first = [
        ("void (anonymous namespace)::k_deposit_cell_mfma<false, false, false, false>(amps::MeshView, double)",
         fetch, 100.0, 10),
        ("(anonymous namespace)::k_ecsim_mass(double const*, int const*, double const*, double)", fetch, 40.0, 20),
        ("(anonymous namespace)::k_ecsim_mass(double const*, int const*, double const*, double)", fetch, 60.0, 25),
        ("k_copy_probe(HIP_vector_type<float, 4u> const*, HIP_vector_type<float, 4u>*, long)", fetch, 2048.0, 30),
        ("void k_push_f64<0, false, 1>(amps::MeshView, amps::Faces, amps::Species)", fetch, 7.0, 40),
        ("k_mig_order_slab(long, int, int, int const*, int const*)", fetch, 3.0, 41),
        ("k_mig_dest(int const*, int const*, int const*, int const*, unsigned long const*, long, int*)",
         fetch, 5.0, 42),
        ("void (anonymous namespace)::k_deposit_cell_mfma<false, false, true, false>(amps::MeshView, double)",
         fetch, 9.0, 43),
    ]
    second = [
        ("void k_push_f64<0, false, 1>(amps::MeshView, amps::Faces, amps::Species)", fetch, 8.0, 50),
        ("k_mig_order_slab(long, int, int, int const*, int const*)", fetch, 4.0, 51),
        ("k_mig_order_big(long, int, int const*, int const*, int*, int*, int*, int const*)", fetch, 999.0, 52),
        ("void k_push_f64<0, false, 1>(amps::MeshView, amps::Faces, amps::Species)", "WRITE_SIZE", 555.0, 53),
        ("k_mig_dest(int const*, int const*, int const*, int const*, unsigned long const*, long, int*)",
         fetch, 6.0, 54),
        ("void (anonymous namespace)::k_deposit_cell_mfma<false, false, true, false>(amps::MeshView, double)",
         fetch, 10.0, 55),
        ("k_deposit_node_staged(long, int, int const*, int const*, double const*, double*)", fetch, 12.0, 56),
        ("k_mig_move(int const*, int const*, long, double const*, double const*, int)", fetch, 11.0, 60),
    ]
    files = [_write(tmp_path / "a_counter_collection.csv", first),
             _write(tmp_path / "b_counter_collection.csv", second)]
    v = bench._pmc_values(files, fetch)
    assert v == {"ecsim_mass": 50.0, "copy": 2048.0, "push": 8.0, "mig_order": 4.0, "mig_dest": 6.0,
                 "deposit_cell": 10.0, "deposit_node": 12.0, "mig_move": 11.0}


def test_traffic_bytes_calibrated_by_the_copy_probe():
    """A fetch counter reading half the copy probe's bytes is scaled by 2,
    a write counter reading them exactly by 1; a key missing from either
    pass is left out."""
    kib = bench.COPY_PROBE_BYTES / 1024.0
    per = {"FETCH_SIZE": {"copy": kib / 2, "push": 1000.0, "mig_dest": 10.0},
           "WRITE_SIZE": {"copy": kib, "push": 500.0}}
    out = bench._traffic_bytes(per)
    assert out["calibration"] == {"fetch": 2.0, "write": 1.0}
    assert out["push"] == 1024 * (1000 * 2 + 500)
    assert out["push_fetch"] == 1024 * 1000 * 2
    assert set(out) == {"calibration", "push", "push_fetch"}


def test_more_than_one_gpu_exits_with_a_message():
    """The benchmark measures one GPU: asked for more it exits with status 2
    and says so, before touching a device."""
    r = subprocess.run([sys.executable, os.path.abspath(bench.__file__), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
    assert r.returncode == 2 and r.stdout == b""
    assert b"one GPU only" in r.stderr and b"--gpus 2" in r.stderr
