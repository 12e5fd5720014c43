"""bench.py bookkeeping, on the CPU: the roofline object of the dominant
kernel takes its HBM traffic from the committed rocprofv3 summary, keyed by
the kernel's name as the profiler prints it."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_lookup_for_every_kernel_the_bench_can_report():
    b = _bench()
    # the dominant kernels of the Ax+gs step: in-kernel geometry on the
    # matrix cores / on the vector ALUs, and the stored-geometry kernel of
    # the operator built from Neko's arrays
    for code in (300, 201, 4):
        name = b.ax_kernel_name(code, 8)
        traffic = b.measured_traffic(name)
        assert isinstance(traffic, int) and traffic > 0, name
        # the in-kernel geometry kernels move their algorithmic bytes
        alg = b.ax_bytes_per_dof(8, code) * 64 ** 3 * 8 ** 3
        assert 0.95 < traffic / alg < 1.05, (name, traffic, alg)


def test_roofline_entry_labels_the_bound():
    b = _bench()
    n = 64 ** 3 * 8 ** 3
    r = b.roofline_entry(300, 8, 2900.0, 31.6, b.ax_bytes_per_dof(8, 300) * n, 0.76, n)
    assert r["bound"] == "fp64" and r["unit"] == "TFLOP/s"
    assert r["matrix_flops_per_dof"] == 96 and r["traffic"] is not None
    r = b.roofline_entry(201, 8, 2100.0, 22.0, b.ax_bytes_per_dof(8, 201) * n, 1.07, n)
    assert r["bound"] == "fp64-valu"
    r = b.roofline_entry(4, 8, 5900.0, 10.0, b.ax_bytes_per_dof(8, 4) * n, 1.46, n)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s"
