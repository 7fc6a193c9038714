"""CPU checks of bench.py that need no GPU: every name its functions read resolves, and the
per-kernel byte model and rocprofv3 names of the mixed methods are consistent."""
import builtins
import os
import symtable
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _unresolved(src, name):
    top = symtable.symtable(src, name, "exec")
    defined = set(top.get_identifiers()) | set(dir(builtins))
    bad = []

    def walk(t, enclosing):
        local = {s.get_name() for s in t.get_symbols() if s.is_assigned() or s.is_parameter() or s.is_imported()}
        for s in t.get_symbols():
            n = s.get_name()
            if s.is_referenced() and (s.is_global() or s.is_free()) and n not in defined and n not in enclosing:
                bad.append((t.get_name(), n))
        for ch in t.get_children():
            walk(ch, enclosing | local)

    for ch in top.get_children():
        walk(ch, set())
    return bad


def test_bench_names_resolve():
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert _unresolved(src, "bench.py") == []


def test_unresolved_name_is_caught():
    src = "def f():\n    def g():\n        return h()\n    return g\n"
    assert _unresolved(src, "t.py") == [("g", "h")]


def test_mixed_byte_model_and_names():
    sys.path[:0] = [ROOT, os.path.join(ROOT, "ffv-hc_amd")]
    import bench
    for m in range(5, 29):
        by = bench.kernel_bytes_mixed(m)
        assert set(by) == {"bi_p", "bi_q", "bi_s", "bi_t", "bi_xr"}
        # q_ and t_ differ by r0 (hashed in 10); p_ and s_ by the p read
        assert by["bi_q"] == by["bi_t"] - (8 if m == 10 else 0) and by["bi_p"] - by["bi_s"] == 8
        for st in by:
            assert "<" in bench.mixed_prof_name(m, st)
    # method 8: FP64 vectors, FP32 bands: 7 bands x 4 bytes in the GS and matvec kernels
    assert bench.kernel_bytes_mixed(8)["bi_q"] == 8 + 3 + 28 + 16
    assert bench.mixed_prof_name(8, "bi_p") == "bi_precond_gs_kernel<0, float, double, false>"
    assert bench.mixed_prof_name(7, "bi_t") == "bi_matvec_kernel<double, 1, float, false, false, 0>"
    assert bench.prof_name("bi_q", 4) == "bi_matvec_kernel<float, 0, float, false, false, 0>"
    assert bench.mixed_prof_name(10, "bi_xr") == "bi_xr_kernel<double, true, false>"
    assert bench.kernel_bytes_mixed(10)["bi_xr"] == 56
    assert abs(bench.kernel_bytes_mixed(12)["bi_p"] - (24 + 16 + 16 + 3)) < 1e-9
    assert bench.mixed_prof_name(13, "bi_s") == "bi_precond_gsk_kernel<double, 1, 3, false, 0>"
    assert bench.mixed_prof_name(21, "bi_p") == "bi_precond_gsk_kernel<double, 0, 5, false, 0>"
    assert bench.mixed_prof_name(25, "bi_s") == "bi_precond_gsk_kernel<double, 1, 3, true, 0>"
    assert bench.mixed_prof_name(25, "bi_s", True) == "bi_precond_gsk_kernel<double, 1, 3, true, 1>"
    # method 29: method 26's kernels on FP32 vectors
    assert bench.mixed_prof_name(29, "bi_p", True) == "bi_precond_gsk_kernel<float, 0, 4, true, 1>"
    assert bench.mixed_prof_name(29, "bi_t", True) == "bi_matvec_kernel<float, 1, float, false, false, 1>"
    assert bench.mixed_prof_name(29, "bi_xr") == "bi_xr_kernel<float, false, false>"
    b26, b29 = bench.kernel_bytes_mixed(26, True), bench.kernel_bytes_mixed(29, True)
    assert all(b29[k] == b26[k] / 2 for k in b26)  # no per-cell matrix: only the vectors, at half width
    assert bench.mixed_prof_name(21, "bi_q") == "bi_matvec_kernel<double, 0, float, false, true, 0>"
    assert bench.mixed_prof_name(21, "bi_q", True) == "bi_matvec_kernel<double, 0, float, false, true, 1>"
    # class-uniform blocks: no per-cell band or record is read
    assert abs(bench.kernel_bytes_mixed(26, True)["bi_p"] - (24 + 16 + 3)) < 1e-9
    assert abs(bench.kernel_bytes_mixed(26, True)["bi_q"] - (8 + 3 + 8 + 8)) < 1e-9
    # multiplicative methods store p_ / s_ with the coarse correction added: plain matvec and XR
    assert bench.mixed_prof_name(26, "bi_t") == "bi_matvec_kernel<double, 1, float, false, false, 0>"
    assert bench.mixed_prof_name(26, "bi_xr") == "bi_xr_kernel<double, false, false>"
    assert bench.mixed_prof_name(19, "bi_xr") == "bi_xr_kernel<double, false, true>"
    assert len(bench.MIXED_METHODS) == 30


def test_step_iteration_bytes():
    sys.path[:0] = [ROOT, os.path.join(ROOT, "ffv-hc_amd")]
    import bench
    # FP64 Jacobi-BiCGSTAB: the five stages only (51 + 83 + 43 + 83 + 64 bytes per cell)
    assert bench.step_iteration_bytes(None, 1.0) == 324.0
    u, o = bench.step_iteration_bytes(26, 1.0), bench.step_iteration_bytes(26, 0.0)
    assert abs(u - (43 + 27 + 35 + 27 + 64)) < 1e-9 and o > u
    assert abs(bench.step_iteration_bytes(26, 0.5) - 0.5 * (u + o)) < 1e-9


def test_cpu_time_step_baseline_runs_one_step():
    """The bench's CPU baseline of the time step (the oracle's Solver::Update on the single-level flow past a sphere) runs
    without a GPU: one step on 4 threads, its pressure solve reaching 1e-8 in the layout of
    data-ils.txt, with the stage times of Solver::times[] filled."""
    sys.path.insert(0, ROOT)
    import bench
    r = bench.cpu_time_step_baseline(steps=1, threads=4)
    assert r["value"] > 0 and r["cores"] == 4 and r["kind"] == "port"
    line = r["data_ils"][0].split()
    assert line[0] == "0000000000" and int(line[7]) > 100 and float(line[8]) <= 1e-8
    assert all(v > 0 for v in r["times_per_step"][0].values())
