"""Benchmark of the OpenStaPLE fermion hot path on one MI355X.

Workload (32^4, the lattice the headline metric is quoted on): random SU(3)
links with eps = 0.5 (generate_Conf_cold rule, seed 1) smeared by 2 stout
steps with rho = 0.15, staggered phases with antiperiodic time, Gaussian
source.  A timed step is one application of the FP64 staggered Dslash pair
(D_oe then D_eo, the two hops of M = m^2 - D_eo D_oe) through the C entry
points on HBM-resident fields, queued back to back (asynchronous entry mode)
between two stream synchronisations; `value` is the Dslash bandwidth in GB/s
under the 928 B / output half-site algorithmic model (8 links x 2 stored rows
x 3 complex x 16 B + 8 phases x 8 B + 48 B in + 48 B out).

After the timed steps, solves to ||r||/||b|| < 1e-10 (FP64 true residual):
(a) the rational multi-shift CG-M over the 19 shifts of the order-19 x^(-1/4)
approximation rescaled to [m^2, 64], m = 0.01, in FP64, and (b) its FP32-inner
variant (FP32 CG-M + per-shift FP32 CG with FP64 reliable updates), on the
field above; (c) plain FP64 CG and FP32-inner CG with FP64 reliable updates at
m = 0.005 on a 2-step stout-smeared eps = 0.7 field (seed 2).  Wall times are
reported in "solves".  A bounded sample of the same Dslash is timed on the
host with the CPU oracle (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solves] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Dslash GB/s; s per 32^4 solve to 1e-10 rel. residual, FP64 vs mixed/learned"
BYTES_PER_SITE_F64 = 928
BYTES_PER_SITE_F32 = 464
HBM_PEAK_GBS = 8000.0


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", type=str, default="32,32,32,32")
    ap.add_argument("--no-solves", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--learned-pairs", type=int, default=512,
                    help="(b, x) pairs from plain FP64 CG to train the preconditioner on (0: skip)")
    ap.add_argument("--learned-steps", type=int, default=300)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solve workload")
    ap.add_argument("--iters-per-graph", type=int, default=16,
                    help="iterations captured per HIP graph in the batched driver and the library's own batched "
                         "solves (0: direct launches, for profilers)")
    ap.add_argument("--no-rhmc", action="store_true", help="skip the RHMC trajectory workload")
    ap.add_argument("--rhmc-only", action="store_true", help="run only the RHMC trajectory workload")
    ap.add_argument("--rhmc-lattice", type=str, default="48,48,48,12")
    ap.add_argument("--rhmc-steps", type=int, default=10, help="Omelyan steps per trajectory")
    ap.add_argument("--rhmc-thermalise", type=int, default=12,
                    help="gauge-only HMC trajectories before the 2+1-flavour ones")
    ap.add_argument("--rhmc-tau", type=float, default=1.0, help="trajectory length")
    ap.add_argument("--rhmc-therm-tau", type=float, default=0.25,
                    help="trajectory length of the gauge-only thermalisation")
    ap.add_argument("--rhmc-warmup", type=int, default=1, help="untimed 2+1-flavour trajectories before the timed one")
    ap.add_argument("--rhmc-no-mixed", action="store_true",
                    help="skip the second timed trajectory with reduced-precision force solves")
    ap.add_argument("--dslash-only", action="store_true",
                    help="only the timed Dslash loop and the copy kernel (for counter runs)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; sharding over GPUs (t-slabs with a "
                                   "one-site spinor halo per Dslash) is not built", "n_gpus": args.gpus}))
        sys.exit(2)
    # the CPU baseline's builds run make: before this process touches the GPU
    cpu_libs = None
    if not args.no_cpu and not args.rhmc_only and not args.dslash_only:
        sys.path.insert(0, str(ROOT / "oracle"))
        import tempfile

        import cpu_baseline as cpub

        cpu_libs = cpub.build(tempfile.mkdtemp(prefix="ostp_cpu_baseline_"))
    import openstaple_amd as ost
    from openstaple_amd import lib

    if args.rhmc_only:
        print(json.dumps({"rhmc": run_rhmc(ost, lib, args)}))
        return
    shape = tuple(int(x) for x in args.lattice.split(","))
    L = ost.Lattice(*shape)
    L.init()
    # the library's own batched solves (rational-solve refinement) follow the same setting
    lib.ostp_set_batched_graph_iterations(args.iters_per_graph)
    sizeh = L.sizeh
    vb, gb, pb = L.vec3_bytes, 8 * L.su3_bytes, 8 * 8 * sizeh

    def dmalloc(n):
        p = lib.ostp_malloc(n)
        if not p:
            raise MemoryError(lib.ostp_last_error())
        lib.ostp_memset(p, 0, n)
        return p

    ph = dmalloc(pb)
    b = dmalloc(vb)
    h = dmalloc(vb)
    out = dmalloc(vb)
    lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, 1)
    u = smeared_field(lib, dmalloc, gb, 0.5, 1)
    lib.calc_u1_phases(ph, ost.BfParam(0, 0, 0, 0, 0, 0), 0.0, 0.0)
    lib.ostp_generate_vec3_soa_gauss(b, 1, 0)
    ost.check(0, "setup")

    # ---------------- timed region: K Dslash pairs (Doe then Deo), HIP events on the library stream
    prev = lib.ostp_set_sync_at_return(0)
    for _ in range(args.warmup):
        lib.acc_Doe(u, h, b, ph)
        lib.acc_Deo(u, out, h, ph)
    lib.ostp_synchronize()
    t0 = time.perf_counter()
    lib.ostp_timer_start()
    for _ in range(args.steps):
        lib.acc_Doe(u, h, b, ph)
        lib.acc_Deo(u, out, h, ph)
    lib.ostp_timer_stop()
    lib.ostp_synchronize()
    t1 = time.perf_counter()
    us_events = lib.ostp_timer_elapsed_us()
    lib.ostp_set_sync_at_return(prev)
    ost.check(0, "dslash loop")
    ms_per_step = 1000.0 * (t1 - t0) / args.steps
    alg_bytes_step = 2 * BYTES_PER_SITE_F64 * sizeh
    value = alg_bytes_step / (ms_per_step * 1e-3) / 1e9
    # the dominant kernel's mean launch duration over the timed region (events bracket the 2K launches)
    us_launch = us_events / (2 * args.steps)
    achieved = BYTES_PER_SITE_F64 * sizeh / (us_launch * 1e-6) / 1e9
    # the same kernel launched alone, back to back (Deo, then Doe)
    reps = max(50, args.steps)
    us_single = 0.5 * (lib.ostp_time_dslash(u, out, h, ph, 0, 0, reps) + lib.ostp_time_dslash(u, h, b, ph, 1, 0, reps))

    # HBM ceilings of the box: 1 GiB streaming read / copy / write, best of a small sweep of launch shapes
    cbytes = 1 << 30
    ca, cb = dmalloc(cbytes), dmalloc(cbytes)
    ceilings = {}
    for kind, name in ((1, "read"), (0, "copy"), (2, "write")):
        best = (0.0, "")
        for unroll in (2, 4, 8):
            for grid in (4096, 8192, 16384):
                us = lib.ostp_time_stream(kind, cb, ca, cbytes, 10, grid, unroll)
                gbs = (2 if kind == 0 else 1) * cbytes / (us * 1e-6) / 1e9
                best = max(best, (gbs, f"{unroll} loads/thread, {grid} blocks"))
        ceilings[name] = {"GBs": round(best[0], 1), "shape": best[1]}
    lib.ostp_free(ca)
    lib.ostp_free(cb)
    ost.check(0, "stream probes")

    # FP32 Dslash
    uf = dmalloc(8 * L.su3f_bytes)
    phf = dmalloc(4 * 8 * sizeh)
    bf = dmalloc(vb // 2)
    hf = dmalloc(vb // 2)
    lib.ostp_convert_double_to_float_gauge(u, uf)
    lib.ostp_phases_to_float(ph, phf)
    lib.convert_double_to_float_vec3_soa(b, bf)
    us_f32 = 0.5 * (lib.ostp_time_dslash(uf, hf, bf, phf, 1, 1, reps) + lib.ostp_time_dslash(uf, bf, hf, phf, 0, 1, reps))
    f32_gbs = BYTES_PER_SITE_F32 * sizeh / (us_f32 * 1e-6) / 1e9
    ost.check(0, "timing")

    # HBM bytes per launch: not measured by this script (the counters need a
    # profiler run of their own); read from the committed profile summary
    traffic = None
    prof = ROOT / "profiles" / f"dslash_fp64_{'x'.join(map(str, shape))}_traffic.json"
    if prof.exists():
        try:
            traffic = json.loads(prof.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{'x'.join(map(str, shape))} staggered Dslash pair (Doe+Deo) FP64 on HBM-resident "
                        "fields; random SU(3) eps=0.5 seed 1, 2 stout steps rho=0.15; then 19-shift rational "
                        "CG-M (m=0.01) and CG at m=0.005 (eps=0.7 smeared) to 1e-10, FP64 vs FP32-inner",
            "lattice": list(shape),
            "sizeh": sizeh,
            "bytes_per_half_site": BYTES_PER_SITE_F64,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (f"{prof.relative_to(ROOT)} (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy "
                               "kernel's counters; not measured in this run)") if traffic else None,
            "kernel": "k_dslash<double> (Doe/Deo, mean over the timed region)",
            "us_per_launch": round(us_launch, 3),
            "us_per_launch_alone": round(us_single, 3),
            "ceilings_GBs": ceilings,
            "frac_of_read_ceiling": round(achieved / ceilings["read"]["GBs"], 4),
            "fp32_dslash_GBs": round(f32_gbs, 1),
            "fp32_us_per_launch": round(us_f32, 3),
        },
    }

    if args.dslash_only:
        print(json.dumps(result))
        return
    if not args.no_solves:
        result["solves"] = run_solves(ost, lib, L, u, ph, b, uf, phf, args)
    if not args.no_batch:
        result["batched"] = run_batched(ost, lib, args)
    if not args.no_rhmc:
        result["rhmc"] = run_rhmc(ost, lib, args)

    if cpu_libs is not None:
        result["cpu_baseline"] = cpu_baseline(cpu_libs, shape, bench_shifts(ost), args.cpu_seconds)

    print(json.dumps(result))


def smeared_field(lib, dmalloc, gb, eps, seed):
    """generate_Conf_cold rule with `eps`, then the configured stout steps
    (stout_wrapper); returns the device pointer of the last level."""
    raw = dmalloc(gb)
    lib.ostp_generate_Conf_cold(raw, eps, seed)
    steps = lib.ostp_stout_steps()
    arr = dmalloc(steps * gb)
    lib.stout_wrapper(raw, arr, 0)
    lib.ostp_free(raw)
    return arr + (steps - 1) * gb  # the smeared levels stay allocated for the whole run


def run_solves(ost, lib, L, u, ph, b, uf, phf, args):
    """Wall time of each solve to ||b - M x|| / ||b|| < 1e-10 (FP64 true residual): the median of three
    runs after any warm-up, the three times listed beside it."""
    sizeh = L.sizeh
    vb = L.vec3_bytes
    mass = 0.01
    ra = ost.read_remez(str(ROOT / "tests" / "golden" / "approx_-1_over_4_order_19_mloglm_6.4.REMEZ"))
    lmax = 64.0
    eps = lmax ** (ra.exponent_num / ra.exponent_den)
    ra.RA_a0 *= eps
    for k in range(ra.approx_order):
        ra.RA_a[k] *= lmax * eps
        ra.RA_b[k] *= lmax
    order = ra.approx_order
    fp = ost.ferm_param(mass, ph, phf)

    def dm(n):
        p = lib.ostp_malloc(n)
        lib.ostp_memset(p, 0, n)
        return p

    d64 = [dm(vb) for _ in range(4)]
    d32 = [dm(vb // 2) for _ in range(5)]
    out = dm(order * vb)
    sf = dm(order * vb)
    sf32 = dm(order * vb // 2)
    x = dm(vb)
    x64 = dm(vb)      # FP64 CG solution: the reference point of the reduced-precision solutions
    out64 = dm(order * vb)
    diff = dm(vb)
    res = {}

    def rel_err(a, ref):
        """||a - ref|| / ||ref|| (FP64, library BLAS)."""
        lib.combine_in1_minus_in2(a, ref, diff)
        return (lib.l2norm2_global(diff) / lib.l2norm2_global(ref)) ** 0.5

    def timed(run, reset=None, reps=3):
        """Median wall time of `reps` calls of run() (each after reset()), the
        stream drained before and after each; returns (median, all, last result)."""
        ts, r = [], None
        for _ in range(reps):
            if reset:
                reset()
            lib.ostp_synchronize()
            t0 = time.perf_counter()
            r = run()
            lib.ostp_synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2], [round(t, 4) for t in ts], r

    # (a) FP64 CG-M
    cg = C.c_int(0)
    sres = (C.c_double * 25)()
    tm, tall, st = timed(lambda: lib.ostp_multishift_solve(u, C.byref(fp), C.byref(ra), out, b, 1e-10, *d64, sf,
                                                           20000, 0, C.byref(cg), sres))
    res["multishift_fp64"] = {"seconds": round(tm, 4), "seconds_runs": tall, "iterations": cg.value, "status": st,
                              "max_shift_true_residual": max(sres[k] for k in range(order)), "shifts": order}
    for k in range(order):
        lib.assign_in_to_out(out + k * vb, out64 + k * vb)

    # (b) FP32-inner: FP32 CG-M then per-shift FP32 CG with FP64 reliable updates
    ip = ost.InverterPackage()
    ip.u, ip.u_f = u, uf
    ip.ferm_shift_temp, ip.ferm_shift_temp_f, ip.nshifts = sf, sf32, order
    ip.loc_r, ip.loc_h, ip.loc_s, ip.loc_p = d64
    ip.loc_r_f, ip.loc_h_f, ip.loc_s_f, ip.loc_p_f, ip.out_f = d32
    lib.ostp_set_inverter_tricks(1, 1, 0.1, 10000)
    lib.ostp_synchronize()
    t0 = time.perf_counter()
    it = lib.inverter_multishift_wrapper(ip, C.byref(fp), C.byref(ra), out, b, 1e-10, 20000, 0)
    lib.ostp_synchronize()
    t1 = time.perf_counter()
    lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
    # true residual of every shift, FP64
    worst = 0.0
    s_tmp, h_tmp = d64[2], d64[1]
    bn = lib.l2norm2_global(b)
    for k in range(order):
        xk = out + k * vb
        lib.fermion_matrix_multiplication_shifted(u, s_tmp, xk, h_tmp, C.byref(fp), ra.RA_b[k])
        lib.combine_in1_minus_in2(b, s_tmp, h_tmp)
        worst = max(worst, (lib.l2norm2_global(h_tmp) / bn) ** 0.5)
    res["multishift_mixed"] = {"seconds": round(t1 - t0, 4), "wrapper_iterations": it,
                               "max_shift_true_residual": worst,
                               "max_shift_err_vs_fp64": max(rel_err(out + k * vb, out64 + k * vb)
                                                            for k in range(order))}
    res["multishift_speedup_mixed_vs_fp64"] = round(res["multishift_fp64"]["seconds"] / (t1 - t0), 3)

    # (b') FP32 CG-M with FP64 reliable updates of the base residual, then the
    # shifts still above 1e-10 refined together (batched, FP16 inner)
    cgm = C.c_int(0)
    cgr, trs_ms = (C.c_int * order)(), (C.c_double * order)()
    lib.ostp_multishift_mixed_solve(ip, C.byref(fp), C.byref(ra), out, b, 1e-10, 20000, 1, C.byref(cgm), cgr,
                                    trs_ms)  # warm-up: scratch, autotune
    tm, tall, conv = timed(lambda: lib.ostp_multishift_mixed_solve(ip, C.byref(fp), C.byref(ra), out, b, 1e-10, 20000,
                                                                   1, C.byref(cgm), cgr, trs_ms))
    t0, t1 = 0.0, tm
    res["multishift_mixed_ru"] = {"seconds": round(t1 - t0, 4), "seconds_runs": tall, "cgm_iterations": cgm.value,
                                  "refine_iterations": [cgr[k] for k in range(order)],
                                  "converged_shifts": conv, "max_shift_true_residual": max(trs_ms[k] for k in range(order)),
                                  "max_shift_err_vs_fp64": max(rel_err(out + k * vb, out64 + k * vb)
                                                               for k in range(order))}
    res["multishift_speedup_mixed_ru_vs_fp64"] = round(res["multishift_fp64"]["seconds"] / (t1 - t0), 3)
    # the same two at the MD force's tolerance, 1e-8
    f64_8, _, _ = timed(lambda: lib.ostp_multishift_solve(u, C.byref(fp), C.byref(ra), out, b, 1e-8, *d64, sf, 20000,
                                                          0, C.byref(cg), sres))
    tm, _, conv8 = timed(lambda: lib.ostp_multishift_mixed_solve(ip, C.byref(fp), C.byref(ra), out, b, 1e-8, 20000, 1,
                                                                 C.byref(cgm), cgr, trs_ms))
    t0, t1 = 0.0, tm
    res["multishift_at_1e-8"] = {"fp64_seconds": round(f64_8, 4), "fp64_iterations": cg.value,
                                 "mixed_ru_seconds": round(t1 - t0, 4), "mixed_ru_cgm_iterations": cgm.value,
                                 "mixed_ru_refine_iterations": [cgr[k] for k in range(order)],
                                 "mixed_ru_converged_shifts": conv8,
                                 "speedup_mixed_ru_vs_fp64": round(f64_8 / (t1 - t0), 3)}

    # (c) plain FP64 CG and FP32-inner CG at m = 0.005 on a 2-step stout-smeared
    # eps = 0.7 field (seed 2), the same source
    gb = 8 * L.su3_bytes
    u4 = smeared_field(lib, dm, gb, 0.7, 2)
    u4f = dm(8 * L.su3f_bytes)
    lib.ostp_convert_double_to_float_gauge(u4, u4f)
    m4 = 0.005
    fp4 = ost.ferm_param(m4, ph, phf)
    ip.u, ip.u_f = u4, u4f
    tr = C.c_double(0)
    zero_x = lambda: lib.ostp_memset(x, 0, vb)
    tm, tall, st = timed(lambda: lib.ostp_cg_solve(u4, C.byref(fp4), x, b, 1e-10, *d64, 40000, 0.0, 10000,
                                                   C.byref(cg), C.byref(tr)), zero_x)
    t0, t1 = 0.0, tm
    res["cg_fp64"] = {"seconds": round(t1 - t0, 4), "seconds_runs": tall, "iterations": cg.value,
                      "true_residual": tr.value,
                      "status": st, "mass": m4, "field": "eps=0.7 seed 2, 2 stout steps rho=0.15"}
    lib.assign_in_to_out(x, x64)
    ru = C.c_int(0)
    tm, tall, st = timed(lambda: lib.ostp_mixed_solve(ip, C.byref(fp4), x, b, 1e-10, 40000, 0.0, 0.1, C.byref(cg),
                                                      C.byref(ru), C.byref(tr)), zero_x)
    t0, t1 = 0.0, tm
    res["cg_mixed"] = {"seconds": round(t1 - t0, 4), "seconds_runs": tall, "iterations": cg.value,
                       "reliable_updates": ru.value,
                       "true_residual": tr.value, "status": st, "err_vs_fp64": rel_err(x, x64)}
    res["cg_speedup_mixed_vs_fp64"] = round(res["cg_fp64"]["seconds"] / (t1 - t0), 3)

    # the same mixed-precision algorithm as one system of the batched solver: its
    # iterations captured in a HIP graph, scalars on the device
    sysd = ost.BatchSystem()
    sysd.nd[:] = L.nd
    sysd.par_temp = 0
    sysd.u, sysd.u_f, sysd.phases, sysd.phases_f = u4, u4f, ph, phf
    sysd.in_, sysd.out, sysd.mass = b, x, m4
    arr = (ost.BatchSystem * 1)(sysd)
    cgs, rus, trs = (C.c_int * 1)(), (C.c_int * 1)(), (C.c_double * 1)()
    lib.ostp_memset(x, 0, vb)
    lib.ostp_batched_mixed_solve(1, arr, 1e-10, 40000, 0.1, args.iters_per_graph, cgs, rus, trs)  # warm-up: autotune, graph
    tm, tall, conv = timed(lambda: lib.ostp_batched_mixed_solve(1, arr, 1e-10, 40000, 0.1, args.iters_per_graph, cgs,
                                                                rus, trs), zero_x)
    t0, t1 = 0.0, tm
    res["cg_mixed_graph"] = {"seconds": round(t1 - t0, 4), "seconds_runs": tall, "iterations": cgs[0],
                             "reliable_updates": rus[0],
                             "true_residual": trs[0], "status": int(conv == 1), "err_vs_fp64": rel_err(x, x64)}
    res["cg_speedup_mixed_graph_vs_fp64"] = round(res["cg_fp64"]["seconds"] / (t1 - t0), 3)
    # FP16 storage in the inner solver (same graph-captured driver)
    lib.ostp_memset(x, 0, vb)
    lib.ostp_batched_half_solve(1, arr, 1e-10, 40000, 0.1, args.iters_per_graph, cgs, rus, trs)  # warm-up
    tm, tall, conv = timed(lambda: lib.ostp_batched_half_solve(1, arr, 1e-10, 40000, 0.1, args.iters_per_graph, cgs,
                                                               rus, trs), zero_x)
    t0, t1 = 0.0, tm
    res["cg_half_graph"] = {"seconds": round(t1 - t0, 4), "seconds_runs": tall, "iterations": cgs[0],
                            "reliable_updates": rus[0],
                            "true_residual": trs[0], "status": int(conv == 1), "err_vs_fp64": rel_err(x, x64)}
    res["cg_speedup_half_graph_vs_fp64"] = round(res["cg_fp64"]["seconds"] / (t1 - t0), 3)

    # learned preconditioner: trained on (b, x) pairs from the plain FP64 CG with
    # the library's kernels, used in flexible CG (FP64 true residual, CG fallback)
    if args.learned_pairs > 0:
        from openstaple_amd import learned
        op = learned.Operator(u4, u4f, fp4)
        t0 = time.perf_counter()
        pairs = learned.Pairs(op, vb, args.learned_pairs, seed=1234)
        t1 = time.perf_counter()
        params, hist = learned.train(op, pairs, learned.init_params(), steps=args.learned_steps, lr=0.02, batch=8)
        t2 = time.perf_counter()
        pairs.free()
        lib.ostp_memset(x, 0, vb)
        learned.solve(op, params, x, b, 1e-10)  # warm-up (scratch allocation)
        tm, tall, (st, its, fb, trr) = timed(lambda: learned.solve(op, params, x, b, 1e-10), zero_x)
        t3, t4 = 0.0, tm
        res["cg_learned_fcg"] = {"seconds": round(t4 - t3, 4), "seconds_runs": tall, "iterations": its,
                                 "fallback_iterations": fb,
                                 "true_residual": trr, "status": st, "err_vs_fp64": rel_err(x, x64),
                                 "training": {"pairs": args.learned_pairs, "pair_seconds": round(t1 - t0, 2),
                                              "steps": args.learned_steps, "train_seconds": round(t2 - t1, 2),
                                              "loss_first": round(float(hist[:5].mean()), 4),
                                              "loss_last": round(float(hist[-20:].mean()), 4),
                                              "loss": "A-norm of N b - x over <x, b>"}}
        res["cg_speedup_learned_vs_fp64"] = round(res["cg_fp64"]["seconds"] / (t4 - t3), 3)
        res["cg_iteration_ratio_learned_vs_fp64"] = round(res["cg_fp64"]["iterations"] / max(1, its + fb), 3)
        # the same network inside the FP32 flexible CG, corrected in FP64
        lib.ostp_memset(x, 0, vb)
        learned.solve_mixed(op, params, x, b, 1e-10)  # warm-up (scratch allocation)
        tm, tall, (st, its, outer, fb, trr) = timed(lambda: learned.solve_mixed(op, params, x, b, 1e-10), zero_x)
        t3, t4 = 0.0, tm
        res["cg_learned_fcg_fp32"] = {"seconds": round(t4 - t3, 4), "seconds_runs": tall, "iterations": its,
                                      "reliable_updates": outer,
                                      "update_every": 0.1, "fallback_iterations": fb, "true_residual": trr,
                                      "status": st, "err_vs_fp64": rel_err(x, x64)}
        res["cg_speedup_learned_fp32_vs_fp64"] = round(res["cg_fp64"]["seconds"] / (t4 - t3), 3)
    ost.check(0, "solves")
    for p in [*d64, *d32, out, sf, sf32, x, x64, out64, diff, u4f]:
        lib.ostp_free(p)
    return res


def run_batched(ost, lib, args):
    """24 Z2 sources on a 32^3x8 field plus 6 parallel-tempering replicas on
    24^3x8 fields (K = c_r on the t-boundary defect, c_r = 0, 0.2, ..., 1),
    m = 0.02, to 1e-9: one batched solve (HIP graph) against the same 30
    mixed-precision solves run one after the other."""
    def dm(n):
        p = lib.ostp_malloc(n)
        lib.ostp_memset(p, 0, n)
        return p

    mass, res = 0.02, 1e-9
    groups = []  # (shape, par_temp, [(u, uf, ph, phf, b, x)])
    # 32^3 x 8, one field, 24 sources
    shape = (32, 32, 32, 8)
    La = ost.Lattice(*shape)
    La.init()
    lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, 1)
    gb = 8 * La.su3_bytes
    u = smeared_field(lib, dm, gb, 0.5, 3)
    uf = dm(8 * La.su3f_bytes)
    lib.ostp_convert_double_to_float_gauge(u, uf)
    ph, phf = dm(8 * 8 * La.sizeh), dm(4 * 8 * La.sizeh)
    lib.calc_u1_phases(ph, ost.BfParam(0, 0, 0, 0, 0, 0), 0.0, 0.0)
    lib.ostp_phases_to_float(ph, phf)
    items = []
    for k in range(24):
        bz, xz = dm(La.vec3_bytes), dm(La.vec3_bytes)
        lib.ostp_generate_vec3_soa_z2noise(bz, 5, k)
        items.append((u, uf, ph, phf, bz, xz))
    groups.append((shape, 0, items))
    # 6 replicas on 24^3 x 8 with the K plane
    shape = (24, 24, 24, 8)
    Lb = ost.Lattice(*shape, par_temp=True)
    Lb.init()
    gb = 8 * Lb.su3_bytes
    ph, phf = dm(8 * 8 * Lb.sizeh), dm(4 * 8 * Lb.sizeh)
    lib.calc_u1_phases(ph, ost.BfParam(0, 0, 0, 0, 0, 0), 0.0, 0.0)
    lib.ostp_phases_to_float(ph, phf)
    items = []
    for r in range(6):
        ur = smeared_field(lib, dm, gb, 0.5, 20 + r)
        lib.init_k(ur, 0.2 * r, 3, (C.c_int * 3)(24, 24, 24), None, 1)
        urf = dm(8 * Lb.su3f_bytes)
        lib.ostp_convert_double_to_float_gauge(ur, urf)
        bz, xz = dm(Lb.vec3_bytes), dm(Lb.vec3_bytes)
        lib.ostp_generate_vec3_soa_z2noise(bz, 6, r)
        items.append((ur, urf, ph, phf, bz, xz))
    groups.append((shape, 1, items))
    ost.check(0, "batched setup")

    systems = []
    for shape, pt, items in groups:
        for (uu, uuf, pp, ppf, bz, xz) in items:
            sy = ost.BatchSystem()
            sy.nd[:] = shape
            sy.par_temp = pt
            sy.u, sy.u_f, sy.phases, sy.phases_f, sy.in_, sy.out, sy.mass = uu, uuf, pp, ppf, bz, xz, mass
            systems.append(sy)
    n = len(systems)
    arr = (ost.BatchSystem * n)(*systems)
    cg, ru, tr = (C.c_int * n)(), (C.c_int * n)(), (C.c_double * n)()
    lib.ostp_batched_mixed_solve(n, arr, res, 40000, 0.1, args.iters_per_graph, cg, ru, tr)  # warm-up: autotune
    for sy in systems:
        lib.ostp_memset(sy.out, 0, 48 * sy.nd[0] * sy.nd[1] * sy.nd[2] * sy.nd[3] // 2)
    lib.ostp_synchronize()
    t0 = time.perf_counter()
    conv = lib.ostp_batched_mixed_solve(n, arr, res, 40000, 0.1, args.iters_per_graph, cg, ru, tr)
    lib.ostp_synchronize()
    t1 = time.perf_counter()
    # algorithmic Dslash bytes: FP32 pairs per iteration, FP64 pairs per reliable update and at start/end
    dbytes = 0
    for k, sy in enumerate(systems):
        sizeh = sy.nd[0] * sy.nd[1] * sy.nd[2] * sy.nd[3] // 2
        dbytes += 2 * sizeh * (BYTES_PER_SITE_F32 * cg[k] + BYTES_PER_SITE_F64 * (ru[k] + 2))
    # FP16-storage inner solver on the same batch
    for sy in systems:
        lib.ostp_memset(sy.out, 0, 48 * sy.nd[0] * sy.nd[1] * sy.nd[2] * sy.nd[3] // 2)
    ch, rh, trh = (C.c_int * n)(), (C.c_int * n)(), (C.c_double * n)()
    lib.ostp_batched_half_solve(n, arr, res, 40000, 0.1, args.iters_per_graph, ch, rh, trh)  # warm-up: autotune
    for sy in systems:
        lib.ostp_memset(sy.out, 0, 48 * sy.nd[0] * sy.nd[1] * sy.nd[2] * sy.nd[3] // 2)
    lib.ostp_synchronize()
    th0 = time.perf_counter()
    convh = lib.ostp_batched_half_solve(n, arr, res, 40000, 0.1, args.iters_per_graph, ch, rh, trh)
    lib.ostp_synchronize()
    th1 = time.perf_counter()
    out = {"systems": n, "converged": conv, "seconds": round(t1 - t0, 4), "iterations_total": sum(cg),
           "iterations_max": max(cg), "max_true_residual": max(tr[k] for k in range(n)),
           "dslash_GBs_aggregate": round(dbytes / (t1 - t0) / 1e9, 1), "mass": mass, "tolerance": res,
           "half": {"converged": convh, "seconds": round(th1 - th0, 4), "iterations_total": sum(ch),
                    "max_true_residual": max(trh[k] for k in range(n))}}
    # the same systems one by one with the single-system mixed solver
    t_seq, it_seq = 0.0, 0
    for (shape, pt, items) in groups:
        Lx = ost.Lattice(*shape, par_temp=bool(pt))
        Lx.init()
        vb = Lx.vec3_bytes
        d64 = [dm(vb) for _ in range(4)]
        d32 = [dm(vb // 2) for _ in range(5)]
        for (uu, uuf, pp, ppf, bz, xz) in items:
            fp = ost.ferm_param(mass, pp, ppf)
            ip = ost.InverterPackage()
            ip.u, ip.u_f = uu, uuf
            ip.loc_r, ip.loc_h, ip.loc_s, ip.loc_p = d64
            ip.loc_r_f, ip.loc_h_f, ip.loc_s_f, ip.loc_p_f, ip.out_f = d32
            lib.ostp_memset(xz, 0, vb)
            c1, r1, tr1 = C.c_int(0), C.c_int(0), C.c_double(0)
            lib.ostp_synchronize()
            ta = time.perf_counter()
            lib.ostp_mixed_solve(ip, C.byref(fp), xz, bz, res, 40000, 0.0, 0.1, C.byref(c1), C.byref(r1), C.byref(tr1))
            lib.ostp_synchronize()
            t_seq += time.perf_counter() - ta
            it_seq += c1.value
        for p_ in d64 + d32:
            lib.ostp_free(p_)
    out["sequential_seconds"] = round(t_seq, 4)
    out["sequential_iterations_total"] = it_seq
    out["speedup_batched_vs_sequential"] = round(t_seq / (t1 - t0), 3)
    rep = C.create_string_buffer(4096)
    lib.ostp_autotune_report(rep, 4096)
    out["autotune"] = rep.value.decode().strip().splitlines()[-1] if rep.value else ""
    ost.check(0, "batched")
    return out


# SU(3) products in the tree-level Symanzik force of one link: per plane
# (3 of them) 2 plaquette staples x 2 products and 6 rectangle staples x 4
# products, plus U * staple; 198 flops per 3x3 complex product (27 complex
# multiplies, 18 complex additions).  Row reconstructions and the staple sums
# are not counted.
GAUGE_FORCE_FLOPS_PER_LINK = (3 * (2 * 2 + 6 * 4) + 1) * 198
FP64_PEAK_TFLOPS = 78.6


def run_rhmc(ost, lib, args):
    """One RHMC trajectory (UPDATE_SOLOACC_UNOSTEP_VERSATILE) on 48^3 x 12:
    tree-level Symanzik gauge action at beta = 3.8, 2 stout steps rho = 0.15,
    Nf = 2+1 staggered with m_l = 0.0015 (degeneracy 2) and m_s = 0.04, one
    pseudofermion each, order-19 approximations (x^(1/4), x^(-1/2) on
    [1e-7, 1] for the light quark, x^(1/8), x^(-1/4) on [4e-7, 1] for the
    strange), args.rhmc_steps Omelyan steps of trajectory length args.rhmc_tau with 5 gauge
    sub-steps, FP64 multishift solves (1e-8 in the force, 1e-10 for the heat
    bath and the acceptance), links thermalised by gauge-only trajectories
    from the generate_Conf_cold rule with eps = 0.5 (seed 7), seeded momenta
    and pseudofermions.  Also the device time of the gauge-force kernel."""
    golden = ROOT / "tests" / "golden"

    def dm(n):
        p = lib.ostp_malloc(n)
        if not p:
            raise MemoryError(lib.ostp_last_error())
        lib.ostp_memset(p, 0, n)
        return p

    shape = tuple(int(x) for x in args.rhmc_lattice.split(","))
    L = ost.Lattice(*shape)
    L.init()
    lib.ostp_set_batched_graph_iterations(args.iters_per_graph)
    lib.ostp_set_stout(0.15, 2, 0.0, 0, 0, 1)
    lib.ostp_set_inverter_tricks(0, 0, 0.1, 10000)
    sizeh, gb, vb, tb = L.sizeh, 8 * L.su3_bytes, L.vec3_bytes, 8 * 64 * L.sizeh
    conf = dm(gb)
    lib.ostp_generate_Conf_cold(conf, 0.5, 7)
    sizes = {"conf_acc_bkp": gb, "aux_conf_acc": gb, "auxbis_conf_acc": gb, "gstout_conf_acc_arr": 2 * gb,
             "momenta": tb, "ipdot_acc": tb, "ferm_phi_acc": 2 * vb, "ferm_chi_acc": 2 * vb,
             "ferm_shiftmulti_acc": 19 * vb, "k_p_shiftferm": 19 * vb, "kloc_r": vb, "kloc_h": vb, "kloc_s": vb,
             "kloc_p": vb}
    bufs = {k: dm(n) for k, n in sizes.items()}
    for k, p in bufs.items():
        C.c_void_p.in_dll(lib, k).value = p
    ph = dm(8 * 8 * sizeh)
    lib.calc_u1_phases(ph, ost.BfParam(0, 0, 0, 0, 0, 0), 0.0, 0.0)
    fps = (ost.FermParam * 2)()
    for i, (mass, deg, fi, md) in enumerate([
            (0.0015, 2, "approx_1_over_4_order_19_mloglm_7.0.REMEZ", "approx_-1_over_2_order_19_mloglm_7.0.REMEZ"),
            (0.04, 1, "approx_1_over_8_order_19_mloglm_6.4.REMEZ", "approx_-1_over_4_order_19_mloglm_6.4.REMEZ")]):
        fp = fps[i]
        fp.ferm_mass, fp.degeneracy, fp.number_of_ps, fp.index_of_the_first_ps = mass, deg, 1, i
        fp.phases = ph
        fp.approx_fi_mother = ost.read_remez(str(golden / fi))
        fp.approx_md_mother = ost.read_remez(str(golden / md))
        fp.approx_li_mother = ost.read_remez(str(golden / md))
    C.c_void_p.in_dll(lib, "fermions_parameters").value = C.addressof(fps)
    ai = ost.AllocSettings.in_dll(lib, "alloc_info")
    ai.NDiffFlavs, ai.NPS_tot, ai.maxApproxOrder, ai.maxNeededShifts = 2, 2, 19, 19
    ost.ActionParam.in_dll(lib, "act_params").beta = 3.8
    mdp = ost.MdParam(no_md=args.rhmc_steps, gauge_scale=5, t=args.rhmc_tau)
    C.memmove(C.addressof(ost.MdParam.in_dll(lib, "md_parameters")), C.addressof(mdp),
                   C.sizeof(mdp))
    lib.initialize_md_global_variables(mdp)
    ost.check(0, "rhmc setup")

    reps = 20
    us_force = lib.ostp_time_gauge_force(conf, None, bufs["ipdot_acc"], None, 0.0, reps)
    us_force_mom = lib.ostp_time_gauge_force(conf, None, bufs["ipdot_acc"], bufs["momenta"], 0.0, reps)
    ost.check(0, "gauge force timing")
    flops = GAUGE_FORCE_FLOPS_PER_LINK * 8 * sizeh
    tflops = flops / (us_force * 1e-6) / 1e12

    # thermalisation: gauge-only trajectories (the same update with no
    # fermion flavour) of length --rhmc-therm-tau, then --rhmc-warmup full
    # 2+1-flavour trajectories, untimed
    def plaquette():
        return lib.calc_plaquette_soloopenacc(conf, None, None) / (6.0 * 2 * sizeh * 3.0)

    lib.ostp_set_rng_seed(2026, 0)
    plaq = [plaquette()]
    ai.NDiffFlavs = 0
    t0 = time.perf_counter()
    acc_g = 0
    therm_plaq = []

    def set_tau(tau):
        m = ost.MdParam(no_md=args.rhmc_steps, gauge_scale=5, t=tau)
        C.memmove(C.addressof(ost.MdParam.in_dll(lib, "md_parameters")), C.addressof(m), C.sizeof(m))
        lib.initialize_md_global_variables(m)

    set_tau(args.rhmc_therm_tau)
    n_metro = args.rhmc_thermalise // 2
    for it in range(args.rhmc_thermalise):
        # the first half: molecular dynamics with fresh momenta and no
        # accept/reject (from the far-from-equilibrium start every Metropolis
        # test would reject); the second half: HMC with the Metropolis test
        metro = 1 if it >= args.rhmc_thermalise - n_metro else 0
        acc_g = lib.UPDATE_SOLOACC_UNOSTEP_VERSATILE(conf, 1e-10, 1e-8, it, acc_g if metro else 0, metro, 50000)
        therm_plaq.append(round(plaquette(), 5))
    ost.check(0, "gauge-only thermalisation")
    set_tau(args.rhmc_tau)
    t_therm = time.perf_counter() - t0
    progress(f"rhmc: {args.rhmc_thermalise} gauge-only trajectories in {t_therm:.1f} s, plaquette {plaquette():.5f}")
    plaq.append(plaquette())
    ai.NDiffFlavs = 2
    warm = ost.RhmcInfo()
    warm_info = []
    # warm-up trajectories run the MD force with the reduced-precision rational
    # solve when that leg is timed below, so that its one-time set-up (the
    # batched refinement's autotuning and graph) stays out of the timed region;
    # the FP64 solves are warmed by the heat bath and the acceptance test
    warm_mode = 0 if args.rhmc_no_mixed else 2
    for it in range(args.rhmc_warmup):
        t0 = time.perf_counter()
        prev_mode = lib.ostp_set_multishift_precision(warm_mode, 1)
        lib.UPDATE_SOLOACC_UNOSTEP_VERSATILE(conf, 1e-10, 1e-8, it, 0, 1, 50000)
        lib.ostp_set_multishift_precision(prev_mode, 0)
        ost.check(0, "UPDATE_SOLOACC_UNOSTEP_VERSATILE (warm-up)")
        lib.ostp_rhmc_last_info(C.byref(warm))
        warm_info.append({"seconds": round(time.perf_counter() - t0, 2), "delta_S": warm.delta_S,
                          "accepted": warm.accepted, "md_multishift_iterations": warm.md_multishift_iterations,
                          "force_solves": "mixed" if warm_mode else "fp64"})
        progress(f"rhmc: warm-up trajectory {it}: {warm_info[-1]}")
    plaq.append(plaquette())
    # the timed trajectory twice from the same links, momenta and pseudofermions:
    # FP64 rational solves, then the reduced-precision rational solve (FP32
    # CG-M with FP64 reliable updates, batched FP16-inner refinement) in the
    # MD force (heat bath and acceptance stay FP64)
    start = dm(gb)
    ost.check(lib.ostp_memcpy_dtod(start, conf, gb), "save links")
    runs = {}
    for mode, name in ((0, "fp64"), (1, "mixed"))[:1 if args.rhmc_no_mixed else 2]:
        ost.check(lib.ostp_memcpy_dtod(conf, start, gb), "restore links")
        prev_mode = lib.ostp_set_multishift_precision(2 * mode, 1)
        lib.ostp_set_rng_seed(4242, 0)
        t0 = time.perf_counter()
        acc = lib.UPDATE_SOLOACC_UNOSTEP_VERSATILE(conf, 1e-10, 1e-8, 1, 0, 1, 50000)
        t1 = time.perf_counter()
        lib.ostp_set_multishift_precision(prev_mode, 0)
        ost.check(0, "UPDATE_SOLOACC_UNOSTEP_VERSATILE")
        inf = ost.RhmcInfo()
        lib.ostp_rhmc_last_info(C.byref(inf))
        runs[name] = (t1 - t0, acc, inf)
        progress(f"rhmc: timed trajectory ({name} force solves) {t1 - t0:.2f} s, dS {inf.delta_S:.4g}")
    t0, t1 = 0.0, runs["fp64"][0]
    acc, info = runs["fp64"][1], runs["fp64"][2]
    plaq.append(plaquette())
    lib.ostp_free(start)
    for k in bufs:
        C.c_void_p.in_dll(lib, k).value = None
    C.c_void_p.in_dll(lib, "fermions_parameters").value = None
    for p in list(bufs.values()) + [conf, ph]:
        lib.ostp_free(p)
    out = {
        "workload": "RHMC trajectory %s, TLSM beta=3.8, 2 stout rho=0.15, Nf=2+1 m_l=0.0015 m_s=0.04, order 19, "
                    "%d Omelyan steps x 5 gauge sub-steps, tau=%g, FP64 multishift (force 1e-8, "
                    "accept 1e-10; the two flavours' solves and power iterations share their hops)"
                    % ("x".join(map(str, shape)), args.rhmc_steps, args.rhmc_tau),
        "s_per_trajectory": round(t1 - t0, 3),
        "accepted": int(acc),
        "delta_S": info.delta_S,
        "actions": {"gauge": [info.action_in, info.action_fin], "momenta": [info.action_mom_in, info.action_mom_fin],
                    "fermions": [info.action_ferm_in, info.action_ferm_fin]},
        "start": (f"generate_Conf_cold eps=0.5, then {args.rhmc_thermalise} gauge-only trajectories of length "
                  f"{args.rhmc_therm_tau:g} ({args.rhmc_thermalise - n_metro} without and {n_metro} with the "
                  f"Metropolis test, {acc_g} of those accepted; {t_therm:.1f} s) and {args.rhmc_warmup} "
                  f"2+1-flavour RHMC trajectories before the timed one"),
        "thermalisation_plaquettes": therm_plaq,
        "warmup": warm_info,
        "plaquette": {"cold": round(plaq[0], 6), "after_gauge_hmc": round(plaq[1], 6),
                      "after_warmup": round(plaq[2], 6), "after_timed": round(plaq[3], 6)},
        "md_multishift_iterations": info.md_multishift_iterations,
    }
    if "mixed" in runs:
        out["mixed_precision_solves"] = {
            "s_per_trajectory": round(runs["mixed"][0], 3),
            "speedup_vs_fp64": round(runs["fp64"][0] / runs["mixed"][0], 3),
            "accepted": int(runs["mixed"][1]),
            "delta_S": runs["mixed"][2].delta_S,
            "md_multishift_iterations": runs["mixed"][2].md_multishift_iterations,
            "solver": "MD force (1e-8): FP32 CG-M with FP64 reliable updates + batched FP16-inner refinement "
                      "of the shifts above the tolerance; heat bath and acceptance (1e-10) FP64 CG-M",
        }
    out["gauge_force"] = {
        "kernel": "k_gauge_force (tree-level Symanzik staples + TA(U S))",
        "us_per_launch": round(us_force, 1),
        "us_per_launch_with_momentum_step": round(us_force_mom, 1),
        "flops_per_link": GAUGE_FORCE_FLOPS_PER_LINK,
        "achieved_TFLOPs": round(tflops, 2),
        "peak_TFLOPs": FP64_PEAK_TFLOPS,
        "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
    }
    return out


def bench_shifts(ost):
    """The 19 shifts of the bench's rational approximation (x^(-1/4) of order
    19 rescaled to [m^2, 64], m = 0.01)."""
    ra = ost.read_remez(str(ROOT / "tests" / "golden" / "approx_-1_over_4_order_19_mloglm_6.4.REMEZ"))
    return [ra.RA_b[k] * 64.0 for k in range(ra.approx_order)]


def cpu_baseline(libs, shape, shifts, seconds):
    """The CPU oracle (a scalar restatement of the reference loops) timed on
    this host: gcc -O3 -march=native, serial and OpenMP; `value` is the
    OpenMP Dslash bandwidth at the bench lattice (928 B / half-site model)."""
    import cpu_baseline as cpub

    d = cpub.run(libs, shape, shifts, dslash_seconds=seconds / 2)
    key = "dslash_%s" % "x".join(map(str, shape))
    if "error" in d:
        return {"value": None, "unit": "GB/s", "cores": 0, "kind": "port", "sample": "build failed", **d}
    omp, ser = d["omp"], d["serial"]
    return {"value": omp[key]["GBs"], "unit": "GB/s", "cores": omp["cores"], "kind": "port",
            "sample": (f"{omp[key]['pairs']} Doe+Deo pairs at {'x'.join(map(str, shape))} in {omp[key]['seconds']} s "
                       f"(OpenMP, {omp['cores']} threads); serial: {ser[key]['pairs']} pairs, {ser[key]['GBs']} GB/s; "
                       "plus the 8^4 Dslash, the 8^4 CG (m=0.05, 1e-10) and 19-shift CG-M iterations per build"),
            **d}


if __name__ == "__main__":
    main()
