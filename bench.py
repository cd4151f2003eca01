#!/usr/bin/env python3
"""SIE-GSTC synthesis iteration benchmark on one MI355X.

Workload (default): the anisotropic GSTC metasurface on a circular-arc
spindle (arc radius 1.5 lambda about its 2 lambda chord), N ~ 72k unknowns,
10 real susceptibility parameters per face drawn uniformly from [-2, 2]/k0
(seeded), one x-polarised plane wave along +z, Dunavant rule 17 (the GSTC
default), objective = far-field power in the 10-degree cone about +z.

One step = one synthesis iteration: surface maps from the parameters ->
FP64 eta0-scaled assembly -> complex-FP32 LU on the matrix cores -> FP64
GMRES-IR to a normwise backward error of IR_TOL -> objective -> adjoint
solve with Z^H (the same FP32 factors + FP64 refinement) -> per-face
gradient contraction -> parameter update. The plain FP64 path (FP64 LU,
FP64 solves) then runs the last iterate once untimed (its solution is the
reference for the errors) and --fp64-steps times timed, for the speed ratio.

Prints one JSON line (rank 0). value = seconds per iteration (lower is better).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sie-gstc-2026_amd"))
# kernel arguments in device memory (HIP runtime option, read when the runtime
# starts): shorter dispatch of the factorisation's ~20k dependent launches,
# ~25 ms per iteration at N = 72k, for the mixed and the FP64 path alike; a
# value already in the environment is kept
os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")

METRIC = "assemble+solve+adjoint-grad s/iter @N≈72k, FP32-MFMA+IR vs FP64, err≤1e-8"
# refinement target of both solves: normwise backward error |r| / (|Z|_F |x| + |b|)
# of 1e-17, the level of the plain FP64 LU solution's own backward error on
# this operator (7e-18 to 2.5e-16 on the trajectory's iterates, reported as
# fp64_backward_error): the mixed and the
# FP64 solutions then differ by about cond(Z) (be_mixed + be_fp64), the order
# of the FP64 solution's own error, also at the near-resonant iterates of the
# trajectory where cond(Z) reaches ~1e8 (a stop at 1e-15 left x 8e-8 and the
# gradient 2e-7 from FP64 there). A solve that stalls at the FP64 residual
# floor above the target ends there (FGMRES-IR's stagnation exit) and is
# accepted up to IR_ACCEPT.
IR_TOL = 1e-17
IR_ACCEPT = 1e-15

WORKLOADS = {
    # name: (mesh kind, size parameter, quadrature rule, parameters per face)
    "gstc-spindle-72k": ("spindle", 91, 17, 10),
    "gstc-spindle-31k": ("spindle", 60, 17, 10),
    "gstc-sphere-30k": ("sphere", 4, 17, 3),
    "gstc-spindle-72k-rule7": ("spindle", 91, 7, 10),
    "gstc-spindle-8k": ("spindle", 30, 7, 10),
}

PROF = {"gemm_f32": 0, "gemm_f64": 1, "far_pairs": 2, "panel": 3, "near_pairs": 4, "gather": 5, "gemv": 6,
        "gemm_f16x3": 7, "gemm_f16x3_wide": 8, "sweeps": 9, "contraction": 10, "matvec_free": 11}
# events inside the timed iterations: only kernels with few launches per iteration
TIMED_PROF = ("gemm_f16x3_wide", "far_pairs", "near_pairs", "gather", "contraction")
# the many-launch kernels: events on an untimed re-solve of the last iterate
UNTIMED_PROF = ("gemm_f32", "panel", "gemm_f16x3", "gemv", "sweeps")
# Published dense peaks. The split-FP16 GEMM issues three FP16 MFMA products per
# algorithmic FP32 product, so its ceiling for algorithmic flops is 2.5 PF / 3.
# "latency": no throughput bound applies (reported as time per unit).
HBM = 8.0e12
PEAK = {"gemm_f32": (157.3e12, "mfma"), "gemm_f64": (78.6e12, "mfma"), "far_pairs": (78.6e12, "valu-fp64"),
        "gemm_f16x3": (2.5e15 / 3, "mfma"), "gemm_f16x3_wide": (2.5e15 / 3, "mfma"),
        "gemv": (HBM, "hbm"), "sweeps": (HBM, "hbm"), "contraction": (HBM, "hbm"), "gather": (HBM, "hbm"),
        "matvec_free": (78.6e12, "valu-fp64"), "panel": (None, "latency"), "near_pairs": (None, "latency")}


DTYPE = {"f16x3": "c64 LU on split-fp16 MFMA (fp32-accurate) + c128 refinement",
         "4m": "c64 LU on fp32 MFMA + c128 refinement", "3m": "c64 LU on fp32 MFMA (3M) + c128 refinement"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prof_read(lib, k):
    n = ctypes.c_int()
    ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib.sie_profile_read(k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by))
    return n.value, ms.value, fl.value, by.value


def cpu_baseline(mesh, o, kind, seconds):
    """Reference host numerics (oracle/_ref) on a bounded sample of the same
    workload, extrapolated to one iteration: the Calderon assembly of a subset
    of test faces against all source faces (rule as configured) and a dense
    complex128 LU (OpenBLAS zgetrf through scipy) at a smaller dimension,
    scaled by F^2 and N^3. The adjoint and refinement sweeps are O(N^2) and
    are added as two triangular-solve pairs at measured zgetrs speed."""
    import numpy as np
    import scipy.linalg
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    if not ol.ref_available():
        return None
    lib = ctypes.CDLL(ol.REF_SO)
    f = lib.ref_sample_pairs
    P = ctypes.c_void_p
    f.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                  ctypes.c_float]
    f.restype = ctypes.c_double
    import sie_mom_amd as sm
    p = sm.calderon_params(o)
    V = mesh.vertices
    F = mesh.faces
    nf = mesh.nf
    # calibrate the number of test-face rows so that the sample takes ~0.75 x seconds
    # (the first timing of a small sample includes thread start-up, so refine twice)
    rows2, dt = 2, 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        f(V.ctypes.data_as(P), F.ctypes.data_as(P), nf, 0, rows2, p.rule, p.k1, p.k2, p.threshold)
        dt = time.perf_counter() - t0
        if dt > 0.5 * seconds or rows2 >= nf:
            break
        rows2 = max(2, min(nf, int(rows2 * (0.75 * seconds) / max(dt, 1e-3))))
    asm = dt * nf / rows2
    # LU at a smaller n
    n_s = 4096
    rng = np.random.default_rng(0)
    A = rng.normal(size=(n_s, n_s)) + 1j * rng.normal(size=(n_s, n_s))
    t0 = time.perf_counter()
    lu = scipy.linalg.lu_factor(A, overwrite_a=True, check_finite=False)
    tlu = time.perf_counter() - t0
    b = rng.normal(size=n_s) + 0j
    t0 = time.perf_counter()
    for _ in range(4):
        scipy.linalg.lu_solve(lu, b, check_finite=False)
    tsolve = (time.perf_counter() - t0) / 4
    N = 4 * mesh.ne
    lu_ext = tlu * (N / n_s) ** 3
    solve_ext = tsolve * (N / n_s) ** 2
    total = asm + lu_ext + 2 * solve_ext
    threads = os.cpu_count()
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return {"value": total, "unit": "s/iter", "cores": threads, "kind": "reference",
            "sample": (f"reference host Calderon assembly of {rows2} of {nf} test faces x {nf} source faces "
                       f"({dt:.1f} s, x{nf / rows2:.1f}) + scipy/OpenBLAS zgetrf n={n_s} ({tlu:.2f} s, x(N/n)^3) "
                       f"+ 2 zgetrs sweeps; extrapolated to N={N}"),
            "assembly_s": asm, "lu_s": lu_ext}


# counter summaries per profiled GEMM kernel: (kernel-name part, summary file)
TRAFFIC = {
    "gemm_f16x3_wide": ("cgemm_f16x3_r3_kernel",
                        os.path.join(ROOT, "profiles", "cgemm_f16x3_r3_gstc-spindle-72k_traffic.txt")),
    "gemm_f16x3": ("cgemm_f16x3_kernel", os.path.join(ROOT, "profiles", "cgemm_f16x3_gstc-spindle-72k_traffic.txt")),
}


def copy_calibration():
    """One 256 MiB device-to-device copy: the reference kernel for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes (the runtime's copyBuffer kernel)."""
    import torch
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4}))


def lu_gemms(n=72048, variant="f16x3"):
    """The split-FP16 products of the FP32 LU at N = n, on their own, in the
    factorisation's order and shapes: per 4096-column level (width w, first
    column c), after NB step s the in-level update C (M x N) -= L (M x 256)
    U (256 x N) with M = n - c - 256 (s + 1), N = w - 256 (s + 1); then the
    level's triangular solve for the columns right of the level (its updates:
    M = w - 256 (s + 1), N = n - c - w, K = 256) and the level's trailing update
    M = N = n - c - w, K = 4096 (the wide kernel). Seeded data; the workload
    of the counter passes (HBM bytes per launch depend on the shapes only)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "sie-gstc-2026_amd"))
    import sie_mom_amd as sm
    g = torch.Generator(device="cuda").manual_seed(7)
    Cbuf = torch.randn(n * n, dtype=torch.complex64, device="cuda", generator=g)
    Lb = torch.randn(n * 4096, dtype=torch.complex64, device="cuda", generator=g) * 0.5
    Ub = torch.randn(4096 * n, dtype=torch.complex64, device="cuda", generator=g)
    launches = 0

    def upd(M, N, K):
        nonlocal launches
        if M <= 0 or N <= 0:
            return
        A = Lb[: M * K].view(K, M).T
        B = Ub[: N * K].view(N, K).T
        C = Cbuf[: M * N].view(N, M).T
        sm.linalg.gemm(A, B, C, alpha=-1.0, beta=1.0, variant=variant)
        launches += 1

    for c in range(0, n, 4096):
        w = min(4096, n - c)
        for st in range(w // 256):
            upd(n - c - 256 * (st + 1), w - 256 * (st + 1), 256)
        right = n - c - w
        for st in range(w // 256):
            upd(w - 256 * (st + 1), right, 256)
        upd(right, right, 4096)
    torch.cuda.synchronize()
    print(json.dumps({"lu_gemm_launches": launches, "n": n, "variant": variant}))


def assembly_only(workload="gstc-spindle-72k"):
    """The FP64 assembly of the workload's GSTC operator (its parameters,
    mesh and rule), twice: the workload of the far-pair counter passes."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "sie-gstc-2026_amd"))
    import sie_mom_amd as sm
    from sie_mom_amd import adjoint as ad
    kind_mesh, size, rule, kind = WORKLOADS[workload]
    assert kind_mesh == "spindle"
    k0 = 2 * math.pi
    mesh = sm.Mesh.spindle(2.0, 1.5, size)
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule)
    rng = np.random.default_rng(20260101)
    params = torch.from_numpy(rng.uniform(-2, 2, size=(mesh.nf, kind)) / k0).cuda()
    rec = ad.records(params, kind, o)
    Z = None
    for r in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Z = sm.ops.assemble_gstc(mesh, o, rec, out=Z)
        torch.cuda.synchronize()
        log(f"assembly {r}: {time.perf_counter() - t0:.3f} s (N = {Z.shape[0]})")
    print(json.dumps({"workload": workload, "faces": mesh.nf, "unknowns": int(Z.shape[0])}))


# FP64 VALU counters of one pass (rocprofv3 --pmc): per dispatch, summed over
# the device's instances
VALU_COUNTERS = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                 "SQ_INSTS_VALU_TRANS_F64", "SQ_INSTS_VALU", "GRBM_GUI_ACTIVE")


def summarize_valu(pmc_csv, stats_csv, out_path, command):
    """Writes the VALU summary of far_pairs_kernel: wave instructions per
    launch by kind, the FP64 operations they execute (64 lanes; an FMA counts
    two), and the share of the SIMDs' cycles their issue takes: a wave's VALU
    instruction occupies its 16-lane SIMD for 4 cycles, 4 SIMDs per CU, over
    the launch's GRBM_GUI_ACTIVE cycles (the CSV sums that counter over the
    8 XCDs; each counts the same clock)."""
    import csv
    name_part = "far_pairs_kernel<1, true>"
    per, span = {}, {}
    for r in csv.DictReader(open(pmc_csv)):
        if name_part in r["Kernel_Name"]:
            d = per.setdefault(r["Dispatch_Id"], {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            span[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    if not per:
        raise SystemExit(f"no {name_part} dispatches in {pmc_csv}")
    mean = {c: sum(d.get(c, 0.0) for d in per.values()) / len(per) for c in VALU_COUNTERS}
    durs = [float(r["AverageNs"]) for r in csv.DictReader(open(stats_csv)) if name_part in r["Name"]]
    if not durs:
        raise SystemExit(f"no {name_part} row in {stats_csv}")
    dur_s = durs[0] * 1e-9
    dur_pmc = sum(span.values()) / len(span) * 1e-9
    f64 = sum(mean[c] for c in VALU_COUNTERS[:4])
    flops = 64.0 * (2.0 * mean["SQ_INSTS_VALU_FMA_F64"] + mean["SQ_INSTS_VALU_MUL_F64"]
                    + mean["SQ_INSTS_VALU_ADD_F64"] + mean["SQ_INSTS_VALU_TRANS_F64"])
    n_xcd, n_simd = 8, 4 * 256
    cycles = mean["GRBM_GUI_ACTIVE"] / n_xcd
    with open(out_path, "w") as out:
        out.write(f"# VALU work of {name_part} (GSTC spindle, N = 72048, rule 17: the bench's assembly, "
                  f"run alone by bench.py --assembly)\n# command: {command}\n")
        out.write(f"launches: {len(per)}\n")
        for c in VALU_COUNTERS:
            out.write(f"{c.lower()}_per_launch: {mean[c]:.6e}\n")
        out.write(f"f64_wave_instructions_per_launch: {f64:.6e}\n")
        out.write(f"f64_share_of_valu_instructions: {f64 / mean['SQ_INSTS_VALU']:.4f}\n")
        out.write(f"executed_f64_flops_per_launch: {flops:.6e}\n")
        out.write(f"avg_duration_ms (kernel trace run): {dur_s * 1e3:.3f}\n")
        out.write(f"avg_duration_ms (counter run): {dur_pmc * 1e3:.3f}\n")
        out.write(f"executed_f64_tflops (kernel trace duration): {flops / dur_s / 1e12:.2f} "
                  f"(published FP64 vector peak 78.6)\n")
        out.write(f"clock_mhz (GRBM_GUI_ACTIVE per XCD / counter-run duration): {cycles / dur_pmc / 1e6:.0f}\n")
        out.write(f"valu_issue_share_of_simd_cycles: {4.0 * mean['SQ_INSTS_VALU'] / (n_simd * cycles):.4f}\n")
        out.write(f"f64_issue_share_of_simd_cycles: {4.0 * f64 / (n_simd * cycles):.4f}\n")
    print(open(out_path).read())


def _pmc_mean_kib(csv_path, counter, name_part):
    import csv
    per = {}
    for r in csv.DictReader(open(csv_path)):
        if name_part in r["Kernel_Name"] and r["Counter_Name"] == counter:
            per[r["Dispatch_Id"]] = per.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} rows for {name_part} in {csv_path}")
    return sum(per.values()) / len(per), len(per)


def summarize_pmc(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, command):
    """Writes the traffic summaries of the split-FP16 GEMM kernels: HBM bytes per
    launch from FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by what
    the counter reads for a 256 MiB copy (elementwise copy kernel)."""
    # the runtime's blit kernel performs the 256 MiB device-to-device copy
    cf, _ = _pmc_mean_kib(cal_fetch_csv, "FETCH_SIZE", "copyBuffer")
    cw, _ = _pmc_mean_kib(cal_write_csv, "WRITE_SIZE", "copyBuffer")
    mib256 = 256 * 1024.0  # KiB
    ff, fw = mib256 / cf, mib256 / cw
    for name_part, path in TRAFFIC.values():
        try:
            f, nf = _pmc_mean_kib(fetch_csv, "FETCH_SIZE", name_part)
            w, nw = _pmc_mean_kib(write_csv, "WRITE_SIZE", name_part)
        except SystemExit as e:
            print(e)
            continue
        traffic = (f * ff + w * fw) * 1024.0
        with open(path, "w") as out:
            out.write(f"# HBM traffic of {name_part} (LU trailing updates at N = 72048: the factorisation's shapes, "
                      f"run alone by bench.py --lu-gemms)\n")
            out.write(f"# command: {command}\n")
            out.write(f"launches_fetch: {nf}\nlaunches_write: {nw}\n")
            out.write(f"fetch_kib_per_launch_raw: {f:.1f}\nwrite_kib_per_launch_raw: {w:.1f}\n")
            out.write(f"copy_256MiB_fetch_kib: {cf:.1f}\ncopy_256MiB_write_kib: {cw:.1f}\n")
            out.write(f"fetch_scale: {ff:.4f}\nwrite_scale: {fw:.4f}\n")
            out.write(f"traffic_bytes_per_launch: {traffic:.6e}\n")
        print(open(path).read())


def summarize_stats(stats_csv, out_path, title, command):
    """Writes a kernel-statistics summary (rocprofv3 --kernel-trace --stats CSV):
    calls, total and average duration per kernel, by total time."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    with open(out_path, "w") as out:
        out.write(f"# rocprofv3 kernel statistics: {title}\n# command: {command}\n\n")
        out.write(f"{'kernel':66s}{'calls':>7s}{'total ms':>11s}{'avg us':>12s}{'%':>7s}\n")
        for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
            name = r["Name"]
            if name.startswith("_Z"):
                import subprocess
                try:
                    name = subprocess.run(["c++filt", name], capture_output=True, text=True).stdout.strip() or name
                except OSError:
                    pass
            name = name.replace("sie::(anonymous namespace)::", "").split("(")[0]
            t = float(r["TotalDurationNs"])
            out.write(f"{name[:65]:66s}{int(r['Calls']):7d}{t / 1e6:11.1f}{float(r['AverageNs']) / 1e3:12.1f}"
                      f"{100 * t / tot:7.2f}\n")
        out.write(f"# total kernel time {tot / 1e6:.1f} ms\n")
    print(open(out_path).read())


def traffic_per_launch(kernel):
    """HBM bytes per launch of a profiled GEMM kernel from its committed counter
    summary, or None."""
    if kernel not in TRAFFIC:
        return None
    try:
        for line in open(TRAFFIC[kernel][1]):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# Dielectric (T-PMCHWT) workloads: name -> (icosphere level, quadrature rule, plane waves)
PMCHWT_WORKLOADS = {"pmchwt-sphere-61k": (5, 17, 256), "pmchwt-sphere-15k": (4, 17, 1)}


def pmchwt_bench(args):
    """T-PMCHWT dielectric sphere (radius 1, eps_r = 4, k0 a = 5) with many
    plane waves on the Fibonacci sphere (alternating theta / phi
    polarisation): FP64 assembly, FP32-MFMA LU, FP64 blocked refinement of all
    columns to a normwise backward error of 1e-15 (timed, `value` = seconds
    per assemble + factor + solve of every column); then the plain FP64 LU
    path for the ratio and the error; then the iterative comparison path:
    one matrix-free operator product (quadrature recomputed) timed, and
    GMRES(50) to 1e-8 on the first wave with the stored operator (same
    operator, so the step count carries over to the matrix-free product)."""
    import numpy as np
    import torch
    import sie_mom_amd as sm
    from sie_mom_amd import _lib
    level, rule, nwaves = PMCHWT_WORKLOADS[args.workload]
    mesh = sm.Mesh.icosphere(1.0, level)
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(5.0), quadrature_rule=rule)
    o.interior.relative_permittivity = 4.0
    n = 2 * mesh.ne
    waves = sm.fibonacci_plane_waves(o, nwaves) if nwaves > 1 else [sm.plane_wave(o, (1, 0, 0), (0, 0, 1))]
    log(f"workload {args.workload}: F={mesh.nf} E={mesh.ne} N={n} rule={rule} waves={nwaves}")
    lib = _lib.lib

    def mixed_once():
        t = {}
        t0 = time.perf_counter()
        Z = sm.ops.assemble_pmchwt(mesh, o)
        B = sm.ops.rhs_plane_waves(mesh, o, waves)
        torch.cuda.synchronize()
        t["assemble"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        LU32, zf = sm.linalg.to_precision(Z, 32, fro=True)
        lu = sm.linalg.DenseLuSolver(variant="f16x3")
        lu.factorize(LU32)
        torch.cuda.synchronize()
        t["factor"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        X, rr, its, be = sm.linalg.ir_solve(Z, LU32, lu.pivots, B if nwaves > 1 else B[:, 0].contiguous(),
                                            tol=1e-15, maxit=40, method="ir" if nwaves > 1 else "fgmres",
                                            znorm=zf)
        torch.cuda.synchronize()
        t["solve"] = time.perf_counter() - t0
        return Z, B, X, its, be, t

    Z, B, X, its, be, _ = mixed_once()  # warm: code objects, workspaces
    del Z, B, X
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    lib.sie_profile_enable(PROF["far_pairs"], 1)
    t0 = time.perf_counter()
    Z, B, X, its, be, ph = mixed_once()
    elapsed = time.perf_counter() - t0
    far = prof_read(lib, PROF["far_pairs"])
    lib.sie_profile_enable(PROF["far_pairs"], 0)
    log(f"mixed: {elapsed:.3f} s  phases {json.dumps({k: round(v, 3) for k, v in ph.items()})}; "
        f"refinement sweeps max {int(np.max(its))}, backward error max {np.max(be):.2e}")
    out = {}
    if not args.no_fp64:
        Zc = Z.clone()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lu64 = sm.linalg.DenseLuSolver()
        lu64.factorize(Zc)
        X64 = B.clone() if nwaves > 1 else B[:, 0].clone()
        lu64.solve(X64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t1 + ph["assemble"]
        Xm = X if X.dim() == 2 else X[:, None]
        X6 = X64 if X64.dim() == 2 else X64[:, None]
        err = float((torch.linalg.norm(Xm - X6, dim=0) / torch.linalg.norm(X6, dim=0)).max())
        log(f"fp64: {t64:.3f} s (assembly counted as in the mixed path); max rel err of the mixed columns {err:.2e}")
        out = {"fp64_s": t64, "speedup_vs_fp64": t64 / elapsed, "rel_err_x_vs_fp64_max": err}
        del Zc, lu64
        torch.cuda.empty_cache()
    # iterative comparison path
    xv = torch.randn(n, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    lib.sie_profile_enable(PROF["matvec_free"], 1)
    sm.ops.pmchwt_matvec(mesh, o, xv)
    torch.cuda.synchronize()
    mv = prof_read(lib, PROF["matvec_free"])
    lib.sie_profile_enable(PROF["matvec_free"], 0)
    t2 = time.perf_counter()
    g = sm.linalg.gmres(Z, B[:, 0].contiguous(), m=50, maxnit=40, tol=1e-8)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t2
    mv_s = mv[1] * 1e-3
    log(f"matrix-free product {mv_s:.3f} s ({mv[2] / mv_s / 1e12:.1f} TFLOP/s algorithmic); GMRES(50) to 1e-8: "
        f"{g.iterations} steps, converged {g.converged}, {tg:.2f} s with the stored operator, "
        f"{g.iterations * mv_s:.1f} s of matrix-free products")
    line = {"metric": f"assemble+factor+solve s for {nwaves} RHS, FP32-MFMA+IR vs FP64, T-PMCHWT",
            "value": elapsed, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": elapsed * 1e3,
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": DTYPE["f16x3"], "data": "synthetic (icosphere, Fibonacci plane waves)",
            "config": {"workload": args.workload, "faces": mesh.nf, "edges": mesh.ne, "unknowns": n,
                       "quadrature_rule": rule, "rhs": nwaves, "eps_r": 4.0, "k0a": 5.0},
            "phases_s": {k: round(v, 4) for k, v in ph.items()},
            "refinement": {"sweeps_max": int(np.max(its)), "backward_error_max": float(np.max(be))},
            "far_pairs": {"ms": far[1], "tflops": far[2] / (far[1] * 1e-3) / 1e12 if far[1] else None},
            "matrix_free": {"product_s": mv_s, "tflops": mv[2] / mv_s / 1e12, "gmres_steps": g.iterations,
                            "gmres_converged": g.converged, "gmres_stored_s": tg,
                            "gmres_matrix_free_s_est": g.iterations * mv_s}}
    line.update(out)
    print(json.dumps(line))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--lu-gemms":
        return lu_gemms(variant=sys.argv[2] if len(sys.argv) > 2 else "f16x3")
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-calibration":
        return copy_calibration()
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-pmc":
        return summarize_pmc(*sys.argv[2:6], " ".join(sys.argv[6:]))
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-stats":
        return summarize_stats(sys.argv[2], sys.argv[3], sys.argv[4], " ".join(sys.argv[5:]))
    if len(sys.argv) > 1 and sys.argv[1] == "--assembly":
        return assembly_only()
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-valu":
        return summarize_valu(sys.argv[2], sys.argv[3], sys.argv[4], " ".join(sys.argv[5:]))
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="gstc-spindle-72k", choices=sorted(WORKLOADS) + sorted(PMCHWT_WORKLOADS))
    ap.add_argument("--variant", default="f16x3", choices=["4m", "3m", "f16x3"])
    ap.add_argument("--ir-method", default="fgmres", choices=["ir", "fgmres"],
                    help="FP64 refinement around the FP32 LU: classic sweeps or FGMRES")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison path")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed FP64 iterations (after one untimed FP64 iteration on the same iterate)")
    ap.add_argument("--ir-tol", type=float, default=IR_TOL,
                    help="refinement target: normwise backward error |r| / (|Z|_F |x| + |b|)")
    ap.add_argument("--no-matvec", action="store_true", help="skip the matrix-free product measurement")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the dense SIE-MoM path is "
              "single-GPU here and its multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    if args.workload in PMCHWT_WORKLOADS:
        return pmchwt_bench(args)

    import numpy as np
    import torch
    import sie_mom_amd as sm
    from sie_mom_amd import _lib
    from sie_mom_amd import adjoint as ad

    kind_mesh, size, rule, kind = WORKLOADS[args.workload]
    lam = 1.0
    k0 = 2 * math.pi / lam
    if kind_mesh == "spindle":
        mesh = sm.Mesh.spindle(2 * lam, 1.5 * lam, size)
    else:
        mesh = sm.Mesh.icosphere(1.0, size)
        k0 = 4.0
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule)
    waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")]
    N = 4 * mesh.ne
    log(f"workload {args.workload}: F={mesh.nf} E={mesh.ne} N={N} rule={rule} params/face={kind}")
    rng = np.random.default_rng(20260101)
    params = torch.from_numpy(rng.uniform(-2, 2, size=(mesh.nf, kind)) / k0).cuda()
    problem = ad.GstcProblem(mesh, o, waves, kind, mode="mixed", ir_tol=args.ir_tol, ir_maxit=40,
                             variant=args.variant, ir_method=args.ir_method, ir_stop="backward",
                             ir_accept=IR_ACCEPT)
    lib = _lib.lib
    # load every kernel once on a tiny problem (code objects, workspaces)
    tiny = sm.Mesh.icosphere(1.0, 1)
    tp = ad.GstcProblem(tiny, sm.GstcOptions(frequency_hz=o.frequency_hz, quadrature_rule=rule), waves, kind)
    tst = tp.solve(torch.from_numpy(np.random.default_rng(1).uniform(-2, 2, size=(tiny.nf, kind)) / k0))
    tJ, tF = tp.objective(tst)
    tp.gradient(tst, tF)
    del tst, tp
    torch.cuda.synchronize()

    ir_log = []  # per iterate: forward / adjoint refinement steps and end state

    def step(p, prob=problem):
        st = prob.solve(p)
        J, F = prob.objective(st)
        g = prob.gradient(st, F)
        if prob is problem:
            ir_log.append((st.ir_iterations, st.adj_iterations, st.ir_backward_error, st.adj_backward_error,
                           st.ir_residual, st.adj_residual))
        # gradient ascent on the cone power, parameter step bounded by 1e-3/k0
        p = p + (1e-3 / k0) * g / g.abs().max()
        return p, st, J, g

    for _ in range(args.warmup):
        params, st, J, g = step(params)
        del st
    ir_log.clear()
    # events inside the timed iteration only on the few-launch kernels (the
    # dominant wide GEMM and the far-pair quadrature): an event pair per launch
    # of the thousands of small panel / in-panel kernels would add host time
    for name in TIMED_PROF:
        lib.sie_profile_enable(PROF[name], 1)
    problem.timings = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p_start = params.clone()
    for _ in range(args.steps):
        p_last = params  # the parameters of this step's solve (step() returns a new tensor)
        params, st, J, g = step(params)
        if _ < args.steps - 1:
            del st
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s_per_iter = elapsed / args.steps
    timed_log = list(ir_log)
    prof = {name: prof_read(lib, PROF[name]) for name in TIMED_PROF}
    for name in TIMED_PROF:
        lib.sie_profile_enable(PROF[name], 0)
    phases = {k: v / args.steps for k, v in problem.timings.items()}
    # the other kernels' rooflines: one more solve of the last iterate, untimed
    timings_keep = dict(problem.timings)
    for name in UNTIMED_PROF:
        lib.sie_profile_enable(PROF[name], 1)
    del st
    st = problem.solve(p_last)  # the timed iterate again: the same solution, now with events
    J_re, F_re = problem.objective(st)
    problem.gradient(st, F_re)
    torch.cuda.synchronize()
    for name in UNTIMED_PROF:
        prof[name] = prof_read(lib, PROF[name])
        lib.sie_profile_enable(PROF[name], 0)
    problem.timings = timings_keep
    ir_log.clear()
    fw = [e[0] for e in timed_log]
    aw = [e[1] for e in timed_log]
    ir_stats = {"stop": "normwise backward error |r|/(|Z|_F |x|+|b|)", "tol": args.ir_tol, "accept": IR_ACCEPT,
                "forward_steps": fw, "adjoint_steps": aw,
                "forward_steps_mean": sum(fw) / len(fw), "forward_steps_max": max(fw),
                "adjoint_steps_mean": sum(aw) / len(aw), "adjoint_steps_max": max(aw),
                "backward_error_max": max(max(e[2], e[3]) for e in timed_log),
                "rel_residual_max": max(max(e[4], e[5]) for e in timed_log)}
    log(f"mixed: {s_per_iter:.3f} s/iter  phases {json.dumps({k: round(v, 3) for k, v in phases.items()})}")
    log(f"{args.ir_method}: steps per iterate forward {fw} adjoint {aw}; max backward error "
        f"{ir_stats['backward_error_max']:.2e}, max rel residual {ir_stats['rel_residual_max']:.2e}; J={float(J):.6e}")

    # re-run the last iterate's solve in FP64 for the error and the speed ratio
    out_extra = {}
    if not args.no_fp64:
        x_mixed = st.xs.clone()
        g_mixed = g.clone()
        J_mixed = float(J)
        p64 = ad.GstcProblem(mesh, o, waves, kind, mode="fp64")
        # one untimed FP64 iteration of the last iterate (allocation, first
        # launches; its solution is the reference for the errors), then the
        # same iteration timed fp64_steps times like the mixed one (whole
        # steps, no per-launch events). The mixed state (its FP64 operator) is
        # kept through the untimed iteration for the backward errors of both
        # solutions against the same operator (208 GB at N = 72k).
        st64 = p64.solve(p_last)
        J64, F64 = p64.objective(st64)
        g64 = p64.gradient(st64, F64)
        ex = float(torch.linalg.norm(x_mixed - st64.xs) / torch.linalg.norm(st64.xs))
        eg = float(torch.linalg.norm(g_mixed - g64) / torch.linalg.norm(g64))
        eJ = abs(J_mixed - float(J64)) / abs(float(J64))
        bvec = sm.ops.rhs_plane_waves(mesh, o, waves)[:, 0].contiguous()
        bn = float(torch.linalg.norm(bvec))

        def backward_error(x):
            r = sm.linalg.zgemv(st.Z, x, bvec.clone(), alpha=-1.0, beta=1.0)
            return float(torch.linalg.norm(r)) / (st.znorm * float(torch.linalg.norm(x)) + bn)

        be64, bem = backward_error(st64.xs), backward_error(x_mixed)
        log(f"backward error on the last iterate's operator: fp64 LU solution {be64:.2e}, mixed {bem:.2e}")
        # freed into torch's cache, not released: the timed FP64 iterations then
        # take their operator from the cache, as the mixed ones do (a release
        # here put a fresh 83 GB allocation into the first timed FP64 assembly)
        del st64, st, bvec
        p64.timings = {}
        t64 = None
        if args.fp64_steps > 0:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.fp64_steps):
                _, st64, _, _ = step(p_last, p64)
                del st64
            torch.cuda.synchronize()
            t64 = (time.perf_counter() - t1) / args.fp64_steps
            log(f"fp64: {t64:.3f} s/iter over {args.fp64_steps} iterations  phases "
                f"{json.dumps({k: round(v / args.fp64_steps, 3) for k, v in p64.timings.items()})}")
        log(f"mixed vs fp64: rel err x {ex:.2e}, grad {eg:.2e}, J {eJ:.2e}")
        out_extra = {"fp64_s_per_iter": t64, "fp64_steps": args.fp64_steps,
                     "speedup_vs_fp64": (t64 / s_per_iter) if t64 else None, "rel_err_x_vs_fp64": ex,
                     "fp64_backward_error": be64, "mixed_backward_error": bem,
                     "rel_err_grad_vs_fp64": eg, "rel_err_J_vs_fp64": eJ,
                     "fp64_phases_s": {k: round(v / max(args.fp64_steps, 1), 4) for k, v in p64.timings.items()}}
        torch.cuda.empty_cache()

    # one matrix-free product of the same operator (the reference's
    # recomputing matvec: its quadrature is redone for every product)
    if not args.no_matvec:
        recs = ad.records(p_last, kind, o)
        xv = torch.randn(N, dtype=torch.complex128, device="cuda")
        sm.ops.gstc_matvec(mesh, o, recs, xv)  # warm (scratch)
        torch.cuda.synchronize()
        lib.sie_profile_enable(PROF["matvec_free"], 1)
        sm.ops.gstc_matvec(mesh, o, recs, xv)
        torch.cuda.synchronize()
        prof["matvec_free"] = prof_read(lib, PROF["matvec_free"])
        lib.sie_profile_enable(PROF["matvec_free"], 0)
        del recs, xv
        torch.cuda.empty_cache()

    rooflines = {}
    for name, (n, ms, fl, by) in prof.items():
        if n == 0 or ms <= 0:
            continue
        peak, bound = PEAK[name]
        if bound == "latency":
            rooflines[name] = {"bound": "latency", "launches": n, "avg_ms": ms / n, "total_ms": ms}
            if name == "panel":
                # one grid-wide pivot exchange per column of the factorisation
                rooflines[name]["us_per_column"] = ms * 1e3 / N
            continue
        if bound == "hbm":
            ach = by / (ms * 1e-3)
            rooflines[name] = {"bound": "hbm", "achieved": ach / 1e9, "peak": peak / 1e9, "unit": "GB/s",
                               "frac": ach / peak, "launches": n, "avg_ms": ms / n, "total_ms": ms,
                               "alg_bytes_per_launch": by / n}
        else:
            ach = fl / (ms * 1e-3)
            rooflines[name] = {"bound": bound, "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                               "frac": ach / peak, "launches": n, "avg_ms": ms / n, "total_ms": ms,
                               "alg_bytes_per_launch": by / n}
    dom = max(("gemm_f32", "gemm_f16x3", "gemm_f16x3_wide", "far_pairs"), key=lambda k: prof[k][1])
    r = dict(rooflines.get(dom, {}))
    roof = {"kernel": dom, "bound": "valu-fp64" if dom == "far_pairs" else "mfma",
            "achieved": r.get("achieved"), "peak": r.get("peak"), "unit": "TFLOP/s", "frac": r.get("frac"),
            "traffic": traffic_per_launch(dom), "avg_ms": r.get("avg_ms"),
            "launches": r.get("launches"), "alg_bytes_per_launch": r.get("alg_bytes_per_launch")}

    # the FP16 matrix cores' sustained rate on this device, under the GEMM's
    # instruction mix, and the clock held meanwhile: the attainable ceiling
    # behind the published peak (which assumes the 2.4 GHz peak clock)
    if roof["kernel"] in ("gemm_f16x3", "gemm_f16x3_wide") and roof["achieved"]:
        tf16, mhz = sm.linalg.mfma_f16_rate(iters=10000, reps=3)
        roof["sustained"] = {"mfma_f16_tflops": tf16, "clock_mhz": mhz, "peak": tf16 / 3,
                             "frac": roof["achieved"] / (tf16 / 3),
                             "method": "48 independent 32x32x16 FP16 MFMAs per wave and slice from registers, "
                                       "one 512-thread workgroup per CU; best of 3 launches"}
        log(f"sustained FP16 MFMA rate {tf16:.0f} TFLOP/s at {mhz:.0f} MHz: split-FP16 ceiling {tf16 / 3:.0f} "
            f"TFLOP/s, wide GEMM at {roof['achieved'] / (tf16 / 3):.2f} of it")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(mesh, o, kind, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline skipped: {e}")
    line = {
        "metric": METRIC, "value": s_per_iter, "unit": "s/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": s_per_iter * 1e3, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": DTYPE[args.variant], "data": "synthetic (seeded parameters, generated mesh)",
        "config": {"workload": args.workload, "faces": mesh.nf, "edges": mesh.ne, "unknowns": N,
                   "quadrature_rule": rule, "params_per_face": kind, "rhs": 1, "lu_variant": args.variant,
                   "objective": "far-field power, 10 deg cone about +z", "ir_tol": args.ir_tol,
                   "ir_stop": "backward error", "ir_method": args.ir_method},
        "phases_s": {k: round(v, 4) for k, v in phases.items()},
        "refinement": ir_stats,
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
    }
    line.update(out_extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
