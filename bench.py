#!/usr/bin/env python3
"""Benchmark of the SSR hot path on one MI355X.

Workload (the constitutive + assembly microbenchmark): 3D box 40x20x20 m,
100x50x40 hexes x 6 P2 tetrahedra = 1.2M elements x 11 integration points =
13.2M points.  Trial strains N(0, 1e-3^2) per Voigt component with a seeded
30% subset scaled x5; c0 = 15 kPa, phi = 30 deg, psi = 0, E = 20 MPa,
nu = 0.3, lambda ~ U[1,2] per point, Davis B.

One step = one Mohr-Coulomb stress + consistent-tangent evaluation (S and DS)
over all 13.2M points through the MEX-layout kernel (point-major E/S/DS, five
per-point material arrays), inputs resident in HBM.  Throughput is reported in
Gpt/s (1e9 points/s).  The Davis reduction producing c_bar/sin_phi runs once
before the timed region (it is a separate, cheaper kernel; its own rate is
reported as an extra field).

Roofline: 424 algorithmic bytes per point (E 48 + 5 x 8 material + S 48 +
DS 288), HBM peak 8 TB/s.  CPU baseline: the reference's own per-point kernel
compiled from its sources (oracle/_ref) with OpenMP, timed on a bounded
sample on this host.

The metric's time-to-FoS half is the extra field `time_to_fos`: the ~1.5M-DOF
3D homogeneous slope solved to its factor of safety in the plain GPU form
(FP64) and in the chosen form (FP32 Krylov/AMG with FP64 refinement), with
the FoS of each, |dFoS|, the time cut and the gate flag (cut >= 30 % at
|dFoS| <= 1e-4), and the 2D plain-method configuration end to end on the
GPU and on this host's CPU cores (`config1_end_to_end`).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_HEX = (100, 50, 40)
N_Q = 11
BYTES_PER_POINT = 424
HBM_PEAK_GBS = 8000.0
METRIC = "SSR time-to-FoS (|ΔFoS|≤1e-4) chosen vs plain GPU; constitutive Gpt/s"


def cpu_baseline(d, cb, sp, budget_s=12.0, sample=1_320_000):
    """Time the reference kernel (or the oracle restatement) on a bounded
    sample of the same inputs with all host OpenMP threads."""
    from oracle import oracle as O

    O.build()
    kind = "reference" if O.ref_available() else "port"
    n = min(sample, d["E"].shape[0])
    E = np.ascontiguousarray(d["E"][:n].T)
    mats = (cb[:n], sp[:n], d["shear"][:n], d["bulk"][:n], d["lame"][:n])
    O.mc3d(E[:, :1000], *[m[:1000] for m in mats], use_ref=(kind == "reference"))
    reps, t0 = 0, time.perf_counter()
    while True:
        O.mc3d(E, *mats, use_ref=(kind == "reference"))
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 2000:
            break
    rate = reps * n / el / 1e9
    return {
        "value": rate,
        "unit": "Gpt/s",
        "cores": O.num_threads(),
        "kind": kind,
        "sample": f"{reps} x {n} points (first {n} of the {d['E'].shape[0]}-point workload), "
                  f"S+DS, OpenMP static schedule, {el:.1f} s",
        "host": _host_cpus(),
    }


def _host_cpus():
    """The CPUs this process may run on and the machine's count, beside the
    OpenMP thread count the CPU baseline used."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = None
    return {"cpus_available_to_process": avail, "cpus_on_machine": os.cpu_count(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def run_constitutive2d(args):
    """The 2D (plane-strain) return mapping with tangent, the kernel of the
    2D configurations: S (4) and DS (3 x 3) per point from E (3) and four
    material fields, MEX layout, on the microbenchmark's strain recipe
    (N(0, 1e-3) per component, a seeded 30 % scaled x5, lambda ~ U[1, 2],
    Davis B).  Timed at the microbenchmark's 13.2M points and at the point
    counts of the 2D configurations (h = 0.5 m: 70k points; the largest
    batch member, ~200k unknowns: 700k points), with HIP events on the
    library stream."""
    from oracle import oracle as O
    from slope_stability_octave_amd import _lib
    from slope_stability_octave_amd import constitutive_mex as cm
    from slope_stability_octave_amd import workloads as W
    from slope_stability_octave_amd._lib import DeviceArray, EventTimer

    _lib.require_gpu()
    n = args.points
    d = W.config2_points(n, dim=2)
    dE = DeviceArray.from_host(d["E"])
    dc0, dphi, dpsi, dlam = (DeviceArray.from_host(d[k]) for k in ("c0", "phi", "psi", "lam"))
    dG, dK, dL = (DeviceArray.from_host(d[k]) for k in ("shear", "bulk", "lame"))
    dcb, dsp = DeviceArray(n), DeviceArray(n)
    dS, dD = DeviceArray((n, 4)), DeviceArray((n, 9))
    cm.reduction_dev(n, "B", dc0, dphi, dpsi, dlam, True, dcb, dsp)
    tim = EventTimer()
    sizes = {}
    for m in sorted({n, min(n, 700_000), min(n, 70_000)}, reverse=True):
        row = {}
        for with_ds in (True, False):
            f = (lambda: cm.constitutive_dev(2, m, dE, dcb, dsp, dG, dK, dL, dS, dD)) if with_ds else \
                (lambda: cm.constitutive_dev(2, m, dE, dcb, dsp, dG, dK, dL, dS, None))
            for _ in range(args.warmup):
                f()
            _lib.synchronize()
            steps = args.steps if m == n else 10 * args.steps
            tim.start()
            for _ in range(steps):
                f()
            tim.stop()
            ms = tim.elapsed_ms() / steps
            nbytes = (24 + 32 + 32 + (72 if with_ds else 0)) * m
            row["S_DS" if with_ds else "S"] = {"ms": ms, "gpts": m / ms / 1e6, "GBs": nbytes / ms / 1e6,
                                               "frac_hbm": nbytes / ms / 1e6 / 8000.0}
        sizes[str(m)] = row
    big = sizes[str(n)]["S_DS"]
    # parity spot check against the oracle (not timed)
    k = min(n, 200_000)
    S0, D0, br0 = O.mc2d(np.ascontiguousarray(d["E"][:k].T), dcb.get()[:k], dsp.get()[:k], d["shear"][:k],
                         d["bulk"][:k], d["lame"][:k])
    cm.constitutive_dev(2, n, dE, dcb, dsp, dG, dK, dL, dS, dD)
    S, DS = dS.get()[:k].T, dD.get()[:k].T
    den = np.maximum(np.abs(D0).max(axis=0), 1e-300)
    parity = {"points_checked": k,
              "max_rel_S": float((np.abs(S - S0).max(axis=0) / np.maximum(np.abs(S0).max(axis=0), 1e-300)).max()),
              "max_rel_DS": float((np.abs(DS - D0).max(axis=0) / den).max()),
              "branch_histogram": np.bincount(br0, minlength=5).tolist()}
    cpu = None
    if not args.no_cpu_baseline:
        O.build()
        kind = "reference" if O.ref_available() else "port"
        ns = min(n, 4_000_000)
        E = np.ascontiguousarray(d["E"][:ns].T)
        cbh, sph = dcb.get()[:ns], dsp.get()[:ns]
        mats = (cbh, sph, d["shear"][:ns], d["bulk"][:ns], d["lame"][:ns])
        O.mc2d(E[:, :1000], *[x[:1000] for x in mats], use_ref=(kind == "reference"))
        reps, t0 = 0, time.perf_counter()
        while True:
            O.mc2d(E, *mats, use_ref=(kind == "reference"))
            reps += 1
            el = time.perf_counter() - t0
            if el >= 10.0 or reps >= 2000:
                break
        cpu = {"value": reps * ns / el / 1e9, "unit": "Gpt/s", "cores": O.num_threads(), "kind": kind,
               "sample": f"{reps} x {ns} points, S+DS, OpenMP static schedule, {el:.1f} s"}
    out = {
        "metric": "2D constitutive S+DS throughput (plane strain, MEX layout)",
        "value": big["gpts"], "unit": "Gpt/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": big["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D return mapping with tangent, {n} points"},
        "roofline": {"bound": "hbm", "achieved": big["GBs"], "peak": 8000.0, "unit": "GB/s",
                     "frac": big["frac_hbm"], "traffic": None},
        "cpu_baseline": cpu,
        "extra": {"sizes": sizes, "parity": parity,
                  "bytes_per_point": {"S_DS": 160, "S": 88}},
    }
    print(json.dumps(out))


def run_spmv(args):
    """The node-block SpMV of the time-to-FoS slope's operator (K_r of a
    plastic state; 60x30x20 m slope at --cells-per-metre, 1.3 by default:
    514,047 node rows, 14.4M 3x3 blocks) in double and in float (the float
    copy of the FP32 Krylov form), timed with HIP events on the library
    stream.  Algorithmic bytes per product: every block's values once, its
    column index, x once and y once."""
    from slope_stability_octave_amd import _lib, problems
    from slope_stability_octave_amd import linear_solvers as LS
    from slope_stability_octave_amd._lib import DeviceArray, EventTimer

    _lib.require_gpu()
    cpm = args.cells_per_metre if args.cells_per_metre != 1.0 else CONFIG3_CPM
    counts = tuple(int(round(c * cpm)) for c in (20, 20, 20, 10, 20, 20))
    prob = problems.slope_3d(counts, *GEOM_3D, problems.HOMO_2D, device=True)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    rng = np.random.default_rng(1)
    U = DeviceArray.from_host(rng.normal(size=(cmb.n_n, 3)) * 1e-3 * cmb.Q.T)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-3)
    P = cmb.pattern
    nv = P.nnzb * P.bs
    K32 = DeviceArray(nv, dtype=np.float32)
    _lib.call("ssr_dev_convert_f64_f32", nv, _lib.ptr(K), _lib.ptr(K32))
    if args.copy_calibration:
        # plain copy of the double operator (2 x 8 x nv bytes): the reference
        # point for scaling FETCH_SIZE / WRITE_SIZE counter readings
        dst = DeviceArray(nv)
        for _ in range(2):
            LS.axpby(1.0, K, 0.0, dst)
        _lib.synchronize()
        del dst
    tim = EventTimer()
    res = {}
    for bits in (64, 32):
        dt = np.float64 if bits == 64 else np.float32
        Kb = K if bits == 64 else K32
        x = DeviceArray.from_host(rng.normal(size=(P.n_n, 3)).astype(dt))
        y = DeviceArray((P.n_n, 3), dtype=dt)
        fn = "ssr_dev_spmv_bsr" if bits == 64 else "ssr_dev_spmv_bsr_f32"
        a = (3, P.n_n, _lib.ptr(P.d_row_ptr), _lib.ptr(P.d_col), _lib.ptr(Kb), _lib.ptr(x), _lib.ptr(y), 1.0, 0.0)
        for _ in range(args.warmup):
            _lib.call(fn, *a)
        _lib.synchronize()
        tim.start()
        for _ in range(args.steps):
            _lib.call(fn, *a)
        tim.stop()
        ms = tim.elapsed_ms() / args.steps
        w = 8 if bits == 64 else 4
        nbytes = P.nnzb * (9 * w + 4) + (P.n_n + 1) * 4 + 2 * 3 * P.n_n * w
        res[f"fp{bits}"] = {"ms": ms, "GBs": nbytes / ms / 1e6, "frac_hbm": nbytes / ms / 1e6 / HBM_PEAK_GBS,
                            "bytes": nbytes}
    out = {
        "metric": "node-block SpMV throughput (time-to-FoS slope operator)",
        "value": res["fp64"]["GBs"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["fp64"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"SpMV of K_r, 3D slope cells {counts}: {P.n_n} node rows, {P.nnzb} 3x3 blocks"},
        "roofline": {"bound": "hbm", "achieved": res["fp64"]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": res["fp64"]["frac_hbm"], "traffic": None},
        "cpu_baseline": None,
        "extra": res,
    }
    print(json.dumps(out))


CONFIG3_CPM = 1.3  # cells per metre: 1,492,582 free unknowns (the ~1.5M-DOF configuration)
GEOM_3D = (20.0, 20.0, 20.0, 10.0, 20.0, 20.0)


def _fos_summary(res):
    its = int(res["newton_iterations"])
    krylov = int(res["counts"].get("DFGMRES.solve.n_iters", 0))
    return {"time_to_fos_s": res["time_to_fos_s"], "FoS": res["FoS"], "steps": len(res["lambda_hist"]),
            "newton_iterations": its, "krylov_iterations": krylov,
            "ms_per_newton_iteration": 1e3 * res["time_to_fos_s"] / max(its, 1),
            "lambda_hist": [float(x) for x in res["lambda_hist"]],
            "omega_hist": [float(x) for x in res["omega_hist"]]}


def time_to_fos_leg(log, cpm=CONFIG3_CPM, lambda_init=0.7, linear_tol=None):
    """The metric's time-to-FoS half: the 3D homogeneous slope (60x30x20 m
    box, 20 m high 1:1 slope, structured P2 tetrahedra, ~1.5M free DOF, the
    material of the 2D slope: c = 6 kPa, phi = 45 deg, E = 40 MPa, nu = 0.3,
    gamma = 20 kN/m^3, Davis B) solved to its factor of safety by the indirect
    continuation with DFGMRES + AMG, in the plain GPU form (FP64 throughout,
    secant predictor) and in the chosen form (FP32 Krylov and AMG with FP64
    outer-residual refinement; constitutive branches, Newton residual and the
    continuation in FP64).  The continuation, Newton and Krylov parameters
    are the reference driver's (slope_stability_3D_homo_SSR.m:118-140: DFGMRES
    tolerance 1e-1), with lambda_init 0.7 below this mesh's factor of safety.
    Continuation time only; the setup (mesh generation on the host, FE data,
    material fields, load and pattern on the device) is reported separately,
    beside the same setup done by the host restatements."""
    from slope_stability_octave_amd import assembly, mesh, problems

    counts = tuple(int(round(c * cpm)) for c in (20, 20, 20, 10, 20, 20))
    t0 = time.perf_counter()
    coord, elem, Q = mesh.mesh_P2_3D(*counts, *GEOM_3D)
    mesh_gen_s = time.perf_counter() - t0
    mat_id = np.zeros(elem.shape[1], dtype=np.int64)
    t0 = time.perf_counter()
    host_prob = problems._finish(coord, elem, Q, 3, problems.HOMO_2D, mat_id)
    assembly.Pattern(elem, Q, 3, device=False)
    host_setup_s = time.perf_counter() - t0
    del host_prob
    prob = problems._finish(coord, elem, Q, 3, problems.HOMO_2D, mat_id, device=True)
    mesh_s = mesh_gen_s + prob["setup_s"]
    p = dict(problems.PARAMS_3D, lambda_init=lambda_init)
    if linear_tol is not None:
        p["linear_tol"] = linear_tol
    log(f"[bench] time-to-FoS: 3D slope cells {counts}, {prob['n_dof_free']} free DOFs (mesh {mesh_s:.1f} s)")
    runs = {}
    for name, prec in (("plain", 64), ("chosen", 32)):
        res = problems.run_ssr(prob, p, preconditioner="amg", precision=prec)
        runs[name] = (res, _fos_summary(res))
        log(f"[bench] {name}: FoS {res['FoS']:.7f}, {runs[name][1]['steps']} steps, "
            f"{runs[name][1]['newton_iterations']} Newton iterations, {res['time_to_fos_s']:.2f} s")
    (rp, sp_), (rc, sc) = runs["plain"], runs["chosen"]
    cut = 1.0 - rc["time_to_fos_s"] / rp["time_to_fos_s"]
    dfos = abs(rc["FoS"] - rp["FoS"])
    per_it_cut = 1.0 - sc["ms_per_newton_iteration"] / sp_["ms_per_newton_iteration"]
    return {
        "workload": f"3D homogeneous slope SSR, 60x30x20 m box with a 20 m high 1:1 slope, structured P2 "
                    f"tetrahedra (cells {counts}), {prob['n_dof_free']} free DOFs, {prob['n_int']} points, "
                    f"lambda_init {lambda_init}, DFGMRES tol {p['linear_tol']} + SA-AMG",
        "plain": sp_, "chosen": sc,
        "chosen_form": "FP32 Krylov/AMG with FP64 refinement, secant predictor",
        "time_cut": cut, "time_cut_per_newton_iteration": per_it_cut,
        "dFoS": dfos, "curve_max_dlambda": _curve_gap(rp, rc),
        "gate_pass": bool(cut >= 0.30 and dfos <= 1e-4),
        "setup_s": {"mesh_generation_host": mesh_gen_s,
                    "fe_materials_load_pattern_device": prob["setup_s"],
                    "fe_materials_load_pattern_host_restatement": host_setup_s,
                    "elastic_operator_and_solver_plain": rp["setup_s"],
                    "elastic_operator_and_solver_chosen": rc["setup_s"]},
    }


def config1_leg(log):
    """The plain method's CPU-runnable configuration (2D homogeneous slope,
    P2 h = 0.5, 20,840 free DOF, lambda_init 1) end to end: the GPU path
    (FP64, DFGMRES tol 1e-2 + SA-AMG) and the CPU restatement of the same
    Newton/continuation path (oracle/newton_path.py: FP64, OpenMP constitutive
    and assembly kernels, sparse direct solves) timed on this host."""
    from oracle import newton_path as NP
    from oracle import oracle as O
    from slope_stability_octave_amd import problems

    prob = problems.slope_2d(h=0.5)
    p = dict(problems.PARAMS_2D, lambda_init=1.0, linear_tol=1e-2)
    res = problems.run_ssr(prob, p, preconditioner="amg", precision=64)
    log(f"[bench] config 1 GPU: FoS {res['FoS']:.7f} in {res['time_to_fos_s']:.2f} s")
    O.build()
    cpu = NP.ConstitutiveCPU(prob["coord"], prob["elem"], prob["Q"], 2, prob["DPhi"], prob["WEIGHT"], prob["mats"])
    t0 = time.perf_counter()
    _, lam, om = NP.SSR_indirect_continuation(
        p["lambda_init"], p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["omega_max_stop"], p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], prob["f"], cpu)
    t_cpu = time.perf_counter() - t0
    log(f"[bench] config 1 CPU: FoS {lam[-1]:.7f} in {t_cpu:.1f} s")
    return {"workload": "2D homogeneous slope SSR, P2 h=0.5 (20,840 free DOF), lambda_init 1.0",
            "gpu": _fos_summary(res),
            "cpu": {"time_to_fos_s": t_cpu, "FoS": float(lam[-1]), "steps": int(len(lam)),
                    "cores": O.num_threads(), "kind": "port",
                    "note": "oracle/newton_path.py: OpenMP C constitutive and assembly kernels on `cores` "
                            "threads, single-threaded sparse LU per Newton system"},
            "dFoS": abs(res["FoS"] - float(lam[-1])), "speedup": t_cpu / res["time_to_fos_s"]}


def run_assembly(args):
    """Tangent assembly on the constitutive + assembly microbenchmark box
    (40x20x20 m, 100x50x40 hexes x 6 P2 tetrahedra = 1.2M elements, 11
    points each, 13.2M points) with DS from the return mapping of the
    benchmark's trial strains (Davis B at lambda ~ U[1,2]).  Times, with HIP
    events on the library stream: the tangent K_t into the node-block values
    (ssr_dev_assemble_bsr: element pair blocks, then the per-block gather)
    and the regularised K_r written once (ssr_dev_assemble_K_r).
    Algorithmic bytes per element: DS 36x11x8 + WEIGHT 11x8 + DPhi 3x10x11x8
    read, the block values written once (8 bytes x 9 x blocks / elements);
    K_r also reads the elastic operator once.  CPU baseline: the oracle's
    OpenMP element assembly (the reference's assemble_K_tangent_vals.c
    restated) on a 12,000-element box of the same element size."""
    from oracle import fe_oracle as FO
    from oracle import oracle as O
    from slope_stability_octave_amd import _lib, assembly as A, fem, mesh
    from slope_stability_octave_amd import constitutive_mex as cm
    from slope_stability_octave_amd import workloads as W
    from slope_stability_octave_amd._lib import DeviceArray, EventTimer

    _lib.require_gpu()
    log = (lambda s: print(s, file=sys.stderr, flush=True))
    nx, ny, nz = N_HEX
    t0 = time.perf_counter()
    coord, elem, Q = mesh.mesh_P2_box(nx, ny, nz, 40.0, 20.0, 20.0)
    DPhi, WEIGHT, _, WF = fem.gradients(coord, elem, 3)
    n_e, n_q, n_int = elem.shape[1], WF.size, WEIGHT.size
    P = A.Pattern(elem, Q, 3)
    log(f"[bench] box mesh, gradients and pattern: {n_e} elements, {P.nnzb} blocks, "
        f"{time.perf_counter() - t0:.1f} s")
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in DPhi]
    del DPhi
    dW = DeviceArray.from_host(WEIGHT)
    d = W.config2_points(n_int)
    dE = DeviceArray.from_host(d["E"])
    dc0, dphi, dpsi, dlam = (DeviceArray.from_host(d[k]) for k in ("c0", "phi", "psi", "lam"))
    dG, dK, dL = (DeviceArray.from_host(d[k]) for k in ("shear", "bulk", "lame"))
    dcb, dsp = DeviceArray(n_int), DeviceArray(n_int)
    cm.reduction_dev(n_int, "B", dc0, dphi, dpsi, dlam, True, dcb, dsp)
    dS, dDS = DeviceArray((n_int, 6)), DeviceArray((n_int, 36))
    cm.constitutive_dev(3, n_int, dE, dcb, dsp, dG, dK, dL, dS, dDS)
    del dE, dc0, dphi, dpsi, dlam, dS, dcb, dsp
    dDSe = DeviceArray((n_int, 36))
    A.ds_elastic_dev(3, n_int, dG, dK, dDSe)
    bsz = P.nnzb * P.bs
    Kraw, Kel = DeviceArray(bsz), DeviceArray(bsz)
    P.assemble(n_q, dD, dDSe, dW, Kraw)
    P.regularize(1.0, Kraw, Kraw, Kel)
    del dDSe
    Kt, Kr = Kraw, DeviceArray(bsz)
    if args.copy_calibration:
        from slope_stability_octave_amd import linear_solvers as LS
        for _ in range(2):
            LS.axpby(1.0, Kel, 0.0, Kr)  # plain copy of 8*bsz bytes: the counter calibration
        _lib.synchronize()
    tim = EventTimer()
    res = {}
    for name, fn in (("K_t", lambda: P.assemble(n_q, dD, dDS, dW, Kt)),
                     ("K_r", lambda: P.assemble_K_r(n_q, dD, dDS, dW, 1e-4, Kel, Kr))):
        fn()
        _lib.synchronize()
        tim.start()
        for _ in range(args.steps):
            fn()
        tim.stop()
        res[name] = tim.elapsed_ms() / args.steps
    blk_bytes = 8.0 * bsz / n_e
    per_elem = {"K_t": 36 * n_q * 8 + n_q * 8 + 3 * 10 * n_q * 8 + blk_bytes}
    per_elem["K_r"] = per_elem["K_t"] + blk_bytes
    # CPU baseline on a 12,000-element box of the same element size
    cpu = None
    if not args.no_cpu_baseline:
        c2, e2, Q2 = mesh.mesh_P2_box(20, 10, 10, 8.0, 4.0, 5.0)
        D2, W2, _, _ = fem.gradients(c2, e2, 3)
        n2 = W2.size
        d2 = W.config2_points(n2)
        cb2, sp2 = O.reduction(d2["c0"], d2["phi"], d2["psi"], d2["lam"], "B")
        _, DS2, _ = O.mc3d(np.ascontiguousarray(d2["E"].T), cb2, sp2, d2["shear"], d2["bulk"], d2["lame"])
        R2 = FO.RefPattern(e2, Q2, 3, D2, W2, d2["shear"], d2["bulk"])
        smap = R2.scatter_map(e2, c2.shape[1])
        O.assemble(3, D2, DS2, W2, smap, n_q, R2.nnz)
        reps, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 10.0:
            O.assemble(3, D2, DS2, W2, smap, n_q, R2.nnz)
            reps += 1
        el = time.perf_counter() - t1
        cpu = {"value": reps * e2.shape[1] / el / 1e6, "unit": "M elements/s", "cores": O.num_threads(),
               "kind": "port", "sample": f"{reps} x {e2.shape[1]} elements (20x10x10 hexes, same element size), "
                                         f"oracle element assembly with the int64 scatter map, {el:.1f} s"}
    out = {
        "metric": "tangent assembly throughput (config-2 box)",
        "value": n_e / (res["K_t"] * 1e-3) / 1e6,
        "unit": "M elements/s",
        "higher_is_better": True,
        "n_gpus": 1,
        "steps": args.steps,
        "ms_per_step": res["K_t"],
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D box 40x20x20 m, 100x50x40 hexes x 6 P2 tetrahedra, 11 points, DS of the "
                               "microbenchmark strains", "elements": n_e, "blocks": P.nnzb},
        "roofline": {"bound": "hbm", "achieved": per_elem["K_t"] * n_e / (res["K_t"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per_elem["K_t"] * n_e / (res["K_t"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "extra": {"K_r_ms": res["K_r"], "K_r_GBs": per_elem["K_r"] * n_e / (res["K_r"] * 1e-3) / 1e9,
                  "K_r_frac": per_elem["K_r"] * n_e / (res["K_r"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "bytes_per_element": per_elem,
                  "element_kernel": _lib.get_tuning(_lib.TUNE_ELEMENT_KERNEL)},
    }
    print(json.dumps(out))


def run_fos(args):
    """Time-to-FoS of the indirect SSR continuation on the GPU (plain FP64
    form): the 2D homogeneous slope of slope_stability_2D_homo_SSR.m or the 3D
    homogeneous slope of slope_stability_3D_homo_SSR.m on its SSR_homo_ada_L1
    mesh (continuation stopped at omega = 3000 as in the recorded run)."""
    from slope_stability_octave_amd import problems
    from slope_stability_octave_amd import _lib

    _lib.require_gpu()
    log = (lambda s: print(s, file=sys.stderr, flush=True)) if args.verbose else None
    if args.workload == "fos3d":
        import numpy as np
        mesh_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                 "mesh_SSR_homo_ada_L1.npz")
        with np.load(mesh_file) as z:
            prob = problems.slope_3d_mesh(dict(z))
        params = dict(problems.PARAMS_3D_NOTEBOOK)
        if args.omega_max != 7e7:
            params["omega_max_stop"] = args.omega_max
        name = "3D homogeneous slope SSR, P2 tetrahedra, mesh SSR_homo_ada_L1 (25591 nodes, 16615 elements)"
    elif args.workload in ("fos3d_slope", "fos3d_hetero"):
        n = args.cells_per_metre
        counts = tuple(int(round(c * n)) for c in (20, 20, 20, 10, 20, 20))
        hetero = args.workload == "fos3d_hetero"
        if log:
            log(f"building the mesh and FE data (cells {counts})")
        if hetero:
            prob = problems.slope_3d(counts, 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HETERO_3D_LAYERS,
                                     layers=problems.HETERO_3D_LEVELS, device=not args.host_setup)
        else:
            prob = problems.slope_3d(counts, 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HOMO_2D,
                                     device=not args.host_setup)
        if log:
            log(f"mesh ready: {prob['n_dof_free']} free DOFs, {prob['n_int']} integration points")
        params = dict(problems.PARAMS_3D)
        if args.omega_max != 7e7:
            params["omega_max_stop"] = args.omega_max
        name = (f"3D {'heterogeneous (4 horizontal layers)' if hetero else 'homogeneous'} slope SSR, 60x30x20 m box "
                f"with a 20 m high 1:1 slope, structured P2 tetrahedra (cells {counts})")
    elif args.workload == "fos2d_family":
        prob, mat = problems.slope_2d_family(args.seed, h=args.h)
        params = dict(problems.PARAMS_FAMILY)
        if args.omega_max != 7e7:
            params["omega_max_stop"] = args.omega_max
        name = f"2D slope family seed {args.seed} (angle, c, phi = {tuple(round(x, 2) for x in mat)}), P2 h={args.h}"
    else:
        prob = problems.slope_2d(h=args.h, device=not args.host_setup)
        params = dict(problems.PARAMS_2D)
        params["lambda_init"] = args.lambda_init if args.lambda_init is not None else 1.0
        params["omega_max_stop"] = args.omega_max
        name = f"2D homogeneous slope SSR, P2 triangles h={args.h}"
    if args.linear_tol is not None:
        params["linear_tol"] = args.linear_tol
    if args.lambda_init is not None:
        params["lambda_init"] = args.lambda_init
    if args.stored_gradients:
        prob["stored_gradients"] = True
    amg_opts = amg_options(args)
    res = problems.run_ssr(prob, params, preconditioner=args.precond, log=log, amg_options=amg_opts,
                           precision=args.precision, driver=args.newton_driver)
    out = {
        "metric": "SSR time-to-FoS (%s)" % ("plain GPU form, FP64" if args.precision == 64
                                             else "FP32 Krylov/AMG + FP64 refinement"),
        "value": res["time_to_fos_s"],
        "unit": "s",
        "higher_is_better": False,
        "n_gpus": 1,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": name,
                   "free_dof": prob["n_dof_free"], "int_points": prob["n_int"],
                   "lambda_init": params["lambda_init"], "omega_max_stop": params["omega_max_stop"],
                   "preconditioner": args.precond, "krylov_precision": args.precision,
                   "linear_tol": params["linear_tol"], "amg_options": dict(amg_opts or {}),
                   "newton_driver": args.newton_driver},
        "FoS": res["FoS"],
        "steps": len(res["lambda_hist"]),
        "newton_iterations": res["newton_iterations"],
        "lambda_hist": [float(x) for x in res["lambda_hist"]],
        "omega_hist": [float(x) for x in res["omega_hist"]],
        "setup_s": res["setup_s"],
        "device_setup_s": prob.get("setup_s"),
        "timings_s": res["timings"],
        "counts": res["counts"],
    }
    print(json.dumps(out))


def run_learned(args):
    """Chosen form (learned predictor + FP32 Krylov/AMG with FP64 refinement)
    against the plain GPU form (secant predictor, FP64) on held-out members of
    the seeded 2D slope family.  The network is trained first on trajectories
    the plain solver generates for other seeds (training time reported, not
    part of time-to-FoS)."""
    from slope_stability_octave_amd import _lib, learned, problems

    _lib.require_gpu()
    log = (lambda s: print(s, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    train_raw = learned.generate(range(args.train_seeds), h=args.h_train, log=log, linear_tol=args.linear_tol or 1e-2)
    val_raw = learned.generate(range(1000, 1004), h=args.h_train, log=log, linear_tol=args.linear_tol or 1e-2)
    gen_s = time.perf_counter() - t0
    net, norm, rep = learned.fit(train_raw, val_raw, epochs=args.epochs, log=log)
    log(f"[mlp] fit {rep}")
    rows = []
    tot = {"plain": 0.0, "chosen": 0.0}
    for seed in range(2000, 2000 + args.eval_seeds):
        prob, mat = problems.slope_2d_family(seed, h=args.h)
        p = dict(problems.PARAMS_FAMILY, linear_tol=args.linear_tol or 1e-2)
        r_plain = problems.run_ssr(prob, p, preconditioner="amg", precision=64)
        pred = learned.predictor_factory(net, norm, mat)
        r_ch = problems.run_ssr(prob, p, preconditioner="amg", precision=32, predictor=pred)
        pr = r_ch["predictor"]
        row = dict(seed=seed, material=[round(x, 3) for x in mat], free_dof=prob["n_dof_free"],
                   FoS_plain=r_plain["FoS"], FoS_chosen=r_ch["FoS"], dFoS=abs(r_plain["FoS"] - r_ch["FoS"]),
                   t_plain=r_plain["time_to_fos_s"], t_chosen=r_ch["time_to_fos_s"],
                   newton_plain=r_plain["newton_iterations"], newton_chosen=r_ch["newton_iterations"],
                   steps_plain=len(r_plain["lambda_hist"]), steps_chosen=len(r_ch["lambda_hist"]),
                   predictor_used=pr.used, predictor_rejected=pr.rejected)
        log(f"[eval] {row}")
        rows.append(row)
        tot["plain"] += row["t_plain"]
        tot["chosen"] += row["t_chosen"]
    out = {
        "metric": "SSR time-to-FoS reduction, learned predictor + FP32 Krylov/AMG vs plain FP64 GPU form",
        "value": 1.0 - tot["chosen"] / tot["plain"],
        "unit": "fraction of plain time-to-FoS saved",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "f64 (Newton, constitutive) / f32 (Krylov, AMG)",
        "data": "synthetic",
        "config": {"workload": f"2D slope family (angle U[26,63] deg, c U[5,30] kPa, phi U[20,40] deg), "
                               f"P2 h={args.h}, held-out seeds 2000..{1999 + args.eval_seeds}",
                   "train": f"{args.train_seeds} seeds at h={args.h_train}", "linear_tol": args.linear_tol or 1e-2},
        "t_plain_s": tot["plain"], "t_chosen_s": tot["chosen"],
        "max_dFoS": max(r["dFoS"] for r in rows),
        "newton_plain": sum(r["newton_plain"] for r in rows), "newton_chosen": sum(r["newton_chosen"] for r in rows),
        "training": dict(rep, data_generation_s=gen_s),
        "instances": rows,
    }
    print(json.dumps(out))


def run_learned3d(args):
    """Chosen form (learned predictor + FP32 Krylov/AMG with FP64 refinement)
    against the plain GPU form (secant predictor, FP64) on the 3D homogeneous
    slope at --cells-per-metre (1.3: about 1.5M DOF).  The network is trained
    first on trajectories of the plain solver on coarser meshes of the same
    slope (--train-cpm); training is reported, not part of time-to-FoS."""
    from slope_stability_octave_amd import _lib, learned, problems

    _lib.require_gpu()
    log = (lambda s: print(s, file=sys.stderr, flush=True))
    geom = (20.0, 20.0, 20.0, 10.0, 20.0, 20.0)
    mat = (45.0, 6.0, 45.0)  # 1:1 slope, c0, phi of the homogeneous material
    p = dict(problems.PARAMS_3D, linear_tol=args.linear_tol or 1e-2,
             lambda_init=args.lambda_init if args.lambda_init is not None else 0.7)

    def slope(cpm):
        counts = tuple(int(round(c * cpm)) for c in (20, 20, 20, 10, 20, 20))
        return counts, problems.slope_3d(counts, *geom, problems.HOMO_2D)

    def cases(cpms):
        for cpm in cpms:
            counts, prob = slope(cpm)
            yield f"3D slope cells {counts}", prob, mat, p
    t0 = time.perf_counter()
    cpms = [float(x) for x in args.train_cpm.split(",")]
    train_raw = learned.collect(cases(cpms[:-1]), log=log)
    val_raw = learned.collect(cases(cpms[-1:]), log=log)
    gen_s = time.perf_counter() - t0
    net, norm, rep = learned.fit(train_raw, val_raw, epochs=args.epochs, log=log)
    log(f"[mlp] fit {rep}")
    counts, prob = slope(args.cells_per_metre)
    log(f"evaluation mesh cells {counts}: {prob['n_dof_free']} free DOFs")
    r_plain = problems.run_ssr(prob, p, preconditioner="amg", precision=64, log=log)
    log(f"plain: FoS {r_plain['FoS']:.7f} time {r_plain['time_to_fos_s']:.2f} s")
    r_ch = problems.run_ssr(prob, p, preconditioner="amg", precision=32, log=log,
                            predictor=learned.predictor_factory(net, norm, mat))
    pr = r_ch["predictor"]
    out = {
        "metric": "SSR time-to-FoS reduction, learned predictor + FP32 Krylov/AMG vs plain FP64 GPU form",
        "value": 1.0 - r_ch["time_to_fos_s"] / r_plain["time_to_fos_s"],
        "unit": "fraction of plain time-to-FoS saved",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "f64 (Newton, constitutive) / f32 (Krylov, AMG)",
        "data": "synthetic",
        "config": {"workload": f"3D homogeneous slope SSR, 60x30x20 m box with a 20 m high 1:1 slope, "
                               f"structured P2 tetrahedra (cells {counts})",
                   "free_dof": prob["n_dof_free"], "lambda_init": p["lambda_init"],
                   "train": f"plain-solver trajectories at {cpms[:-1]} cells/m, validation at {cpms[-1:]}",
                   "linear_tol": p["linear_tol"]},
        "t_plain_s": r_plain["time_to_fos_s"], "t_chosen_s": r_ch["time_to_fos_s"],
        "FoS_plain": r_plain["FoS"], "FoS_chosen": r_ch["FoS"], "dFoS": abs(r_plain["FoS"] - r_ch["FoS"]),
        "newton_plain": r_plain["newton_iterations"], "newton_chosen": r_ch["newton_iterations"],
        "steps_plain": len(r_plain["lambda_hist"]), "steps_chosen": len(r_ch["lambda_hist"]),
        "predictor_used": pr.used, "predictor_rejected": pr.rejected,
        "curve_max_dlambda": _curve_gap(r_plain, r_ch),
        "lambda_plain": [float(x) for x in r_plain["lambda_hist"]],
        "omega_plain": [float(x) for x in r_plain["omega_hist"]],
        "lambda_chosen": [float(x) for x in r_ch["lambda_hist"]],
        "omega_chosen": [float(x) for x in r_ch["omega_hist"]],
        "training": dict(rep, data_generation_s=gen_s),
    }
    print(json.dumps(out))


def _curve_gap(ra, rb):
    """Largest |lambda_b(omega) - lambda_a(omega)| over b's accepted steps
    inside a's omega range (a's curve interpolated linearly): how far the two
    continuation curves are apart, independent of where each run stopped."""
    oa, la = np.asarray(ra["omega_hist"]), np.asarray(ra["lambda_hist"])
    ob, lb = np.asarray(rb["omega_hist"]), np.asarray(rb["lambda_hist"])
    sel = (ob >= oa[0]) & (ob <= oa[-1])
    return float(np.abs(np.interp(ob[sel], oa, la) - lb[sel]).max()) if sel.any() else float("nan")


def amg_options(args):
    """AMG options named on the command line (None: the defaults)."""
    opts = None
    for key, val in (("rebuild_every", args.amg_rebuild_every), ("theta", args.amg_theta),
                     ("cheb_degree", args.amg_cheb_degree), ("max_coarse", args.amg_max_coarse),
                     ("reuse_hierarchy", False if args.amg_per_step else None)):
        if val is not None:
            opts = dict(opts or {}, **{key: val})
    return opts


def run_ll(args):
    """Time to the limit load of the LL continuation (2D homogeneous slope,
    slope_stability_2D_homo_LL.m), plain GPU form."""
    from slope_stability_octave_amd import _lib, problems

    _lib.require_gpu()
    prob = problems.slope_2d(h=args.h)
    log = (lambda s: print(s, file=sys.stderr, flush=True)) if args.verbose else None
    amg_opts = amg_options(args)
    res = problems.run_ll(prob, {"linear_tol": args.linear_tol or 1e-2}, preconditioner=args.precond, log=log,
                          precision=args.precision, amg_options=amg_opts)
    out = {
        "metric": "LL time to limit load",
        "value": res["time_to_limit_s"],
        "unit": "s",
        "higher_is_better": False,
        "n_gpus": 1,
        "dtype": "f64" if args.precision == 64 else "f64 (Newton) / f32 (Krylov, AMG)",
        "data": "synthetic",
        "config": {"workload": f"2D homogeneous slope LL, P2 triangles h={args.h}", "free_dof": prob["n_dof_free"],
                   "amg_options": dict(amg_opts or {})},
        "limit_load": res["limit_load"],
        "t_hist": [float(x) for x in res["t_hist"]],
        "omega_hist": [float(x) for x in res["omega_hist"]],
        "newton_iterations": res["newton_iterations"],
        "setup_s": res["setup_s"],
        "timings_s": res["timings"],
        "counts": res["counts"],
    }
    print(json.dumps(out))


def run_batch(args):
    """Many independent 2D slope instances of uneven size (seeded family,
    h drawn from {1, 0.75, 0.5}) solved concurrently on per-thread streams,
    against the same instances one after another."""
    from slope_stability_octave_amd import _lib, problems

    _lib.require_gpu()
    rng = np.random.default_rng(7)
    if args.h_range:
        lo, hi = (float(x) for x in args.h_range.split(","))
        hs = rng.uniform(lo, hi, size=args.batch)
        sizes = f"h ~ U[{lo}, {hi}]"
    else:
        hs = rng.choice([1.0, 0.75, 0.5], size=args.batch)
        sizes = "h in {1, 0.75, 0.5}"
    probs = [problems.slope_2d_family(3000 + i, h=float(hs[i]))[0] for i in range(args.batch)]
    print(f"{args.batch} instances, {min(p['n_dof_free'] for p in probs)}..{max(p['n_dof_free'] for p in probs)} "
          f"free DOFs", file=sys.stderr, flush=True)
    p = dict(problems.PARAMS_FAMILY, linear_tol=args.linear_tol or 1e-2)
    log = (lambda s: print(s, file=sys.stderr, flush=True))
    drv = args.newton_driver
    if args.no_sequential:
        seq, t_seq = None, float("nan")
    else:
        seq, t_seq = problems.run_batch(probs, workers=1, params=p, precision=args.precision, log=log, driver=drv)
        log(f"one after another: {t_seq:.1f} s")
    par, t_par = problems.run_batch(probs, workers=args.workers, params=p, precision=args.precision, log=log,
                                    driver=drv)
    if seq is None:
        seq = par
    ok = [i for i in range(len(probs)) if "error" not in seq[i] and "error" not in par[i]]
    dfos = max((abs(seq[i]["FoS"] - par[i]["FoS"]) for i in ok), default=float("nan"))
    out = {
        "metric": "SSR instances per second, concurrent per-thread streams vs one after another",
        "value": args.batch / t_par,
        "unit": "instances/s",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "f64" if args.precision == 64 else "f64 (Newton) / f32 (Krylov, AMG)",
        "data": "synthetic",
        "config": {"workload": f"{args.batch} 2D slope-family instances, {sizes}",
                   "free_dof": [int(pr["n_dof_free"]) for pr in probs], "workers": args.workers,
                   "driver": drv},
        "wall_sequential_s": t_seq, "wall_concurrent_s": t_par, "speedup": t_seq / t_par,
        "max_dFoS_concurrent_vs_sequential": dfos,
        "stopped_with_error": [i for i in range(len(probs)) if "error" in seq[i] or "error" in par[i]],
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a kernel-form / launch-shape knob (ssr_set_tuning) before the run; repeatable")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=N_HEX[0] * N_HEX[1] * N_HEX[2] * 6 * N_Q)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-fos", action="store_true",
                    help="skip the time-to-FoS leg (plain vs chosen form on the ~1.5M-DOF 3D slope)")
    ap.add_argument("--workload", default="constitutive", choices=["constitutive", "constitutive2d", "spmv", "assembly", "fos2d", "fos3d", "fos3d_slope", "fos3d_hetero", "fos2d_family", "learned",
                             "learned3d", "batch", "ll2d"])
    ap.add_argument("--element-kernel", type=int, default=None, choices=[0, 1, 2],
                    help="element tangent kernel form (tuning key 8)")
    ap.add_argument("--h", type=float, default=0.5)
    ap.add_argument("--lambda-init", type=float, default=None,
                    help="first lambda of the continuation (default: 1.0 for fos2d, the parameter set's otherwise)")
    ap.add_argument("--omega-max", type=float, default=7e7)
    ap.add_argument("--precond", default="amg", choices=["amg", "block_jacobi"])
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--newton-driver", default="native", choices=["python", "native", "library"],
                    help="FoS workloads: Newton loop inside the library (default), in Python over the C ABI, "
                         "or the whole continuation inside the library (library)")
    ap.add_argument("--variant", type=int, default=None, help="3D constitutive kernel form (tuning key 0; 11 is the default)")
    ap.add_argument("--linear-tol", type=float, default=None)
    ap.add_argument("--amg-rebuild-every", type=int, default=None,
                    help="rebuild the AMG hierarchy every n-th Newton iteration (0: once per continuation step)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy of known size first (counter calibration)")
    ap.add_argument("--host-setup", action="store_true",
                    help="FoS workloads: FE data, material fields, load and pattern from the host restatements "
                         "instead of the device setup kernels")
    ap.add_argument("--stored-gradients", action="store_true",
                    help="FoS workloads with the device setup: strain and force read the stored DPhi instead of "
                         "computing each point's gradients from the node coordinates (same bits)")
    ap.add_argument("--precision", type=int, default=64, choices=[32, 64],
                    help="Krylov/preconditioner arithmetic of the FoS workloads (32: float with double refinement)")
    ap.add_argument("--amg-theta", type=float, default=None, help="AMG strength threshold (FoS workloads)")
    ap.add_argument("--amg-cheb-degree", type=int, default=None, help="AMG Chebyshev smoothing degree")
    ap.add_argument("--amg-max-coarse", type=int, default=None, help="AMG coarsest-level size")
    ap.add_argument("--seed", type=int, default=2000, help="member of the 2D slope family (fos2d_family)")
    ap.add_argument("--cells-per-metre", type=float, default=1.0, help="mesh density of fos3d_slope")
    ap.add_argument("--batch", type=int, default=16, help="instances of the batch workload")
    ap.add_argument("--h-range", default=None, help="batch: mesh sizes h ~ U[lo, hi] as 'lo,hi' (0.16,0.5: 20k-200k DOF)")
    ap.add_argument("--workers", type=int, default=8, help="host threads (streams) of the batch workload")
    ap.add_argument("--no-sequential", action="store_true", help="batch: skip the one-after-another baseline")
    ap.add_argument("--train-seeds", type=int, default=16)
    ap.add_argument("--train-cpm", default="0.5,0.6,0.7,0.8,0.9",
                    help="learned3d: cells per metre of the training meshes (the last one validates)")
    ap.add_argument("--eval-seeds", type=int, default=4)
    ap.add_argument("--h-train", type=float, default=1.0)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--amg-block-smooth", type=int, default=None, choices=[0, 1],
                    help="finest-level Chebyshev scaled by node-block (1) or scalar (0) diagonal inverses")
    ap.add_argument("--amg-per-step", action="store_true",
                    help="a new AMG hierarchy for every continuation step (default: one hierarchy shared by all "
                         "steps, rebuilt every --amg-rebuild-every setups)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the 3D kernel forms on this workload first and use the fastest")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="also time write-only, copy and read-only streaming kernels of the DS size")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the hot path is measured on one "
              "MI355X; the multi-GPU decomposition is a design note, not a launch mode)",
              file=sys.stderr)
        sys.exit(2)
    for kv in args.tune:
        from slope_stability_octave_amd import _lib as L
        k, v = (int(t) for t in kv.split("="))
        L.set_tuning(k, v)
    if args.variant is not None:
        from slope_stability_octave_amd import _lib as L
        L.set_tuning(L.TUNE_CONSTITUTIVE_VARIANT, args.variant)
    if args.amg_block_smooth is not None:
        from slope_stability_octave_amd import _lib as L
        L.set_tuning(L.TUNE_AMG_BLOCK_SMOOTH, args.amg_block_smooth)
    if args.workload in ("fos2d", "fos3d", "fos3d_slope", "fos3d_hetero", "fos2d_family"):
        return run_fos(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload == "learned3d":
        return run_learned3d(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "ll2d":
        return run_ll(args)
    if args.element_kernel is not None:
        from slope_stability_octave_amd import _lib as L
        L.set_tuning(L.TUNE_ELEMENT_KERNEL, args.element_kernel)
    if args.workload == "assembly":
        return run_assembly(args)
    if args.workload == "constitutive2d":
        return run_constitutive2d(args)
    if args.workload == "spmv":
        return run_spmv(args)

    from slope_stability_octave_amd import _lib
    from slope_stability_octave_amd import constitutive_mex as cm
    from slope_stability_octave_amd import workloads as W
    from slope_stability_octave_amd._lib import DeviceArray, EventTimer

    _lib.require_gpu()
    n = args.points
    t_gen = time.perf_counter()
    d = W.config2_points(n)
    print(f"[bench] generated {n} points in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    dE = DeviceArray.from_host(d["E"])
    dc0, dphi, dpsi, dlam = (DeviceArray.from_host(d[k]) for k in ("c0", "phi", "psi", "lam"))
    dG, dK, dL = (DeviceArray.from_host(d[k]) for k in ("shear", "bulk", "lame"))
    dcb, dsp = DeviceArray(n), DeviceArray(n)
    dS, dD = DeviceArray((n, 6)), DeviceArray((n, 36))

    if args.copy_calibration:
        # plain copy of 36*n doubles (algorithmic traffic 2*8*36*n bytes): the
        # reference point for scaling FETCH_SIZE / WRITE_SIZE counter readings
        from slope_stability_octave_amd import linear_solvers as LS
        src = dD.view(0, (36 * n,))
        dst = DeviceArray(36 * n)
        for _ in range(2):
            LS.axpby(1.0, src, 0.0, dst)
        _lib.synchronize()
        del dst

    tim = EventTimer()
    cm.reduction_dev(n, "B", dc0, dphi, dpsi, dlam, True, dcb, dsp)  # first launch loads the code object
    _lib.synchronize()
    tim.start()
    for _ in range(5):
        cm.reduction_dev(n, "B", dc0, dphi, dpsi, dlam, True, dcb, dsp)
    tim.stop()
    red_ms = tim.elapsed_ms() / 5
    tuned = None
    if args.autotune:
        from slope_stability_octave_amd.autotune import Autotuner
        tuned = Autotuner().constitutive(n, dE, dcb, dsp, dG, dK, dL, dS, dD)
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        cm.constitutive_dev(3, n, dE, dcb, dsp, dG, dK, dL, dS, dD)
    _lib.synchronize()

    tw0 = time.perf_counter()
    tim.start()
    for _ in range(args.steps):
        cm.constitutive_dev(3, n, dE, dcb, dsp, dG, dK, dL, dS, dD)
    tim.stop()
    kern_ms = tim.elapsed_ms() / args.steps
    _lib.synchronize()
    wall_ms = (time.perf_counter() - tw0) * 1e3 / args.steps
    print(f"[bench] {args.steps} steps: {kern_ms:.3f} ms/step (events), {wall_ms:.3f} ms/step (wall)",
          file=sys.stderr, flush=True)
    # stress-only variant (the S evaluations of the line search), 136 B/point
    tim.start()
    for _ in range(5):
        cm.constitutive_dev(3, n, dE, dcb, dsp, dG, dK, dL, dS, None)
    tim.stop()
    s_only_ms = tim.elapsed_ms() / 5

    # streaming ceilings of this device with plain kernels over the DS buffer
    # (36n doubles): write-only fill, copy (read + write) and read-only dot
    probe = None
    if args.hbm_probe:
        from slope_stability_octave_amd import linear_solvers as LS
        nb = 36 * n
        src = dD.view(0, (nb,))
        dst = DeviceArray(nb)
        res = {}
        res_dev = DeviceArray(1)

        def read_probe():
            _lib.call("ssr_dev_dot", nb, _lib.ptr(src), _lib.ptr(src), _lib.ptr(None), _lib.ptr(res_dev))

        for name, fn, nbytes in (("write", lambda: LS.fill(dst, 1.0), 8 * nb),
                                 ("copy", lambda: LS.axpby(1.0, src, 0.0, dst), 16 * nb),
                                 ("read", read_probe, 8 * nb)):
            fn()
            _lib.synchronize()
            tim.start()
            for _ in range(5):
                fn()
            tim.stop()
            res[name + "_GBs"] = nbytes / (tim.elapsed_ms() / 5 * 1e-3) / 1e9
        del dst
        probe = res

    # parity spot check on a slice against the oracle (not timed)
    parity = None
    try:
        from oracle import oracle as O
        m = min(n, 200_000)
        S = dS.get()[:m].T
        DS = dD.get()[:m].T
        cbh, sph = dcb.get()[:m], dsp.get()[:m]
        S0, D0, br0 = O.mc3d(np.ascontiguousarray(d["E"][:m].T), cbh, sph, d["shear"][:m],
                             d["bulk"][:m], d["lame"][:m])
        den = np.maximum(np.abs(D0).max(axis=0), 1e-300)
        parity = {
            "points_checked": m,
            "max_rel_S": float((np.abs(S - S0).max(axis=0) / np.abs(S0).max(axis=0)).max()),
            "max_rel_DS": float((np.abs(DS - D0).max(axis=0) / den).max()),
            "branch_histogram": np.bincount(br0, minlength=5).tolist(),
        }
        cpu = None if args.no_cpu_baseline else cpu_baseline(d, cbh if m == n else dcb.get(),
                                                             sph if m == n else dsp.get())
    except Exception as exc:  # the checker must not hide the measurement
        print(f"[bench] oracle check skipped: {exc}", file=sys.stderr)
        cpu = None

    ttf = None
    if not args.no_time_to_fos:
        del dE, dc0, dphi, dpsi, dlam, dG, dK, dL, dcb, dsp, dS, dD  # free the 13.2M-point arrays
        ttf = time_to_fos_leg(lambda s: print(s, file=sys.stderr, flush=True))
        if not args.no_cpu_baseline:
            try:
                ttf["config1_end_to_end"] = config1_leg(lambda s: print(s, file=sys.stderr, flush=True))
            except Exception as exc:  # the checker must not hide the measurement
                print(f"[bench] config-1 CPU leg skipped: {exc}", file=sys.stderr)

    gpts = n / (kern_ms * 1e-3) / 1e9
    achieved = BYTES_PER_POINT * n / (kern_ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "constitutive_sds_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": gpts,
        "unit": "Gpt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kern_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "constitutive microbench: 3D box 40x20x20 m, 100x50x40 hexes x 6 P2 tets, "
                        "11 points each (13.2M points), Mohr-Coulomb S+DS, Davis B, lambda~U[1,2]",
            "points": n,
            "layout": "point-major E/S/DS (MEX interface), per-point material arrays",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "extra": {
            "wall_ms_per_step": wall_ms,
            "reduction_ms": red_ms,
            "reduction_gpts": n / (red_ms * 1e-3) / 1e9,
            "stress_only_ms": s_only_ms,
            "stress_only_gpts": n / (s_only_ms * 1e-3) / 1e9,
            "stress_only_hbm_frac": 136 * n / (s_only_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "parity": parity,
            "hbm_probe": probe,
            "kernel_variant": _lib.get_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT),
            "autotune": tuned,
            "time_to_fos": ttf,
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
