"""Problem setup and the time-to-FoS driver on the GPU.

``slope_2d`` / ``slope_3d`` build the homogeneous or layered slope problems of
the reference drivers (slope_stability_2D_homo_SSR.m:27-110,
slope_stability_3D_homo_SSR.m:27-110 with the structured generator):
P2 mesh, gradients, per-point materials (heterogenous_materials.m), gravity
load f_V = [0; -gamma(; 0)] (vector_volume_{2D,3D}.m).

``run_ssr`` runs the indirect continuation on the GPU with the DFGMRES
(or DCG) solver and returns the factor of safety, histories and timings.
"""
from __future__ import annotations

import time

import numpy as np

from . import fem, mesh
from ._lib import DeviceArray

# continuation / Newton / linear-solver parameters of slope_stability_2D_homo_SSR.m:114-137
PARAMS_2D = dict(lambda_init=0.9, d_lambda_init=0.05, d_lambda_min=1e-5, d_lambda_diff_scaled_min=0.001,
                 omega_max_stop=7e7, step_max=100, it_newt_max=30, it_damp_max=10, tol=1e-4, r_min=1e-4,
                 linear_tol=1e-1, linear_maxit=100, deflation_tol=1e-3)
# slope_stability_3D_homo_SSR.m:118-140
PARAMS_3D = dict(lambda_init=0.9, d_lambda_init=0.1, d_lambda_min=1e-3, d_lambda_diff_scaled_min=0.001,
                 omega_max_stop=7e7, step_max=100, it_newt_max=100, it_damp_max=10, tol=1e-4, r_min=1e-4,
                 linear_tol=1e-1, linear_maxit=100, deflation_tol=1e-3)

HOMO_2D = [dict(c0=6, phi=45, psi=0, young=40000, poisson=0.3, gamma_sat=20, gamma_unsat=20)]
# four horizontal layers of the 3D heterogeneous slope workload, bottom to
# top (kPa, degrees); equal thickness over the 30 m height
HETERO_3D_LAYERS = [dict(c0=50, phi=40, psi=0, young=100000, poisson=0.3, gamma_sat=20, gamma_unsat=20),
                    dict(c0=18, phi=32, psi=0, young=50000, poisson=0.3, gamma_sat=20, gamma_unsat=20),
                    dict(c0=10, phi=35, psi=0, young=20000, poisson=0.3, gamma_sat=20, gamma_unsat=20),
                    dict(c0=15, phi=30, psi=0, young=10000, poisson=0.3, gamma_sat=20, gamma_unsat=20)]
HETERO_3D_LEVELS = [7.5, 15.0, 22.5]


def _finish(coord, elem, Q, dim, mats, mat_id, device=False):
    if device:
        from .device_setup import finish_device
        return finish_device(coord, elem, Q, dim, mats, mat_id)
    DPhi, WEIGHT, HatP, WF = fem.gradients(coord, elem, dim)
    n_q = WF.size
    m = fem.materials_per_point(mat_id, n_q, mats)
    n_int = WEIGHT.size
    f_int = np.zeros((dim, n_int))
    f_int[1] = -m["gamma"]
    f = fem.volume_force(elem, coord.shape[1], HatP, WEIGHT, f_int)
    from .device_setup import material_table
    return dict(coord=coord, elem=elem, Q=Q, dim=dim, DPhi=DPhi, WEIGHT=WEIGHT, HatP=HatP, n_q=n_q,
                mats=m, f=f, n_int=n_int, n_dof_free=int(Q.sum()), material_table=material_table(mats),
                mat_id=np.asarray(mat_id))


def slope_2d(h=0.5, x1=15.0, x2=10.0, x3=15.0, y1=10.0, y2=10.0, materials=None, device=False):
    """2D homogeneous slope (10 m foundation, 10 m high 1:1 slope, 15 m crest
    and toe berms by default).  ``device``: gradients, fields, load and
    pattern built on the GPU (device_setup.finish_device)."""
    coord, elem, Q = mesh.mesh_P2_2D(h, x1, x2, x3, y1, y2)
    mats = materials or HOMO_2D
    return _finish(coord, elem, Q, 2, mats, np.zeros(elem.shape[1], dtype=np.int64), device=device)


def slope_3d(counts, x1, x2, x3, y1, y2, z, materials, layers=None, device=False):
    """3D slope from mesh_P2_3D with cell counts (nx1, nx2, nx3, ny1, ny2, nz).
    ``layers``: list of y-levels splitting the body into horizontal layers
    (material id 0 at the bottom), assigned by element centroid."""
    coord, elem, Q = mesh.mesh_P2_3D(*counts, x1, x2, x3, y1, y2, z)
    mat_id = np.zeros(elem.shape[1], dtype=np.int64)
    if layers:
        yc = coord[1, elem[:4]].mean(axis=0)
        for lv in sorted(layers):
            mat_id += (yc > lv)
    return _finish(coord, elem, Q, 3, materials, mat_id, device=device)


def make_constitutive(prob, Davis_type="B"):
    from .constitutive import CONSTITUTIVE
    m = prob["mats"]
    dim = prob["dim"]
    pairs = None
    if prob.get("device"):
        tab = prob["material_table"]
        pairs = (tab[:, 1], tab[:, 2])
    materials = None
    if prob.get("material_table") is not None and prob.get("mat_id") is not None:
        materials = (prob["mat_id"], prob["material_table"])
    cmb = CONSTITUTIVE(prob["elem"], prob["DPhi"], m["c0"], m["phi"], m["psi"], Davis_type, m["shear"], m["bulk"],
                       m["lame"], prob["WEIGHT"], 6 if dim == 3 else 3, prob["n_int"], dim, friction_pairs=pairs,
                       materials=materials)
    cmb.init_K_r_pattern(prob["Q"], pattern=prob.get("pattern"))
    if prob.get("coord_dev") is not None and not prob.get("stored_gradients"):
        # DPhi came from the device setup on these coordinates: strain and
        # force compute the gradients from them (same bits, less traffic)
        cmb.set_node_coordinates(prob["coord_dev"])
    return cmb


def load_vector(prob):
    """f as a device field (n_n, dim)."""
    f = prob["f"]
    return f if isinstance(f, DeviceArray) else DeviceArray.from_host(f.T)


def run_ssr(prob, params=None, solver="dfgmres", preconditioner="block_jacobi", log=None, predictor=None,
            amg_options=None, precision=64, recorder=None, driver="native"):
    """``predictor`` may also be a factory called as predictor(cmb, f) once the
    device problem exists (see learned.LearnedPredictor); ``recorder`` likewise
    (factory(cmb) or a hook).  ``driver``: "native" (the Newton loop inside
    libssr_hip.so, default) or "python" (newton.py); both give the same bits."""
    from .continuation import SSR_indirect_continuation
    from .linear_solvers import DCG, DFGMRES
    p = dict(PARAMS_2D if prob["dim"] == 2 else PARAMS_3D)
    p.update(params or {})
    t0 = time.perf_counter()
    cmb = make_constitutive(prob)
    if log:
        log(f"pattern and constitutive data ready ({time.perf_counter() - t0:.1f} s)")
    cls = DFGMRES if solver == "dfgmres" else DCG
    if preconditioner == "amg":
        amg_options = dict(amg_options or {})
        amg_options.setdefault("coord", prob["coord"])
    lin = cls(cmb.pattern, p["linear_tol"], p["linear_maxit"], p["deflation_tol"], preconditioner,
              amg_options=amg_options, precision=precision)
    f = load_vector(prob)
    if predictor is not None and getattr(predictor, "is_factory", False):
        predictor = predictor(cmb, f)
    if recorder is not None and getattr(recorder, "is_factory", False):
        recorder = recorder(cmb)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    U, lam_h, om_h, umax_h, stats = SSR_indirect_continuation(
        p["lambda_init"], p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["omega_max_stop"], p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], None, prob["Q"], f, cmb,
        lin, log=log, predictor=predictor, recorder=recorder, driver=driver)
    t_run = time.perf_counter() - t1
    timings = dict(cmb.timings)
    timings.update(lin.timings)
    counts = dict(cmb.counts)
    counts.update(lin.counters)
    return dict(FoS=float(lam_h[-1]), lambda_hist=lam_h, omega_hist=om_h, Umax_hist=umax_h, stats=stats,
                U=U, setup_s=t_setup, time_to_fos_s=t_run, timings=timings, counts=counts,
                newton_iterations=int(sum(stats["attempt_newton_iterations"]) + sum(stats["init_newton_iterations"])),
                predictor=predictor, recorder=recorder)


# slope_stability_3D_homo_SSR.m:118-140 (recorded run: ss/slope_stability_3D_homo_SSR.ipynb:366-410)
PARAMS_3D_NOTEBOOK = dict(PARAMS_3D, omega_max_stop=3000.0)


def slope_3d_mesh(data, materials=None, mat_id=None):
    """3D problem on a mesh in the reference's HDF5 layout (see
    mesh.load_mesh_P2), homogeneous unless ``mat_id`` is given."""
    coord, elem, surf, Q, material = mesh.load_mesh_P2(data)
    mats = materials or HOMO_2D
    mid = np.zeros(elem.shape[1], dtype=np.int64) if mat_id is None else np.asarray(mat_id)
    return _finish(coord, elem, Q, 3, mats, mid)


# the 2D parameters with a lower starting lambda: weak, steep members of the
# family have factors of safety below the reference drivers' 0.9
PARAMS_FAMILY = dict(PARAMS_2D, lambda_init=0.5)


def slope_2d_family(seed, h=1.0):
    """2D homogeneous slope of a seeded family: slope angle ~ U[26, 63] deg,
    c0 ~ U[5, 30] kPa, phi ~ U[20, 40] deg, psi = 0, E = 40 MPa, nu = 0.3,
    gamma = 20 kN/m^3; 10 m foundation, 10 m high slope, 15 m berms.
    Returns (problem, (slope angle, c0, phi))."""
    rng = np.random.default_rng(seed)
    beta = float(rng.uniform(26.0, 63.0))
    c0 = float(rng.uniform(5.0, 30.0))
    phi = float(rng.uniform(20.0, 40.0))
    y2 = 10.0
    x2 = y2 / np.tan(np.deg2rad(beta))
    mats = [dict(c0=c0, phi=phi, psi=0.0, young=40000.0, poisson=0.3, gamma_sat=20.0, gamma_unsat=20.0)]
    return slope_2d(h=h, x1=15.0, x2=x2, x3=15.0, y1=10.0, y2=y2, materials=mats), (beta, c0, phi)


def run_ssr_direct(prob, params=None, preconditioner="amg", log=None, amg_options=None, precision=64,
                   driver="native"):
    """Direct SSR continuation (SSR_direct_continuation.m) on the GPU: the
    second factor-of-safety estimate of the reference drivers."""
    from .continuation import SSR_direct_continuation
    from .linear_solvers import DFGMRES
    p = dict(PARAMS_2D if prob["dim"] == 2 else PARAMS_3D)
    p.update(params or {})
    t0 = time.perf_counter()
    cmb = make_constitutive(prob)
    if preconditioner == "amg":
        amg_options = dict(amg_options or {})
        amg_options.setdefault("coord", prob["coord"])
    lin = DFGMRES(cmb.pattern, p["linear_tol"], p["linear_maxit"], p["deflation_tol"], preconditioner,
                  amg_options=amg_options, precision=precision)
    f = load_vector(prob)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    U, lam_h, om_h, umax_h, work_h = SSR_direct_continuation(
        p["lambda_init"], p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], None, prob["Q"], f, cmb, lin, log=log,
        driver=driver)
    t_run = time.perf_counter() - t1
    timings = dict(cmb.timings)
    timings.update(lin.timings)
    return dict(FoS=float(lam_h[-1]), lambda_hist=lam_h, omega_hist=om_h, Umax_hist=umax_h, work_hist=work_h, U=U,
                setup_s=t_setup, time_to_fos_s=t_run, timings=timings, counts=dict(cmb.counts))


def run_batch(problems, workers=8, params=None, preconditioner="amg", precision=64, predictor_factories=None,
              log=None, driver="native"):
    """Independent SSR instances (of any sizes) solved concurrently: each
    worker thread drives its instances on a HIP stream of its own, so the
    kernels of different instances overlap on the device (the library
    releases the Python lock inside every call).  Returns the per-instance
    results (in input order; an instance whose continuation stops with an
    error gives FoS nan and the message) and the wall time of the whole
    batch.  ``driver`` as for run_ssr ("library": each instance's whole
    continuation is one library call, so a worker holds the Python lock only
    between instances; not with predictors)."""
    from concurrent.futures import ThreadPoolExecutor
    from ._lib import ThreadStream
    import threading

    local = threading.local()
    streams = []
    lock = threading.Lock()

    def solve(i):
        if not hasattr(local, "ts"):
            local.ts = ThreadStream().__enter__()
            with lock:
                streams.append(local.ts)
        pf = predictor_factories[i] if predictor_factories else None
        from .continuation import ContinuationError
        try:
            res = run_ssr(problems[i], params, preconditioner=preconditioner, precision=precision, predictor=pf,
                          driver=driver if pf is None else "native")
        except ContinuationError as exc:  # the instance's own outcome, reported like the reference's error
            res = dict(FoS=float("nan"), error=str(exc))
        if log:
            log(f"instance {i}: FoS {res['FoS']:.7f}" + (f" ({res['error']})" if "error" in res else ""))
        return res

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        results = list(ex.map(solve, range(len(problems))))
    wall = time.perf_counter() - t0
    for ts in streams:
        ts.__exit__(None, None, None)
    return results, wall


# slope_stability_2D_homo_LL.m:151-176
PARAMS_LL_2D = dict(d_t_min=1e-3, step_max=100, omega_max=2000.0, it_newt_max=100, it_damp_max=10, tol=1e-4,
                    r_min=1e-4, linear_tol=1e-1, linear_maxit=100, deflation_tol=1e-3)


def run_ll(prob, params=None, preconditioner="amg", log=None, amg_options=None, precision=64, elastic_tol=1e-10,
           driver="native"):
    """Limit-load analysis (slope_stability_2D_homo_LL.m:205-223): the elastic
    field at the unreduced material, d_omega_ini = omega_el / 5, the elastic
    field scaled by 1/5 as the start, then the LL indirect continuation.  The
    limit load is t_hist[-1].  ``driver``: "native" or "python" Newton loop."""
    from .continuation import LL_indirect_continuation
    from .linear_solvers import DFGMRES, dot, lincomb
    p = dict(PARAMS_LL_2D)
    p.update(params or {})
    t0 = time.perf_counter()
    cmb = make_constitutive(prob)
    if preconditioner == "amg":
        amg_options = dict(amg_options or {})
        amg_options.setdefault("coord", prob["coord"])
    lin = DFGMRES(cmb.pattern, p["linear_tol"], p["linear_maxit"], p["deflation_tol"], preconditioner,
                  amg_options=amg_options, precision=precision)
    f = load_vector(prob)
    cmb.reduction(1.0)
    # elastic solve (the reference solves it with its Krylov solver; here to elastic_tol)
    el = DFGMRES(cmb.pattern, elastic_tol, 1000, p["deflation_tol"], preconditioner, amg_options=amg_options)
    el.setup_preconditioner(cmb.K_elast)
    U_el = DeviceArray(f.shape)
    fQ = DeviceArray(f.shape)  # K_elast(Q,Q) U(Q) = f(Q): the right-hand side restricted to Q
    lincomb(1.0, f, 0.0, f, fQ)
    from ._lib import call as _call, ptr as _ptr
    _call("ssr_dev_mul", int(np.prod(f.shape)), _ptr(cmb.mask), _ptr(fQ))
    el.solve(cmb.K_elast, fQ, U_el)
    lin.expand_deflation_basis(U_el)
    omega_el = dot(f, U_el, cmb.mask)
    if not omega_el > 0:
        from .continuation import ContinuationError
        raise ContinuationError(f"elastic solve failed: f'U_elast = {omega_el:g} (relative residual "
                                f"{el.last_relres:g} after {el.last_iterations} iterations)")
    lincomb(0.2, U_el, 0.0, U_el, U_el)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    U, t_h, om_h, umax_h, its = LL_indirect_continuation(
        omega_el / 5, p["d_t_min"], p["step_max"], p["omega_max"], p["it_newt_max"], p["it_damp_max"], p["tol"],
        p["r_min"], None, U_el, prob["Q"], f, cmb, lin.copy(), log=log, driver=driver)
    t_run = time.perf_counter() - t1
    timings = dict(cmb.timings)
    timings.update(lin.timings)
    counts = dict(cmb.counts)
    counts.update(lin.counters)
    return dict(limit_load=float(t_h[-1]), t_hist=t_h, omega_hist=om_h, Umax_hist=umax_h, omega_el=omega_el, U=U,
                setup_s=t_setup, time_to_limit_s=t_run, newton_iterations=its, timings=timings, counts=counts)
