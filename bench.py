"""Projection throughput on one MI355X.

A step is one full pressure projection (GFM assembly, multigrid-preconditioned
CG to max|div u| <= 1e-10, velocity correction) of the 256^3 two-phase case
(rho_l/rho_g = 1000, breaking-wave interface, no solid) with every input
already resident in HBM.  The headline value is mixed-precision (FP32 V-cycle
inside FP64 CG) projection solves/s; the all-FP64 solver is timed the same
way in the same run.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--no-cpu]
    python bench.py --all-configs     # configs 0-4 at full size, mixed and FP64,
                                      # plus a forward+adjoint pass (one JSON line each)
    python bench.py --config 4 --adjoint   # adds a timed forward+adjoint pass and its
                                           # per-kernel GB/s and fraction of HBM peak
"""

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    0: "64^3 two-phase, sphere solid (config 0)",
    1: "128^3 z-stretched two-phase, cylinder solid (config 1)",
    2: "256^3 two-phase rho 1000:1 breaking-wave interface, no solid (config 2)",
    3: "128^3 two-phase, sphere solid (config 3)",
    4: "512x256x256 wave tank, three submerged boxes (config 4)",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--size", type=int, nargs=3, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--levels", type=int, default=0)
    ap.add_argument("--no-levels4", action="store_true",
                    help="skip the timing with the configuration's stated 4 multigrid levels")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-array (PCIe-inclusive) timing")
    ap.add_argument("--all-configs", action="store_true",
                    help="configs 0-4 at full size: mixed and FP64 solves/s, the mixed "
                         "vs FP64 pressure error and a timed forward+adjoint pass")
    ap.add_argument("--adjoint", action="store_true",
                    help="time forward + adjoint passes of the mixed solver and report the "
                         "per-kernel GB/s of one profiled pair")
    ap.add_argument("--vtr", type=str, default=None, metavar="DIR",
                    help="after the timed run, write the mixed solution as PrintResultVTR3D's "
                         "VTR file and PrintResultVTK's binary file under DIR and report "
                         "the writers' times")
    ap.add_argument("--stream-probe", action="store_true",
                    help="measure the practical HBM bandwidth of plain streaming kernels "
                         "(torch copy and a*x+y on 2 GiB FP64 arrays) and report it")
    ap.add_argument("--geometry", action="store_true",
                    help="also time GridPreprocess (cut-cell geometry of the P, U, V, W grids) "
                         "at the same grid size on a three-sphere solid, with its CPU baseline")
    ap.add_argument("--predictor", action="store_true",
                    help="also time the momentum predictor (sp_predict_uvw) at the same grid "
                         "size, with its CPU baseline")
    ap.add_argument("--predictor-only", action="store_true",
                    help="time only the momentum predictor and print its record")
    ap.add_argument("--chebyshev", action="store_true",
                    help="also time the headline solve with the Chebyshev-Jacobi smoother "
                         "(degrees 2 and 4) instead of red-black Gauss-Seidel")
    ap.add_argument("--clsvof", action="store_true",
                    help="also time the interface transport (sp_clsvof_scheme) at the same grid "
                         "size, with its CPU baseline")
    ap.add_argument("--clsvof-only", action="store_true",
                    help="time only the interface transport and print its record")
    ap.add_argument("--geometry-only", action="store_true",
                    help="time only GridPreprocess (no projection) and print its record")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 1 GiB device-to-device copy (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    return ap.parse_args()


def to_device(case):
    import torch
    dev = {}
    for n, a in case.fields.items():
        # Fortran-ordered (I+2,J+2,K+2) == C-ordered (K+2,J+2,I+2): same bytes
        dev[n] = torch.from_numpy(a.T.copy()).cuda()
    return dev


def make_outputs(case, dev_like):
    import torch
    I, J, K = case.shape
    hs = (K + 2, J + 2, I + 2)
    return {n: torch.zeros(hs, dtype=torch.float64, device="cuda") for n in ("p", "u", "v", "w")}


def run_solves(P, dev, out, case, count):
    import torch
    stats = []
    for _ in range(count):
        st = P.project_dev(dev, case.bc_flag, case.dt, out)
        stats.append(st)
    torch.cuda.synchronize()
    return stats


def timed(P, dev, out, case, steps, live_kernel=None):
    """Wall time of `steps` solves.  With `live_kernel`, HIP events on the
    engine's stream bracket every launch of that one kernel inside the timed
    region (two event records per launch); returns its per-launch record."""
    import torch
    if live_kernel:
        P.set_profiling(True)
        P.set_profiling_filter(live_kernel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = run_solves(P, dev, out, case, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    live = None
    if live_kernel:
        rows = [r for r in P.kernel_times() if r["name"] == live_kernel]
        P.set_profiling_filter(None)
        P.set_profiling(False)
        if rows and rows[0]["launches"]:
            r = rows[0]
            # The host polls the convergence flag one CG iteration behind, so
            # each solve enqueues one more iteration whose kernels return at
            # once (a few us).  For a kernel launched once per iteration those
            # no-op launches are left out of the per-launch average.
            iters = sum(st.iterations for st in stats)
            work = working_launches(r["launches"], iters, len(stats))
            per_launch = r["bytes"] / r["launches"]
            avg_s = r["total_ms"] * 1e-3 / work
            live = dict(name=live_kernel, launches=r["launches"], work_launches=work,
                        avg_us=avg_s * 1e6, bytes_per_launch=per_launch,
                        gbs=per_launch / avg_s / 1e9)
    return (t1 - t0), stats, live


def working_launches(launches, iters, solves=1):
    """Launches that did work.  The host polls the convergence flag one CG
    iteration behind, so each solve enqueues one more iteration whose kernels
    return at once: a kernel launched m times per iteration shows
    m * (iters + solves) launches, of which m * iters worked."""
    per = iters + solves
    if iters > 0 and launches >= per and launches % per == 0:
        return launches // per * iters
    return launches


def kernel_table(P, dev, out, case):
    """Per-kernel device time from HIP events on the engine's stream (one
    profiled solve, untimed).  avg_us and gbs are per working launch (the
    same rule as the live roofline); the no-op launches' time is a few us."""
    P.set_profiling(True)
    st = run_solves(P, dev, out, case, 1)[0]
    rows = P.kernel_times()
    P.set_profiling(False)
    table = []
    for r in rows:
        if r["launches"] == 0:
            continue
        work = working_launches(r["launches"], st.iterations)
        avg_ms = r["total_ms"] / work
        per_launch = r["bytes"] / r["launches"]
        gbs = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 and per_launch > 0 else 0.0
        table.append(dict(name=r["name"], launches=r["launches"], work_launches=work,
                          total_ms=r["total_ms"], avg_us=avg_ms * 1e3,
                          bytes_per_launch=per_launch, gbs=gbs))
    table.sort(key=lambda t: -t["total_ms"])
    return table


def median(v):
    v = sorted(v)
    n = len(v)
    return 0.5 * (v[(n - 1) // 2] + v[n // 2]) if n else 0.0


def host_arrays(P, case, steps=3):
    """End-to-end rate through the C ABI with host arrays (sp_project): every
    input is copied host-to-device and p, u, v, w back, from pageable numpy
    arrays as a Fortran host would pass them.  Not the headline value."""
    import numpy as np
    I, J, K = case.shape
    hs = (I + 2, J + 2, K + 2)
    fields = {n: case.fields[n] for n in case.fields}
    out = {n: np.zeros(hs, order="F") for n in ("p", "u", "v", "w")}
    P.project(fields, case.bc_flag, case.dt, out=out)  # stages the device buffers
    t0 = time.perf_counter()
    sts = [P.project(fields, case.bc_flag, case.dt, out=out)[1] for _ in range(steps)]
    el = time.perf_counter() - t0
    # a host whose solid has not moved since the previous step: the geometry
    # stays on the device (sp_inputs.static_geometry)
    t1 = time.perf_counter()
    stg = [P.project(fields, case.bc_flag, case.dt, out=out, static_geometry=True)[1]
           for _ in range(steps)]
    el_s = time.perf_counter() - t1
    return dict(solves_per_s=round(steps / el, 3), ms_per_step=round(el / steps * 1e3, 3),
                device_ms_per_step=round(median([s.t_solve_ms + s.t_assemble_ms + s.t_correct_ms
                                                 for s in sts]), 3),
                static_geometry_solves_per_s=round(steps / el_s, 3),
                static_geometry_ms_per_step=round(el_s / steps * 1e3, 3),
                steps=steps, iterations=sts[-1].iterations, final_div=stg[-1].final_div,
                note="pageable host arrays in and out of sp_project; includes PCIe copies; "
                     "static_geometry: apertures, vof, vofO, cell types and MoExCell left on "
                     "the device from the previous call")


def face_masks(case):
    """1 on the faces a predictor may carry (interior, not wall-normal)."""
    import torch
    I, J, K = case.shape
    hs = (K + 2, J + 2, I + 2)
    ms = []
    for d in range(3):
        m = torch.zeros(hs, dtype=torch.float64, device="cuda")
        m[1:-1, 1:-1, 1:-1] = 1.0
        if d == 0:
            m[:, :, I] = 0.0
        elif d == 1:
            m[:, J, :] = 0.0
        else:
            m[K, :, :] = 0.0
        ms.append(m)
    return ms


def adjoint_buffers(case, out):
    """Cotangents U(0,1) on the corrected faces (a source with a large net
    flux) and the gradient outputs, as device tensors."""
    import torch
    I, J, K = case.shape
    gen = torch.Generator(device="cuda").manual_seed(11)
    g = {n: m * torch.rand(m.shape, generator=gen, device="cuda", dtype=torch.float64)
         for n, m in zip(("gu", "gv", "gw"), face_masks(case))}
    hs = out["p"].shape
    res = {n: torch.zeros(hs, dtype=torch.float64, device="cuda")
           for n in ("gu_star", "gv_star", "gw_star")}
    for n in ("gk_e", "gk_n", "gk_t"):
        res[n] = torch.zeros((K, J, I), dtype=torch.float64, device="cuda")
    return g, res


def forward_adjoint(P, dev, out, case, steps):
    """`steps` timed forward + adjoint pairs, then one profiled pair whose
    engine HIP-event timers give each kernel's device time, algorithmic
    GB/s and fraction of the HBM peak."""
    import torch
    g, res = adjoint_buffers(case, out)
    P.project_dev(dev, case.bc_flag, case.dt, out)
    P.adjoint_dev(g, res)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = P.project_dev(dev, case.bc_flag, case.dt, out)
        ast = P.adjoint_dev(g, res)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    P.set_profiling(True)
    P.project_dev(dev, case.bc_flag, case.dt, out)
    P.adjoint_dev(g, res)
    rows = P.kernel_times()
    P.set_profiling(False)
    # Each of the two solves enqueues one CG iteration past convergence whose
    # kernels return at once; kernels launched a whole number of times per
    # iteration count those no-op launches' bytes, which are taken out here.
    work = st.iterations + ast.iterations
    table = []
    for r in sorted(rows, key=lambda r: -r["total_ms"]):
        if not r["launches"]:
            continue
        nbytes = r["bytes"]
        if r["launches"] % (work + 2) == 0:
            nbytes *= work / (work + 2)
        gbs = nbytes / (r["total_ms"] * 1e-3) / 1e9 if nbytes and r["total_ms"] > 0 else 0.0
        table.append(dict(name=r["name"], launches=r["launches"], total_ms=round(r["total_ms"], 3),
                          avg_us=round(r["total_ms"] * 1e3 / r["launches"], 2),
                          gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4)))
    return dict(pairs_per_s=round(steps / el, 3), ms_per_pair=round(el / steps * 1e3, 3),
                forward_iterations=st.iterations, adjoint_iterations=ast.iterations,
                adjoint_converged=bool(ast.converged), kernels=table)


def stream_probe(reps=10):
    """Practical HBM bandwidth on this GPU: device-side time of torch's copy
    (16 B moved per element) and y = a*x + y (24 B) on 2^28-element FP64
    arrays, the best of `reps` launches each."""
    import torch
    n = 1 << 28
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.rand(n, dtype=torch.float64, device="cuda")
    z = torch.empty_like(x)
    res = {}
    for name, fn, nbytes in (("copy", lambda: z.copy_(x), 16 * n),
                             ("axpy", lambda: y.add_(x, alpha=0.5), 24 * n)):
        best = float("inf")
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        res[name + "_gbs"] = round(nbytes / best / 1e9, 1)
    del x, y, z
    torch.cuda.empty_cache()
    return res


def write_vtr(P, dev, out, case, directory):
    """sp_write_vtr and sp_write_vtk of the device-resident solution: wall time
    of each call and the GPU kernel's algorithmic GB/s (VTR: 8 B read, 17.2 B
    of text written per value)."""
    import numpy as np
    import torch
    x = torch.from_numpy(np.cumsum(case.dxp) - 0.5 * case.dxp).cuda()
    y = torch.from_numpy(np.cumsum(case.dyp) - 0.5 * case.dyp).cuda()
    z = torch.from_numpy(np.cumsum(case.dzp) - 0.5 * case.dzp).cuda()
    I, J, K = case.shape
    mres = torch.zeros((K, J, I), dtype=torch.float64, device="cuda")
    fields = dict(x=x, y=y, z=z, u=out["u"], v=out["v"], w=out["w"], p=out["p"],
                  phi=dev["phi_l"], phi_l=dev["phi_l"], vof=dev["vof"], vof_l=dev["vof_l"],
                  mres=mres)
    P.set_profiling(True)
    P.set_profiling_filter("vtr_format")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    path = P.write_vtr(directory, "case", 1, fields, device=True)
    el = time.perf_counter() - t0
    rows = [r for r in P.kernel_times() if r["name"] == "vtr_format"]
    P.set_profiling_filter(None)
    P.set_profiling(False)
    r = rows[0]
    size = os.path.getsize(path)
    res = dict(path=path, bytes=size, seconds=round(el, 3), mb_per_s=round(size / el / 1e6, 1),
               format_launches=r["launches"], format_ms=round(r["total_ms"], 3),
               format_gbs=round(r["bytes"] / (r["total_ms"] * 1e-3) / 1e9, 1))
    # PrintResultVTK's binary XML file of the same fields (sp_write_vtk):
    # base64 kernel GB/s over 6 B of doubles read and 4 B of text written per
    # 3 stream bytes
    dx, dy, dz = (torch.from_numpy(np.ascontiguousarray(a)).cuda()
                  for a in (case.dxp, case.dyp, case.dzp))
    kfields = dict(x=x, y=y, z=z, dx=dx, dy=dy, dz=dz, p=out["p"], u=out["u"], v=out["v"],
                   w=out["w"], phi_l=dev["phi_l"], phi=dev["phi_l"], vof_l=dev["vof_l"],
                   vof=dev["vof"])
    P.set_profiling(True)
    P.set_profiling_filter("vtk_base64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kpath = P.write_vtk(directory if directory.endswith("/") else directory + "/", 1, kfields,
                        device=True)
    el = time.perf_counter() - t0
    rows = [r for r in P.kernel_times() if r["name"] == "vtk_base64"]
    P.set_profiling_filter(None)
    P.set_profiling(False)
    r = rows[0]
    size = os.path.getsize(kpath)
    res["vtk"] = dict(path=kpath, bytes=size, seconds=round(el, 3),
                      mb_per_s=round(size / el / 1e6, 1), base64_launches=r["launches"],
                      base64_ms=round(r["total_ms"], 3),
                      base64_gbs=round(r["bytes"] / (r["total_ms"] * 1e-3) / 1e9, 1))
    return res


def predictor(shape, steps, warmup, cpu=True):
    """The momentum predictor (sp_predict_uvw, PredictingUVW with SpaceOrder
    1 and TimeOrder 1, then SpaceOrder 2) on device-resident arrays: the geometry of three
    spherical solids on a z-stretched grid is built on the GPU by the
    engine's own geometry update, velocities and BC values are seeded.
    Timed calls of the whole step (four launches), per-kernel GB/s from HIP
    events over the timed calls, and the oracle's predictor on the same
    inputs (copied to the host) as the CPU baseline."""
    import numpy as np
    import torch

    from simcoflow3d_amd import configs
    from simcoflow3d_amd.engine import GEOM_FIELDS, INTERP_FIELDS, Projector, geom_shapes

    I, J, K = shape
    tg = time.perf_counter()
    cells = configs.geometry_case(shape, "bodies", seed=5, stretch=True)
    P = Projector(shape)
    P0 = cells[0]
    P.set_grid_1d(P0["dx"][:, 0, 0], P0["dy"][0, :, 0], P0["dz"][0, 0, :],
                  cells[1]["dx"][:, 0, 0], cells[2]["dy"][0, :, 0], cells[3]["dz"][0, 0, :])
    sh = geom_shapes(I, J, K)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to("cuda")  # noqa: E731
    dev = []
    for c in cells:
        d = {}
        for n in GEOM_FIELDS:
            if n in c:
                d[n] = T(c[n])
            else:
                dt = torch.int32 if n == "cell_type" else torch.float64
                d[n] = torch.zeros(sh[n][::-1], dtype=dt, device="cuda")
        dev.append(d)
    del cells
    zi = lambda shp: torch.zeros(shp, dtype=torch.int32, device="cuda")  # noqa: E731
    hshape = (K + 2, J + 2, I + 2)
    interp = [dict(moex=zi((K, J, I)), msce=zi((3, K, J, I)),
                   **{n: torch.zeros(hshape, dtype=torch.float64, device="cuda")
                      for n in INTERP_FIELDS[2:]}) for _ in range(3)]
    P.grid_preprocess(dev, device=True)
    P.define_moex(*(d["cell_type"] for d in dev), *(x["moex"] for x in interp), device=True)
    P.new_cell_face(dev, interp, device=True)
    f = configs.predictor_fields(shape, seed=3, time_order=1, itt=2)
    geo_np = None
    if cpu:
        back = lambda t: np.asfortranarray(t.cpu().numpy().T)  # noqa: E731
        geo_np = ([{n: back(t) for n, t in d.items()} for d in dev],
                  [{n: back(t) for n, t in x.items()} for x in interp])
    # the predictor's view of the cells (device tensors): UCellO = UCell here
    pc = [dict() for _ in range(3)]
    for q in range(3):
        for n in configs.PRED_CELL_FIELDS:
            src = {"ee": "ee_area", "ne": "ne_area", "te": "te_area", "cent": "cell_cent"}.get(n, n)
            pc[q][n] = dev[q + 1][src] if src in dev[q + 1] else interp[q][src]
    pa = [dev[0]["ee_area"], dev[0]["ne_area"], dev[0]["te_area"]]
    fd = dict(f)
    fd["vel"] = [T(a) for a in f["vel"]]
    fd["vel_n"] = None
    fd["flux_div_old"] = T(f["flux_div_old"])
    fd["bcv"] = [T(a) for a in f["bcv"]]
    pred = [torch.zeros(hshape, dtype=torch.float64, device="cuda") for _ in range(3)]
    dp = [torch.zeros(hshape, dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    print(f"[bench] predictor inputs on the device ({time.perf_counter() - tg:.1f} s)",
          file=sys.stderr, flush=True)
    N = I * J * K

    def timed(space_order):
        fd["params"] = dict(f["params"], space_order=space_order)
        for _ in range(max(1, warmup)):
            P.predict_uvw(pa, pc, pc, fd, pred, dp, device=True)
        P.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            P.predict_uvw(pa, pc, pc, fd, pred, dp, device=True)  # synchronises its stream
        el = time.perf_counter() - t0
        rows = P.kernel_times()
        P.set_profiling(False)
        kern = []
        for r in rows:
            ms = r["total_ms"]
            gbs = r["bytes"] / (ms * 1e-3) / 1e9 if ms > 0 and r["bytes"] > 0 else None
            kern.append(dict(name=r["name"], launches=r["launches"],
                             avg_us=round(ms / max(r["launches"], 1) * 1e3, 1),
                             bytes_per_launch=r["bytes"] / max(r["launches"], 1),
                             gbs=round(gbs, 1) if gbs else None,
                             frac=round(gbs / HBM_PEAK_GBS, 4) if gbs else None))
        kern.sort(key=lambda t: -(t["avg_us"] * t["launches"]))
        # device time of the timed kernels per call (the wall time adds the
        # host's launch and synchronisation overhead)
        dev_ms = sum(r["total_ms"] for r in rows) / steps
        return el, kern, dev_ms

    el, kern, dev_ms = timed(1)
    rec = dict(workload=f"PredictingUVW, SpaceOrder 1, TimeOrder 1, {I}x{J}x{K}, three spheres, "
                        "z stretched 2% per cell, geometry from the engine's geometry update",
               calls=steps, ms_per_call=round(el / steps * 1e3, 3), cells_per_s=round(N * steps / el, 1),
               timed_kernel_ms_per_call=round(dev_ms, 3), kernels=kern)
    # SpaceOrder 2 (ModifiedConvectiveFlux) on the same inputs, second step
    el2, kern2, dev_ms2 = timed(2)
    rec["space_order_2"] = dict(calls=steps, ms_per_call=round(el2 / steps * 1e3, 3),
                                cells_per_s=round(N * steps / el2, 1),
                                timed_kernel_ms_per_call=round(dev_ms2, 3), kernels=kern2)
    P.close()
    del dev, interp, pc, pa, fd, pred, dp
    torch.cuda.empty_cache()
    if cpu:
        from oracle import orc

        class _G:
            pass
        g = _G()
        g.I, g.J, g.K = I, J, K
        geo, xs = geo_np
        g.dxp, g.dyp, g.dzp = (np.ascontiguousarray(geo[0][n][:, 0, 0] if n == "dx" else
                                                    (geo[0][n][0, :, 0] if n == "dy" else
                                                     geo[0][n][0, 0, :])) for n in ("dx", "dy", "dz"))
        g.dxu = np.ascontiguousarray(geo[1]["dx"][:, 0, 0])
        g.dyv = np.ascontiguousarray(geo[2]["dy"][0, :, 0])
        g.dzw = np.ascontiguousarray(geo[3]["dz"][0, 0, :])
        cc = configs.pred_cells(geo, xs)
        pa_h = [geo[0]["ee_area"], geo[0]["ne_area"], geo[0]["te_area"]]
        runs, secs = 0, 0.0
        while runs == 0 or (secs < 10.0 and runs < 20):
            t0 = time.perf_counter()
            orc.predict(g, pa_h, cc, cc, f["vel"], f["vel_n"], f["flux_div_old"], f["bc_flag"],
                        f["bcv"], f["params"])
            secs += time.perf_counter() - t0
            runs += 1
        rec["cpu_baseline"] = dict(value=round(runs * N / secs, 1), unit="cells/s",
                                   cores=orc.num_threads(), kind="port",
                                   sample=f"{runs} x the same {I}x{J}x{K} step",
                                   seconds=round(secs, 3))
        # SpaceOrder 2: the flux loop carries delh in loop order, one thread
        p2 = dict(f["params"], space_order=2)
        runs, secs = 0, 0.0
        while runs == 0 or (secs < 10.0 and runs < 20):
            t0 = time.perf_counter()
            orc.predict(g, pa_h, cc, cc, f["vel"], f["vel_n"], f["flux_div_old"], f["bc_flag"],
                        f["bcv"], p2)
            secs += time.perf_counter() - t0
            runs += 1
        rec["space_order_2"]["cpu_baseline"] = dict(
            value=round(runs * N / secs, 1), unit="cells/s", cores=orc.num_threads(),
            kind="port", sample=f"{runs} x the same {I}x{J}x{K} step (fluxes on one thread)",
            seconds=round(secs, 3))
    return rec


def clsvof_fields(shape):
    """A sphere of liquid (radius 0.15 at (0.35, 0.35, 0.35) in the unit
    cube; volume fractions from 2^3 sub-samples, level set its signed
    distance) in the deformation field of the reference's AdvectionTest
    (Test/AdvectionTest/AdvectionTest.f90:186-196) at t = 0, with Neumann
    level-set and vof boundary conditions.  numpy arrays in the layouts of
    sp_clsvof_scheme."""
    import numpy as np
    I, J, K = shape
    e = [np.linspace(0.0, 1.0, n + 1) for n in shape]
    c = [0.5 * (q[1:] + q[:-1]) for q in e]
    d = [np.diff(q) for q in e]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    DX, DY, DZ = np.meshgrid(*d, indexing="ij")
    f = np.asfortranarray
    hs = (I + 2, J + 2, K + 2)

    def halo(a):
        h = np.zeros(hs, order="F")
        h[1:-1, 1:-1, 1:-1] = a
        return h
    r = np.sqrt((X - 0.35) ** 2 + (Y - 0.35) ** 2 + (Z - 0.35) ** 2)
    vof = np.zeros(shape)
    for a in (-0.25, 0.25):
        for b in (-0.25, 0.25):
            for g in (-0.25, 0.25):
                rr = np.sqrt((X + a * DX - 0.35) ** 2 + (Y + b * DY - 0.35) ** 2
                             + (Z + g * DZ - 0.35) ** 2)
                vof += (rr < 0.15) * 0.125
    hv = halo(vof)
    hv[0], hv[-1] = hv[1], hv[-2]
    hv[:, 0], hv[:, -1] = hv[:, 1], hv[:, -2]
    hv[:, :, 0], hv[:, :, -1] = hv[:, :, 1], hv[:, :, -2]
    u = 2.0 * np.sin(np.pi * X) ** 2 * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    v = -np.sin(2 * np.pi * X) * np.sin(np.pi * Y) ** 2 * np.sin(2 * np.pi * Z)
    w = -np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(np.pi * Z) ** 2
    hu, hvv, hw = halo(u), halo(v), halo(w)
    dt = 0.4 * d[0][0] / 2.0
    faces = [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]
    inp = dict(u=hu, v=hvv, w=hw, x=f(X), y=f(Y), z=f(Z), dx=f(DX), dy=f(DY), dz=f(DZ),
               nx=np.zeros(hs, order="F"), ny=np.zeros(hs, order="F"),
               nz=np.ones(hs, order="F"), bcu_w=f(hu[0, 1:-1, 1:-1]),
               bcv_s=f(hvv[1:-1, 0, 1:-1]), bcw_b=f(hw[1:-1, 1:-1, 0]),
               lvs_flag=[1] * 6, lvs_const=[0.0] * 6, vof_flag=[1] * 6, vof_const=[0.0] * 6,
               dt=dt)
    state = dict(vof_l=hv, phi_l=halo(r - 0.15), nx_l=np.zeros(hs, order="F"),
                 ny_l=np.zeros(hs, order="F"), nz_l=np.zeros(hs, order="F"),
                 lvs_var=[np.zeros(q, order="F") for q in faces],
                 vof_var=[np.zeros(q, order="F") for q in faces])
    return inp, state


def clsvof(shape, steps, warmup, cpu=True):
    """Clsvof_Scheme (sp_clsvof_scheme) on device-resident arrays: one call is
    the whole step of six sub-steps (60 launches).  Timed calls, per-kernel
    GB/s from HIP events over the timed calls, and the oracle's
    Clsvof_Scheme on a bounded sample (the same workload at 96^3 at most) as
    the CPU baseline."""
    import numpy as np
    import torch

    from simcoflow3d_amd.engine import Projector

    I, J, K = shape
    tg = time.perf_counter()
    inp, st = clsvof_fields(shape)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda")  # noqa: E731
    dinp = {n: (T(a) if isinstance(a, np.ndarray) else a) for n, a in inp.items()}
    dst = {n: T(st[n]) for n in ("vof_l", "phi_l", "nx_l", "ny_l", "nz_l")}
    dst["lvs_var"] = [T(a) for a in st["lvs_var"]]
    dst["vof_var"] = [T(a) for a in st["vof_var"]]
    del inp, st
    P = Projector(shape)
    torch.cuda.synchronize()
    print(f"[bench] clsvof inputs on the device ({time.perf_counter() - tg:.1f} s)",
          file=sys.stderr, flush=True)
    for _ in range(max(1, warmup)):
        P.clsvof_scheme(dinp, dst, device=True)
    P.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        P.clsvof_scheme(dinp, dst, device=True)  # synchronises its stream
    el = time.perf_counter() - t0
    rows = P.kernel_times()
    P.set_profiling(False)
    P.close()
    N = I * J * K
    kern = []
    for r in rows:
        ms = r["total_ms"]
        gbs = r["bytes"] / (ms * 1e-3) / 1e9 if ms > 0 and r["bytes"] > 0 else None
        kern.append(dict(name=r["name"], launches=r["launches"],
                         avg_us=round(ms / max(r["launches"], 1) * 1e3, 1),
                         bytes_per_launch=r["bytes"] / max(r["launches"], 1),
                         gbs=round(gbs, 1) if gbs else None,
                         frac=round(gbs / HBM_PEAK_GBS, 4) if gbs else None))
    kern.sort(key=lambda t: -(t["avg_us"] * t["launches"]))
    interface = float((dst["vof_l"] > 0).logical_and(dst["vof_l"] < 1).sum().item())
    rec = dict(workload=f"Clsvof_Scheme, {I}x{J}x{K}, a sphere of radius 0.15 in the "
                        "AdvectionTest deformation field, CFL 0.4, six sub-steps per call",
               calls=steps, ms_per_call=round(el / steps * 1e3, 3),
               cells_per_s=round(N * steps / el, 1), interface_cells=int(interface),
               kernels=kern)
    del dinp, dst
    torch.cuda.empty_cache()
    if cpu:
        from oracle import orc
        cs = tuple(min(n, 96) for n in shape)
        inp, st = clsvof_fields(cs)
        runs, secs = 0, 0.0
        while runs == 0 or (secs < 10.0 and runs < 20):
            t0 = time.perf_counter()
            st = orc.clsvof_scheme(inp, st)
            secs += time.perf_counter() - t0
            runs += 1
        n = cs[0] * cs[1] * cs[2]
        rec["cpu_baseline"] = dict(value=round(runs * n / secs, 1), unit="cells/s",
                                   cores=orc.num_threads(), kind="port",
                                   sample=f"{runs} consecutive steps at {cs[0]}x{cs[1]}x{cs[2]}",
                                   seconds=round(secs, 3))
    return rec


def geometry(shape, steps, warmup, cpu=True):
    """GridPreprocess (sp_grid_preprocess_dev) on device-resident arrays of
    the four staggered grids, a stretched grid with three spherical solids.
    Timed calls of the whole preprocessing (all launches of the four grids),
    then of the whole geometry update after a solid move; per-kernel GB/s
    from one profiled update; the oracle's GridPreprocess timed on a bounded
    sample of the same workload as the CPU baseline."""
    import numpy as np
    import torch

    from simcoflow3d_amd import configs
    from simcoflow3d_amd.engine import GEOM_FIELDS, Projector, geom_shapes

    I, J, K = shape
    tg = time.perf_counter()
    cells = configs.geometry_case(shape, "bodies", seed=5, stretch=True)
    sh = geom_shapes(I, J, K)
    dev = []
    for c in cells:
        d = {}
        for n in GEOM_FIELDS:
            if n in c:
                d[n] = torch.from_numpy(np.ascontiguousarray(c[n].T)).to("cuda")
            else:
                dt = torch.int32 if n == "cell_type" else torch.float64
                d[n] = torch.zeros(sh[n][::-1], dtype=dt, device="cuda")
        dev.append(d)
    # pristine copies of the arrays the call updates, restored before each call
    keep = [{n: d[n].clone() for n in ("vof", "vof_l", "cell_type")} for d in dev]
    del cells
    torch.cuda.synchronize()
    print(f"[bench] geometry inputs on the device ({time.perf_counter() - tg:.1f} s)",
          file=sys.stderr, flush=True)

    def restore():
        for d, k in zip(dev, keep):
            for n, t in k.items():
                d[n].copy_(t)

    P = Projector(shape)
    for _ in range(warmup):
        restore()
        P.grid_preprocess(dev, device=True)
    times = []
    for _ in range(steps):
        restore()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.grid_preprocess(dev, device=True)  # synchronises its stream
        times.append(time.perf_counter() - t0)
    # the whole geometry update after a solid move (SolidBody.f90:815-824):
    # GridPreprocess, DefineMomentumExchangeCell, NumberExternalCell x 4,
    # NewCellFace, all on device-resident arrays
    from simcoflow3d_amd.engine import INTERP_FIELDS
    zi = lambda shp: torch.zeros(shp, dtype=torch.int32, device="cuda")
    hshape = (K + 2, J + 2, I + 2)
    interp = [dict(moex=zi((K, J, I)), msce=zi((3, K, J, I)),
                   **{n: torch.zeros(hshape, dtype=torch.float64, device="cuda")
                      for n in INTERP_FIELDS[2:]}) for _ in range(3)]
    posnu = [zi((K - iw, J - iv, I - iu)) for iu, iv, iw in
             ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))]

    def update():
        P.grid_preprocess(dev, device=True)
        for x in interp:
            x["moex"].zero_()
        P.define_moex(*(d["cell_type"] for d in dev), *(x["moex"] for x in interp), device=True)
        for q, (iu, iv, iw) in enumerate(((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))):
            P.number_external(dev[q]["cell_type"], iu, iv, iw, posnu=posnu[q], device=True)
        P.new_cell_face(dev, interp, device=True)

    for _ in range(max(1, warmup)):
        restore()
        update()
    utimes = []
    for _ in range(steps):
        restore()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        update()
        torch.cuda.synchronize()
        utimes.append(time.perf_counter() - t0)
    restore()
    P.set_profiling(True)
    update()
    rows = P.kernel_times()
    P.set_profiling(False)
    P.close()
    del interp, posnu
    N = I * J * K
    kern = []
    for r in rows:
        ms = r["total_ms"]
        gbs = r["bytes"] / (ms * 1e-3) / 1e9 if ms > 0 and r["bytes"] > 0 else None
        kern.append(dict(name=r["name"], launches=r["launches"], total_ms=round(ms, 3),
                         avg_us=round(ms / max(r["launches"], 1) * 1e3, 1),
                         bytes_per_launch=r["bytes"] / max(r["launches"], 1),
                         gbs=round(gbs, 1) if gbs else None))
    kern.sort(key=lambda t: -t["total_ms"])
    el = sum(times) / len(times)
    rec = dict(workload=f"GridPreprocess of the P, U, V, W grids, {I}x{J}x{K}, three spheres, "
                        "z stretched 2% per cell (kernel table: one profiled solid update)",
               calls=steps, ms_per_call=round(el * 1e3, 3),
               median_ms=round(median(times) * 1e3, 3),
               cells_per_s=round(N / el, 1),
               solid_update=dict(what="GridPreprocess + DefineMomentumExchangeCell + "
                                      "NumberExternalCell x 4 + NewCellFace",
                                 ms_per_call=round(sum(utimes) / len(utimes) * 1e3, 3),
                                 median_ms=round(median(utimes) * 1e3, 3)),
               kernels=kern)
    del dev, keep
    torch.cuda.empty_cache()
    rec["liquid_field"] = liquid_field(shape, steps, warmup, cpu)
    if cpu:
        from oracle import orc
        # bounded sample: the oracle on a 96^3 case of the same workload,
        # repeated until about 10 s of CPU work
        sample = (96, 96, 96)
        sc = configs.geometry_case(sample, "bodies", seed=5, stretch=True)
        runs, secs = 0, 0.0
        while runs == 0 or (secs < 10.0 and runs < 20):
            t0 = time.perf_counter()
            orc.grid_preprocess(sc)
            secs += time.perf_counter() - t0
            runs += 1
        ns = sample[0] * sample[1] * sample[2]
        rec["cpu_baseline"] = dict(value=round(runs * ns / secs, 1), unit="cells/s", cores=1,
                                   kind="port", seconds=round(secs, 3),
                                   sample=f"{runs} oracle GridPreprocess calls on the same "
                                          f"workload at {sample[0]}^3")
    return rec


def liquid_field(shape, steps, warmup, cpu=True):
    """ComputeUVWLiquidField (sp_compute_uvw_liquid_field, device arrays) on
    the same grid with a liquid level set of three spheres: timed calls,
    the kernel's algorithmic GB/s, and the oracle on a bounded sample."""
    import numpy as np
    import torch

    from simcoflow3d_amd import configs
    from simcoflow3d_amd.engine import LIQUID_FIELDS, Projector

    I, J, K = shape

    def case(sz):
        dzp = 1.02 ** np.arange(sz[2])
        grids = configs.staggered_spacings(np.full(sz[0], 1.0 / sz[0]),
                                           np.full(sz[1], 1.0 / sz[1]), dzp / dzp.sum())
        g = grids[0]
        phi, nx, ny, nz = configs.solid_level_set("bodies", g["x"], g["y"], g["z"],
                                                  np.random.default_rng(5))
        P = dict(dx=g["dx"], dy=g["dy"], dz=g["dz"], phi_l=phi, nx_l=nx, ny_l=ny, nz_l=nz,
                 vof_l=np.random.default_rng(6).uniform(size=sz))
        return [P] + [dict(dx=q["dx"], dy=q["dy"], dz=q["dz"]) for q in grids[1:]]

    cells = case(shape)
    dev = [{n: (torch.from_numpy(np.ascontiguousarray(np.asarray(c[n]).T)).to("cuda")
                if n in c else torch.zeros((K, J, I), dtype=torch.float64, device="cuda"))
            for n in LIQUID_FIELDS} for c in cells]
    del cells
    P = Projector(shape)
    for _ in range(max(1, warmup)):
        P.uvw_liquid_field(dev, device=True)
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.uvw_liquid_field(dev, device=True)
        times.append(time.perf_counter() - t0)
    P.set_profiling(True)
    P.uvw_liquid_field(dev, device=True)
    r = [x for x in P.kernel_times() if x["name"] == "liquid_uvw"][0]
    P.set_profiling(False)
    P.close()
    del dev
    torch.cuda.empty_cache()
    N = I * J * K
    el = sum(times) / len(times)
    rec = dict(workload=f"ComputeUVWLiquidField, {I}x{J}x{K}, three spherical drops, "
                        "z stretched 2% per cell",
               calls=steps, ms_per_call=round(el * 1e3, 3), cells_per_s=round(N / el, 1),
               kernel_us=round(r["total_ms"] * 1e3 / r["launches"], 1),
               kernel_gbs=round(r["bytes"] / (r["total_ms"] * 1e-3) / 1e9, 1),
               bytes_per_cell=256)
    if cpu:
        from oracle import orc
        sample = (96, 96, 96)
        sc = case(sample)
        runs, secs = 0, 0.0
        while runs == 0 or (secs < 10.0 and runs < 40):
            t0 = time.perf_counter()
            orc.uvw_liquid_field(sc)
            secs += time.perf_counter() - t0
            runs += 1
        rec["cpu_baseline"] = dict(value=round(runs * 96 ** 3 / secs, 1), unit="cells/s",
                                   cores=1, kind="port", seconds=round(secs, 3),
                                   sample=f"{runs} oracle calls on the same workload at 96^3")
    return rec


def all_configs(args):
    """Configs 0-4 at full size: one JSON line each."""
    import torch

    from simcoflow3d_amd import configs
    from simcoflow3d_amd.engine import Projector, SP_PREC_FP64, SP_PREC_MIXED

    steps = max(1, args.steps)
    for cfg in range(5):
        case = configs.generate(cfg)
        I, J, K = case.shape
        dev = to_device(case)
        out = make_outputs(case, dev)
        common = dict(rho_l=case.rho_l, rho_g=case.rho_g, rho_ref=case.rho_l, pin_i=1, pin_j=1,
                      pin_k=1, div_tol=1e-10)
        rec = {"config": cfg, "workload": WORKLOADS[cfg], "grid": [I, J, K]}
        ps = {}
        for name, prec in (("mixed", SP_PREC_MIXED), ("fp64", SP_PREC_FP64)):
            P = Projector(case.shape, precision=prec, **common)
            P.set_grid_1d(case.dxp, case.dyp, case.dzp, case.dxu, case.dyv, case.dzw)
            run_solves(P, dev, out, case, 2)
            el, sts, _ = timed(P, dev, out, case, steps)
            ps[name] = out["p"].clone()
            s = sts[-1]
            rec[name] = dict(solves_per_s=round(steps / el, 3), ms_per_step=round(el / steps * 1e3, 3),
                             median_ms_per_step=round(median([x.t_total_ms for x in sts]), 3),
                             iterations=s.iterations, final_div=s.final_div,
                             converged=bool(s.converged))
            # one forward + adjoint pass (reverse mode of the same projection)
            g, res = adjoint_buffers(case, out)
            P.adjoint_dev(g, res)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                P.project_dev(dev, case.bc_flag, case.dt, out)
                ast = P.adjoint_dev(g, res)
            torch.cuda.synchronize()
            ela = time.perf_counter() - t0
            rec[name]["forward_adjoint_per_s"] = round(steps / ela, 3)
            rec[name]["adjoint_iterations"] = ast.iterations
            rec[name]["adjoint_converged"] = bool(ast.converged)
            P.close()
        pm, p64 = ps["mixed"], ps["fp64"]
        rec["mixed_p_rel_err_vs_fp64"] = float((pm - p64).abs().max() / p64.abs().max())
        rec["mixed_speedup"] = round(rec["mixed"]["solves_per_s"] / rec["fp64"]["solves_per_s"], 3)
        print(json.dumps(rec), flush=True)
        del dev, out, ps
        torch.cuda.empty_cache()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only in this build)", file=sys.stderr)
        sys.exit(2)
    if args.all_configs:
        all_configs(args)
        return
    import numpy as np
    import torch

    from simcoflow3d_amd import configs
    from simcoflow3d_amd.engine import Projector, SP_PREC_FP64, SP_PREC_MIXED

    if args.copy_probe:
        # plain copy of known size: reads 2^30 B, writes 2^30 B
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cfg = args.config
    size = tuple(args.size) if args.size else None
    if args.predictor_only:
        pshape = size or configs.config_dims(cfg)
        print(json.dumps(dict(predictor=predictor(pshape, args.steps, args.warmup,
                                                  cpu=not args.no_cpu))), flush=True)
        return
    if args.clsvof_only:
        cshape = size or configs.config_dims(cfg)
        print(json.dumps(dict(clsvof=clsvof(cshape, args.steps, args.warmup,
                                            cpu=not args.no_cpu))), flush=True)
        return
    if args.geometry_only:
        gshape = size or configs.config_dims(cfg)
        print(json.dumps(dict(geometry=geometry(gshape, args.steps, args.warmup,
                                                cpu=not args.no_cpu))), flush=True)
        return
    tg = time.perf_counter()
    case = configs.generate(cfg, size=size)
    gen_s = time.perf_counter() - tg
    I, J, K = case.shape
    N = I * J * K
    print(f"[bench] generated {I}x{J}x{K} in {gen_s:.1f} s", file=sys.stderr, flush=True)
    dev = to_device(case)
    out = make_outputs(case, dev)
    torch.cuda.synchronize()
    print("[bench] inputs on the device", file=sys.stderr, flush=True)

    common = dict(rho_l=case.rho_l, rho_g=case.rho_g, rho_ref=case.rho_l, pin_i=1, pin_j=1,
                  pin_k=1, div_tol=1e-10, mg_levels=args.levels)
    Pm = Projector(case.shape, precision=SP_PREC_MIXED, **common)
    Pm.set_grid_1d(case.dxp, case.dyp, case.dzp, case.dxu, case.dyv, case.dzw)
    run_solves(Pm, dev, out, case, args.warmup)
    # per-kernel table from one untimed profiled solve: names the dominant kernel
    ktab = [] if args.no_profile else kernel_table(Pm, dev, out, case)
    dom_name = ktab[0]["name"] if ktab else None
    print("[bench] warm-up done", file=sys.stderr, flush=True)
    el, st_m, live = timed(Pm, dev, out, case, args.steps, dom_name)
    sm = st_m[-1]
    p_mixed = out["p"].cpu().numpy()
    value = args.steps / el
    ms_per_step = el / args.steps * 1e3
    median_ms = median([st.t_total_ms for st in st_m])
    host = None if args.no_host else host_arrays(Pm, case)
    fwd_adj = forward_adjoint(Pm, dev, out, case, max(1, args.steps // 2)) if args.adjoint else None
    vtr = write_vtr(Pm, dev, out, case, args.vtr) if args.vtr else None
    stream = stream_probe() if args.stream_probe else None
    Pm.close()
    geo = pred_rec = cls_rec = None
    if args.geometry or args.predictor or args.clsvof:
        del dev, out
        torch.cuda.empty_cache()
        if args.geometry:
            geo = geometry(case.shape, args.steps, args.warmup, cpu=not args.no_cpu)
        if args.predictor:
            pred_rec = predictor(case.shape, args.steps, args.warmup, cpu=not args.no_cpu)
        if args.clsvof:
            cls_rec = clsvof(case.shape, args.steps, args.warmup, cpu=not args.no_cpu)
        dev = to_device(case)
        out = make_outputs(case, dev)

    fp64 = {}
    if not args.no_fp64:
        P64 = Projector(case.shape, precision=SP_PREC_FP64, **common)
        P64.set_grid_1d(case.dxp, case.dyp, case.dzp, case.dxu, case.dyv, case.dzw)
        run_solves(P64, dev, out, case, max(1, args.warmup))
        el64, st_64, _ = timed(P64, dev, out, case, args.steps)
        p64 = out["p"].cpu().numpy()
        ktab64 = [] if args.no_profile else kernel_table(P64, dev, out, case)
        P64.close()
        s64 = st_64[-1]
        fp64 = dict(solves_per_s=args.steps / el64, ms_per_step=el64 / args.steps * 1e3,
                    median_ms_per_step=median([st.t_total_ms for st in st_64]),
                    iterations=s64.iterations, final_div=s64.final_div,
                    dominant=(ktab64[0]["name"] if ktab64 else None))
        err = float(np.max(np.abs(p_mixed - p64)) / max(np.max(np.abs(p64)), 1e-300))
        fp64["mixed_p_rel_err_vs_fp64"] = err
        fp64["mixed_speedup"] = value / fp64["solves_per_s"]
        fp64["kernels"] = [dict(name=k["name"], work_launches=k["work_launches"],
                                total_ms=round(k["total_ms"], 3), avg_us=round(k["avg_us"], 2),
                                gbs=round(k["gbs"], 1)) for k in ktab64]

    # the preconditioner as the configuration states it: 4 levels (256^3 ->
    # 32^3), 2 + 2 red-black Gauss-Seidel sweeps; the headline coarsens
    # while every dimension is >= 4 (8 levels at 256^3)
    lv4 = None
    if not args.no_levels4 and args.levels == 0:
        c4 = dict(common, mg_levels=4)
        P4 = Projector(case.shape, precision=SP_PREC_MIXED, **c4)
        P4.set_grid_1d(case.dxp, case.dyp, case.dzp, case.dxu, case.dyv, case.dzw)
        run_solves(P4, dev, out, case, max(1, args.warmup))
        el4, st_4, _ = timed(P4, dev, out, case, max(1, args.steps // 2))
        P4.close()
        lv4 = dict(mg_levels=st_4[-1].levels, solves_per_s=round(len(st_4) / el4, 3),
                   ms_per_step=round(el4 / len(st_4) * 1e3, 3), iterations=st_4[-1].iterations,
                   final_div=st_4[-1].final_div, converged=bool(st_4[-1].converged))

    # the Chebyshev-Jacobi smoother (SP_FLAG_CHEBYSHEV) in the same V-cycle,
    # degrees 2 and 4, beside the fused red-black Gauss-Seidel headline
    cheb = None
    if args.chebyshev:
        cheb = []
        for deg in (2, 4):
            cc = dict(common, flags=2, mg_pre=deg, mg_post=deg)
            PC = Projector(case.shape, precision=SP_PREC_MIXED, **cc)
            PC.set_grid_1d(case.dxp, case.dyp, case.dzp, case.dxu, case.dyv, case.dzw)
            run_solves(PC, dev, out, case, max(1, args.warmup))
            elc, st_c, _ = timed(PC, dev, out, case, max(1, args.steps // 2))
            PC.set_profiling(True)
            run_solves(PC, dev, out, case, 1)
            rows = PC.kernel_times()
            PC.set_profiling(False)
            PC.close()
            fine = [r for r in rows if r["name"] == "mg_cheb_fine"]
            rec = dict(degree=deg, solves_per_s=round(len(st_c) / elc, 3),
                       ms_per_step=round(elc / len(st_c) * 1e3, 3), iterations=st_c[-1].iterations,
                       final_div=st_c[-1].final_div, converged=bool(st_c[-1].converged))
            if fine and fine[0]["total_ms"] > 0:
                f = fine[0]
                rec["cheb_fine_avg_us"] = round(f["total_ms"] / f["launches"] * 1e3, 1)
                rec["cheb_fine_gbs"] = round(f["bytes"] / (f["total_ms"] * 1e-3) / 1e9, 1)
            cheb.append(rec)

    roof = None
    if live:
        dom = live
        roof = dict(bound="hbm", kernel=dom["name"], achieved=round(dom["gbs"], 1),
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=round(dom["gbs"] / HBM_PEAK_GBS, 4),
                    traffic=None, avg_us=round(dom["avg_us"], 2), launches=dom["launches"],
                    work_launches=dom["work_launches"], bytes_per_launch=dom["bytes_per_launch"])
        tfile = os.path.join(REPO, "profiles", "traffic.json")
        if os.path.exists(tfile):
            with open(tfile) as fh:
                tr = json.load(fh)
            key = f"{dom['name']}@{I}x{J}x{K}"
            if key in tr:
                roof["traffic"] = tr[key]

    cpu = None
    if not args.no_cpu:
        from oracle import orc
        ncores = orc.num_threads()
        if cfg == 2 and size is None:
            sample_case = case
            sample = "one full FP64 projection of the same 256^3 case"
        else:
            sample_case = case
            sample = f"one full FP64 projection of the same {I}x{J}x{K} case"
        # repeat until about 10 s of CPU work (at least one projection); the
        # oracle's multigrid with the engine's own number of levels
        runs, total = 0, 0.0
        while runs == 0 or (total < 10.0 and runs < 20):
            secs, info = orc.baseline_project(sample_case, levels=sm.levels)
            runs += 1
            total += secs
        cpu = dict(value=round(runs / total, 4), unit="projection solves/s", cores=ncores,
                   kind="port", sample=f"{runs} x {sample}", seconds=round(total, 3),
                   iterations=info["iterations"], div_inf=info["div_inf"])
        # the mixed GPU pressure against the FP64 oracle's of the same case
        # (both pinned at cell (1,1,1)); halo order (K+2, J+2, I+2) on the GPU
        p_orc = info["Pp"].T
        cpu["mixed_p_rel_err_vs_oracle"] = float(
            np.max(np.abs(p_mixed - p_orc)) / max(np.max(np.abs(p_orc)), 1e-300))

    line = {
        "metric": "projection solves/s at 256^3 to ||div u||<=1e-10, mixed vs FP64 GPU; HBM GB/s",
        "value": round(value, 3),
        "unit": "projection solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "median_ms_per_step": round(median_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer CG + f32 V-cycle",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": (WORKLOADS.get(cfg, f"config {cfg}") if size is None else
                                f"{I}x{J}x{K} variant of: " + WORKLOADS.get(cfg, f"config {cfg}")),
                   "grid": [I, J, K],
                   "cells": N, "precision": "mixed", "mg_levels": sm.levels,
                   "iterations": sm.iterations, "final_div": sm.final_div,
                   "converged": bool(sm.converged)},
        "fp64": fp64,
        "levels4": lv4,
        "chebyshev": cheb,
        "host_arrays": host,
        "forward_adjoint": fwd_adj,
        "vtr": vtr,
        "stream_probe": stream,
        "geometry": geo,
        "predictor": pred_rec,
        "clsvof": cls_rec,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": [dict(name=k["name"], launches=k["launches"], work_launches=k["work_launches"],
                         total_ms=round(k["total_ms"], 3), avg_us=round(k["avg_us"], 2),
                         gbs=round(k["gbs"], 1)) for k in ktab],
        "gen_s": round(gen_s, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
