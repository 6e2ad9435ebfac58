"""Benchmark: steady MH Darcy solves/s on one MI355X (+ explicit transport steps).

Workload (the mixed 1D/2D/3D configuration): unit cube, 55^3 cells x 6
tetrahedra (998,250 tets), 8 whole-plane fractures meshed conformingly by the
tetrahedron faces (48,400 triangles, cross-section 1e-3, K x 1e3), 200 random
channels on fracture-triangle sides (~4.6k segments, cross-section 1e-5, K x 1e3);
ln K Gaussian (mean 0, sigma 1, exponential covariance, correlation length 0.1,
seeded); Dirichlet head 1 at x=0 and 0 at x=1, no flow elsewhere; FP64.
One step = one full steady solve (element assembly + condensation to the
edge system, Jacobi-CG to relative residual 1e-12 from a zero start,
reconstruction of pressures and side fluxes), inputs resident in HBM.
After the timed solves: 4 substances of explicit upwind transport for 1000
CFL-limited steps on the computed flux field.

Prints one JSON line (see the repository's contract).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "svn-export_amd"))

import flowgpu as G  # noqa: E402
import meshgen  # noqa: E402

HBM_PEAK_GBS = 8000.0


def build_problem(n=55, seed=0, n_features=512):
    mesh = meshgen.cube_mesh(n, n_fractures=8, n_channels=200, channel_len=25, seed=seed)
    cen, bulk = meshgen.bulk_centroids(mesh)
    reg = mesh["elem_region"][bulk]
    K = meshgen.lognormal_field(cen, sigma=1.0, corr_len=0.1, seed=seed, n_features=n_features)
    cs = np.ones(len(K))
    fr = reg == meshgen.REGION_FRACTURE
    ch = reg == meshgen.REGION_CHANNEL
    K[fr] *= 1e3
    K[ch] *= 1e3
    cs[fr] = 1e-3
    cs[ch] = 1e-5
    # file boundary elements (.left then .right) come first; implicit ones have no BC
    bmask = ~bulk
    bnodes = mesh["elem_nodes"][bmask][:, :3]
    bcen = mesh["nodes"][bnodes].mean(1)
    breg = mesh["elem_region"][bmask]
    return mesh, K, cs, bcen, breg


def bc_arrays(n_bc, bcen, breg):
    bct = np.zeros(n_bc, np.int32)
    bcp = np.zeros(n_bc)
    nf = len(breg)
    left, right = breg == meshgen.REGION_LEFT, breg == meshgen.REGION_RIGHT
    bct[:nf][left | right] = 1
    # piezometric head h -> pressure h - z (FieldAddPotential with gravity (0,0,-1))
    bcp[:nf][left] = 1.0 - bcen[left, 2]
    bcp[:nf][right] = 0.0 - bcen[right, 2]
    return bct, bcp


def spmv_bytes(F, nnz):
    # values + column indices + row pointers + z, p_in read + p_out, q written
    return 12.0 * nnz + 4.0 * (F + 1) + 32.0 * F


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=55)
    ap.add_argument("--rtol", type=float, default=1e-12)
    ap.add_argument("--transport-steps", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cg-iters", type=int, default=50)
    ap.add_argument("--no-accuracy", action="store_true", help="skip the tightly converged reference solve")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.set_device(0)

    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    mesh, K, cs, bcen, breg = build_problem(args.n)
    gm = G.Mesh(arrays=mesh)
    E, F, S, NB = gm.n_elements, gm.n_edges, gm.n_sides, gm.n_bc_elements
    bct, bcp = bc_arrays(NB, bcen, breg)
    sigma, f = np.ones(E), np.zeros(E)
    zeros_b = np.zeros(NB)
    darcy = G.DarcyFlowMH_Steady(gm)
    darcy.set_data(K, cs, sigma, f, bct, bcp, zeros_b, zeros_b)
    nnz = G.lib().fl_darcy_edge_nnz(darcy.h)
    print("setup %.1fs: E=%d F=%d S=%d nnz=%d" % (time.time() - t0, E, F, S, nnz), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        darcy.update_solution(r_tol=args.rtol, a_tol=0.0, max_it=100000)
    torch.cuda.synchronize()
    infos = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        infos.append(darcy.update_solution(r_tol=args.rtol, a_tol=0.0, max_it=100000))
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    elapsed = t_end - t_start
    sol = darcy.get_solution_vector()
    info = infos[-1]
    print("solve: %.1f ms/solve, iters %d, assembly %.2f ms, cg %.1f ms, resolve %.2f ms, spmv %.1f us, update %.1f us"
          % (1e3 * elapsed / args.steps, info["iterations"], info["ms_assembly"], info["ms_solve"],
             info["ms_resolve"], 1e3 * info["ms_spmv_avg"], 1e3 * info["ms_update_avg"]), file=sys.stderr, flush=True)

    # accuracy: against a tightly converged FP64 solve of the same system
    flux_err, rinfo = None, {"iterations": 0}
    if not args.no_accuracy:
        ref = G.DarcyFlowMH_Steady(gm)
        ref.set_data(K, cs, sigma, f, bct, bcp, zeros_b, zeros_b)
        rinfo = ref.update_solution(r_tol=1e-14, a_tol=0.0, max_it=30000)
        rsol = ref.get_solution_vector()
        flux_err = float(np.linalg.norm(sol[:S] - rsol[:S]) / np.linalg.norm(rsol[:S]))
        del ref

    # roofline of the dominant kernel (CG SpMV + direction update), live event samples
    spmv_ms = float(np.mean([i["ms_spmv_avg"] for i in infos]))
    upd_ms = float(np.mean([i["ms_update_avg"] for i in infos]))
    b_spmv = spmv_bytes(F, nnz)
    achieved = b_spmv / (spmv_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "k_cg_spmv", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": b_spmv, "avg_launch_us": round(spmv_ms * 1e3, 2),
                "update_kernel": {"avg_launch_us": round(upd_ms * 1e3, 2), "bytes_per_launch": 64.0 * F,
                                  "achieved_GBs": round(64.0 * F / (upd_ms * 1e-3) / 1e9, 1)}}

    # transport: 4 substances, 1000 CFL-limited steps on the solved flux field
    transport = None
    if args.transport_steps > 0:
        nsub = 4
        bcc = np.ones((nsub, NB))
        por = np.ones(E)
        probe = G.ConvectionTransport(gm, por, cs, bcc, np.zeros((nsub, E)), 1.0, 1.0)
        probe.set_velocity_field(darcy=darcy)
        cfl = probe.prepare()  # transport matrix of this flux field: CFL limit
        del probe
        T_end = args.transport_steps * cfl
        tr = G.ConvectionTransport(gm, por, cs, bcc, np.zeros((nsub, E)), T_end, T_end / 10)
        tr.set_velocity_field(darcy=darcy)
        torch.cuda.synchronize()
        tt0 = time.perf_counter()
        nsteps = 0
        for k in range(1, 11):
            nsteps += tr.update_solution(T_end * k / 10)
        torch.cuda.synchronize()
        tt = time.perf_counter() - tt0
        ms_dev, _ = tr.stats()
        tnnz = G.lib().fl_transport_nnz(tr.h)
        step_bytes = tnnz * 12.0 + 4.0 * (E + 1) + E * nsub * 8.0 * 3
        c = tr.conc()
        mb = tr.mass_balance()
        tot = mb["totals"]
        transport = {"substances": nsub, "steps": nsteps, "steps_per_s": round(nsteps / tt, 1),
                     "ms_per_step_device": round(ms_dev / max(nsteps, 1), 4), "nnz": tnnz,
                     "bytes_per_step": step_bytes,
                     "achieved_GBs": round(step_bytes / (ms_dev / max(nsteps, 1) * 1e-3) / 1e9, 1),
                     "frac_of_peak": round(step_bytes / (ms_dev / max(nsteps, 1) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "conc_range": [float(c.min()), float(c.max())],
                     "mass_balance_rel_err": float(abs(tot[0, 0] + tot[0, 2] - tot[0, 1] - tot[0, 3]) /
                                                   max(abs(tot[0, 3]), 1e-300))}
        del tr

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mesh, K, cs, bct, bcp, info["iterations"], args.cpu_cg_iters)

    value = args.steps / elapsed
    out = {
        "metric": "Darcy solves/s at 1e-8 rel flux err: learned-PCG vs plain CG, 1 GPU",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal K, generated fractured cube mesh)",
        "config": {"workload": "mixed 1D/2D/3D cube: %d tets + %d fracture tris + %d channel segments; steady MH "
                               "Darcy FP64, plain Jacobi-CG (the drop-in path) to rel residual %g"
                               % (int((gm.elem_dim == 3).sum()), int((gm.elem_dim == 2).sum()),
                                  int((gm.elem_dim == 1).sum()), args.rtol),
                   "elements": E, "edges": F, "sides": S, "edge_nnz": nnz, "cg_iterations": info["iterations"],
                   "method": "plain FP64 Jacobi-CG", "flux_rel_err": flux_err,
                   "flux_err_reference": "FP64 GPU solve to rel residual 1e-14 (%d iterations)" % rinfo["iterations"],
                   "ms_assembly": round(info["ms_assembly"], 3), "ms_cg": round(info["ms_solve"], 3),
                   "ms_resolve": round(info["ms_resolve"], 3), "transport": transport},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(mesh, K, cs, bct, bcp, gpu_iters, cg_iters):
    """Times the FP64 oracle (a single-threaded restatement of the reference
    algorithm: full MH assembly, two-stage sparse Schur, Jacobi-CG, resolve)
    on the same mesh and fields: everything except CG runs in full, CG for a
    bounded number of iterations; the solve time is extrapolated to the GPU's
    iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    om = O.Mesh(arrays=mesh)
    E, NB = om.n_elements, om.n_bc_elements
    z = np.zeros(NB)
    D = O.Darcy(om, K, cs, np.ones(E), np.zeros(E), bct, bcp, z, z)
    D.solve(0, rtol=1e-30, atol=0.0, max_it=cg_iters)
    st = D.stats
    per_it = st["t_solve"] / max(st["iters"], 1)
    t_solve = st["t_assembly"] + st["t_schur"] + st["t_resolve"] + per_it * gpu_iters
    return {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle on the full mesh: assembly %.2fs + Schur %.2fs + resolve %.2fs in full, %d Jacobi-CG "
                      "iterations (%.1f ms each) extrapolated to the GPU's %d iterations; wall clock, 1 thread"
                      % (st["t_assembly"], st["t_schur"], st["t_resolve"], st["iters"], 1e3 * per_it, gpu_iters)}


if __name__ == "__main__":
    main()
