"""Benchmark of the MG-PCG / SpMV hot path on one MI355X.

One step = one full FP64 MG-PCG solve of the 64^3 hexahedral cantilever (811,200 free dofs,
63.7M nonzeros; element moduli log-uniform in [1e-2, 1e2]; x = 0 clamped; total load -1 in z
on the x = 1 face) to relative residual `--tol`, starting from x = 0, with the matrix, the
multigrid hierarchy and the right-hand side resident in HBM.  The timed region covers
exactly K solves.

Printed JSON line: value = solves/s; roofline of the dominant kernel (the level-0 CSR
SpMV p -> A p, timed with HIP events on the solver's stream during the timed solves);
cpu_baseline = the oracle's single-core C restatement of the same FP64 solve (a bounded
sample of its iterations, scaled to the GPU solve's iteration count).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED_MODULI = 20261203 + 2          # config 2: log-uniform moduli


def csr_spmv_bytes(nnz, nrows, ncols):
    """Algorithmic bytes of one FP64 CSR SpMV: values + column indices, row offsets,
    x read once, y written once."""
    return 12 * nnz + 4 * (nrows + 1) + 8 * ncols + 8 * nrows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="elements per axis")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--smoother", default="gs")
    ap.add_argument("--nu", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0: skip the CPU leg")
    ap.add_argument("--graph", type=int, default=1, help="replay PCG iterations as a captured hipGraph")
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 SpMV / V-cycle inside FP64 iterative refinement")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the solver's multi-GPU "
              f"sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from compressiontests_amd import _lib
    from compressiontests_amd.device import DeviceArray, rng_uniform
    from compressiontests_amd.model import FemCantilever3D
    from compressiontests_amd.solvers import HipMgPcgSolver

    assert _lib.device_count() > 0, "no GPU visible to libmgx.so"
    n = args.n
    dens = 10 ** (4 * rng_uniform(SEED_MODULI, n ** 3) - 2)
    t0 = time.time()
    model = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    solver = HipMgPcgSolver(model, levels=args.levels, smoother=args.smoother, nu=args.nu, tolerance=args.tol,
                            max_iterations=500, use_graph=bool(args.graph), check_every=args.check_every,
                            precision=args.precision)
    f = model.rhs()
    solver.Initialize(f)
    # the same solve with HIP events around every level-0 A*p SpMV (no graph): the roofline pass
    prof = HipMgPcgSolver(model, levels=args.levels, smoother=args.smoother, nu=args.nu, tolerance=args.tol,
                          max_iterations=500, profile=True, check_every=args.check_every)
    prof.Initialize(f)
    setup_s = time.time() - t0
    K = model.device_matrix()
    nrows, nnz = K.nrows, K.nnz

    b_dev = DeviceArray.from_numpy(f)
    x_dev = DeviceArray(nrows)
    for _ in range(args.warmup):
        solver.solve_device(b_dev, x_dev)
    _lib.call("mgx_device_synchronize")
    iters, rels, outer = [], [], []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        st = solver.solve_device(b_dev, x_dev)[0]
        iters.append(st.iterations)
        outer.append(st.outer_iterations)
        rels.append(st.rel_residual)
        if st.converged != 1:
            raise SystemExit(f"solve did not converge: {st.converged}")
    _lib.call("mgx_device_synchronize")
    elapsed = time.perf_counter() - t1
    solves_per_s = args.steps / elapsed
    x = x_dev.numpy()
    spmv_ms, spmv_n = 0.0, 0
    xp = DeviceArray(nrows)
    for _ in range(max(1, min(args.steps, 3))):
        st = prof.solve_device(b_dev, xp)[0]
        spmv_ms += st.spmv_ms
        spmv_n += st.spmv_launches

    # FP64 check outside the timed region: energy-norm error against a 1e-13 solve
    tight = HipMgPcgSolver(model, levels=args.levels, smoother=args.smoother, nu=args.nu, tolerance=1e-13,
                           max_iterations=2000)
    tight.Initialize(f)
    xt, stt = tight.Solve()
    e = DeviceArray.from_numpy(x - xt)
    Ke = DeviceArray(nrows)
    from compressiontests_amd import kernels
    kernels.matrix_vector_product(K, e, Ke)
    xd = DeviceArray.from_numpy(xt)
    Kx = DeviceArray(nrows)
    kernels.matrix_vector_product(K, xd, Kx)
    energy_err = float(np.sqrt(max((x - xt) @ Ke.numpy(), 0.0) / (xt @ Kx.numpy())))

    bytes_per_launch = csr_spmv_bytes(nnz, nrows, nrows)
    avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"spmv_csr_fp64_{n}cubed_pmc.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(n, dens, f, args, int(np.median(iters)))

    line = {
        "metric": "MG-PCG solves/s to 1e-8 energy-norm (FP64-checked); SpMV HBM GB/s vs peak",
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"config2: {n}^3 hex cantilever, {nrows} free dofs, {nnz} nnz, moduli 10^U(-2,2) "
                        f"seed {SEED_MODULI}, tip load -1 on x=1, "
                        + ("FP64 MG-PCG" if args.precision == "fp64" else
                           "FP64 iterative refinement around FP32 MG-PCG")
                        + f" ({args.levels} levels Galerkin, {args.smoother} nu={args.nu}) to rel. residual "
                          f"{args.tol:g} (FP64 true residual) from x=0",
            "precision": args.precision,
            "dofs": nrows, "nnz": nnz, "iterations": int(np.median(iters)),
            "outer_iterations": int(np.median(outer)),
            "dof_per_s": round(nrows * solves_per_s, 1),
            "energy_norm_error": energy_err,
            "rel_residual": float(max(rels)),
            "setup_s": round(setup_s, 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv_dot_alpha<16> (level-0 CSR SpMV q = A p with p.q), HIP events on the solver "
                      "stream around each launch in profiled solves of the same system",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_ms, 5),
            "launches": spmv_n,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(n, dens, f, args, gpu_iterations):
    """The oracle's single-threaded C restatement of the same FP64 MG-PCG (same matrix, Galerkin
    hierarchy, smoother and right-hand side).  Bounded sample: the time of 4 PCG iterations
    (a 6-iteration run minus a 2-iteration run, so the per-solve set-up cancels), scaled to
    the GPU solve's iteration count."""
    from oracle import oracle as O
    A = O.cantilever(n, n, n, dens=dens)
    lv, m = [], n
    for _ in range(args.levels - 1):
        P = O.p_trilinear(m, m, m)
        lv.append(dict(A=A, R=O.transpose(P), P=P))
        A = O.galerkin(A, P)
        m //= 2
    lv.append(dict(A=A))
    sm = 1 if args.smoother == "gs" else 0
    times = {}
    for it in (2, 6):
        t = time.perf_counter()
        O.mgpcg(lv, f, smoother=sm, omega=0.0, nu=args.nu, tol=0.0, max_it=it)
        times[it] = time.perf_counter() - t
    per_it = max(times[6] - times[2], 1e-9) / 4
    solve_s = per_it * gpu_iterations
    return {"value": round(1.0 / solve_s, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"4 FP64 MG-PCG iterations of the same {n}^3 system (oracle C restatement, -O2, "
                      f"1 thread; 6-iteration minus 2-iteration run), {per_it:.3f} s/iteration, scaled to "
                      f"{gpu_iterations} iterations = {solve_s:.1f} s per solve"}


if __name__ == "__main__":
    main()
