"""Benchmark of the MG-PCG hot path on one MI355X.

Default workload (the headline): the 128^3 hexahedral cantilever (6,390,144 free dofs,
509,597,550 nonzeros; element moduli log-normal, sigma 1.5, clipped to a 1e3 contrast; x = 0
clamped) with 32 load cases (iid N(0,1) forces on the top-face nodes, each case of unit norm),
solved by the batching + precision form: the right-hand sides advance in blocks through one
block PCG (block width and node visiting order chosen by the library's timing at initialize)
whose 5-level Galerkin V-cycle (Gauss-Seidel, one forward / one backward sweep), SpMM and
vectors are FP32, inside FP64 iterative refinement (FP64 true residual, dots, alpha/beta and
coarse solve), to relative residual 1e-8 on every FP64 true residual; the energy-norm error
of every load case against an FP64 solve to 1e-12 is measured outside the timed region.  The matrices, hierarchy and right-hand
sides are resident in HBM; one step = one solve of all 32 load cases; the timed region covers
exactly K steps.

Printed JSON line: value = solves/s (load cases solved per second); config carries the plain
FP64 GPU path (one right-hand side at a time) timed on the same system, the largest energy-norm
error of the batched solutions against FP64 solves to 1e-12 of every load case (the run exits
with status 3 after printing when it exceeds 1e-8), and the autotuner's timings; roofline = the
kernel with the most device time in profiled solves of the same batch (HIP events on the solver
stream around each SpMM launch and around each level-0 Gauss-Seidel sweep of launches, one per
colour pair or per colour; a sweep's algorithmic bytes count X read once), roofline_next = the
other of the level-0 Gauss-Seidel and SpMM, roofline_solve = the algorithmic bytes of every kernel
of the solve per second (compressiontests_amd.traffic); cpu_baseline = the oracle's C MG-PCG
on the same hierarchy, all cores (OpenMP build) and one core, a bounded sample of its iterations.

--n 64 --rhs 1 --precision fp64 --moduli loguniform gives the single-right-hand-side FP64
solve of the 64^3 system (moduli log-uniform in [1e-2, 1e2], tip load).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ENERGY_BOUND = 1e-8                 # relative energy-norm error every load case must meet
SEED_LOGUNIFORM = 20261203 + 2      # moduli log-uniform in [1e-2, 1e2]
SEED_LOGNORMAL = 20261203 + 3       # moduli log-normal, sigma 1.5, contrast 1e3
SEED_LOADS = 20261203 + 4           # top-face load cases


def roofline(name, bytes_per_launch, total_ms, launches, pmc_tag, timing):
    """HBM roofline object of one kernel: algorithmic bytes per launch / average launch time (HIP
    events on the solver stream, placed as `timing` says), and the rocprofv3 PMC traffic per launch
    if profiles/<pmc_tag>_pmc.json holds it."""
    avg_ms = total_ms / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"{pmc_tag}_pmc.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    return {"bound": "hbm",
            "kernel": name + ", HIP events on the solver stream " + timing + " in profiled solves of the same batch",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": int(bytes_per_launch),
            "traffic_ratio": round(traffic / bytes_per_launch, 3) if traffic else None,
            "avg_launch_ms": round(avg_ms, 5), "launches": launches, "total_ms": round(total_ms, 3)}


def moduli(kind, n):
    import numpy as np
    from compressiontests_amd.device import rng_normal, rng_uniform
    if kind == "loguniform":
        return 10 ** (4 * rng_uniform(SEED_LOGUNIFORM, n ** 3) - 2)
    g = np.clip(1.5 * rng_normal(SEED_LOGNORMAL, n ** 3), -0.5 * np.log(1e3), 0.5 * np.log(1e3))
    return np.exp(g)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="elements per axis")
    ap.add_argument("--levels", type=int, default=0, help="0: coarsest level 8^3")
    ap.add_argument("--rhs", type=int, default=32, help="load cases (1: the tip load)")
    ap.add_argument("--moduli", default="lognormal", choices=["lognormal", "loguniform"])
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--block-width", type=int, default=0, help="0: chosen by the library's timing")
    ap.add_argument("--strip-window", type=float, default=-1,
                    help="X bytes per strip of the level-0 tile order (-1: chosen by the library's timing)")
    ap.add_argument("--smoother", default="gs")
    ap.add_argument("--nu", type=int, default=1)
    ap.add_argument("--gs-pairs", type=int, default=1,
                    help="1: two Gauss-Seidel colours per launch where the colouring allows it")
    ap.add_argument("--tol", type=float, default=1e-8, help="relative residual of every FP64 true residual")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0: skip the CPU leg")
    ap.add_argument("--fp64-compare", type=int, default=2, help="load cases timed on the plain FP64 path")
    ap.add_argument("--check", type=int, default=32, help="load cases (the first ones) checked against FP64 1e-12 solves")
    ap.add_argument("--graph", type=int, default=1, help="replay PCG iterations as a captured hipGraph")
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--hierarchy-file", default="",
                    help="on-disk hierarchy: loaded when the file exists, else built and written there")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the solver's multi-GPU "
              f"sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from compressiontests_amd import _lib, kernels
    from compressiontests_amd import traffic as T
    from compressiontests_amd.device import DeviceArray, load_hierarchy, save_hierarchy
    from compressiontests_amd.model import FemCantilever3D
    from compressiontests_amd.solvers import HipMgPcgSolver

    assert _lib.device_count() > 0, "no GPU visible to libmgx.so"
    n = args.n
    levels = args.levels or max(2, int(round(np.log2(n))) - 2)
    nrhs = args.rhs
    dens = moduli(args.moduli, n)
    t0 = time.time()
    model = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    F = model.rhs()[None, :] if nrhs == 1 else model.load_cases(nrhs, SEED_LOADS)
    common = dict(levels=levels, smoother=args.smoother, nu=args.nu, check_every=args.check_every,
                  gs_pairs=bool(args.gs_pairs))
    phases = {}
    t_h = time.time()
    hier_source = "built"
    if args.hierarchy_file and os.path.exists(args.hierarchy_file):
        hier = load_hierarchy(args.hierarchy_file)
        hier_source = "file"
    else:
        hier = "galerkin"
    solver = HipMgPcgSolver(model, tolerance=args.tol, max_iterations=500, use_graph=bool(args.graph), nrhs=nrhs,
                            precision=args.precision, block_width=args.block_width,
                            strip_window=args.strip_window, hierarchy=hier, **common)
    hier = solver.hierarchy
    phases["hierarchy_s"] = round(time.time() - t_h, 2)
    if args.hierarchy_file and hier_source == "built":
        t_w = time.time()
        save_hierarchy(args.hierarchy_file, hier)
        phases["hierarchy_write_s"] = round(time.time() - t_w, 2)
    t_i = time.time()
    solver.Initialize(F)
    phases["initialize_s"] = round(time.time() - t_i, 2)
    width, tune_ms = solver.BlockWidth
    strip = solver.StripWindow
    print(f"# setup done in {time.time() - t0:.1f} s, block width {width}", file=sys.stderr, flush=True)
    # the same solve with HIP events around every level-0 A*p launch (no graph): the roofline pass
    prof = HipMgPcgSolver(model, tolerance=args.tol, max_iterations=500, profile=True, nrhs=nrhs,
                          precision=args.precision, block_width=width, strip_window=strip[0], hierarchy=hier,
                          **common)
    t_p = time.time()
    prof.Initialize(F)
    phases["profile_initialize_s"] = round(time.time() - t_p, 2)
    setup_s = time.time() - t0
    print(f"# setup phases {phases}", file=sys.stderr, flush=True)
    K = model.device_matrix()
    ndof, nnz = K.nrows, K.nnz

    b_dev = DeviceArray.from_numpy(F.reshape(-1))
    x_dev = DeviceArray(F.size)
    for _ in range(args.warmup):
        solver.solve_device(b_dev, x_dev)
    _lib.call("mgx_device_synchronize")
    iters, outer, rels, conv = [], [], [], True
    t1 = time.perf_counter()
    for _ in range(args.steps):
        st = solver.solve_device(b_dev, x_dev)
        iters += [s.iterations for s in st]
        outer += [s.outer_iterations for s in st]
        rels += [s.rel_residual for s in st]
        conv = conv and all(s.converged == 1 for s in st)
    _lib.call("mgx_device_synchronize")
    elapsed = time.perf_counter() - t1
    if not conv:
        raise SystemExit("a load case did not converge")
    solves_per_s = nrhs * args.steps / elapsed
    print(f"# timed {args.steps} steps: {solves_per_s:.2f} solves/s", file=sys.stderr, flush=True)
    X = x_dev.numpy().reshape(nrhs, -1)

    spmv_ms, spmv_n, gs_ms, gs_n = 0.0, 0, 0.0, 0
    xp = DeviceArray(F.size)
    nprof = max(1, min(args.steps, 2))
    nblocks = (nrhs + max(width, 1) - 1) // max(width, 1)
    for _ in range(nprof):
        st = prof.solve_device(b_dev, xp)
        for j in range(0, nrhs, max(width, 1)):   # one entry per block: the block's launches
            spmv_ms += st[j].spmv_ms
            spmv_n += st[j].spmv_launches
            gs_ms += st[j].gs_ms
            gs_n += st[j].gs_launches

    # the plain FP64 GPU path (one right-hand side at a time) on the first load cases
    fp64 = None
    if args.fp64_compare > 0 and not (args.precision == "fp64" and width == 1):
        m = min(args.fp64_compare, nrhs)
        one = HipMgPcgSolver(model, tolerance=args.tol, max_iterations=500, nrhs=m, block_width=1,
                             hierarchy=hier, use_graph=bool(args.graph), **common)
        one.Initialize(F[:m])
        bd = DeviceArray.from_numpy(F[:m].reshape(-1))
        xd = DeviceArray(m * ndof)
        one.solve_device(bd, xd)
        _lib.call("mgx_device_synchronize")
        t = time.perf_counter()
        st1 = one.solve_device(bd, xd)
        _lib.call("mgx_device_synchronize")
        dt = time.perf_counter() - t
        fp64 = {"solves_per_s": round(m / dt, 4), "load_cases": m,
                "iterations": int(np.median([s.iterations for s in st1]))}
        one.ReleaseResources()
        print(f"# plain FP64 path: {fp64['solves_per_s']} solves/s", file=sys.stderr, flush=True)

    # FP64 check outside the timed region: energy-norm error against FP64 solves to 1e-12
    cols = list(range(max(1, min(args.check, nrhs))))
    tw = max([w for w in (1, 4, 8, 16) if w <= len(cols)])
    tight = HipMgPcgSolver(model, tolerance=1e-12, max_iterations=2000, nrhs=len(cols), block_width=tw,
                           strip_window=strip[0], hierarchy=hier, **common)
    tight.Initialize(F[cols])
    Xt, stt = tight.Solve()
    Xt = Xt.reshape(len(cols), -1)
    energy_err = 0.0
    for i, j in enumerate(cols):
        e = DeviceArray.from_numpy(X[j] - Xt[i])
        Ke = DeviceArray(ndof)
        kernels.matrix_vector_product(K, e, Ke)
        xd = DeviceArray.from_numpy(Xt[i])
        Kx = DeviceArray(ndof)
        kernels.matrix_vector_product(K, xd, Kx)
        energy_err = max(energy_err, float(np.sqrt(max((X[j] - Xt[i]) @ Ke.numpy(), 0.0) / (Xt[i] @ Kx.numpy()))))
    tight.ReleaseResources()

    vb = 4 if args.precision == "mixed" else 8
    layout = solver.LevelLayout(0)
    tname = "float" if vb == 4 else "double"
    if width > 1:
        spmm_per_launch = (T.nodeblock_spmm_bytes(nnz, ndof, width, vb) if layout == "node-block" else
                           T.csr_spmv_bytes(nnz, ndof, ndof, vb) - vb * 2 * ndof + 2 * vb * width * ndof)
        spmm_name = {"node-block": f"k_nb_spmm<{tname}, K={width}, DOT> (level-0 node-block (3x3) SpMM Q = A P with "
                                   f"per-column P.Q)",
                     "node-triple": f"k_spmm_node<{tname}, K={width}, DOT> (level-0 node-triple CSR SpMM Q = A P "
                                    f"with per-column P.Q)"}.get(layout, f"k_spmm<{tname}, K={width}, DOT> (level-0 "
                                                                         f"CSR SpMM)")
    else:
        spmm_per_launch = T.csr_spmv_bytes(nnz, ndof, ndof, vb)
        spmm_name = f"k_csr_stream<{tname}, DOT> (level-0 CSR-stream SpMV q = A p with p.q)"
    kind = ("spmm_nodeblock" if layout == "node-block" else "spmm") if width > 1 else "spmv"
    rooflines = [roofline(spmm_name, spmm_per_launch, spmv_ms, spmv_n,
                          f"{kind}_{'f32' if vb == 4 else 'f64'}_k{width}_{n}cubed", "around each launch")]
    ncol, npair = solver.LevelColors(0), solver.GsPairs(0)
    nl = ncol - npair                                   # launches per sweep
    if gs_n > 0 and ncol > 0 and npair in (0, ncol // 2) and ncol % 2 == 0:
        tag = "gs_pair_nodeblock" if npair else "gs_nodeblock"
        what = (f"k_nb_gs_pair<{tname}, K={width}> (one launch of the level-0 node-block multicolour Gauss-Seidel "
                f"post-smoothing sweep, two colours line by line; {nl} launches of a sweep, bytes averaged over them"
                if npair else
                f"k_nb_gs<{tname}, K={width}> (one colour launch of the level-0 node-block multicolour Gauss-Seidel "
                f"post-smoothing sweep; bytes averaged over the {ncol} colours of a sweep")
        rooflines.append(roofline(what + "; a sweep's algorithmic bytes are the blocks, X read once, B read once "
                                  "and X written once; the pre-smoothing sweep from x = 0 reads only earlier colours "
                                  "and is not counted)",
                                  T.nodeblock_gs_sweep_bytes(nnz, ndof, width, vb) / nl, gs_ms, gs_n,
                                  f"{tag}_{'f32' if vb == 4 else 'f64'}_k{width}_{n}cubed",
                                  "around the back-to-back launches of each sweep"))
    # the dominant kernel: the most device time in the profiled solves
    rooflines.sort(key=lambda r: -r["total_ms"])
    solve_roof = solve_roofline(solver, hier, args, width, nrhs, solves_per_s, spmv_n / max(1, nprof * nblocks),
                                int(max(outer)), T)

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(hier, F[0], args, fp64["iterations"] if fp64 else int(np.median(iters)))

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
            "workload": f"{n}^3 hex cantilever, {ndof} free dofs, {nnz} nnz, moduli "
                        + ("log-normal sigma 1.5 clipped to 1e3 contrast" if args.moduli == "lognormal"
                           else "log-uniform [1e-2, 1e2]")
                        + (f", {nrhs} top-face load cases (unit norm)" if nrhs > 1 else ", tip load -1 on x=1")
                        + f"; {'batched ' if width > 1 else ''}MG-PCG, {levels}-level Galerkin V-cycle, "
                          f"{args.smoother} nu={args.nu}, "
                        + ("FP32 V-cycle/SpMM inside FP64 iterative refinement" if args.precision == "mixed"
                           else "FP64")
                        + f", to rel. residual {args.tol:g} on every FP64 true residual, from x=0",
            "dofs": ndof, "nnz": nnz, "load_cases": nrhs, "levels": levels,
            "precision": args.precision, "block_width": width,
            "autotune_ms_per_iteration": {f"k{k}": round(v, 4) for k, v in zip((4, 8, 16, 32), tune_ms) if v > 0},
            "strip_window_bytes": strip[0],
            "autotune_strip_ms_per_iteration": {f"{w:g}": round(v, 4) for w, v in zip((6e6, 3e6, 1.5e6, 0), strip[1])
                                                if v > 0},
            "iterations_median": int(np.median(iters)), "iterations_max": int(max(iters)),
            "outer_iterations_max": int(max(outer)),
            "dof_per_s": round(ndof * solves_per_s, 1),
            "energy_norm_error": energy_err,
            "energy_norm_bound": ENERGY_BOUND,
            "energy_norm_bound_met": bool(energy_err <= ENERGY_BOUND),
            "energy_norm_checked_cases": cols,
            "rel_residual_max": float(max(rels)),
            "fp64_single_rhs_path": fp64,
            "speedup_vs_fp64_single_rhs": round(solves_per_s / fp64["solves_per_s"], 3) if fp64 else None,
            "setup_s": round(setup_s, 2),
            "setup_phases_s": phases,
            "hierarchy": hier_source,
        },
        "roofline": rooflines[0],
        "roofline_next": rooflines[1] if len(rooflines) > 1 else None,
        "roofline_solve": solve_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if energy_err > ENERGY_BOUND:
        # the metric is solves/s to the energy-norm bound: a line whose batch missed it is not a result
        print(f"# INVALID: energy-norm error {energy_err:.3e} above {ENERGY_BOUND:g}", file=sys.stderr, flush=True)
        sys.exit(3)


def solve_roofline(solver, hier, args, width, nrhs, solves_per_s, iterations, outer_steps, T):
    """Whole-solve HBM roofline of the batching + precision form: the algorithmic bytes of every
    kernel of one block solve (compressiontests_amd.traffic.block_mixed_solve_bytes, from the
    level sizes, the block width, the block PCG iterations the profiled solves ran and the
    refinement steps) times the blocks solved per second.  None for other configurations."""
    L = len(hier)
    if (width <= 1 or args.precision != "mixed" or args.smoother != "gs" or args.nu != 1
            or any(solver.LevelLayout(l) != "node-block" for l in range(L - 1))
            or any(solver.TransferLayout(l) != "node-transfer" for l in range(L - 1))):
        return None
    levels = []
    for A, R, P in hier:
        d = {"n": A.nrows, "nnz": A.nnz}
        if R is not None:
            d["r_nnz"], d["p_nnz"] = R.nnz, P.nnz
        levels.append(d)
    per_block = T.block_mixed_solve_bytes(levels, width, iterations, outer_steps)
    achieved = per_block * (solves_per_s / width) / 1e9
    return {"bound": "hbm", "scope": "every kernel of a block solve (block width %d, %.1f block PCG iterations, %d "
                                     "refinement steps), algorithmic bytes, blocks of load cases solved per second"
                                     % (width, iterations, outer_steps),
            "bytes_per_load_case": int(per_block / width), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hier, f, args, iterations):
    """The oracle's C restatement of the FP64 MG-PCG (one right-hand side, the same Galerkin
    hierarchy, smoother and first load case), built -O3: the serial build on one core and the
    OpenMP build on every core the process may use (bit-identical results).  Bounded sample: one
    warm-up iteration, then the median of 5 timed iterations (each timed inside the oracle),
    scaled to the FP64 GPU solve's iteration count."""
    import numpy as np
    from oracle import oracle as O
    lv = []
    for A, R, P in hier:
        d = {"A": _oracle_csr(O, A)}
        if R is not None:
            d["R"], d["P"] = _oracle_csr(O, R), _oracle_csr(O, P)
        lv.append(d)
    sm = 1 if args.smoother == "gs" else 0
    res = {}
    for omp in (False, True):
        _, _, t = O.mgpcg_timed(lv, f, 6, smoother=sm, omega=0.0, nu=args.nu, omp=omp)
        per_it = float(np.median(t[1:]))
        res[omp] = (per_it, O.num_threads(omp))
        print(f"# cpu baseline ({res[omp][1]} threads): {per_it:.3f} s/iteration", file=sys.stderr, flush=True)
    per_it, cores = res[True]
    one = res[False][0]
    return {"value": round(1.0 / (per_it * iterations), 6), "unit": "solves/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": len(os.sched_getaffinity(0)),
            "single_core": {"value": round(1.0 / (one * iterations), 6), "unit": "solves/s", "cores": 1,
                            "s_per_iteration": round(one, 4)},
            "s_per_iteration": round(per_it, 4),
            "sample": f"FP64 MG-PCG on one load case of the same {args.n}^3 system and hierarchy (oracle C "
                      f"restatement, -O3 -ffp-contract=off, OpenMP on {cores} threads; the single-core figure is the "
                      f"serial build): 1 warm-up + median of 5 timed iterations, {per_it:.3f} s/iteration, scaled to "
                      f"the FP64 GPU path's {iterations} iterations = {per_it * iterations:.1f} s per solve"}


def _oracle_csr(O, M):
    h = M.to_host()
    return O.Csr(h.rowptr, h.col, h.val, h.nrows, h.ncols)


if __name__ == "__main__":
    main()
