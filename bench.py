#!/usr/bin/env python3
"""Time-to-rtol of the assemble-and-solve path on one MI355X.

Workload (default): the heterogeneous plate -- rectangle 2x1, 1413x1413 nodes (3,993,138
DOFs), interior nodes jittered +-0.2h, per-element ln E ~ N(0,1), nu=0.3, plane strain,
ux=0 on x=0, uy=0 on y=0, traction tx=1 on x=2, body force (0,-1).  One step = stiffness
pattern + assembly + loads + Dirichlet elimination + Jacobi setup + solve to a true FP64
relative residual below 1e-10, all on the device; the mesh and material are resident in
HBM before the timed region starts.

Methods: fp64 (the reference's FP64 Jacobi-CG recurrence with FP64 true-residual
replacement), ir32 (FP32 Jacobi-CG under FP64 refinement), learned1 (flexible CG with the
learned single-level preconditioner, FP64 matrix and vectors), learned32 (the same in FP32 with
FP64 refreshes), learned_mg (flexible CG with the aggregation multilevel preconditioner whose
fine-level smoother is the learned M) and mg_cheb (the same hierarchy with a Chebyshev fine
smoother instead of the network).  The learned methods train the network once, before the
timed steps, on a small plate of the same family (its time is reported as train_s); every
timed step rebuilds M and the hierarchy for the freshly assembled matrix (reported as
precond_s, part of the step).

Prints one JSON line.  The roofline object covers the FP64 Jacobi-CG's SpMV (k_cg_spmv, the
dominant kernel of the baseline method, whose bandwidth the target is stated for), timed
with HIP events on the library stream; roofline_fcg gives the same for the flexible CG's SpMV
(q = A p plus p'r), which every iteration of the learned methods runs;
cpu_baseline is the oracle (the CPU restatement of the reference's element loop and PCG,
bitwise equal to the reference on uniform-E inputs) timed on a bounded sample, its PCG on
all host threads (OMP_NUM_THREADS, else every core up to 16); the one-core figures are kept
beside it.

--config 3 times the batched workload instead: 512 independent plates (578 to 51,842 DOFs)
solved by one persistent launch (FP64, and FP32 under FP64 refinement), the launch shape
picked by the autotuner; its CPU baseline runs the oracle on every 8th plate, plates in
parallel on the host threads.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mpi-cst-fem-solver_amd"))

import numpy as np  # noqa: E402

METRIC = "time-to-rtol 1e-10 (s) learned/mixed FCG vs FP64 Jacobi-CG; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


# arithmetic of each method: the outer iteration, its preconditioner, the network
DTYPES = {"fp64": "f64", "mg_cheb": "f64", "ir32": "f32 + f64 refinement",
          "learned1": "f64 outer + f32 M + bf16 network",
          "learned32": "f32 + f64 refinement + f32 M + bf16 network",
          "learned_mg": "f64 outer + f32 cycle + bf16 network"}
PRECISION = {
    "fp64": "FP64 matrix, vectors and sums",
    "mg_cheb": "FP64 flexible CG; FP64 multilevel cycle (Chebyshev smoothing)",
    "ir32": "FP32 matrix copy and vectors with FP64 dot products; FP64 x and true-residual refreshes",
    "learned1": "FP64 flexible CG; M stored in FP32, applied with FP64 sums; network on BF16 MFMA",
    "learned32": "FP32 flexible CG with FP64 refreshes; M in FP32; network on BF16 MFMA",
    "learned_mg": ("FP64 flexible CG, SpMV and true residual; cycle row sums read FP32 copies of "
                   "the level-0 and coarse matrices and M's FP32 values, accumulating in FP64; "
                   "network forward on BF16 MFMA with FP32 accumulation"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, nnz):
    """Algorithmic bytes of one FP64 SpMV on 2x2 blocks: values, block column indices,
    block row pointers, x read once, y written once."""
    return 8 * nnz + nnz + 4 * (n // 2 + 1) + 16 * n


def cpu_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set (the GPU box sets it to
    its CPU share), else every core, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def cpu_model():
    """The host CPU's model name and logical CPU count (from /proc/cpuinfo)."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=name, logical_cpus=os.cpu_count())


def cpu_baseline(prob, iterations, sample_iters, full_limit_s=40.0):
    """The oracle's restatement of the reference path on the host: pattern + assembly + loads
    + Dirichlet in full (sequential, as the reference), then Jacobi-PCG iterations of the same
    recurrence, timed on one core (`sample_iters`) and on all threads (10x as many).
    - When the threaded solve to rtol 1e-10 is predicted to take at most `full_limit_s`, it
      is run to convergence and `value` is the measured set-up + solve time (the prediction
      is reported beside it); otherwise `value` is the set-up time plus the threaded time per
      iteration times the GPU FP64 iteration count.
    - Where the reference's own serial core is built (oracle/_ref), its pattern + assembly
      and PCG iterations run on the same system on one core: the calibration of the
      restatement against the reference on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    oc = oracle_lib.Oracle(oracle_lib.ORACLE_PATH, "oc")
    m = prob.mesh
    strain = 1 if prob.plane == 1 else 0
    t0 = time.perf_counter()
    rp, c = oc.build_pattern(m.num_nodes, m.tri)
    v = oc.assemble(m.xy, m.tri, rp, c, E=prob.E, nu=prob.nu, strain=strain, t=prob.thickness,
                    E_elem=prob.E_elem)
    f = np.zeros(m.num_dofs)
    for g, kw in prob.tractions:
        oc.add_traction(m.xy, m.tri, m.groups[g], f, t=prob.thickness, **kw)
    if prob.body is not None:
        oc.add_body_force(m.xy, m.tri, f, prob.thickness, *prob.body)
    d, val = prob.dirichlet_dofs()
    K, b = oc.apply_dirichlet(rp, c, v, f, d, val)
    t1 = time.perf_counter()
    log(f"cpu baseline: assembly+bcs {t1 - t0:.2f} s")
    _, st = oc.pcg_jacobi(rp, c, K, b, 1e-300, sample_iters)
    t2 = time.perf_counter()
    per_it1 = (t2 - t1) / max(1, st["iterations"])
    threads = cpu_threads()
    _, st_mt = oc.pcg_jacobi_mt(rp, c, K, b, 1e-300, 10 * sample_iters, threads=threads)
    t3 = time.perf_counter()
    per_it = (t3 - t2) / max(1, st_mt["iterations"])
    est = (t1 - t0) + per_it * iterations
    log(f"cpu baseline: 1 core {per_it1 * 1e3:.1f} ms/it, {threads} threads "
        f"{per_it * 1e3:.2f} ms/it, predicted {est:.1f} s")
    out = dict(value=est, unit="s", cores=threads, kind="port", measured=False,
               sample=(f"oracle restatement: pattern+assembly+loads+Dirichlet on 1 core "
                       f"({t1 - t0:.2f} s) + {st_mt['iterations']} Jacobi-PCG iterations on "
                       f"{threads} OpenMP threads ({per_it * 1e3:.2f} ms/it), extrapolated to "
                       f"the GPU FP64 count of {iterations} iterations"),
               host=cpu_model(), assembly_s=t1 - t0, s_per_iteration=per_it, predicted_s=est,
               one_core=dict(s_per_iteration=per_it1, iterations_timed=st["iterations"],
                             value=(t1 - t0) + per_it1 * iterations))
    if per_it * iterations <= full_limit_s:
        t4 = time.perf_counter()
        _, st_full = oc.pcg_jacobi_mt(rp, c, K, b, 1e-10, 20 * iterations + 1000, threads=threads)
        t5 = time.perf_counter()
        out.update(value=(t1 - t0) + (t5 - t4), measured=True,
                   sample=(f"oracle restatement: pattern+assembly+loads+Dirichlet on 1 core "
                           f"({t1 - t0:.2f} s) + Jacobi-PCG to rtol 1e-10 on {threads} OpenMP "
                           f"threads, run to convergence ({st_full['iterations']} iterations, "
                           f"{t5 - t4:.2f} s)"),
                   iterations=st_full["iterations"], converged=bool(st_full["converged"]))
        log(f"cpu baseline: measured {out['value']:.2f} s to convergence "
            f"({st_full['iterations']} iterations) against {est:.2f} s predicted")
    if os.path.exists(oracle_lib.REF_PATH):
        # the reference's own core on the same system: pattern + assembly (with the uniform
        # modulus, as the reference's assembly takes one material; the element loop's cost is
        # the same), and PCG iterations on the oracle's eliminated system
        ref = oracle_lib.Oracle(oracle_lib.REF_PATH, "ref")
        t_o = time.perf_counter()
        rp_o, c_o = oc.build_pattern(m.num_nodes, m.tri)
        oc.assemble(m.xy, m.tri, rp_o, c_o, E=prob.E, nu=prob.nu, strain=strain, t=prob.thickness)
        t_asm_uniform = time.perf_counter() - t_o
        t6 = time.perf_counter()
        rp_r, c_r = ref.build_pattern(m.num_nodes, m.tri)
        ref.assemble(m.xy, m.tri, rp_r, c_r, E=prob.E, nu=prob.nu, strain=strain,
                     t=prob.thickness)
        t7 = time.perf_counter()
        ref_iters = max(5, min(sample_iters, int(3.0 / max(per_it1, 1e-6))))
        _, st_r = ref.pcg_jacobi(rp, c, K, b, 1e-300, ref_iters)
        t8 = time.perf_counter()
        ref_it = (t8 - t7) / max(1, st_r["iterations"])
        out["reference_one_core"] = dict(
            pattern_assembly_s=t7 - t6, s_per_iteration=ref_it, iterations_timed=st_r["iterations"],
            oracle_over_reference_assembly=t_asm_uniform / (t7 - t6),
            oracle_over_reference_per_iteration=per_it1 / ref_it)
        log(f"cpu baseline: reference core, 1 core: pattern+assembly {t7 - t6:.2f} s "
            f"(oracle {t_asm_uniform:.2f} s), {ref_it * 1e3:.1f} ms/it (oracle "
            f"{per_it1 * 1e3:.1f})")
    return out


def oracle_plate_solve(oc, p, rtol):
    """One plate through the oracle's restatement of the reference path (pattern, assembly,
    loads, Dirichlet, Jacobi-PCG), as the reference solves one case; returns iterations."""
    m = p.mesh
    rp, c = oc.build_pattern(m.num_nodes, m.tri)
    v = oc.assemble(m.xy, m.tri, rp, c, E=p.E, nu=p.nu, strain=0, t=p.thickness)
    f = np.zeros(m.num_dofs)
    for g, kw in p.tractions:
        oc.add_traction(m.xy, m.tri, m.groups[g], f, t=p.thickness, **kw)
    d, val = p.dirichlet_dofs()
    K, b = oc.apply_dirichlet(rp, c, v, f, d, val)
    _, st = oc.pcg_jacobi(rp, c, K, b, rtol, 200000)
    return st["iterations"], len(K)


def run_batch(args):
    """The batched workload (512 independent plates of 578 to 51,842 DOFs): one persistent
    launch solves every plate to an FP64 true relative residual below 1e-10 (FP64 Jacobi-CG,
    and FP32 Jacobi-CG under FP64 refinement), with the launch shape picked by the autotuner.
    The plates are assembled and constrained on the device before the timed region; a step
    is one solve of the whole batch (device time between HIP events)."""
    import cstfem_hip
    from cstfem_hip import pipeline
    t = time.perf_counter()
    bp = pipeline.batch_problem()
    log(f"batch: {bp.count} plates, {bp.mesh.num_dofs} DOFs (generated in "
        f"{time.perf_counter() - t:.1f} s)")
    dev = cstfem_hip.Device(0)
    device_name, cus = dev.info()
    pipeline.load_batch(dev, bp)
    best, t_tune = dev.batch_autotune()
    log(f"autotuned launch shape {best} ({t_tune * 1e3:.1f} ms)")
    methods = {"fp64": cstfem_hip.METHOD_CG_FP64_REFINED, "fp32": cstfem_hip.METHOD_IR_FP32}
    results = {}
    for name, meth in methods.items():
        for _ in range(args.warmup):
            dev.batch_solve(method=meth, config=best, download=False)
        times, res = [], None
        for _ in range(args.steps):
            _, res, secs = dev.batch_solve(method=meth, config=best, download=False)
            times.append(secs)
        its = [r.iterations for r in res]
        ok = all(r.converged and r.true_rel_residual < 1e-10 for r in res)
        results[name] = dict(s=sum(times) / len(times), verified=bool(ok),
                             iterations_min=min(its), iterations_max=max(its),
                             iterations_total=int(sum(its)),
                             worst_true_rel_residual=max(r.true_rel_residual for r in res))
        log(f"batch {name}: {results[name]['s']:.3f} s, iterations {min(its)}..{max(its)}, "
            f"verified {ok}")
    fp64 = results["fp64"]
    # effective bandwidth of the FP64 solve: per plate and iteration, the SpMV bytes of its
    # 2x2-block matrix (8 nnz values, nnz / 4 block columns of 4 bytes, x read and y written:
    # 16 n) plus 11 vector passes of 8 n, summed over the plates' own iteration counts, over
    # the makespan
    _, res64, _ = dev.batch_solve(method=methods["fp64"], config=best, download=False)
    rp, _ = dev.download_pattern()
    moved = 0.0
    for k, p in enumerate(bp.plates):
        n = p.mesh.num_dofs
        lo = 2 * int(bp.node_off[k])
        nnz = int(rp[lo + n] - rp[lo])
        moved += res64[k].iterations * (9.0 * nnz + 104.0 * n)
    achieved = moved / fp64["s"] / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        from concurrent.futures import ThreadPoolExecutor
        oc = oracle_lib.Oracle(oracle_lib.ORACLE_PATH, "oc")
        sample = list(range(0, bp.count, 8))
        threads = cpu_threads()
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as pool:
            out = list(pool.map(lambda k: oracle_plate_solve(oc, bp.plates[k], 1e-10), sample))
        t_cpu = time.perf_counter() - t0
        # scale by the GPU's iterations x DOFs (the work per plate) of all plates over the sample
        w_all = sum(res64[k].iterations * bp.plates[k].mesh.num_dofs for k in range(bp.count))
        w_smp = sum(res64[k].iterations * bp.plates[k].mesh.num_dofs for k in sample)
        cpu = dict(value=t_cpu * w_all / w_smp, unit="s", cores=threads, kind="port",
                   sample=(f"oracle restatement of the reference path (pattern, assembly, "
                           f"loads, Dirichlet, Jacobi-PCG to 1e-10) on every 8th plate "
                           f"({len(sample)} plates, {t_cpu:.1f} s on {threads} threads, plates "
                           f"in parallel), scaled by the GPU iterations x DOFs of all plates"))
        log(f"cpu baseline: {len(sample)} plates in {t_cpu:.1f} s -> {cpu['value']:.1f} s")
    line = {
        "metric": "batched solve: 512 plates to FP64 true rel residual 1e-10 each (s)",
        "value": fp64["s"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fp64["s"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batch512: {bp.count} plates, {bp.mesh.num_dofs} DOFs, one "
                               f"persistent launch", "dofs": int(bp.mesh.num_dofs),
                   "plates": bp.count, "launch_shape": list(best),
                   "device": device_name, "cus": cus},
        "methods": results,
        "roofline": {"bound": "hbm", "kernel": "k_batch_cg (persistent, FP64)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "effective: per plate and iteration 9 nnz + 104 n bytes "
                             "(SpMV and 11 vector passes), summed over the plates, over the "
                             "makespan; the largest plate on one workgroup sets the makespan"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    dev.close()



def method_codes():
    import cstfem_hip
    return {"fp64": cstfem_hip.METHOD_CG_FP64_REFINED, "ir32": cstfem_hip.METHOD_IR_FP32,
            "learned1": cstfem_hip.METHOD_FCG_LEARNED,
            "learned32": cstfem_hip.METHOD_FCG_LEARNED_FP32,
            "learned_mg": cstfem_hip.METHOD_FCG_LEARNED,
            "mg_cheb": cstfem_hip.METHOD_FCG_LEARNED}


def time_methods(dev, prob, args, chosen, thetas, kind=None):
    """Times `args.steps` steps (after `args.warmup`) of every method in `chosen` on the
    problem loaded in `dev`.  thetas: network parameters per method name (methods using the
    network); kind: the solver of each name when it is not the name itself (e.g. two
    learned_mg variants with different networks)."""
    from cstfem_hip import pipeline
    codes = method_codes()
    kind = kind or {}

    def step(name):
        method = kind.get(name, name)
        times = pipeline.PhaseTimes()
        pipeline.assemble_and_constrain(dev, prob, times,
                                        multilevel=method in ("learned_mg", "mg_cheb"))
        t0 = time.perf_counter()
        dev.setup()
        if name in thetas:
            dev.mpnn_set_params(thetas[name])
            dev.mpnn_build()
        if method in ("learned_mg", "mg_cheb"):
            learned_fine = method == "learned_mg"
            dev.mg_setup(args.mg_levels, args.mg_coarse, 0 if learned_fine else 1,
                         args.mg_kcycle if learned_fine else args.mg_kcycle_cheb)
        t1 = time.perf_counter()
        _, st = dev.solve(rtol=1e-10, max_iter=args.max_iter, method=codes[method],
                          download=False)
        times.solve += time.perf_counter() - t0
        st.precond_s = t1 - t0
        return st, times

    results = {}
    for meth in chosen:
        for i in range(args.warmup):
            st, times = step(meth)
            log(f"{meth} warmup {i}: {st.message}, {st.outer_iterations} refinements, "
                f"true rel {st.true_rel_residual:.3e}")
        step_times = []
        for i in range(args.steps):
            t0 = time.perf_counter()
            st, times = step(meth)
            dt = time.perf_counter() - t0
            step_times.append(dt)
            log(f"{meth} step {i}: {dt:.3f} s, {st.message}, {st.outer_iterations} refinements, "
                f"true rel {st.true_rel_residual:.3e}")
        verified = st.converged and st.true_rel_residual < 1e-10
        if not verified:
            log(f"WARNING: {meth} not verified: {st.message}; true rel {st.true_rel_residual:.3e}")
        results[meth] = dict(s=sum(step_times) / len(step_times), iterations=st.iterations,
                             refinements=st.outer_iterations, verified=bool(verified),
                             true_rel_residual=st.true_rel_residual,
                             fallback=int(st.fallback),
                             s_per_iteration=(times.solve - st.precond_s) / max(1, st.iterations),
                             phases_s={"pattern": times.pattern, "assembly": times.assembly,
                                       "bcs": times.bcs, "solve": times.solve,
                                       "precond": st.precond_s})
    return results


def run_protocol(args):
    """The learned preconditioner's training protocol and its deployment: pairs harvested at
    plain Jacobi-CG iterations 1..50 of 2,000 plates of 128x128 cells (ln E ~ N(0,1) per
    element, nu = 0.3, plane strain, unit traction of random direction), solved in batches on
    the device, plus white-noise pairs (training on the harvested pairs alone gives a
    smoother the set-up check rejects); the network then runs as the fine smoother of the
    multilevel flexible CG (FP32 values in the cycle, FP64 outer iteration) on the 4M-DOF
    heterogeneous plate.  Beside it: the same hierarchy with the default network (white
    noise on one plate of the family, --train-n nodes per side, --train-steps Adam steps)
    and FP64 Jacobi-CG."""
    import cstfem_hip
    from cstfem_hip import learned, pipeline
    t0 = time.perf_counter()
    with cstfem_hip.Device(0) as tdev:
        theta_h, losses_h, info = learned.train_harvested(
            tdev, plates=args.h_plates, chunk=args.h_chunk, cells=args.h_cells, k_first=1,
            k_last=50, steps_per_chunk=args.h_steps, project=False, white=args.h_white, log=log)
    protocol_s = time.perf_counter() - t0
    log(f"protocol: {args.h_plates} plates in {protocol_s:.1f} s ({info})")
    t0 = time.perf_counter()
    with cstfem_hip.Device(0) as tdev:
        theta_w, losses_w = learned.train(tdev, learned.training_problem(2, args.train_n),
                                          steps=args.train_steps)
    white_s = time.perf_counter() - t0
    prob = pipeline.config_problem(2, args.scale)
    dev = cstfem_hip.Device(0)
    device_name, cus = dev.info()
    pipeline.load_problem(dev, prob)
    chosen = ["learned_mg_protocol", "learned_mg_white"] + (["fp64"] if args.with_fp64 else [])
    results = time_methods(dev, prob, args, chosen,
                           {"learned_mg_protocol": theta_h, "learned_mg_white": theta_w},
                           kind={"learned_mg_protocol": "learned_mg",
                                 "learned_mg_white": "learned_mg"})
    dev.setup()
    dev.mpnn_set_params(theta_h)
    dev.mpnn_build()
    lam = None
    dev.mg_setup(args.mg_levels, args.mg_coarse, 0, args.mg_kcycle)
    lam = dev.mg_level_info(0)[3]
    n, nnz = dev.n, dev.nnz
    t_fcg = dev.profile_kernel(6, reps=50)
    b_fcg = spmv_bytes(n, nnz) + 8 * n
    res = results["learned_mg_protocol"]
    line = {
        "metric": METRIC, "value": res["s"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["s"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer + f32 cycle + bf16 network",
        "data": "synthetic",
        "config": {"workload": (f"learned-preconditioner protocol: {args.h_plates} plates of "
                                f"{args.h_cells}x{args.h_cells} cells, pairs at Jacobi-CG "
                                f"iterations 1..50 (+{args.h_white} white-noise rows per batch), "
                                f"deployed on {prob.name} ({prob.mesh.num_dofs} DOFs) to FP64 "
                                f"true rel residual 1e-10"),
                   "dofs": prob.mesh.num_dofs, "method": "learned_mg (protocol network)",
                   "device": device_name, "cus": cus},
        "protocol": dict(info, plates=args.h_plates, chunk=args.h_chunk, cells=args.h_cells,
                         steps=len(losses_h), protocol_s=protocol_s,
                         loss_first=losses_h[0], loss_last=losses_h[-1],
                         lambda_max_MK=lam),
        "white_network": dict(train_s=white_s, steps=len(losses_w), loss_first=losses_w[0],
                              loss_last=losses_w[-1]),
        "methods": results,
        "speedup_vs_fp64": (results["fp64"]["s"] / res["s"]) if "fp64" in results else None,
        "roofline": {"bound": "hbm", "kernel": "k_fcg_spmv<F64>", "bytes_per_launch": b_fcg,
                     "avg_launch_s": t_fcg, "achieved": b_fcg / t_fcg / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_fcg / t_fcg / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    dev.close()



def measured_traffic(dofs):
    """HBM bytes per launch of k_cg_spmv<F64> from the committed rocprofv3 --pmc measurement
    (FETCH_SIZE and WRITE_SIZE runs, scaled by the copy-kernel calibration) when it was taken
    on this workload; (None, None) otherwise."""
    path = os.path.join(ROOT, "profiles", "k_cg_spmv_hetero4M_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, None
    if rec.get("dofs") != dofs:
        return None, None
    return rec["traffic_bytes"], "profiles/k_cg_spmv_hetero4M_traffic.json (rocprofv3 --pmc)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--max-iter", type=int, default=200000)
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only the hot kernels (for hardware-counter profiles)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--methods", default="fp64,learned_mg,mg_cheb,learned1",
                    help="comma list of fp64 (FP64 Jacobi-CG, FP64 true-residual refinement), "
                         "ir32 (FP32 Jacobi-CG under FP64 refinement), learned1 (FP64 flexible "
                         "CG with the learned preconditioner), learned32 (its FP32 variant), "
                         "learned_mg (flexible CG, multilevel with the learned smoother), "
                         "mg_cheb (multilevel with a Chebyshev smoother)")
    ap.add_argument("--mg-levels", type=int, default=12)
    ap.add_argument("--mg-coarse", type=int, default=500)
    ap.add_argument("--mg-kcycle", type=int, default=3,
                    help="coarse levels on the K-cycle (learned fine smoother)")
    ap.add_argument("--mg-kcycle-cheb", type=int, default=16,
                    help="coarse levels on the K-cycle (Chebyshev fine smoother)")
    ap.add_argument("--train-n", type=int, default=128, help="training plate nodes per side")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--train-lr", type=float, default=1e-3, help="Adam step size")
    ap.add_argument("--train-lr-final", type=float, default=None,
                    help="step size at the end of a half-cosine decay (constant when absent)")
    ap.add_argument("--train-pairs", type=int, default=32, help="white-noise pairs per set")
    ap.add_argument("--train-seed", type=int, default=0, help="network and pair seed")
    ap.add_argument("--train-resample", type=int, default=0,
                    help="draw a new set of pairs every this many steps (0: one set)")
    ap.add_argument("--h-plates", type=int, default=2000, help="--config 4: training plates")
    ap.add_argument("--h-chunk", type=int, default=250, help="--config 4: plates per batch")
    ap.add_argument("--h-cells", type=int, default=128, help="--config 4: cells per side")
    ap.add_argument("--h-steps", type=int, default=20, help="--config 4: Adam steps per batch")
    ap.add_argument("--h-white", type=int, default=16,
                    help="--config 4: white-noise pair rows added per batch")
    ap.add_argument("--with-fp64", action="store_true",
                    help="--config 4: also time FP64 Jacobi-CG on the deployment plate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the whole plate sits in one MI355X's HBM and this "
              f"build has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    if args.config == 3:
        run_batch(args)
        return
    if args.config == 4:
        run_protocol(args)
        return

    import cstfem_hip
    from cstfem_hip import pipeline

    t = time.perf_counter()
    prob = pipeline.config_problem(args.config, args.scale)
    m = prob.mesh
    log(f"{prob.name}: {m.num_nodes} nodes, {m.num_elements} elements, {m.num_dofs} DOFs "
        f"(generated in {time.perf_counter() - t:.1f} s)")
    dev = cstfem_hip.Device(0)
    name, cus = dev.info()
    pipeline.load_problem(dev, prob)

    if args.kernels_only:
        # hot kernels only (for hardware-counter runs): one assembly, then `reps` launches of
        # each kernel class on the library stream
        pipeline.assemble_and_constrain(dev, prob)
        dev.setup()
        n, nnz = dev.n, dev.nnz
        out = {}
        for kind, name_k in ((0, "k_cg_spmv<F64>"), (5, "k_cg_spmv<F32>"), (2, "k_cg_update"),
                             (3, "k_cg_direction<F64>")):
            out[name_k] = dev.profile_kernel(kind, reps=args.reps)
        # flexible-CG kernels, with M at the polynomial starting point of the network
        from cstfem_hip import learned
        dev.mpnn_init(0)
        dev.mpnn_set_params(learned.polynomial_start(dev.mpnn_get_params()))
        dev.mpnn_build()
        for kind, name_k in ((6, "k_fcg_spmv<F64>"), (7, "k_fcg_precond<F64>"),
                             (8, "k_fcg_update<F64>"), (9, "k_fcg_direction<F64>")):
            out[name_k] = dev.profile_kernel(kind, reps=args.reps)
        out["k_copy(256MiB)"] = dev.profile_kernel(4, reps=args.reps, nbytes=256 << 20)
        out["spmv_bytes"] = spmv_bytes(n, nnz)
        print(json.dumps(out), flush=True)
        return

    uses_net = ("learned1", "learned32", "learned_mg")
    chosen = [m for m in args.methods.split(",") if m]
    theta, train_info = None, None
    if any(m in uses_net for m in chosen):
        from cstfem_hip import learned
        t0 = time.perf_counter()
        with cstfem_hip.Device(0) as tdev:
            theta, losses = learned.train(tdev, learned.training_problem(args.config,
                                                                         args.train_n),
                                          steps=args.train_steps, log=log,
                                          lr=args.train_lr, lr_final=args.train_lr_final,
                                          pairs=args.train_pairs, resample=args.train_resample,
                                          seed=args.train_seed)
        train_info = dict(train_s=time.perf_counter() - t0, plate_nodes=args.train_n ** 2,
                          steps=args.train_steps, lr=args.train_lr, lr_final=args.train_lr_final,
                          pairs=args.train_pairs, resample=args.train_resample,
                          seed=args.train_seed, loss_first=losses[0], loss_last=losses[-1])
        log(f"trained the preconditioner in {train_info['train_s']:.2f} s "
            f"(loss {losses[0]:.4f} -> {losses[-1]:.4f})")

    results = time_methods(dev, prob, args, chosen,
                           {m_: theta for m_ in chosen if m_ in uses_net})
    ok = {k: v for k, v in results.items() if v["verified"]} or results
    best = min(ok, key=lambda k: ok[k]["s"])
    per_step = results[best]["s"]
    fp64 = results.get("fp64")
    st_it = results[best]["iterations"]
    times = pipeline.PhaseTimes(**{k: v for k, v in results[best]["phases_s"].items()
                                   if k != "precond"})

    # dominant kernel: fused SpMV + p'Ap, average launch duration with HIP events
    n, nnz = dev.n, dev.nnz
    t_spmv = dev.profile_kernel(0, reps=50)
    t_upd = dev.profile_kernel(2, reps=50)
    t_dir = dev.profile_kernel(3, reps=50)
    copy_bytes = 1 << 30
    t_copy = dev.profile_kernel(4, reps=20, nbytes=copy_bytes)
    t_read = dev.profile_kernel(10, reps=20, nbytes=copy_bytes)
    t_spmv32 = dev.profile_kernel(5, reps=50)
    mg_levels = None
    if "learned_mg" in chosen or "mg_cheb" in chosen:
        dev.setup()
        if theta is not None:
            dev.mpnn_set_params(theta)
            dev.mpnn_build()
        dev.mg_setup(args.mg_levels, args.mg_coarse, 0 if theta is not None else 1,
                     args.mg_kcycle)
        mg_levels = [dict(zip(("dofs", "aggregates", "entries", "lam"), dev.mg_level_info(l)))
                     for l in range(dev.mg_levels())]
    fcg_kernels = None
    if theta is not None:
        dev.setup()
        dev.mpnn_set_params(theta)
        dev.mpnn_build()
        fcg_kernels = {name_k: dev.profile_kernel(kind, reps=50) for kind, name_k in
                       ((6, "k_fcg_spmv<F64>"), (7, "k_fcg_precond<F64>"),
                        (8, "k_fcg_update<F64>"), (9, "k_fcg_direction<F64>"))}
        log("flexible CG kernels (us): " + ", ".join(f"{k} {v * 1e6:.1f}"
                                                     for k, v in fcg_kernels.items()))
    b_spmv = spmv_bytes(n, nnz)
    achieved = b_spmv / t_spmv / 1e9
    copy_gbs = 2 * copy_bytes / t_copy / 1e9
    read_gbs = copy_bytes / t_read / 1e9
    vec_bytes_upd = 8 * n * 7     # reads x p r Ap invd, writes x r
    vec_bytes_dir = 8 * n * 4     # reads r invd p, writes p
    iter_bytes = b_spmv + 8 * n * 11
    log(f"spmv {t_spmv * 1e6:.1f} us ({achieved:.0f} GB/s), update {t_upd * 1e6:.1f} us "
        f"({vec_bytes_upd / t_upd / 1e9:.0f} GB/s), direction {t_dir * 1e6:.1f} us "
        f"({vec_bytes_dir / t_dir / 1e9:.0f} GB/s), copy {copy_gbs:.0f} GB/s, "
        f"read {read_gbs:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, (fp64 or results[best])["iterations"], args.cpu_iters)

    traffic, traffic_src = measured_traffic(m.num_dofs)
    out = {
        "metric": METRIC,
        "value": per_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[best],
        "precision": PRECISION[best],
        "mg_levels": mg_levels,
        "train": train_info,
        "fcg_kernel_s": fcg_kernels,
        "data": "synthetic",
        "config": {
            "workload": (f"{prob.name}: {m.num_dofs} DOFs, {m.num_elements} CST elements, "
                         f"assembly+BCs+Jacobi-CG to FP64 true rel residual 1e-10"),
            "dofs": m.num_dofs, "nnz": int(nnz), "elements": m.num_elements,
            "method": best, "device": name, "cus": cus,
        },
        "methods": results,
        "speedup_vs_fp64": (fp64["s"] / per_step) if fp64 else None,
        "iterations": st_it,
        "dofs_assembled_per_s": m.num_dofs / times.assembly if times.assembly > 0 else None,
        "spmv_gbs": achieved,
        "spmv_gbs_scalar_csr_model": (12 * nnz + 4 * (n + 1) + 16 * n) / t_spmv / 1e9,
        "copy_gbs": copy_gbs,
        "read_gbs": read_gbs,
        "spmv32_gbs": (4 * nnz + nnz + 4 * (n // 2 + 1) + 8 * n) / t_spmv32 / 1e9,
        "fp64_iteration_gbs": (iter_bytes / fp64["s_per_iteration"] / 1e9) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "frac_of_read_stream": achieved / read_gbs,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_cg_spmv", "bytes_per_launch": b_spmv,
                     "avg_launch_s": t_spmv},
        "roofline_fcg": ({"kernel": "k_fcg_spmv<F64>", "bound": "hbm",
                          "bytes_per_launch": b_spmv + 8 * n,
                          "avg_launch_s": fcg_kernels["k_fcg_spmv<F64>"],
                          "achieved": (b_spmv + 8 * n) / fcg_kernels["k_fcg_spmv<F64>"] / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": (b_spmv + 8 * n) / fcg_kernels["k_fcg_spmv<F64>"] / 1e9
                          / HBM_PEAK_GBS} if fcg_kernels else None),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
