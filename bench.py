"""Benchmark of the PIGFEM assemble-and-solve path on one MI355X.

Workload (default, the 3D cohesive configuration): unit cube, 128^3 cubes x 6
TET4 (12.6 M elements), 200 seeded spheres (RSA, radii U[0.03, 0.06]) with
cohesive interfaces (bilinear law, sigma_c = 30 MPa, delta_c = 1e-3,
delta_f = 1e-2), matrix E=3.2 GPa nu=0.35, inclusion E=70 GPa nu=0.22, left
face fixed, right face u_x = 0.02 applied in 20 equal increments.

A step is one full FP64 Jacobian + internal-force assembly over every element
(the reference's assemble_nonlinear(func=true, jac=true)), cohesive law and
damage update included, scattered into the CSR matrix with the reference's
sparsity pattern; inputs are resident in HBM.  value = elements assembled per
second.  The same JSON line carries the roofline of the dominant assembly
kernel (and of the other pass), the time to a 1e-8 FP64 residual for the
first two load increments in four modes (plain FP64 PCG; previous increment
only; learned network + previous increment; the same with FP32 matrix
values), each mode's displacement error against the plain FP64 solve, the
time-to-residual roofline fraction, SpMV bandwidth and a CPU baseline (the
oracle's reference-structured element loop and its Jacobi-PCG, timed on this
host on 1 core and on all cores).  --workload c1 / c3 run the 2D single-mesh
and the batched-RVE workloads instead.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_elem(dim, nn=4):
    """Per element: connectivity, coordinates + displacements (FP64), K_e
    read-modify-write, f_e read-modify-write (FP64)."""
    m = dim * nn
    return 4 * nn + 16 * m + 16 * m * m + 16 * m


def gather_bytes(dim, nnz_rows, n_rows, n_all, n_elem):
    """Row-gather assembly, minimum HBM traffic per call: every K entry of the
    assembled rows written once (FP64), the internal force written once,
    coordinates and displacements of every node read once (FP64), element
    connectivity (int32 x 4) and material (1 B) read once."""
    return 8 * nnz_rows + 8 * n_rows + 2 * 8 * dim * n_all + 17 * n_elem


def tune_gather_layout(ctx, flags, reps=5, log=print):
    """Time the gather pass for every layout, 1, 2, 4 or 8 threads per unit
    and a unit being one matrix row or all rows of a node (results are
    bit-identical), and keep the fastest; untimed set-up, like an autotuner."""
    ms = {}
    layouts = [(slots, node_rows) for node_rows in (0, 1) for slots in (1, 2, 4, 8)]
    names = {}
    for slots, node_rows in layouts:
        ctx.gather_layout = (slots, node_rows)
        ctx.assemble(flags)
        name = "%s%d" % ("node" if node_rows else "row", slots)
        names[name] = (slots, node_rows)
        ms[name] = ctx.time_assemble_split(flags, reps=reps)[2]
    best = min(ms, key=ms.get)
    ctx.gather_layout = names[best]
    log("# gather pass ms by layout (threads per row / per node): %s -> %s" % (
        ", ".join("%s: %.3f" % (k, v) for k, v in ms.items()), best))
    return best, ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--n-incl", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-learn", action="store_true", help="skip the learned-guess load steps")
    ap.add_argument("--workload", choices=["c1", "c2", "c3"], default="c2")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device five times (calibrates the HBM byte counters)")
    ap.add_argument("--c2-n", type=int, default=128)
    ap.add_argument("--c2-incl", type=int, default=200)
    ap.add_argument("--c2-steps", type=int, default=2, help="load increments timed (of 20 to 2%% strain)")
    ap.add_argument("--c2-modes", default="plain_fp64,prev_fp64_forcing,learned_fp64_forcing,learned_fp32_forcing",
                    help="comma list of " + ",".join(ALL_MODES))
    ap.add_argument("--c3-rves", type=int, default=256)
    ap.add_argument("--c3-nmin", type=int, default=64)
    ap.add_argument("--c3-nmax", type=int, default=512)
    ap.add_argument("--c3-slots", type=int, default=1, help="inclusion radii per RVE with direct sensitivities")
    ap.add_argument("--c3-persistent", action="store_true",
                    help="also time the load-case solve as one persistent cooperative launch (small batches only)")
    ap.add_argument("--setup", choices=["device", "host"], default="device",
                    help="where the IGFEM preprocessing (detection, enrichment, integration tables) runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        return copy_probe()
    if args.workload == "c2":
        return bench_c2(args)
    if args.workload == "c3":
        return bench_c3(args)

    from pigfem_assignment_amd import configs, _lib as L
    from pigfem_assignment_amd.api import DeviceContext

    t0 = time.time()
    prob = configs.c1_problem(n=args.n, n_incl=args.n_incl)
    setup_s = time.time() - t0
    info = prob.info
    print("# setup: %d elements (%d cut), %d dofs, host set-up %.2f s" % (info.n_elem, info.n_cut, info.n_free,
                                                                       setup_s), flush=True)
    ctx = DeviceContext(prob)
    print("# device context ready: nnz %d, %d colours" % (ctx.csr.nnz, ctx.csr.n_colors), flush=True)

    flags = L.PGF_ASM_JAC | L.PGF_ASM_FUNC
    slots, slot_ms = tune_gather_layout(ctx, flags, log=lambda m: print(m, flush=True))
    for _ in range(args.warmup):
        ctx.assemble(flags)
    ms = ctx.time_assemble(flags, reps=args.steps)
    value = info.n_elem / (ms * 1e-3)
    # dominant kernel: the row-gather pass (k_gather); cut elements are a
    # separate, much smaller pass (k_cut_elem_mfma)
    ms_call, ms_cut, ms_gather = ctx.time_assemble_split(flags, reps=args.steps)
    nnz_free_rows = int(ctx.pattern()[0][info.n_free])
    bytes_gather = gather_bytes(2, nnz_free_rows, info.n_free, info.n_all_nodes, info.n_elem)
    achieved = bytes_gather / (ms_gather * 1e-3) / 1e9
    print("# assembly K+f FP64 (row gather): %.3f ms/step (cut pass %.3f ms, gather pass %.3f ms), %.3e elem/s, "
          "gather pass %.1f GB/s algorithmic" % (ms, ms_cut, ms_gather, value, achieved), flush=True)
    ms_k32 = ctx.time_assemble(L.PGF_ASM_JAC | L.PGF_ASM_FP32, reps=args.steps)
    ms_f64 = ctx.time_assemble(L.PGF_ASM_FUNC, reps=args.steps)
    ms_col = ctx.time_assemble(flags | L.PGF_ASM_COLORED, reps=args.steps)
    print("# assembly K+f FP64 (coloured RMW): %.3f ms/step, %.1f GB/s algorithmic (RMW bytes)" % (
        ms_col, info.n_elem * algorithmic_bytes_per_elem(2) / (ms_col * 1e-3) / 1e9), flush=True)

    extra = {"gather_layout": slots, "gather_pass_ms_by_layout": slot_ms,
             "assembly_gather_ms": ms, "assembly_gather_cut_pass_ms": ms_cut,
             "assembly_gather_pass_ms": ms_gather, "assembly_gather_bytes": bytes_gather,
             "assembly_colored_ms": ms_col,
             "assembly_colored_elem_per_s": info.n_elem / (ms_col * 1e-3),
             "assembly_colored_GBs_rmw_bytes": info.n_elem * algorithmic_bytes_per_elem(2) / (ms_col * 1e-3) / 1e9}
    # SpMV (K_ff) bandwidth: 12 B/nnz + 4 B/row + 16 B/row (x once, y write)
    ctx.assemble(L.PGF_ASM_JAC)
    ctx.assemble(L.PGF_ASM_JAC | L.PGF_ASM_FP32)
    nnz_ff, nf = ctx.csr.nnz_ff, ctx.csr.n_free
    ms_spmv = ctx.time_spmv(64, reps=50)
    ms_spmv32 = ctx.time_spmv(32, reps=50)
    extra["spmv_fp64_GBs"] = (12 * nnz_ff + 4 * (nf + 1) + 16 * nf) / (ms_spmv * 1e-3) / 1e9
    # FP32 values with FP64 vectors (the product the FP32 solve mode runs)
    extra["spmv_fp32_GBs"] = (8 * nnz_ff + 4 * (nf + 1) + 16 * nf) / (ms_spmv32 * 1e-3) / 1e9
    extra["assembly_fp32K_ms"] = ms_k32
    extra["assembly_fp64f_ms"] = ms_f64
    extra["spmv_fp64_ms"] = ms_spmv
    extra["spmv_fp32_ms"] = ms_spmv32
    extra["spmv_fp64_roofline_frac"] = extra["spmv_fp64_GBs"] / HBM_PEAK_GBS
    extra["spmv_fp32_roofline_frac"] = extra["spmv_fp32_GBs"] / HBM_PEAK_GBS

    if not args.no_solve:
        # time to a 1e-8 FP64 residual for the load step
        for prec, lin in ((64, 1e-9), (32, 1e-5)):
            ctx.set_solution(np.zeros((2, info.n_all_nodes)))
            si = ctx.newton_step(1.0, rel_tol=1e-8, precision=prec, lin_rtol=lin, lin_maxit=200000)
            extra["t_to_1e-8_fp%d_ms" % prec] = si.ms_total
            extra["pcg_iters_fp%d" % prec] = si.pcg_iterations
            extra["newton_iters_fp%d" % prec] = si.newton_iterations
            extra["res_rel_fp%d" % prec] = si.res / si.res0 if si.res0 else 0.0
            print("# solve fp%d: %.1f ms, newton %d, pcg %d, rel res %.2e" % (prec, si.ms_total, si.newton_iterations,
                                                                           si.pcg_iterations, si.res / si.res0),
                  flush=True)

    if not args.no_learn:
        extra.update(learned_vs_plain(n=64, n_incl=40, steps=3, log=lambda m: print(m, flush=True)))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        n_sample = min(info.n_elem, 500_000)
        o = orc.Oracle(prob, prob.bc_list(), sample=(0, n_sample))
        _, _, cpu_ms = o.assemble()
        cpu = {"value": n_sample / (cpu_ms * 1e-3), "unit": "element assemblies/s", "cores": 1,
               "kind": "port",
               "sample": "K+f assembly of elements [0, %d) of the same %dx%d mesh (%d cut) by the CPU oracle: "
                         "reference-structured element loop with per-QP dense kernels and a map-based "
                         "MatSetValue scatter, one core" % (n_sample, args.n, args.n,
                                                            int(np.sum(prob.cut()[:, 0] < n_sample)))}
        print("# cpu oracle assembly of %d elements: %.1f ms" % (n_sample, cpu_ms), flush=True)

    line = {
        "metric": "element assemblies/s; time-to-1e-8 residual per load step, learned vs plain",
        "value": value,
        "unit": "element assemblies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D unit square %dx%d QUAD4, %d RSA circular inclusions r~U[0.005,0.02], "
                               "linear elastic plane strain, FP64 K+f assembly" % (args.n, args.n, args.n_incl),
                   "elements": info.n_elem, "cut_elements": info.n_cut, "dofs": info.n_free,
                   "nnz": ctx.csr.nnz, "colors": ctx.csr.n_colors},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic()},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


# (use_guess, precision, PCG tolerance floor, forcing term, FP32 iteration
# budget).  The FP32 mode solves with FP32 tangent values (FP64 Krylov
# vectors) to 1e-3 relative per Newton iteration and the FP64 Newton residual
# refines, as in iterative refinement, down to the 1e-8 target; a solve that
# has not converged after 2000 FP32 iterations continues in FP64.
ALL_MODES = {"plain_fp64": (0, 64, 1e-10, False, 0), "plain_fp64_forcing": (0, 64, 1e-10, True, 0),
             "prev_fp64_forcing": (3, 64, 1e-10, True, 0),
             "learned_fp64_forcing": (2, 64, 1e-10, True, 0), "learned_fp32_forcing": (2, 32, 1e-3, True, 2000)}


def copy_probe():
    """Plain device copy of a known size, for FETCH_SIZE / WRITE_SIZE calibration."""
    import torch
    n = 64 * 1024 * 1024  # float32 -> 256 MiB
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied 5 x %d bytes" % (4 * n), flush=True)


def measured_traffic(workload="c1", kernel="k_gather"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    f = os.path.join(ROOT, "profiles", "%s_%s_hbm_traffic.txt" % (kernel, workload))
    if workload == "c1" and kernel == "k_gather":
        f = os.path.join(ROOT, "profiles", "assembly_c1_hbm_traffic.txt")
    if not os.path.exists(f):
        return None
    for line in open(f):
        line = line.strip()
        if line.startswith("traffic ") and "MB per launch" in line:
            return float(line.split()[1]) * 1e6
    return None


def cut_pass_bytes(prob):
    """Algorithmic bytes of one cut-element pass (k_cut_elem_mfma): per cut element
    its integration-point tables (gradients of every slot, J*W, material),
    per cohesive point N, R, dA*W, node map, law, delta_max read and written
    and the flags written, the element's displacements read, and its K_e and
    P_e written once (FP64)."""
    d = prob.desc()
    nc = prob.info.n_cut
    if nc == 0:
        return 0
    dim = prob.dim
    ne = np.ctypeslib.as_array(d.cut_n_enrich, shape=(nc,)).astype(np.int64)
    vp = np.ctypeslib.as_array(d.vpt_ptr, shape=(nc + 1,)).astype(np.int64)
    cp = np.ctypeslib.as_array(d.cpt_ptr, shape=(nc + 1,)).astype(np.int64)
    M = dim * (4 + ne)
    nv, ncp = np.diff(vp), np.diff(cp)
    per = nv * (8 * M + 8 + 4) + ncp * (4 * 8 + 9 * 8 + 8 + 8 + 1 + 4 + 8 + 8 + 1) + 8 * M + 8 * M * M + 8 * M
    return int(per.sum())


def cut_scratch_bytes(prob):
    """K_e and P_e of every cut element, written by the cut pass and read once
    by the gather pass"""
    nc = prob.info.n_cut
    if nc == 0:
        return 0
    d = prob.desc()
    ne = np.ctypeslib.as_array(d.cut_n_enrich, shape=(nc,)).astype(np.int64)
    M = prob.dim * (4 + ne)
    return int((8 * M * M + 8 * M).sum())


def pcg_iteration_bytes(nnz_ff, n, value_bytes=8):
    """Algorithmic bytes of one Jacobi-PCG iteration: SpMV (values, int32
    column indices, row pointers, x read once, y written) plus two dot
    products, three axpys and the preconditioner as separate vector passes
    (the unfused sum; the fused kernels are reported against it)."""
    return (value_bytes + 4) * nnz_ff + 4 * (n + 1) + 16 * n + 128 * n


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def learned_vs_plain(n, n_incl, steps, prob=None, train_n=64, train_incl=20, train_steps=5, epochs=30, log=print,
                     ctx=None, modes=tuple(ALL_MODES), spmv_info=None):
    """Time-to-1e-8 residual per load step of a 3D cohesive inclusion problem
    in several modes: plain FP64 PCG from a zero start; the Galerkin-scaled
    previous increment only (the method without the network); the network's
    columns combined with the previous increment; the same with FP32 matrix
    values.  The network is trained first on plain solves of two smaller
    problems (its voxel grid does not depend on the mesh).  Every mode's
    displacement after every increment is compared with the plain FP64 one."""
    from pigfem_assignment_amd import learn, configs
    from pigfem_assignment_amd.api import GuessNet
    t0 = time.time()
    net = GuessNet(seed=3)
    out = {"learn_workload": "unit cube %d^3 x 6 TET4, %d spheres, bilinear cohesive, first %d of 20 increments" % (
        n, n_incl, steps)}
    if any(ALL_MODES[m][0] == 2 for m in modes):  # the network is trained only when a mode uses it
        train = [learn.cohesive_problem(train_n, train_incl, seed) for seed in (501, 502)]
        loss = learn.train_net(net, train, train_steps, 0.05, epochs=epochs, lr=1e-3, log=log)
        out.update({"learn_train_s": time.time() - t0, "learn_loss_first": float(loss[0]),
                    "learn_loss_last": float(loss[-1])})
    if prob is None:
        prob = (configs.c2_problem(n=n, n_incl=n_incl), configs.c2_inclusions(n_incl, n))
    dfrac = 0.05  # 20 increments to 2 % strain
    modes = [(m,) + ALL_MODES[m] for m in modes]
    flags, sols = {}, {}
    for name, ug, prec, lin, adapt, m32 in modes:
        U_steps = []
        recs, fl, ctx = learn.run_steps(prob[0], prob[1], steps, dfrac, use_guess=ug, net=net, lin_rtol=lin,
                                        precision=prec, ctx=ctx, adaptive_lin=adapt, fp32_maxit=m32,
                                        keep_solutions=U_steps)
        flags[name] = fl
        sols[name] = U_steps
        out["%s_ms_per_step" % name] = [r["ms"] for r in recs]
        out["%s_pcg_per_step" % name] = [r["pcg"] for r in recs]
        out["%s_newton_per_step" % name] = [r["newton"] for r in recs]
        out["%s_res_rel_per_step" % name] = [r["res"] / r["res0"] if r["res0"] else 0.0 for r in recs]
        if spmv_info is not None:
            nnz_ff, nf, b_asm = spmv_info
            fr = []
            for r in recs:
                b = r["pcg"] * pcg_iteration_bytes(nnz_ff, nf, 4 if prec == 32 else 8) + (r["newton"] + 1) * b_asm
                fr.append(b / (r["ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9))
            out["%s_time_to_residual_roofline_frac" % name] = fr
        if prec == 32:
            out["%s_fp64_fallback_per_step" % name] = [r["fp64_fallback"] for r in recs]
        log("# %s: ms per step %s, PCG %s, Newton %s%s" % (
            name, ["%.1f" % r["ms"] for r in recs], [r["pcg"] for r in recs], [r["newton"] for r in recs],
            ", FP64 fallback at Newton iteration %s" % [r["fp64_fallback"] for r in recs] if prec == 32 else ""))
    ref_name = "plain_fp64" if "plain_fp64" in sols else modes[0][0]
    for name in sols:
        errs = [float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(sols[name], sols[ref_name])]
        out["%s_U_rel_err_vs_%s" % (name, ref_name)] = errs
        log("# %s: displacement error vs %s per step %s" % (name, ref_name, ["%.2e" % e for e in errs]))
    ref = flags[modes[0][0]]
    same = all(np.array_equal(a, b) for m in flags for a, b in zip(ref, flags[m]))
    out["damage_flags_identical"] = bool(same)
    tot = {m[0]: sum(out["%s_ms_per_step" % m[0]]) for m in modes}
    for a, b in (("plain_fp64", "learned_fp32_forcing"), ("plain_fp64", "learned_fp64_forcing"),
                 ("plain_fp64", "prev_fp64_forcing"), ("prev_fp64_forcing", "learned_fp64_forcing"),
                 ("prev_fp64_forcing", "learned_fp32_forcing")):
        if a in tot and b in tot:
            out["speedup_%s_vs_%s" % (b, a)] = tot[a] / tot[b]
            log("# %s vs %s: %.2fx" % (b, a, tot[a] / tot[b]))
    log("# damage flags identical across modes: %s" % same)
    return out


def cpu_baselines(prob, ctx, n_sample, log, threads):
    """The oracle's reference-structured element loop on a bounded sample of
    elements (1 core and `threads` cores), and its Jacobi-PCG iteration on the
    device-assembled K_ff (1 core and `threads` cores)."""
    from oracle import oracle as orc
    info = prob.info
    o = orc.Oracle(prob, prob.bc_list(), sample=(0, n_sample))
    n_cut_sample = int(np.sum(prob.cut()[:, 0] < n_sample))
    _, _, ms1 = o.assemble(dt=0.05)
    orc.set_threads(threads)
    _, _, msN = o.assemble(dt=0.05)
    orc.set_threads(1)
    del o
    log("# cpu oracle assembly of %d elements (%d cut): %.1f ms on 1 core, %.1f ms on %d cores" % (
        n_sample, n_cut_sample, ms1, msN, threads))
    rp, rs, ci = ctx.pattern()
    K = ctx.matrix()
    nf = info.n_free
    b = np.ones(nf)
    it1, itN = 3, 10
    pcg1 = orc.pcg_ms_per_iter(rp[: nf + 1], rs[:nf], ci, K, b, it1, 1)
    pcgN = orc.pcg_ms_per_iter(rp[: nf + 1], rs[:nf], ci, K, b, itN, threads)
    log("# cpu Jacobi-PCG iteration on the same K_ff: %.1f ms on 1 core (%d iterations), %.1f ms on %d cores "
        "(%d iterations)" % (pcg1, it1, pcgN, threads, itN))
    return {"assembly_elem_per_s_1core": n_sample / (ms1 * 1e-3), "assembly_elem_per_s_all": n_sample / (msN * 1e-3),
            "assembly_sample_elements": n_sample, "assembly_sample_cut": n_cut_sample,
            "pcg_ms_per_iteration_1core": pcg1, "pcg_ms_per_iteration_all": pcgN, "threads_all": threads}


def bench_c2(args):
    from pigfem_assignment_amd import configs, _lib as L
    from pigfem_assignment_amd.api import DeviceContext
    t0 = time.time()
    prob = configs.c2_problem(n=args.c2_n, n_incl=args.c2_incl, device_preprocess=args.setup == "device")
    incl = configs.c2_inclusions(args.c2_incl, args.c2_n)
    info = prob.info
    setup_s = time.time() - t0
    print("# C2 setup: %d tets (%d cut), %d dofs, %d cohesive points, set-up %.1f s (IGFEM phases on the %s: "
          "nodes %.1f, elements %.1f, enrichment %.1f, tables %.1f ms)" % (
              info.n_elem, info.n_cut, info.n_free, info.n_cpt, setup_s, args.setup, *info.phase_ms), flush=True)
    ctx = DeviceContext(prob)
    flags = L.PGF_ASM_JAC | L.PGF_ASM_FUNC
    slots, slot_ms = tune_gather_layout(ctx, flags, log=lambda m: print(m, flush=True))
    for _ in range(args.warmup):
        ctx.assemble(flags, dt=0.05)
    ms = ctx.time_assemble(flags, reps=args.steps)
    value = info.n_elem / (ms * 1e-3)
    ms_call, ms_cut, ms_gather = ctx.time_assemble_split(flags, reps=args.steps)
    nnz_free_rows = int(ctx.pattern()[0][info.n_free])
    scratch = cut_scratch_bytes(prob)
    bytes_gather = gather_bytes(3, nnz_free_rows, info.n_free, info.n_all_nodes, info.n_elem) + scratch
    bytes_cut = cut_pass_bytes(prob)
    gb_gather = bytes_gather / (ms_gather * 1e-3) / 1e9
    gb_cut = bytes_cut / (ms_cut * 1e-3) / 1e9
    print("# C2 assembly K+f FP64: %.2f ms (cut pass %.2f ms, %.0f GB/s; gather pass %.2f ms, %.0f GB/s), "
          "%.3e elem/s" % (ms, ms_cut, gb_cut, ms_gather, gb_gather, value), flush=True)
    # the 3D cut pass is the matrix-core kernel k_cut_elem_mfma
    if ms_cut >= ms_gather:
        dom = ("k_cut_elem_mfma", gb_cut, bytes_cut, ms_cut)
        other = {"kernel": "k_gather", "achieved": gb_gather, "frac": gb_gather / HBM_PEAK_GBS, "ms": ms_gather,
                 "algorithmic_bytes": bytes_gather, "traffic": measured_traffic("c2", "k_gather")}
    else:
        dom = ("k_gather", gb_gather, bytes_gather, ms_gather)
        other = {"kernel": "k_cut_elem_mfma", "achieved": gb_cut, "frac": gb_cut / HBM_PEAK_GBS, "ms": ms_cut,
                 "algorithmic_bytes": bytes_cut, "traffic": measured_traffic("c2", "k_cut_elem_mfma")}
    extra = {"assembly_ms": ms, "cut_pass_ms": ms_cut, "gather_pass_ms": ms_gather, "cut_pass_bytes": bytes_cut,
             "gather_pass_bytes": bytes_gather, "gather_layout": slots, "gather_pass_ms_by_layout": slot_ms,
             "roofline_other_pass": other, "setup_s": setup_s, "setup_igfem_on": args.setup,
             "setup_igfem_phase_ms": list(info.phase_ms), "cohesive_points": info.n_cpt}
    # K_ff SpMV of the Newton solves: FP64 values, and FP32 values with FP64 vectors
    ctx.assemble(L.PGF_ASM_JAC, dt=0.05)
    ctx.assemble(L.PGF_ASM_JAC | L.PGF_ASM_FP32, dt=0.05)
    nnz_ff, nf = ctx.csr.nnz_ff, ctx.csr.n_free
    for prec, vb in ((64, 12), (32, 8)):
        t = ctx.time_spmv(prec, reps=20)
        gbs = (vb * nnz_ff + 4 * (nf + 1) + 16 * nf) / (t * 1e-3) / 1e9
        extra["spmv_fp%d_ms" % prec] = t
        extra["spmv_fp%d_GBs" % prec] = gbs
        extra["spmv_fp%d_roofline_frac" % prec] = gbs / HBM_PEAK_GBS
        print("# C2 SpMV K_ff (%d nonzeros) %s: %.3f ms, %.0f GB/s (%.2f of 8 TB/s)" % (
            nnz_ff, "FP64" if prec == 64 else "FP32 values / FP64 vectors", t, gbs, gbs / HBM_PEAK_GBS), flush=True)
    cpu = None
    if not args.no_cpu:
        model, avail = cpu_info()
        threads = max(1, min(16, avail))
        c = cpu_baselines(prob, ctx, min(info.n_elem, 100_000), lambda m: print(m, flush=True), threads)
        c.update({"cpu_model": model, "cpus_available": avail})
        extra["cpu"] = c
        cpu = {"value": c["assembly_elem_per_s_1core"], "unit": "element assemblies/s", "cores": 1, "kind": "port",
               "sample": "K+f assembly of elements [0, %d) (%d cut) of the same 128^3 x 6 TET4 cohesive mesh by "
                         "the CPU oracle: reference-structured element loop, per-QP dense kernels, map-based "
                         "MatSetValue scatter, one core of %s; %d-core figures and the CPU Jacobi-PCG iteration "
                         "time on the same K_ff are in extra.cpu" % (
                             c["assembly_sample_elements"], c["assembly_sample_cut"], model, threads)}
    if not args.no_learn:
        b_asm = bytes_gather + bytes_cut
        extra.update(learned_vs_plain(args.c2_n, args.c2_incl, args.c2_steps, prob=(prob, incl),
                                      log=lambda m: print(m, flush=True), ctx=ctx, modes=args.c2_modes.split(","),
                                      spmv_info=(nnz_ff, nf, b_asm)))
        if cpu is not None and "plain_fp64_pcg_per_step" in extra:
            c = extra["cpu"]
            est = [p * c["pcg_ms_per_iteration_all"] for p in extra["plain_fp64_pcg_per_step"]]
            c["plain_fp64_time_to_residual_ms_estimate_all"] = est
            print("# cpu plain FP64 time to 1e-8 per increment (GPU PCG iteration counts x CPU ms per iteration "
                  "on %d cores): %s ms" % (c["threads_all"], ["%.0f" % e for e in est]), flush=True)
    line = {
        "metric": "element assemblies/s; time-to-1e-8 residual per load step, learned vs plain",
        "value": value, "unit": "element assemblies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D unit cube %d^3 x 6 TET4, %d RSA spheres r~U[0.03,0.06], bilinear cohesive "
                               "interfaces (sigma_c 30 MPa, delta_c 1e-3, delta_f 1e-2), FP64 K+f assembly; load "
                               "increments of 0.1%% strain to a 1e-8 residual" % (args.c2_n, args.c2_incl),
                   "elements": info.n_elem, "cut_elements": info.n_cut, "dofs": info.n_free,
                   "cohesive_points": info.n_cpt, "nnz_ff": nnz_ff},
        "roofline": {"bound": "hbm", "achieved": dom[1], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom[1] / HBM_PEAK_GBS, "traffic": measured_traffic("c2", dom[0]),
                     "kernel": dom[0], "algorithmic_bytes": dom[2], "ms": dom[3]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


def bench_c3(args):
    """Batch of periodic 2D RVEs (the reference's linear subscale model): one
    batched assembly, the three macro-strain load cases of every RVE solved by
    one segmented multi-RHS PCG to 1e-8, C_eff from the volume-averaged
    stress, then the direct sensitivities dC_eff/dr of the first --c3-slots
    inclusion radii of every RVE (each slot: pseudo-loads of 3 cases x all
    RVEs in one more batched solve with the same K)."""
    from pigfem_assignment_amd import batch, sens, _lib as L
    from pigfem_assignment_amd.api import DeviceContext
    rng = np.random.RandomState(20261202)
    sizes = rng.randint(args.c3_nmin, args.c3_nmax + 1, size=args.c3_rves)
    t0 = time.time()
    parts = []
    for k, N in enumerate(sizes):
        parts.append(batch.periodic_rve_problem(int(N), 1000 + k, device_preprocess=args.setup == "device")[0])
        if k % 64 == 63:
            print("# C3 set-up: %d of %d RVEs built, %.1f s" % (k + 1, len(sizes), time.time() - t0), flush=True)
    B = batch.BatchProblem(parts)
    setup_s = time.time() - t0
    phases = np.sum([p.info.phase_ms[:] for p in parts], axis=0)
    print("# C3 set-up: %d periodic RVEs, %d elements (%d cut), %d dofs, set-up %.1f s (IGFEM phases on the %s, "
          "summed: nodes %.0f, elements %.0f, enrichment %.0f, tables %.0f ms)" % (
              len(parts), B.n_elem, B.n_cut, B.n_free, setup_s, args.setup, *phases), flush=True)
    ctx = DeviceContext(B)
    flags = L.PGF_ASM_JAC | L.PGF_ASM_ALL_ROWS
    slots, _ = tune_gather_layout(ctx, flags, log=lambda m: print(m, flush=True))
    for _ in range(args.warmup):
        ctx.assemble(flags)
    ms_asm = ctx.time_assemble(flags, reps=args.steps)
    F = batch.macro_strain_rhs(ctx, B)
    ctx.assemble(L.PGF_ASM_JAC)
    X, its, rel, ms_solve, tune = ctx.batch_pcg(B.sys_ptr, F, rtol=1e-8)
    print("# C3 batched PCG (graph replays): %.1f ms" % ms_solve, flush=True)
    ms_persist, same = None, None
    if args.c3_persistent:
        # the same solve as one persistent cooperative launch (bit-identical)
        try:
            Xp, itp, _, ms_persist, _ = ctx.batch_pcg(B.sys_ptr, F, rtol=1e-8, tune=tune, persistent=True)
            same = bool(np.array_equal(Xp, X) and np.array_equal(itp, its))
            del Xp
            print("# C3 batched PCG as one persistent launch: %.1f ms (graph replays %.1f ms), identical results %s"
                  % (ms_persist, ms_solve, same), flush=True)
        except RuntimeError as e:
            print("# C3 persistent launch refused: %s" % e, flush=True)
    Ce = batch.homogenized_stiffness(ctx, B, X, B.elem_ptr)
    print("# C3 assembly %.2f ms (%.3e elem/s); batched PCG 3 cases x %d RVEs: %.1f ms, iterations %d..%d, "
          "tuned (threads/row, rows/tile) %s" % (ms_asm, B.n_elem / (ms_asm * 1e-3), len(parts), ms_solve,
                                                 its.min(), its.max(), tune), flush=True)
    extra = {"batched_solve_ms": ms_solve, "batched_solve_persistent_ms": ms_persist,
             "persistent_equals_graph": same, "iterations_min": int(its.min()), "iterations_max": int(its.max()),
             "iterations_total": int(its.sum()), "autotuned": list(tune), "gather_layout": slots,
             "setup_s": setup_s, "setup_igfem_on": args.setup, "setup_igfem_phase_ms": phases.tolist(),
             "C_eff_mean": Ce.mean(axis=0).tolist()}
    if args.c3_slots > 0:
        t1 = time.time()
        sens.attach(ctx, *sens.batch_tables(list(parts)))
        dC, info = sens.stiffness_sensitivities(ctx, X, B.elem_ptr, B.sys_ptr, args.c3_slots, batch.VOIGT_CASES,
                                                rtol=1e-8, log=lambda m: print(m, flush=True))
        wall = time.time() - t1
        extra.update({"sensitivity_slots": args.c3_slots,
                      "sensitivity_solve_ms": [float(i[2]) for i in info],
                      "sensitivity_iterations_max": [int(i[0].max()) for i in info],
                      "sensitivity_wall_s": wall,
                      "dC_dr0_mean": dC[:, 0].mean(axis=0).tolist()})
        print("# C3 sensitivities: %d slot(s) x 3 cases x %d RVEs, batched solves %s ms, %.1f s wall" % (
            args.c3_slots, len(parts), ["%.0f" % i[2] for i in info], wall), flush=True)
    line = {
        "metric": "element assemblies/s; time-to-1e-8 residual per load step, learned vs plain",
        "value": B.n_elem / (ms_asm * 1e-3), "unit": "element assemblies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_asm, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d periodic 2D RVEs N~U[%d,%d] QUAD4, 20-40%% inclusions, 3 macro-strain load "
                               "cases + radius sensitivities of %d slot(s), one batched assembly + segmented "
                               "multi-RHS PCG to 1e-8" % (len(parts), args.c3_nmin, args.c3_nmax, args.c3_slots),
                   "elements": B.n_elem, "dofs": B.n_free},
        "roofline": None, "cpu_baseline": None, "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
