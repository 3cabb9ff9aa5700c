"""Benchmark of the field-evaluation hot path on one MI355X.

Workload: the log (line-charge) FMM at N = 1,048,576 -- half the particles uniform on
[-1, 1]^2, half in 16 Gaussian clusters (sigma 0.02, centres uniform on [-0.8, 0.8]^2),
charges +-1, order 12, theta 0.5, softening 1e-4, leaf capacity 8 -- with inputs resident
on the GPU.  One step = FlatQuadtree(...) (tree build + moments) + .fmm() (traversal, M2L,
L2L, L2P + P2P): the whole job for one set of positions.  The mixed-precision step is the
metric; the all-FP64 step is timed beside it, and the three errors against an FP64 direct
sum are reported on 4096 sampled targets.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
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

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
# measured (rocprofv3 --pmc, calibrated by a 256 MiB copy) for this workload's M2L launch
M2L_TRAFFIC_BYTES = 923.2e6
KERNEL_NAMES = ("m2l_mfma", "m2l_mfma64", "m2l", "m2l_fixup", "eval", "l2l", "rows_count", "rows_scan",
                "rows_write",
                "traverse",
                "row_sort", "lists_fill",
                "lists_degree", "sort_scatter", "sort_hist", "scan", "tree_keys", "tree_count",
                "tree_nodes", "tree_leaf_keys", "tree_gather", "moments_p2m", "moments_m2m",
                "scale_moments")
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def make_inputs(n: int, seed: int = 7):
    """The benchmark's synthetic particles (positions (2, n), charges (n,))."""
    rng = np.random.default_rng(seed)
    half = n // 2
    centres = rng.uniform(-0.8, 0.8, (2, 16))
    per = (n - half) // 16
    clusters = (centres[:, :, None] + rng.normal(0.0, 0.02, (2, 16, per))).reshape(2, -1)
    rest = n - half - 16 * per
    pos = np.concatenate((rng.uniform(-1, 1, (2, half + rest)), clusters), axis=1).clip(-1, 1)
    q = rng.choice([-1.0, 1.0], n)
    return np.ascontiguousarray(pos), q


def relative(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _timed(fn, repeats=3):
    import torch

    fn()
    torch.cuda.synchronize()
    best = float("inf")
    out = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, out


def extra_config1(td3):
    """Plain Barnes-Hut, Coulomb, FP64: N = 5000 uniform, charges +-1, theta 0.5,
    eps 1e-3, own charge excluded, against the FP64 direct sum."""
    import torch

    rng = np.random.default_rng(1)
    pos = torch.tensor(rng.uniform(-1, 1, (2, 5000)), device="cuda")
    q = torch.tensor(rng.choice([-1.0, 1.0], 5000), device="cuda")
    own = torch.arange(5000, device="cuda")

    def run():
        tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1))
        return tree.evaluate(pos, exclude=own, theta=0.5, softening=1e-3)

    t, (phi, field) = _timed(run)
    dphi, dfield = td3.direct_field(pos, pos, q, softening=1e-3, exclude=own)
    return {"workload": "BH walk, Coulomb, N=5000, theta 0.5, eps 1e-3 (build + walk)",
            "ms": 1e3 * t, "targets_per_s": 5000 / t,
            "err_vs_direct": {"phi": relative(phi.cpu().numpy(), dphi.cpu().numpy()),
                              "field": relative(field.cpu().numpy(), dfield.cpu().numpy())}}


def extra_config3(td3, n_src=1 << 22, grid=1024):
    """Coulomb fmm_at, order 10: 4,194,304 sources from a 2D Plummer-like profile (scale
    0.1, inside [-1, 1]^2), charges U[-1, 1], at a 1024 x 1024 grid over [-1, 1]^2;
    mixed vs FP64, and both against the FP64 direct sum at 2048 sampled grid points."""
    import torch

    rng = np.random.default_rng(3)
    pts = np.empty((2, 0))
    while pts.shape[1] < n_src:
        u = rng.uniform(0, 1, n_src)
        r = 0.1 * np.sqrt(u / (1 - u))
        a = rng.uniform(0, 2 * np.pi, n_src)
        cand = np.stack((r * np.cos(a), r * np.sin(a)))
        cand = cand[:, np.all(np.abs(cand) <= 1, axis=0)]
        pts = np.concatenate((pts, cand), axis=1)
    pos = torch.tensor(np.ascontiguousarray(pts[:, :n_src]), device="cuda")
    q = torch.tensor(rng.uniform(-1, 1, n_src), device="cuda")
    g = np.linspace(-1, 1, grid)
    gx, gy = np.meshgrid(g, g)
    targets = torch.tensor(np.stack((gx.ravel(), gy.ravel())), device="cuda")
    out = {"workload": f"Coulomb fmm_at, order 10, theta 0.5, eps 0: {n_src} Plummer sources, "
                       f"{grid}x{grid} grid targets (build + plan + FMM)"}
    res = {}
    for prec in ("mixed", "fp64"):
        def run():
            tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), order=10, precision=prec)
            return tree.fmm_at(targets, theta=0.5)

        t, (phi, field) = _timed(run, repeats=2)
        res[prec] = (phi.cpu().numpy(), field.cpu().numpy())
        out[prec] = {"ms": 1e3 * t, "targets_per_s": grid * grid / t}
    idx = np.sort(rng.choice(grid * grid, 2048, replace=False))
    dphi, dfield = td3.direct_field(targets[:, idx], pos, q)
    dphi, dfield = dphi.cpu().numpy(), dfield.cpu().numpy()
    out["err"] = {
        "mixed_vs_fp64": {"phi": relative(res["mixed"][0], res["fp64"][0]),
                          "field": relative(res["mixed"][1], res["fp64"][1])},
        "fp64_vs_direct": {"phi": relative(res["fp64"][0][idx], dphi),
                           "field": relative(res["fp64"][1][:, idx], dfield)},
        "mixed_vs_direct": {"phi": relative(res["mixed"][0][idx], dphi),
                            "field": relative(res["mixed"][1][:, idx], dfield)},
    }
    return out


def pmc_copies():
    """Two copies of a 256 MiB device buffer: the known traffic that calibrates the
    FETCH_SIZE / WRITE_SIZE counters of a --pmc run."""
    import torch

    src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()


def extra_config4(td3, systems=512, steps=100, graph=True):
    """Batched ensemble: 512 Coulomb Barnes-Hut systems, N log-uniform in [1024, 65536]
    (seed 4), Gaussian clouds (sigma 0.1) of unit charges, B_z 1, k_e 1, dt 1e-4,
    theta 0.5, eps 1e-3; every step one forest build, one walk launch, two push launches.
    graph=False leaves out the recorded steps (counter runs take a few direct steps)."""
    rng = np.random.default_rng(4)
    sizes = np.round(np.exp(rng.uniform(np.log(1024), np.log(65536), systems))).astype(int)
    n = int(sizes.sum())
    pos = rng.normal(0.0, 0.1, (2, n)).clip(-0.999, 0.999)
    roots = np.tile([0.0, 0.0, 1.0], (systems, 1))
    ens = td3.Ensemble.from_arrays(pos, np.zeros((2, n)), np.ones(n), np.ones(n), sizes, roots,
                                   theta=0.5, softening=1e-3, k_e=1.0, dt=1e-4, B_z=1.0)
    import torch

    e0 = ens.total_energies()
    waves = ens.autotune()  # the walk's launch shape, timed on this ensemble
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    drift = np.abs(ens.total_energies() - e0) / np.abs(e0)
    direct = {"workload": f"{systems} systems, sum N = {n}, {steps} steps, method tree",
              "ms_per_step": 1e3 * t / steps, "particle_steps_per_s": n * steps / t,
              "walk_waves_per_block": waves,
              "energy_drift": {"median": float(np.median(drift)), "max": float(np.max(drift))}}
    if not graph:
        return direct
    # the same steps as one recorded HIP graph per step (Ensemble.run): the first step
    # direct, the second recording and replaying, the rest timed replays
    rec = td3.Ensemble.from_arrays(pos, np.zeros((2, n)), np.ones(n), np.ones(n), sizes, roots,
                                   theta=0.5, softening=1e-3, k_e=1.0, dt=1e-4, B_z=1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.run(2, graph=True)
    torch.cuda.synchronize()
    t_record = time.perf_counter() - t0
    t0 = time.perf_counter()
    rec.run(steps - 2, graph=True)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    same = all(torch.equal(getattr(rec, k), getattr(ens, k))
               for k in ("positions", "velocities", "forces"))
    return {**direct,
            "graph": {"ms_per_step": 1e3 * t_graph / (steps - 2),
                      "particle_steps_per_s": n * (steps - 2) / t_graph,
                      "first_two_steps_ms": 1e3 * t_record,
                      "node_capacity": rec.graph_capacity,
                      "states_equal_step": bool(same)}}


def extra_config5(td3, n=262144, probes=8192):
    """Differentiable log FMM: N = 262,144 uniform, charges +-1, order 12; loss = mean
    squared potential at 8192 probes + 1/2 sum q phi; forward + adjoint backward."""
    import torch

    rng = np.random.default_rng(5)
    pos = torch.tensor(rng.uniform(-1, 1, (2, n)), device="cuda", requires_grad=True)
    q = torch.tensor(rng.choice([-1.0, 1.0], n), device="cuda", requires_grad=True)
    pr = torch.tensor(rng.uniform(-1, 1, (2, probes)), device="cuda")

    def run():
        phi, phi_p = td3.log_potentials(pos, q, pr, root=td3.Quad(0, 0, 1))
        loss = (phi_p ** 2).mean() + 0.5 * torch.dot(q, phi)
        return torch.autograd.grad(loss, (pos, q))

    def run_fields():
        # the same with the probe fields in the loss: dipole sources in the transposed FMM
        # and the field's Jacobian at the probes
        phi, _, phi_p, field_p = td3.log_fields(pos, q, pr, root=td3.Quad(0, 0, 1))
        loss = (phi_p ** 2).mean() + (field_p ** 2).sum(0).mean() + 0.5 * torch.dot(q, phi)
        return torch.autograd.grad(loss, (pos, q))

    t, _ = _timed(run, repeats=2)
    t_fields, _ = _timed(run_fields, repeats=2)
    return {"workload": f"log FMM N={n}, {probes} probes, order 12: loss + dL/dq + dL/dpos",
            "ms": 1e3 * t,
            "fields_loss_ms": 1e3 * t_fields,
            "fields_loss": "+ mean squared field at the probes (log_fields)"}


def extras(td3):
    out = {}
    for name, fn in (("config1", extra_config1), ("config3", extra_config3),
                     ("config4", extra_config4), ("config5", extra_config5)):
        try:
            out[name] = fn(td3)
        except Exception as exc:  # reported, not raised: the metric above stands alone
            out[name] = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"extra {name} done", file=sys.stderr, flush=True)
    return out


def cpu_baseline(n_sample: int, seconds: float):
    """The CPU oracle (a C/OpenMP restatement of the reference's Pyccel kernels) on a
    bounded sample of the same workload: build + moments + FMM, all host cores."""
    from oracle import oracle

    pos, q = make_inputs(n_sample, seed=11)
    runs, t_total = 0, 0.0
    while t_total < seconds or runs < 1:
        t0 = time.perf_counter()
        tree = oracle.OracleTree(pos, q, (0.0, 0.0, 1.0), kernel="log", order=12)
        tree.fmm(theta=0.5, softening=1e-4)
        t_total += time.perf_counter() - t0
        runs += 1
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": n_sample * runs / t_total,
        "unit": "targets/s",
        "cores": cores,
        "kind": "port",
        "sample": f"log FMM N={n_sample} (same distribution, seed 11), order 12, theta 0.5, "
        f"eps 1e-4, leaf 8: build + moments + traversal + M2L + L2L + L2P/P2P, {runs} runs, "
        f"{t_total:.1f} s, OpenMP threads={cores}",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--pmc", action="store_true",
                    help="one warm mixed step and a 256 MiB device copy only (with --extra "
                         "config4: three direct ensemble steps and the copy); for rocprofv3 "
                         "--pmc runs: the copy calibrates FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other configurations (ensemble, fmm_at, adjoint, walk)")
    ap.add_argument("--extra", choices=["config1", "config3", "config4", "config5"],
                    help="run only this other configuration and print its JSON (for "
                         "profiling it); the metric line is not printed")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    td3 = __graft_entry__.load_package()
    from tree_d_3d_amd import _abi

    if args.extra:
        if args.pmc and args.extra == "config4":
            # counter runs: three direct ensemble steps, then the copies that calibrate
            # FETCH_SIZE / WRITE_SIZE
            print(json.dumps({"config4": extra_config4(td3, steps=3, graph=False)}), flush=True)
            pmc_copies()
            return 0
        fn = {"config1": extra_config1, "config3": extra_config3, "config4": extra_config4,
              "config5": extra_config5}[args.extra]
        print(json.dumps({args.extra: fn(td3)}), flush=True)
        return 0

    n = args.n
    pos_h, q_h = make_inputs(n)
    pos = torch.from_numpy(pos_h).cuda()
    q = torch.from_numpy(q_h).cuda()
    quad = td3.Quad(0.0, 0.0, 1.0)
    params = dict(theta=0.5, softening=1e-4)

    def step(precision):
        tree = td3.FlatQuadtree(pos, q, quad, kernel="log", order=12, precision=precision)
        phi, field = tree.fmm(**params)
        return tree, phi, field

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = step(precision)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, out

    if args.pmc:
        step("mixed")
        torch.cuda.synchronize()
        step("mixed")
        pmc_copies()
        print(json.dumps({"pmc": "one mixed step, then two 256 MiB copies"}))
        return 0
    t_mixed, (tree, phi_m, field_m) = timed("mixed", args.steps, args.warmup)
    result = {
        "metric": "targets/s at N=1M FMM: mixed vs FP64 on GPU at <=1e-6 rel err vs direct",
        "value": n / t_mixed,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (M2L below depth 5 and P2P; tree, moments, top M2L, L2L, L2P and sums f64)",
        "data": "synthetic",
        "config": {
            "workload": "log-kernel FMM, build + moments + traversal + M2L + L2L + L2P/P2P",
            "n": n,
            "distribution": "half uniform [-1,1]^2, half 16 Gaussian clusters sigma 0.02",
            "order": 12, "theta": 0.5, "softening": 1e-4, "leaf_capacity": 8,
            "node_count": tree.node_count, "fmm_pairs": list(tree.fmm_pairs),
            "levels": int(len(tree.levels) - 1),
        },
    }
    if not args.no_fp64:
        t_fp64, (_, phi_d, field_d) = timed("fp64", args.steps, args.warmup)
        result["fp64"] = {"ms_per_step": 1e3 * t_fp64, "value": n / t_fp64,
                          "mixed_speedup": t_fp64 / t_mixed}
        # errors on 4096 sampled targets against the FP64 direct sum (GPU, FP64)
        idx = np.sort(np.random.default_rng(0).choice(n, 4096, replace=False))
        tidx = torch.from_numpy(idx).cuda()
        dphi, dfield = td3.direct_field(pos[:, tidx], pos, q, exclude=tidx, kernel="log",
                                        softening=1e-4)
        dphi, dfield = dphi.cpu().numpy(), dfield.cpu().numpy()
        pm, fm = phi_m[tidx].cpu().numpy(), field_m[:, tidx].cpu().numpy()
        pd, fd = phi_d[tidx].cpu().numpy(), field_d[:, tidx].cpu().numpy()
        result["accuracy"] = {
            "metric": "max-norm relative error max|a-b|/max|b| on 4096 sampled targets",
            "mixed_vs_fp64_fmm": {"phi": relative(pm, pd), "field": relative(fm, fd)},
            "fp64_fmm_vs_direct": {"phi": relative(pd, dphi), "field": relative(fd, dfield)},
            "mixed_vs_direct": {"phi": relative(pm, dphi), "field": relative(fm, dfield)},
            "note": "the FMM softens only near pairs (reference semantics); far cells are "
                    "unsoftened, so FMM-vs-softened-direct carries an O(eps^2/d^2) model error",
        }

    # per-kernel times of one more step of each precision (HIP events on the launch stream)
    def kernel_times(precision):
        _abi.timing_enable(True)
        tree, _, _ = step(precision)
        torch.cuda.synchronize()
        out = {}
        for name in KERNEL_NAMES:
            ms, count = _abi.timing_get(name)
            if count:
                out[name] = {"ms": ms, "launches": count}
        _abi.timing_enable(False)
        return tree, out

    if not args.no_fp64:
        result["fp64"]["kernels_ms"] = kernel_times("fp64")[1]
    tree, kernels = kernel_times("mixed")
    result["kernels_ms"] = kernels
    # roofline of the dominant kernel: algorithmic flops per launch / average launch time
    (mrow, mcol, _), (prow, pcol, _), counts = tree._lists(0.5)
    node_n = (tree._d["node_end"] - tree._d["node_start"]).cpu().numpy()
    prow_h = prow.cpu().numpy().astype(np.int64)
    pcol_h = pcol.cpu().numpy()[: prow_h[-1]].astype(np.int64)
    dst = np.repeat(np.arange(len(prow_h) - 1), np.diff(prow_h))
    near_pairs = int(np.sum(node_n[dst] * node_n[pcol_h]) - np.sum(node_n[node_n > 0]))
    m2l_directed = 2 * counts[0]
    # the M2L on the matrix cores: destinations deeper than the FP64 top levels
    fp64_top = tree.FP64_TOP_DEPTH
    mrow_h = mrow.cpu().numpy().astype(np.int64)
    first_fp32 = int(tree.levels[min(fp64_top + 1, len(tree.levels) - 1)])
    mfma_pairs = int(mrow_h[-1] - mrow_h[first_fp32])
    ms = kernels["m2l_mfma"]["ms"] / kernels["m2l_mfma"]["launches"]
    achieved = 910.0 * mfma_pairs / (ms * 1e-3) / 1e12
    result["roofline"] = {
        "kernel": "m2l_mfma_kernel (FP32 M2L on v_mfma_f32_16x16x4_f32, destinations below "
                  "the FP64 top levels)",
        "bound": "mfma",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        # HBM bytes of one m2l_mfma_kernel launch, rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE
        # x 1.000 (factors from a 256 MiB copy in the same run)
        "traffic": M2L_TRAFFIC_BYTES,
        "algorithmic": {"m2l_directed_pairs": m2l_directed, "mfma_directed_pairs": mfma_pairs,
                        "launch_ms": ms, "near_directed_pairs": near_pairs,
                        "flops_per_m2l_pair": 910, "flops_per_near_pair": 17,
                        "traffic_source": "profiles/fmm_log_n1m_hbm_traffic.md"},
    }
    if not args.no_extras:
        result["extras"] = extras(td3)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(200_000, args.cpu_seconds)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
