"""Benchmark of the MI355X particle FMM on the headline workload.

Workload (one "step" = one FMM evaluation of the hot path, tree/lists/operators prebuilt,
densities already resident in HBM): Laplace potential + gradient (laplace_grad), N = 2^21
points on a strongly adaptive octree (50% on the sphere |x-0.5|=0.4, 30% in a Gaussian blob of
sigma 0.02 at (0.3,0.3,0.7), 20% uniform), densities N(0,1), m = 10, max_pts = 256.

The reported value is the mixed-precision evaluator's Mtargets/s (FP32 P2P, DFTs and up/down
GEMMs, the M2L Hadamard as scaled two-part fp16 splits on the matrix cores; FP64 operator
precompute and per-target merge).  The FP64 GPU evaluator is timed the same way and the ratio
reported; both errors are relative L2 against the FP64 all-pairs direct sum on the GPU (which
the tests check against the CPU direct sum) on 4096 sampled targets.  The CPU baseline is the
FP64 oracle (a restatement of the reference algorithm) on the same workload (--cpu-sample
points, by default all of them), on the host cores; it is the only use of the oracle here.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--n N] [--workload NAME]
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

MI355X_FP32_MFMA_TFLOPS = 157.3
MI355X_F16_MFMA_TFLOPS = 2500.0  # dense fp16 (and bf16)
F16_PRODUCTS_PER_FP32 = 3  # two-way fp16 split of both operands, three partial products kept
MI355X_FP32_VALU_TFLOPS = 78.6  # non-packed FP32 FMA (packed FP32 is not used)
MI355X_FP64_TFLOPS = 78.6
MI355X_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def time_steps(tree, dq, steps, warmup, timing):
    import torch

    for _ in range(warmup):
        tree.evaluate_device(dq, sync=True)
    tree.set_timing(timing)
    phase_sum = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = tree.evaluate_device(dq, sync=True)
        if timing:
            for k, v in tree.phase_times().items():
                phase_sum[k] = phase_sum.get(k, 0.0) + v
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tree.set_timing(False)
    phases = {k: v / steps for k, v in phase_sum.items()}
    return (t1 - t0) / steps, out, phases


def cpu_model() -> str:
    """/proc/cpuinfo's model name with its family, model and stepping numbers (a virtual machine
    may show only a generic name)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (t.strip() for t in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                    info[k] = v
            elif info:
                break
    except OSError:
        pass
    if "model name" not in info:
        return "unknown"
    return (f"{info['model name']} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')})")


def cpu_baseline(n_sample: int):
    """FP64 oracle on the workload (test infrastructure, baseline only): the repository's
    unoptimised C++/OpenMP restatement of the reference algorithm (per-node mat-vecs, its own
    DFT, no FFTW, BLAS or SIMD intrinsics), so it understates the reference's own CPU speed."""
    import oracle
    from pvfmm_amd import workloads as W

    c = W.adaptive_grad_2m(n_sample)
    ops = oracle.precompute(c["kernel"], c["m"])
    tr = oracle.OracleTree(c["src"], c["trg"], c["max_pts"])
    tr.set_operators(ops)
    threads = oracle.nthreads()
    tr.evaluate(c["den"], threads=threads)  # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        tr.evaluate(c["den"], threads=threads)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"value": n_sample / t / 1e6, "unit": "Mtargets/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"oracle (unoptimised FP64 C++/OpenMP restatement of the reference algorithm) FMM "
                      f"evaluation with tree/operators prebuilt, same distribution and m/max_pts at N={n_sample}, "
                      f"median of 3 after 1 warm-up, {t:.2f} s per evaluation"}


def read_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE runs, calibrated against a copy kernel of known size), or None if absent."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_adaptive_grad_2m_m10_mixed.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch") and f"[{kernel}]" in line:
            return float(line.split("=")[1])
    return None


def roofline(ph: dict, st: dict, m: int, real_bytes: int):
    """Roofline of the dominant kernel of the mixed evaluator from algorithmic counts.

    V-list Hadamard (scalar kernels): 512*k0*k1 flop per (V1 pair, frequency) (the reference's
    own count, include/fmm_pts.txx:3336) in FP32, which the kernel computes on the fp16 matrix
    cores as three fp16 products per FP32 product (both operands scaled and split exactly into
    two fp16 parts).  At the dense fp16 peak those products take less time than the kernel's
    compulsory HBM traffic, every source spectrum read once and every target spectrum written
    once (16*M_dim*k0 reals per source parent, 16*M_dim*k1 per target parent, the 26 V1
    matrices once), so with t_bound = max(bytes / HBM peak, flops / MFMA peak) the bound is
    HBM: `achieved` is those bytes over the kernel's time.  Beside it: the FP32-equivalent rate
    against the FP32 matrix peak, the fp16 product rate against the fp16 peak, and the
    target-stationary model's bytes (each pair re-reading its partner's spectrum), which the
    kernel avoids by staging each tile's partners in LDS once; `traffic` is the measured HBM
    bytes per launch.  U-list P2P: 16 flop per pair for laplace_grad, VALU-bound."""
    n1 = 2 * m
    M_dim = n1 * n1 * (n1 // 2 + 1)
    dom = max(ph, key=ph.get)
    t = ph[dom] * 1e-3
    if dom == "V-List:Hadamard":
        name = "k_hadamard_split<250, 496>"
        bytes_model = (16.0 * M_dim * real_bytes * st["v_pairs"] + 16.0 * M_dim * real_bytes * st["v_trg_slots"]
                       + 26.0 * M_dim * 128 * real_bytes)
        bytes_alg = (16.0 * M_dim * real_bytes * (st["v_src_slots"] + st["v_trg_slots"]) + 26.0 * M_dim * 128 * real_bytes)
        flops = 512.0 * st["v_pairs"] * M_dim
        achieved = bytes_alg / t / 1e9
        return {"kernel": name + " (V-list frequency-domain M2L, FP32 via split fp16 MFMA)", "bound": "hbm",
                "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS,
                "traffic": read_traffic(name), "ms_per_launch": ph[dom], "algorithmic_bytes": bytes_alg,
                "algorithmic_flops": flops,
                "t_bound_ms": {"hbm": 1e3 * bytes_alg / (MI355X_HBM_GBS * 1e9),
                               "fp16_mfma": 1e3 * F16_PRODUCTS_PER_FP32 * flops / (MI355X_F16_MFMA_TFLOPS * 1e12)},
                "fp32_equivalent_tflops": flops / t / 1e12,
                "fp32_equivalent_frac_of_fp32_mfma_peak": flops / t / 1e12 / MI355X_FP32_MFMA_TFLOPS,
                "fp16_tflops": F16_PRODUCTS_PER_FP32 * flops / t / 1e12,
                "fp16_frac_of_fp16_mfma_peak": F16_PRODUCTS_PER_FP32 * flops / t / 1e12 / MI355X_F16_MFMA_TFLOPS,
                "target_stationary_model_bytes": bytes_model,
                "target_stationary_model_gbs": bytes_model / t / 1e9,
                "work": f"{st['v_src_slots']} source and {st['v_trg_slots']} target parents x {M_dim} frequencies x "
                        f"16 reals, read and written once; 512 flop x {st['v_pairs']} V1 pairs x {M_dim} frequencies "
                        f"in FP32, each FP32 product as {F16_PRODUCTS_PER_FP32} fp16 products on dense 16-target MFMA "
                        f"groups (offset pairs without partners skipped) from LDS-staged spectra"}
    if dom == "U-List":
        name = "k_p2p_leaf<float, 1, 1>"
        flops = 16.0 * st["u_pairs"]
        achieved = flops / t / 1e12
        return {"kernel": name + " (U-list P2P)", "bound": "valu", "achieved": achieved, "peak": MI355X_FP32_VALU_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / MI355X_FP32_VALU_TFLOPS, "traffic": read_traffic(name),
                "ms_per_launch": ph[dom], "work": f"16 flop x {st['u_pairs']} pairs"}
    return {"kernel": dom, "bound": "hbm", "achieved": None, "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": None,
            "traffic": None, "ms_per_launch": ph[dom]}


def copy_probe(nbytes: int, reps: int = 3):
    """A plain device copy of known size, run under rocprofv3 --pmc to calibrate FETCH_SIZE and
    WRITE_SIZE (their reading for nbytes read and nbytes written)."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").uniform_()
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: {reps} x {nbytes} bytes")


def bwd_fwd(steps: int, warmup: int, n: int):
    """Energy E = 1/2 q . phi (laplace_poten, uniform points, m = 10, max_pts = 256, mixed) with
    the adjoint-FMM backward: per-step times of E and of dE/dx, dE/dq."""
    import torch

    from pvfmm_amd import workloads as W
    from pvfmm_amd.autograd import LaplaceFMM

    c = W.uniform_energy_1m(n)
    fmm = LaplaceFMM(c["src"], m=c["m"], max_pts=c["max_pts"], precision="mixed")
    x = torch.tensor(c["src"], device="cuda:0", requires_grad=True)
    q = torch.tensor(c["den"], device="cuda:0", requires_grad=True)
    tf = tb = 0.0
    for it in range(warmup + steps):
        x.grad = q.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = fmm.energy(x, q)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        E.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= warmup:
            tf += t1 - t0
            tb += t2 - t1
    tf /= steps
    tb /= steps
    return {"workload": f"laplace_poten energy, N={len(c['src'])} uniform, m={c['m']}, max_pts={c['max_pts']}, mixed",
            "fwd_ms": 1e3 * tf, "bwd_ms": 1e3 * tb, "ratio": tb / tf}


def bench_workload(name: str, steps: int, warmup: int, check: bool):
    """Secondary workloads (one JSON line each): uniform-poten-64k (laplace_poten FP64, 65,536
    uniform points); ellipsoid-stokes-1m (stokes_vel on an ellipsoid, mixed and FP64);
    densities-32 (32 density vectors on one tree, mixed); clouds-64 (64 independent clouds as one
    graph launch, mixed); periodic-poten-1m (laplace_poten in the periodic unit cube, mixed and
    FP64); energy-1m (energy forward/backward)."""
    import torch

    import pvfmm_amd as P
    from pvfmm_amd import workloads as W

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    def err(kernel, src, den, out, k1):
        if not check:
            return None
        rng = np.random.default_rng(4096)
        idx = np.sort(rng.choice(len(src), min(4096, len(src)), replace=False))
        ud = P.direct(kernel, src, den, src[idx])  # FP64 all-pairs sum on the GPU
        return rel_l2(out.reshape(len(src), k1)[idx].ravel(), ud)

    line = {"n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic"}
    if name in ("uniform-poten-64k", "ellipsoid-stokes-1m"):
        c = W.uniform_poten_64k() if name == "uniform-poten-64k" else W.ellipsoid_stokes_1m()
        n = len(c["src"])
        k0, k1 = P.kernel_dims(c["kernel"])
        tree = P.PtFMMTree(c["src"], c["trg"], c["max_pts"])
        dq = torch.from_numpy(c["den"]).to("cuda:0")
        res = {}
        for prec in (["fp64"] if name == "uniform-poten-64k" else ["mixed", "fp64"]):
            tree.setup(P.PtFMM(c["kernel"], c["m"], prec))
            shapes = tree.autotune(dq)
            t = timed(lambda: tree.evaluate_device(dq))
            out = tree.evaluate_device(dq).cpu().numpy()
            tree.set_timing(True)
            tree.evaluate_device(dq)
            phases = {k: round(v, 3) for k, v in tree.phase_times().items()}
            tree.set_timing(False)
            res[prec] = {"ms_per_step": 1e3 * t, "mtargets_s": n / t / 1e6, "rel_l2": err(c["kernel"], c["src"], c["den"], out, k1),
                         "launch_shapes": shapes, "phases_ms": phases}
            log(f"{name} {prec}: {1e3 * t:.2f} ms ({n / t / 1e6:.1f} Mtargets/s), rel-L2 {res[prec]['rel_l2']}")
            log("   phases (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()))
        main_prec = "fp64" if name == "uniform-poten-64k" else "mixed"
        line.update(metric=f"{c['kernel']} FMM Mtargets/s", value=res[main_prec]["mtargets_s"], unit="Mtargets/s",
                    ms_per_step=res[main_prec]["ms_per_step"], dtype="f64" if main_prec == "fp64" else "f32",
                    config={"workload": f"{name}: {c['kernel']}, N={n}, m={c['m']}, max_pts={c['max_pts']}"},
                    results=res)
    elif name == "densities-32":
        c = W.uniform_poten_1m_32_densities()
        n, nd = c["den"].shape[1], c["den"].shape[0]
        tree = P.PtFMMTree(c["src"], c["trg"], c["max_pts"]).setup(P.PtFMM(c["kernel"], c["m"], "mixed"), n_vec_max=nd)
        dq = torch.from_numpy(c["den"]).to("cuda:0")
        shapes = tree.autotune(dq)  # tuned on the batched pass of all nd vectors
        t = timed(lambda: tree.evaluate_device(dq))
        out = tree.evaluate_device(dq)[0].cpu().numpy()
        e = err(c["kernel"], c["src"], c["den"][0], out, 1)
        tree.set_timing(True)
        tree.evaluate_device(dq)
        phases = tree.phase_times()
        log(f"   phases of the {nd}-vector pass (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()))
        tree.set_timing(False)
        # the same densities one at a time on a tree set up for one vector
        tree1 = P.PtFMMTree(c["src"], c["trg"], c["max_pts"]).setup(P.PtFMM(c["kernel"], c["m"], "mixed"), n_vec_max=1)
        t1 = timed(lambda: tree1.evaluate_device(dq))
        log(f"{name}: {1e3 * t:.2f} ms for {nd} densities in one pass ({n * nd / t / 1e6:.1f} Mtarget-densities/s), "
            f"{1e3 * t1:.2f} ms one at a time, rel-L2 {e}")
        line.update(metric="laplace_poten FMM Mtarget-densities/s", value=n * nd / t / 1e6, unit="Mtargets/s",
                    ms_per_step=1e3 * t, dtype="f32", rel_l2=e, one_at_a_time_ms=1e3 * t1, launch_shapes=shapes,
                    phases_ms={k: round(v, 3) for k, v in phases.items()},
                    config={"workload": f"{name}: laplace_poten, N={n} uniform, {nd} density vectors in one batched "
                                        f"pass, m={c['m']}, max_pts={c['max_pts']}, mixed"})
    elif name == "clouds-64":
        c = W.uniform_clouds_64()
        fmm = P.PtFMM(c["kernel"], c["m"], "mixed")
        trees = [P.PtFMMTree(x, x, c["max_pts"]).setup(fmm) for x, _ in c["clouds"]]
        dens = [torch.from_numpy(q).to("cuda:0") for _, q in c["clouds"]]
        outs = [torch.empty_like(d) for d in dens]
        ntot = sum(len(q) for _, q in c["clouds"])

        def run_streams():
            for tr, d, o in zip(trees, dens, outs):
                tr.evaluate_device(d, out=o, sync=False)  # each tree has its own stream
            for tr in trees:
                torch.cuda.ExternalStream(tr.stream_handle()).synchronize()

        t_streams = timed(run_streams)
        # launch shapes per cloud from the autotuner, then one graph launch for all clouds
        t0 = time.perf_counter()
        tunes = [tr.autotune(d) for tr, d in zip(trees, dens)]
        t_tune = time.perf_counter() - t0
        batch = P.PtFMMBatch(trees)
        t = timed(lambda: batch.evaluate_device(dens, outs))
        e = err(c["kernel"], c["clouds"][0][0], c["clouds"][0][1], outs[0].cpu().numpy(), 1)
        shapes = {}
        for tu in tunes:
            key = f"{tu['hadamard_tiles_per_block']}/{tu['gemm_split_blocks']}"  # the one-vector shapes
            shapes[key] = shapes.get(key, 0) + 1
        log(f"{name}: {1e3 * t:.2f} ms for {len(trees)} clouds, {ntot} targets ({ntot / t / 1e6:.1f} Mtargets/s) "
            f"as one graph launch; {1e3 * t_streams:.2f} ms with one stream per cloud before tuning; "
            f"autotune {t_tune:.1f} s, shapes (tiles per block / split blocks: clouds) {shapes}")
        line.update(metric="laplace_poten FMM Mtargets/s over 64 clouds", value=ntot / t / 1e6, unit="Mtargets/s",
                    ms_per_step=1e3 * t, dtype="f32", rel_l2_cloud0=e, streams_ms=1e3 * t_streams,
                    autotune_s=t_tune, autotuned_shapes=shapes, graph_records=batch.records(),
                    config={"workload": f"{name}: 64 uniform clouds, {ntot} points, m={c['m']}, max_pts={c['max_pts']}, "
                                        "mixed, autotuned launch shapes, one HIP graph launch"})
    elif name == "periodic-poten-1m":
        # periodic boundaries: both evaluators at m = 8; the error is against the FP64 evaluator
        # at m = 10 on the same tree, with the mean difference removed (a periodic potential's
        # constant is convention); the Ewald-sum checks of the periodic path are in the tests
        c = W.uniform_periodic_1m()
        n = len(c["src"])
        tree = P.PtFMMTree(c["src"], c["trg"], c["max_pts"], bndry=c["bndry"])
        dq = torch.from_numpy(c["den"]).to("cuda:0")
        t0 = time.perf_counter()
        tree.setup(P.PtFMM(c["kernel"], 10, "fp64"))
        t_setup10 = time.perf_counter() - t0
        uref = tree.evaluate_device(dq).cpu().numpy()
        res = {}
        for prec in ("mixed", "fp64"):
            t0 = time.perf_counter()
            tree.setup(P.PtFMM(c["kernel"], c["m"], prec))
            t_setup = time.perf_counter() - t0
            shapes = tree.autotune(dq)
            t = timed(lambda: tree.evaluate_device(dq))
            u = tree.evaluate_device(dq).cpu().numpy()
            d = u - uref
            e = float(np.linalg.norm(d - d.mean()) / np.linalg.norm(uref - uref.mean()))
            tree.set_timing(True)
            tree.evaluate_device(dq)
            phases = {k: round(v, 3) for k, v in tree.phase_times().items()}
            tree.set_timing(False)
            res[prec] = {"ms_per_step": 1e3 * t, "mtargets_s": n / t / 1e6, "rel_l2_vs_fp64_m10": e,
                         "setup_s": t_setup, "launch_shapes": shapes, "phases_ms": phases}
            log(f"{name} {prec}: {1e3 * t:.2f} ms ({n / t / 1e6:.1f} Mtargets/s), rel-L2 vs FP64 m=10 {e:.3e}, "
                f"setup {t_setup:.2f} s")
            log("   phases (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()))
        line.update(metric="laplace_poten periodic FMM Mtargets/s", value=res["mixed"]["mtargets_s"], unit="Mtargets/s",
                    ms_per_step=res["mixed"]["ms_per_step"], dtype="f32", results=res, setup_fp64_m10_s=t_setup10,
                    config={"workload": f"{name}: laplace_poten, periodic (PXYZ), N={n} uniform, m={c['m']}, "
                                        f"max_pts={c['max_pts']}"})
    elif name == "energy-1m":
        bf = bwd_fwd(steps, warmup, 1048576)
        line.update(metric="laplace energy t_bwd / t_fwd", value=bf["ratio"], unit="ratio", higher_is_better=False,
                    ms_per_step=bf["fwd_ms"] + bf["bwd_ms"], dtype="f32", config={"workload": f"{name}: {bf['workload']}"},
                    bwd_fwd=bf)
    else:
        raise SystemExit(f"unknown workload {name}")
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2097152)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the direct-sum error check (profiling runs)")
    ap.add_argument("--cpu-sample", type=int, default=2097152)
    ap.add_argument("--no-equal-error", action="store_true", help="skip the FP64 equal-error point")
    ap.add_argument("--prec", choices=["both", "mixed", "fp64"], default="both",
                    help="evaluators to time (profiling runs); the JSON line needs both")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a device copy of this many bytes")
    ap.add_argument("--no-bwd", action="store_true", help="skip the energy forward/backward timing")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--workload", default="adaptive-grad-2m", choices=["adaptive-grad-2m"] + sorted(
        k for k in ("uniform-poten-64k", "ellipsoid-stokes-1m", "densities-32", "clouds-64", "periodic-poten-1m",
                    "energy-1m")),
                    help="adaptive-grad-2m (headline, default) or a secondary workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the evaluation of one MPI_COMM_SELF share); --gpus {args.gpus} "
              "is not supported: the multi-GPU sharding is a design, not built or measured", file=sys.stderr)
        sys.exit(2)

    import torch

    import pvfmm_amd as P
    from pvfmm_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload != "adaptive-grad-2m":
        bench_workload(args.workload, args.steps, args.warmup, not args.no_check)
        return
    c = W.adaptive_grad_2m(args.n)
    n = len(c["src"])
    log(f"workload: {c['kernel']} N={n} m={c['m']} max_pts={c['max_pts']}")
    t0 = time.perf_counter()
    tree = P.PtFMMTree(c["src"], c["trg"], c["max_pts"])
    t_tree = time.perf_counter() - t0
    dq = torch.from_numpy(c["den"]).to("cuda:0")
    rng = np.random.default_rng(4096)
    idx = np.sort(rng.choice(n, 4096, replace=False))
    # FP64 direct sum on the sampled targets (error reference; the GPU all-pairs sum, which the
    # tests compare with the CPU direct sum)
    ud = None if args.no_check else P.direct(c["kernel"], c["src"], c["den"], c["src"][idx]).reshape(-1, 3)

    res = {}
    for prec in (["fp64", "mixed"] if args.prec == "both" else [args.prec]):
        t0 = time.perf_counter()
        fmm = P.PtFMM(c["kernel"], c["m"], prec)
        t_ops = time.perf_counter() - t0
        t0 = time.perf_counter()
        tree.setup(fmm)
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        shapes = tree.tuning() if args.no_tune else tree.autotune(dq)
        t_tune = time.perf_counter() - t0
        step, out, phases = time_steps(tree, dq, args.steps, args.warmup, timing=True)
        # un-instrumented timing of the same steps
        step_plain, out, _ = time_steps(tree, dq, args.steps, 0, timing=False)
        u = out.cpu().numpy().reshape(n, 3)[idx]
        res[prec] = dict(step=step_plain, step_timed=step, phases=phases, err=None if ud is None else rel_l2(u, ud), t_ops=t_ops,
                         t_setup=t_setup, t_tune=t_tune, shapes=shapes, stats=tree.stats())
        log(f"{prec}: {1e3 * step_plain:.2f} ms/eval ({n / step_plain / 1e6:.1f} Mtargets/s), rel-L2 {res[prec]['err'] if res[prec]['err'] is None else format(res[prec]['err'], '.3e')}, "
            f"operators {t_ops:.1f}s, setup {t_setup:.1f}s, launch shapes {shapes} ({'default' if args.no_tune else f'autotuned in {t_tune:.1f}s'})")
        log("   phases (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()))
        st = res[prec]["stats"]
        log(f"   V-list tiles {st['v_tiles']}, partners {st['v_tile_partners']} (largest union "
            f"{st['v_tile_max_partners']}), staged {st.get('v_tile_staged', st['v_tile_partners'])}")

    if args.prec != "both":
        return
    mixed, fp64 = res["mixed"], res["fp64"]

    # Equal-error comparison.  The error of each evaluator at m = 6, 8, 10 (FP64: truncation in
    # m; mixed: truncation or FP32 rounding, whichever is larger) and its time; for an error
    # level eps each evaluator takes the smallest m whose error is at most eps (the cheapest),
    # and the ratio of those times is the speedup at equal error.  Reported for eps = the mixed
    # evaluator's error at each m, and eps = 1e-6.
    eq = None
    if not args.no_equal_error and mixed["err"] is not None:
        frontier = {"fp64": {c["m"]: (fp64["step"], fp64["err"])}, "mixed": {c["m"]: (mixed["step"], mixed["err"])}}
        for prec in ("fp64", "mixed"):
            for m_eq in (6, 8, 10):
                if m_eq in frontier[prec]:
                    continue
                tree.setup(P.PtFMM(c["kernel"], m_eq, prec))
                if not args.no_tune:
                    tree.autotune(dq)
                step_eq, out, _ = time_steps(tree, dq, max(3, args.steps // 2), 1, timing=False)
                e_eq = rel_l2(out.cpu().numpy().reshape(n, 3)[idx], ud)
                frontier[prec][m_eq] = (step_eq, e_eq)
                log(f"{prec} m={m_eq}: {1e3 * step_eq:.2f} ms/eval, rel-L2 {e_eq:.3e}")

        def cheapest(prec, eps):
            ok = [(t, mm, e) for mm, (t, e) in frontier[prec].items() if e <= eps]
            return min(ok) if ok else None

        levels = []
        for eps in sorted({e for _, e in frontier["mixed"].values()} | {1e-6}, reverse=True):
            f64, mx = cheapest("fp64", eps), cheapest("mixed", eps)
            levels.append({"rel_l2_bound": eps,
                           "fp64": None if f64 is None else {"m": f64[1], "ms_per_step": 1e3 * f64[0], "rel_l2": f64[2]},
                           "mixed": None if mx is None else {"m": mx[1], "ms_per_step": 1e3 * mx[0], "rel_l2": mx[2]},
                           "mixed_over_fp64": None if (f64 is None or mx is None) else f64[0] / mx[0]})
            log(f"equal error <= {eps:.2e}: fp64 m={f64 and f64[1]}, mixed m={mx and mx[1]}, "
                f"ratio {levels[-1]['mixed_over_fp64']}")
        eq = {"frontier": {prec: {str(mm): {"ms_per_step": 1e3 * t, "rel_l2": e} for mm, (t, e) in sorted(d.items())}
                           for prec, d in frontier.items()},
              "levels": levels}
    st = mixed["stats"]
    m = c["m"]
    roof = roofline(mixed["phases"], st, m, real_bytes=4)

    bf = None
    if not args.no_bwd:
        bf = bwd_fwd(max(3, args.steps // 2), 1, 1048576)
        log(f"energy fwd {bf['fwd_ms']:.2f} ms, bwd {bf['bwd_ms']:.2f} ms, ratio {bf['ratio']:.2f}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(min(args.cpu_sample, n))
            log(f"cpu baseline: {cpu['value']:.4f} Mtargets/s on {cpu['cores']} threads")
        except Exception as e:  # noqa: BLE001 - the baseline is reported, not required
            log(f"cpu baseline failed: {e}")

    value = n / mixed["step"] / 1e6
    line = {
        "metric": "Laplace FMM Mtargets/s, N=2M, mixed vs FP64 GPU at equal L2 err; bwd/fwd ratio",
        "value": value,
        "unit": "Mtargets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"laplace_grad FMM evaluation, N={n} adaptive (50% sphere r=0.4, 30% gaussian "
                               f"sigma=0.02, 20% uniform), m={m}, max_pts={c['max_pts']}, mixed FP32 bulk (V-list "
                               f"Hadamard products as scaled 2-part fp16 splits on the matrix cores) / FP64 "
                               f"precompute and merge", "n_targets": n, "m": m, "max_pts": c["max_pts"]},
        "fp64_mtargets_s": n / fp64["step"] / 1e6,
        "fp64_ms_per_step": 1e3 * fp64["step"],
        "mixed_over_fp64": fp64["step"] / mixed["step"],
        "rel_l2_mixed": mixed["err"],
        "rel_l2_fp64": fp64["err"],
        "equal_error": eq,
        "bwd_fwd_ratio": None if bf is None else bf["ratio"],
        "bwd_fwd": bf,
        "setup_s": {"tree": t_tree, "operators_mixed": mixed["t_ops"], "plan_mixed": mixed["t_setup"],
                    "autotune_mixed": mixed["t_tune"]},
        "launch_shapes": {"mixed": mixed["shapes"], "fp64": fp64["shapes"], "autotuned": not args.no_tune},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
