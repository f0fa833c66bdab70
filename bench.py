#!/usr/bin/env python3
"""LDL forward+backward throughput on one MI355X.

Workload (the headline configuration): 512^3 particles on a 512^3 mesh, box
L = 409.6 Mpc/h (0.8 Mpc/h cells), positions held as FP32 displacements from
the Lagrangian lattice (seeded Zel'dovich field, rms 2 cells), 2 sm-kind LDL
layers, parameters [alpha, gamma, kh, kl, nu] = [0.01, 1, 0.1, 1, 0] per layer
and [mu, w, b] = [1, 1, 0]; target = the model at seed-perturbed (+-10%)
parameters plus 1% Gaussian noise; loss = L2 of the residual smoothed with a
1-cell Gaussian.  One step = loss and gradient w.r.t. all 13 parameters.

Prints one JSON line (rank 0) with the metric, the dominant kernel's roofline
(HIP-event timing of its launches inside the timed region), the plain FP64
GPU path on the same inputs, the FP32-vs-FP64 relative errors and a CPU
baseline (the FP64 oracle restatement on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="particles per side = mesh per side")
    ap.add_argument("--nstep", type=int, default=2, help="LDL layers")
    ap.add_argument("--save", default="positions", choices=["all", "positions"],
                    help="positions: checkpoint positions at layer boundaries and recompute the layers in the backward")
    ap.add_argument("--variant-steps", type=int, default=5,
                    help="timed steps of the other memory modes (save-all, every layer recomputed; 0: skip)")
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed steps of the plain FP64 path (0: skip)")
    ap.add_argument("--cpu-n", type=int, default=512, help="per-side size of the CPU baseline's one timed iteration (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: every core this process may use)")
    ap.add_argument("--cpu1-n", type=int, default=128, help="per-side size of the 1-thread CPU sample (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="size of the plain device copy timed during setup (its HBM counters calibrate traffic)")
    ap.add_argument("--fft-reps", type=int, default=5, help="timed standalone r2c and c2r transforms of the mesh (0: skip)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ inputs
def zeldovich_gpu(n, box, rms_cells, seed, device):
    """Lattice + Zel'dovich displacement (box units, float64); the recipe of
    oracle.ldl_oracle.zeldovich_displacement, built on the GPU."""
    from jerald_amd.synthetic import zeldovich_gpu as z
    return z(n, box, rms_cells, seed, device)


def params0(nstep):
    import numpy as np
    return np.array([0.01, 1.0, 0.1, 1.0, 0.0] * nstep + [1.0, 1.0, 0.0])


# --------------------------------------------------------- algorithmic bytes
def kernel_bytes(name, n, b):
    """Algorithmic (compulsory) HBM bytes of one launch of an instrumented
    kernel at n^3 particles on an n^3 mesh, b bytes per real.  Np = Nm = n^3,
    Nh = n^2 (n/2+1) complex values of 2b bytes.  The paint's kernels (sort
    and deposit) have no bytes of their own: the paint is measured as one
    operation (``op_bytes``)."""
    Np = Nm = n ** 3
    Nh = n * n * (n // 2 + 1)
    return {
        "readout3_update": 6 * b * Np + 3 * b * Nm,      # pos in + out, 3 fields
        "sm_bwd_gather": 9 * b * Np + 4 * b * Nm,        # pos, g_out in, g_in out, 4 fields
        "paint_vjp_pos": 6 * b * Np + b * Nm,            # pos in, grad out, 1 field
        "paint_absmax": b * Nm,                          # largest cell of the density (the g maxima arrive with g)
        "zpass_r2c": b * Nm + 2 * b * Nh,
        "zpass_r2c3": 3 * (b * Nm + 2 * b * Nh),
        "zpass_c2r": b * Nm + 2 * b * Nh,
        "zpass_c2r3": 3 * (b * Nm + 2 * b * Nh),
        "zpass_c2r_gamma": 2 * b * Nh + 2 * b * Nm,     # spectrum in, density in, its cotangent out
        "zpass_c2r_loss": 2 * b * Nh + b * Nm,          # smoothed residual in, g_ds out (no masks)
        "zpass_c2r_resp": 2 * b * Nh + 2 * b * Nm,      # smoothed g_ds in, density in, g_rho out
        "zpass_c2r_loss2": 2 * b * Nh + 3 * b * Nm,     # S^2-smoothed residual in, target and density in, g_rho out
        "colpass": 4 * b * Nh,
        "colpass3": 4 * b * Nh,                          # one force spectrum in and out
        "sm_x_fwd": 6 * b * Nh,                          # spectrum in, U out, P's x-inverse out
        "sm_x_bwd": 6 * b * Nh,                          # spectrum + U in, 1 out
        "grad6": 4 * b * Nm,                             # potential in, interleaved forces (3 per cell) out
        "div6": 4 * b * Nm,                              # 3 gradient meshes in, divergence out
        "x_realfilter": 4 * b * Nh,
        "gauss_line": 2 * b * Nm,                        # real mesh in and out (complex pairs)
    }.get(name)


# Operations made of several instrumented kernels: name -> (kernels, the kernel
# counted once per operation, algorithmic bytes at b bytes per real).
def op_bytes(n, b):
    Np = Nm = n ** 3
    return {
        # unit-mass paint: positions in, mesh out; its counting sort (bin, scan,
        # scatter) and the tile deposit
        "paint1": (("paint_bin", "paint_scan", "paint_scatter", "paint_deposit1"), "paint_deposit1", 3 * b * Np + b * Nm),
        # 3-component paint of the position cotangents on the forward's sort
        "paint3": (("paint_deposit3", "paint_absmax"), "paint_deposit3", 6 * b * Np + 3 * b * Nm),
    }


KERNELS = ("paint_absmax", "paint_bin", "paint_scan", "paint_scatter", "paint_deposit1", "paint_deposit3",
           "paint_deposit", "paint_convert", "zpass_r2c", "zpass_r2c3", "zpass_c2r", "zpass_c2r3", "zpass_c2r_gamma", "zpass_c2r_loss",
           "zpass_c2r_resp", "zpass_c2r_loss2", "colpass",
           "colpass3", "grad6", "div6", "sm_x_fwd", "sm_x_bwd", "x_realfilter", "gauss_line", "readout3_update", "sm_bwd_gather", "paint_vjp_pos",
           "loss")


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_ldl512_fp32.txt")


def measured_traffic(kernel, n):
    """HBM bytes per launch of an instrumented kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-probe
    calibration), or None when the summary does not cover this kernel at this
    size.  Data lines: <timer name> <n> <dispatches> <fetch bytes> <write bytes>."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            t = line.split()
            if len(t) == 5 and not line.startswith("#") and t[0] == kernel and int(t[1]) == n:
                return int(float(t[3]) + float(t[4]))
    return None


def copy_probe(mib, dev):
    """One plain device-to-device copy of a known size, timed with HIP events:
    the achievable-bandwidth reference and the calibration dispatch for the
    HBM byte counters."""
    import torch
    if mib <= 0:
        return None
    a = torch.ones(mib << 18, device=dev, dtype=torch.float32)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    by = 2 * (mib << 20)
    del a, b
    return {"bytes": by, "ms": round(ms, 4), "GBps": round(by / ms / 1e6, 1)}


def fft_op(n, dev, reps):
    """The FFT as one operation: standalone FP32 r2c (ldl_pfft) and c2r
    (ldl_pifft) of an n^3 mesh, timed with HIP events on the launch stream,
    against the operation's algorithmic bytes (the real mesh and the half
    spectrum, each moved once).  The transform runs as three passes (z rows,
    y columns, x columns), each reading and writing the whole array, so the
    operation's rate is about a third of a pass's."""
    import torch
    from jerald_amd import jaxpops as jpo
    if reps <= 0:
        return None
    g = torch.Generator(device=dev).manual_seed(87)
    x = torch.randn((n, n, n), device=dev, dtype=torch.float32, generator=g)
    X = jpo.pfft(x)
    jpo.pifft(X, (n, n, n))
    torch.cuda.synchronize()
    by = 4 * n ** 3 + 8 * n * n * (n // 2 + 1)
    out = {"passes": 3, "algorithmic_bytes": by}
    for name, fn in (("r2c", lambda: jpo.pfft(x)), ("c2r", lambda: jpo.pifft(X, (n, n, n)))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"avg_ms": round(ms, 4), "GBps_alg": round(by / ms / 1e6, 1), "frac": round(by / ms / 1e6 / HBM_PEAK_GBS, 4)}
    del x, X
    return out


def timed(model, p, steps, warmup):
    """(seconds for ``steps`` loss+gradient steps after ``warmup``, last result)."""
    import torch
    for _ in range(warmup):
        model.loss_and_grad(p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = model.loss_and_grad(p)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the LDL path here is a "
              f"single-GPU build; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from jerald_amd import _lib as L
    from jerald_amd.model import JERALDModel
    from jerald_amd.pmesh import PMesh

    dev = torch.device("cuda:0")
    n, nstep = args.n, args.nstep
    box = 0.8 * n
    cell = box / n
    p = params0(nstep)
    probe = copy_probe(args.copy_probe_mib, dev)
    fftop = fft_op(n, dev, args.fft_reps)
    log(f"[bench] building inputs n={n} box={box}")
    X = zeldovich_gpu(n, box, 2.0, 83, dev)
    pm32 = PMesh(n, box, device=dev, dtype=torch.float32)
    kw = dict(Nstep=nstep, kind="sm", L1=False, smoothing=("gaussian", cell))
    gen = JERALDModel(X, None, pm32, lattice=(n, n, n), **kw)
    rng = np.random.default_rng(84)
    pt = p * (1 + 0.1 * (2 * rng.random(p.shape) - 1))
    target = gen.evaluate(torch.tensor(pt)).double()
    gnoise = torch.Generator(device=dev).manual_seed(85)
    target += 0.01 * target.std() * torch.randn(target.shape, generator=gnoise, device=dev, dtype=torch.float64)
    del gen
    model = JERALDModel(X, target, pm32, lattice=(n, n, n), save=args.save, **kw)
    log(f"[bench] FP32 warmup + {args.steps} timed steps (save={args.save})")
    L.timer_enable(False)
    for _ in range(args.warmup):
        model.loss_and_grad(p)
    torch.cuda.synchronize()
    L.timer_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss32, grad32 = model.loss_and_grad(p)
    torch.cuda.synchronize()
    dt32 = time.perf_counter() - t0
    kern = {}
    for name in KERNELS:
        tot, cnt = L.timer_read(name)
        if cnt:
            kern[name] = (tot, cnt)
    L.timer_enable(False)
    loss32 = loss32.item()
    grad32 = grad32.cpu().numpy()
    ms = 1e3 * dt32 / args.steps
    log(f"[bench] FP32: {ms:.2f} ms/step, loss {loss32:.8g}")
    # roofline of the dominant instrumented kernel with a defined byte count
    total_k = sum(t for t, _ in kern.values())
    ranked = sorted(kern.items(), key=lambda kv: -kv[1][0])
    dom = next((k for k, _ in ranked if kernel_bytes(k, n, 4) is not None), None)
    roof = None
    if dom is not None:
        tot, cnt = kern[dom]
        avg_s = tot / cnt / 1e3
        by = kernel_bytes(dom, n, 4)
        ach = by / avg_s / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, n),
                "avg_launch_ms": round(avg_s * 1e3, 4),
                "algorithmic_bytes": by, "share_of_kernel_time": round(tot / total_k, 3)}
    kernel_table = {}
    for k, (t, c) in ranked:
        row = {"ms_per_step": round(t / args.steps, 3), "launches_per_step": c / args.steps,
               "avg_ms": round(t / c, 4)}
        by = kernel_bytes(k, n, 4)
        if by is not None:
            row["GBps_alg"] = round(by / (t / c / 1e3) / 1e9, 1)
        tr = measured_traffic(k, n)
        if tr is not None:
            row["traffic_bytes"] = tr
        kernel_table[k] = row
    ops = {}
    for name, (members, anchor, by) in op_bytes(n, 4).items():
        if anchor not in kern:
            continue
        nop = kern[anchor][1]
        t = sum(kern[k][0] for k in members if k in kern)
        avg = t / nop
        ops[name] = {"kernels": list(members), "ops_per_step": nop / args.steps, "avg_ms": round(avg, 4),
                     "algorithmic_bytes": by, "GBps_alg": round(by / (avg / 1e3) / 1e9, 1),
                     "frac": round(by / (avg / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the other memory modes on the same inputs: their loss and gradients must be
    # bitwise those of the headline mode
    variants = {}
    if args.variant_steps > 0:
        for vname, vkw in (("save_all", dict(save="all")), ("recompute_every_layer", dict(save="positions", recompute_last=True))):
            if vkw["save"] == args.save and not vkw.get("recompute_last"):
                continue
            log(f"[bench] FP32 variant {vname}")
            mv = JERALDModel(X, target, pm32, lattice=(n, n, n), **vkw, **kw)
            tv, (lv, gv) = timed(mv, p, args.variant_steps, 1)
            variants[vname] = {"value": round(args.variant_steps / tv, 4), "unit": "it/s",
                               "ms_per_step": round(1e3 * tv / args.variant_steps, 3),
                               "bitwise_equal_to_headline": bool(lv.item() == loss32 and np.array_equal(gv.cpu().numpy(), grad32))}
            del mv
    rho32 = model.density(torch.tensor(p)).double()
    # plain FP64 GPU path: FP64 absolute positions, FP64 atomics, same kernels, same memory mode
    fp64 = None
    if args.fp64_steps > 0:
        log("[bench] FP64 plain path")
        del model
        L.workspaces.clear()
        torch.cuda.empty_cache()
        pm64 = PMesh(n, box, device=dev, dtype=torch.float64)
        m64 = JERALDModel(X, target, pm64, paint_mode="float", save=args.save, **kw)
        t64, (l64, g64) = timed(m64, p, args.fp64_steps, 1)
        l64 = l64.item()
        g64 = g64.cpu().numpy()
        rho64 = m64.density(torch.tensor(p))
        drho = (rho32 - rho64).abs()
        rel = np.abs(grad32 - g64) / np.maximum(np.abs(g64), 1e-300)
        fp64 = {"value": round(args.fp64_steps / t64, 4), "unit": "it/s", "ms_per_step": round(1e3 * t64 / args.fp64_steps, 2),
                "save": args.save, "loss": l64, "loss_rel_err_fp32": abs(loss32 - l64) / abs(l64),
                "grad_rel_err_fp32_max": float(rel.max()), "grad_rel_err_fp32": [float(v) for v in rel],
                "density_rel_err_fp32_max": drho.max().item() / rho64.abs().max().item(),
                "density_rel_err_fp32_rms": (drho.pow(2).mean().sqrt() / rho64.pow(2).mean().sqrt()).item()}
        del m64, rho64, drho
        L.workspaces.clear()
        torch.cuda.empty_cache()
    del rho32
    # CPU baseline: the FP64 oracle restatement (C CIC loops with OpenMP,
    # pocketfft) timed for one iteration at the full size on the host's cores,
    # on the same inputs; plus a 1-thread sample
    cpu = None
    if args.cpu_n > 0:
        from oracle import ldl_oracle as O
        cores = args.cpu_threads or len(os.sched_getaffinity(0))
        nc = args.cpu_n
        if nc == n:
            Xc, tgtc, boxc = X.cpu().numpy(), target.cpu().numpy(), box
        else:
            boxc = 0.8 * nc
            Xc = zeldovich_gpu(nc, boxc, 2.0, 83, dev).cpu().numpy()
            tgtc = 1.0 + 0.1 * np.random.default_rng(86).standard_normal((nc, nc, nc))
        log(f"[bench] CPU baseline: one FP64 iteration at {nc}^3 on {cores} threads")
        orc = O.SMModelOracle(nc, boxc, nstep, tgtc, L1=False, smoothing=("gaussian", boxc / nc), backend="native",
                              threads=cores)
        t0 = time.perf_counter()
        lc, _ = orc.loss_and_grad(p, Xc)
        tc = time.perf_counter() - t0
        del orc, Xc, tgtc
        scale = (n / nc) ** 3
        cpu = {"value": round(1.0 / (tc * scale), 6), "unit": "it/s", "cores": cores, "kind": "port",
               "sample": (f"one FP64 fwd+bwd iteration of the oracle restatement (C CIC loops with OpenMP, pocketfft) at "
                          f"{nc}^3 particles/mesh on the same inputs, measured: {tc:.2f} s on {cores} threads"
                          + ("" if scale == 1 else f"; value scaled by ({n}/{nc})^3 = {scale:.0f}")),
               "loss": lc}
        if nc == n:
            cpu["loss_rel_err_fp32"] = abs(loss32 - lc) / abs(lc)
        if args.cpu1_n > 0:
            n1 = args.cpu1_n
            b1 = 0.8 * n1
            X1 = zeldovich_gpu(n1, b1, 2.0, 83, dev).cpu().numpy()
            t1 = 1.0 + 0.1 * np.random.default_rng(86).standard_normal((n1, n1, n1))
            log(f"[bench] CPU 1-thread sample at {n1}^3")
            o1 = O.SMModelOracle(n1, b1, nstep, t1, L1=False, smoothing=("gaussian", b1 / n1), backend="native", threads=1)
            t0 = time.perf_counter()
            o1.loss_and_grad(p, X1)
            ts = time.perf_counter() - t0
            s1 = (n / n1) ** 3
            cpu["one_thread"] = {"value": round(1.0 / (ts * s1), 7), "unit": "it/s", "cores": 1,
                                 "sample": f"one iteration at {n1}^3 on 1 thread: {ts:.2f} s; value scaled by ({n}/{n1})^3 = {s1:.0f}"}
    line = {
        "metric": "LDL fwd+bwd iters/s @512^3 FP32 (2 layers, 1 GPU)",
        "value": round(args.steps / dt32, 4),
        "unit": "it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Zel'dovich lattice, generated on the GPU)",
        "config": {"workload": f"{n}^3 particles on {n}^3 mesh, {nstep} sm LDL layers, L2 loss with 1-cell "
                               f"Gaussian smoothing, loss+grad of {5 * nstep + 3} params",
                   "n": n, "nstep": nstep, "box": box, "positions": "fp32 displacements from lattice",
                   "paint": "fixed point (integer deposits, deterministic)", "save": args.save,
                   "recompute": ("layers 0..Nstep-2 recomputed from their checkpointed input positions in the backward; "
                                 "the last layer's state is consumed by the backward first"
                                 if args.save == "positions" else "none: every layer's state stored")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64_plain_gpu": fp64,
        "memory_modes": variants,
        "copy_probe": probe,
        "fft_op": fftop,
        "loss_fp32": loss32,
        "ops": ops,
        "kernels": kernel_table,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
