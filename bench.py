"""GW self-energy throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|split] [--no-cpu]

Workload: full 4x4x4 k mesh (nk = 64), nao = 26, naux = 300, ns = 1, beta = 100, IR
Lambda = 1e4 (nts = 100, nw_b = 99); G from a seeded random Fock operator, V seeded
N(0, 0.1^2) generated on the device.  One step = one Sigma evaluation
(P0 -> tau->Omega -> Dyson -> Omega->tau -> Sigma over all q, k, tau) with G, V and Sigma
resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "green-mbpt_amd"))
sys.path.insert(0, ROOT)

METRIC = "GW Σ-evals/s @nk64 nao26 naux300; split-prec vs FP64 at rel err ≤1e-8"
PEAK_FP64_TFLOPS = 78.6
PEAK_INT8_TOPS = 5000.0

WORKLOAD = dict(nk=64, mesh=(4, 4, 4), nao=26, NQ=300, ns=1, beta=100.0, lam=1e4, seed=0, sigma=0.1)

# HBM bytes per launch of the main kernels on this workload, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE in KiB per dispatch) of `bench.py --steps 1 --warmup 0 --no-cpu
# [--no-check | --precision fp64] --no-autotune` with the same build; each scaled by what the
# counter reads for a plain 2 GiB device copy (bench.py --copy-calibration): FETCH_SIZE reads 0.5x
# the bytes the copy reads, WRITE_SIZE 1.0x the bytes it writes.
# Source: profiles/gw_sigma_nk64_nao26_naux300_hbm_traffic.md
PMC_PER_LAUNCH_KIB = {
    "split": {"Sigma GEMM": (3.5381e6, 1.0229e7), "Sigma epilogue": (5.5368e6, 5.1620e4),
              "P0 GEMM": (7.7133e6, 6.8067e5), "P0 prep": (6.9709e5, 5.3279e6),  # P0 GEMM: 16 launches per q
              "Dyson inverse": (5.7157e7, 6.7429e7)},  # every kernel of the inverse, one evaluation
    "fp64": {"Sigma GEMM": (2.9702e7, 1.0307e7), "Sigma epilogue": (5.5368e6, 5.1620e4),
             "P0 GEMM": (2.2308e7, 4.2419e4), "P0 prep": (7.4680e5, 1.0163e7),
             "Dyson inverse": (5.7126e7, 6.7429e7)},
}
PMC_CALIBRATION = (0.5, 1.0)
PMC_SOURCE = "profiles/gw_sigma_nk64_nao26_naux300_hbm_traffic.md"
# the int8 GEMM kernel of each block tile (tuning keys p0_split_tile / sigma_split_tile)
OZAKI_KERNEL = {2: "zgemm_ozaki_pipe_kernel", 4: "zgemm_ozaki32_kernel"}
# the register Sigma epilogue of each form (tuning key epi_form) at nao = 26; the profiled traffic
# figures belong to form 1
EPI_KERNEL = {1: "sigma_epi4_kernel<7> (Sigma += (V G) R, FP64 4x4x4 MFMA)",
              2: "sigma_epi4w_kernel<7,2,4> (Sigma += (V G) R, FP64 4x4x4 MFMA, one i group per wave)",
              3: "sigma_epi_mfma_kernel<2,double> (Sigma += (V G) R, FP64 16x16x4 MFMA)"}


def algorithmic_flops(nk, nao, NQ, nts, nw_b, pseudo_ns=1):
    """Reference-order flop count of one Sigma evaluation (8 flops per complex MAC)."""
    U = 2 * nao ** 3 * NQ + nao ** 2 * NQ ** 2
    f_p0 = 8 * nk * nk * (nts // 2) * pseudo_ns * U
    f_s = 8 * nk * nk * nts * pseudo_ns * U
    f_dys = 8 * nk * nw_b * (7.0 / 3.0) * NQ ** 3
    f_ir = 8 * nk * 2 * nts * nw_b * NQ ** 2
    return dict(P0=f_p0, Sigma=f_s, Dyson=f_dys, IR=f_ir, total=f_p0 + f_s + f_dys + f_ir)


def stage_work(w, T, slices):
    """Work of each kernel family over one Sigma evaluation of the full mesh (nq = nk):
    algorithmic flops (the reference's GEMM count, 8 flops per complex MAC; for the two big GEMMs on
    the half tau grid, which is all the reference computes for P0 and all that R = P-tilde conj(V)
    needs), int8 operations issued by the split-precision GEMMs, and algorithmic bytes (each operand
    read once and each result written once, in the representation the kernel reads and writes)."""
    nk, nao, NQ, nts, nwb = w["nk"], w["nao"], w["NQ"], T.nts, T.nw_b
    half, nao2, B = nts // 2, nao * nao, 16.0
    pairs = slices * (slices + 1) // 2
    kc_s = -(-NQ // 64) * 64
    kc_p = -(-nao2 // 64) * 64
    sl = 2 * slices                                   # int8 bytes per element and K step: planes x slices
    V = nk * NQ * nao2 * B                            # V(k, k +- q) of one q
    out = {
        "Sigma GEMM": dict(flops=8.0 * NQ * NQ * nk * nao2 * half * nk,
                           issued=2.0 * 4 * pairs * NQ * kc_s * nk * nao2 * half * nk,
                           bytes_fp64=nk * (half * NQ * NQ * B + V + half * nk * NQ * nao2 * B),
                           bytes_split=nk * (half * NQ * sl * kc_s + nk * nao2 * sl * kc_s + half * nk * NQ * nao2 * B)),
        "P0 GEMM": dict(flops=8.0 * NQ * NQ * nk * nao2 * half * nk,
                        issued=2.0 * 4 * pairs * NQ * NQ * nk * kc_p * half * nk,
                        bytes_fp64=nk * (half * nk * NQ * nao2 * B + V + half * NQ * NQ * B),
                        bytes_split=nk * (half * NQ * nk * sl * kc_p + NQ * nk * sl * kc_p + half * NQ * NQ * B)),
        "Sigma epilogue": dict(flops=16.0 * NQ * nao ** 3 * nk * nts * nk, issued=None,
                               bytes_fp64=nk * (half * nk * NQ * nao2 * B + V + 3 * nts * nk * nao2 * B)),
        "P0 prep": dict(flops=16.0 * NQ * nao ** 3 * nk * half * nk, issued=None,
                        bytes_fp64=nk * (V + 2 * half * nk * nao2 * B + half * nk * NQ * nao2 * B),
                        bytes_split=nk * (V + 2 * half * nk * nao2 * B + half * NQ * nk * sl * kc_p)),
        "Dyson inverse": dict(flops=8.0 * (4.0 / 3.0) * NQ ** 3 * nk * nwb, issued=None,
                              bytes_fp64=2.0 * nk * nwb * NQ * NQ * B),
    }
    return out


def roofline(stats, precision, w, T, slices, tuning, steps=1):
    """Roofline of the kernel family with the largest share of the step, from the per-launch times
    measured with HIP events on the library's stream: achieved = algorithmic flops per launch / mean
    launch time, frac = achieved / the published peak of the arithmetic the kernel uses (int8 5.0 POP/s
    for the split GEMMs, FP64 78.6 TFLOP/s otherwise); issued_frac = int8 operations issued / time /
    peak; traffic = the rocprofv3 counter bytes per launch of this build, traffic_ratio = traffic /
    algorithmic bytes.  Every family's figures are listed under all_stages.  stats holds launch counts
    and times summed over `steps` evaluations; the work figures are per evaluation."""
    work = stage_work(w, T, slices)
    split = precision == "split"
    int8 = {"Sigma GEMM", "P0 GEMM"} if split else set()
    s_kernel = OZAKI_KERNEL.get(tuning.get("sigma_split_tile", 2), "?")
    p_kernel = OZAKI_KERNEL.get(tuning.get("p0_split_tile", 2), "?")
    names = {"Sigma GEMM": (f"{s_kernel}<{slices}> (R = P-tilde conj(V), int8 slices)"
                            if split else "zgemm_f64_kernel<64,64,true,false> (R = P-tilde conj(V), FP64 MFMA)"),
             "P0 GEMM": (f"{p_kernel}<{slices}> (P0 = Z conj(V)^T, int8 slices)"
                         if split else "zgemm_f64_kernel<64,64,true,true> (P0 = Z conj(V)^T, FP64 MFMA)"),
             "Sigma epilogue": EPI_KERNEL.get(tuning.get("epi_form", 1) or 1, "?"),
             "P0 prep": (f"z_prep_slice_kernel<2,{slices},1>" if split else "z_prep_mfma_kernel<2,double>") +
                        " (Z = Gb V Gq, FP64 MFMA)",
             "Dyson inverse": "block_inverse: Hermitian block Gauss-Jordan, gj_inverse_kernel pivots + "
                              "zgemm_f64_kernel upper-tile updates ((I - P0)^-1 per (q, Omega), FP64)"}
    rows = {}
    for name, (n, ms) in stats.items():
        if not n or name not in work:
            continue
        wk = work[name]
        t = ms / n * 1e-3
        peak = PEAK_INT8_TOPS if name in int8 else PEAK_FP64_TFLOPS
        per = steps / n  # evaluations per launch
        flops = wk["flops"] * per
        alg_bytes = (wk.get("bytes_split") if split and "bytes_split" in wk else wk["bytes_fp64"]) * per
        r = {"kernel": names[name], "launches": n / steps, "avg_launch_ms": t * 1e3,
             "achieved": flops / t / 1e12, "peak": peak, "unit": "TFLOP/s",
             "peak_unit": "TOPS (int8)" if name in int8 else "TFLOP/s (FP64)",
             "frac": flops / t / 1e12 / peak, "algorithmic_flops_per_launch": flops,
             "algorithmic_bytes": alg_bytes}
        if name in int8 and wk["issued"]:
            r["issued_tops"] = wk["issued"] * per / t / 1e12
            r["issued_frac"] = r["issued_tops"] / peak
        pmc = PMC_PER_LAUNCH_KIB.get(precision, {}).get(name)
        if name == "Sigma epilogue" and (tuning.get("epi_form", 1) or 1) != 1:
            pmc = None  # profiled for the 4x4x4 all-groups form only
        if pmc is not None:
            tr = 1024.0 * (pmc[0] / PMC_CALIBRATION[0] + pmc[1] / PMC_CALIBRATION[1])
            r.update({"traffic": tr, "traffic_ratio": tr / alg_bytes, "traffic_source": PMC_SOURCE})
        else:
            r["traffic"] = None
        rows[name] = r
    dom = max(rows, key=lambda nm: rows[nm]["avg_launch_ms"] * rows[nm]["launches"])
    roof = {"bound": "mfma", "stage": dom}
    roof.update({k_: v_ for k_, v_ in rows[dom].items()})
    roof["all_stages"] = rows
    return roof


def device_gm_energy(G_d, S_d, T, nk, nao, ns):
    """Galitskii-Migdal correlation energy of device G and Sigma (full mesh) by the library's
    compute_energy (gwhip_dyson_energy_device)."""
    import numpy as np
    import torch

    from gwhip import lib as gl

    wn = (2 * T.n_fermi + 1) * np.pi / T.beta
    d = gl.Dyson(T.nts, len(wn), ns, nk, nk, nao, nao)
    try:
        d.set_grid(T.T_ft, T.T_tf, wn)
        z = np.zeros((ns, nk, nao, nao), complex)
        d.set_system(z, z + np.eye(nao))
        s1 = torch.zeros(2 * z.size, dtype=torch.float64, device=G_d.device)
        return d.energy_device(G_d.data_ptr(), s1.data_ptr(), S_d.data_ptr())[2]
    finally:
        d.close()


def host_cores():
    """Threads for the CPU baseline: the host share of one GPU (OMP_NUM_THREADS where the
    environment sets it), never above 16."""
    return min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)


def cpu_baseline(w, T, budget_s=20.0):
    """Time the oracle's per-unit contractions on the host and scale to one Sigma evaluation."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import gw_oracle as go

    nao, NQ, nk, nts, nwb = w["nao"], w["NQ"], w["nk"], T.nts, T.nw_b
    rng = np.random.default_rng(0)
    cores = host_cores()
    with threadpool_limits(limits=cores):
        v = (rng.standard_normal((NQ, nao, nao)) + 1j * rng.standard_normal((NQ, nao, nao))) * 0.1
        Gb = rng.standard_normal((nao, nao)) + 1j * rng.standard_normal((nao, nao))
        Gq = rng.standard_normal((nao, nao)) + 1j * rng.standard_normal((nao, nao))
        P0t = np.zeros((NQ, NQ), complex)
        Pt = rng.standard_normal((NQ, NQ)) + 0j
        vm, vmm = v.reshape(NQ, nao * nao), v.reshape(NQ * nao, nao)
        vs = v.reshape(NQ * nao, nao)

        def timed_units(fn, tmax):
            n, t0 = 0, time.perf_counter()
            while True:
                fn()
                n += 1
                dt = time.perf_counter() - t0
                if dt > tmax:
                    return dt / n, n
        t_p0, n_p0 = timed_units(lambda: go.P0_contraction(Gb, Gq, vm, vmm, P0t, 2.0, nao, NQ), budget_s * 0.4)
        t_s, n_s = timed_units(lambda: go.selfenergy_contraction(Gq, vs, Pt, nao, NQ), budget_s * 0.4)
        A = np.eye(NQ) * 2.0 + 0.01 * Pt
        t_d, n_d = timed_units(lambda: go.eval_P_tilde(np.stack([Pt] * 2), np.ones((1, 2)), np.ones((2, 1))),
                               budget_s * 0.1)
        del A
    t_eval = t_p0 * nk * nk * (nts // 2) + t_s * nk * nk * nts + t_d * nk * nwb
    return dict(value=1.0 / t_eval, unit="Σ-evals/s", cores=cores, kind="port",
                sample=(f"oracle per-unit contractions at nao={nao} naux={NQ}: {n_p0} P0 units "
                        f"({t_p0 * 1e3:.2f} ms each), {n_s} Sigma units ({t_s * 1e3:.2f} ms each), "
                        f"{n_d} Dyson solves ({t_d * 1e3:.2f} ms each); scaled by the unit counts of one "
                        f"Sigma evaluation (nq*nk*nts/2, nq*nk*nts, nq*nw_b); OpenBLAS zgemm, {cores} threads"))


def bench_capacity(args):
    """Capacity check: the split-precision evaluation at nk = 64, nao = 64, naux = 300 (V of 80 GB held
    once, its two int8 slice stores, P-tilde and the workspace on one MI355X); prints the device
    memory the handle holds, what is left free, and the time of one evaluation."""
    import numpy as np
    import torch

    from gwhip import inputs, irgrid
    from gwhip import lib as gl

    nk, mesh, nao, NQ, beta = 64, (4, 4, 4), 64, 300, 100.0
    T = irgrid.Transformer.from_grid(irgrid.load_grid(1e4), beta)
    G = inputs.generate_G(0, nk, nao, 1, T.tau, beta)
    k = gl.GwHipKernel(nk, nao, nao, 1, NQ, T.nts, T.nw_b, beta, mesh=mesh)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.generate_V(0, 0.1)
        dev = torch.device("cuda:0")
        G_d = torch.from_numpy(G.view(np.float64)).to(dev)
        S_d = torch.empty_like(G_d)
        torch.cuda.synchronize()
        print("[bench] capacity: system built", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        k.solve_device(G_d.data_ptr(), S_d.data_ptr(), gl.PREC_SPLIT)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        free_b, total_b = torch.cuda.mem_get_info()
        print(json.dumps({"workload": "GW Sigma, 4x4x4 k mesh, nao 64, naux 300, ns 1, split precision",
                          "ms_first_eval": ms, "handle_device_gb": k.device_bytes() / 1e9,
                          "free_gb_after": free_b / 1e9, "total_gb": total_b / 1e9,
                          "kernel_ms": {k_: v_[1] for k_, v_ in k.kernel_stats().items()},
                          "sigma_finite": bool(torch.isfinite(S_d).all().item())}), flush=True)
    finally:
        k.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "split"], default="split")
    ap.add_argument("--slices", type=int, default=4, help="int8 slices per operand in split precision")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 reference evaluation")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vjp", action="store_true",
                    help="time the Sigma[G] vector-Jacobian product on the nk=27 nao=26 naux=250 system instead")
    ap.add_argument("--batch", action="store_true",
                    help="time the 8 uneven ns=2 systems evaluated as one captured graph instead")
    ap.add_argument("--dyson", action="store_true",
                    help="time the Dyson G step and the chemical-potential search at nk=64 nso=26 instead")
    ap.add_argument("--star", action="store_true",
                    help="the star-of-q evaluation on a made-up 8-point irreducible wedge of the same workload")
    ap.add_argument("--hf", action="store_true",
                    help="time the Hartree-Fock Fock build on the same nk=64 nao=26 naux=300 integrals instead")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch parameters (the autotuner's choice for this workload)")
    ap.add_argument("--tuning", default="",
                    help="launch parameters to set after autotuning, e.g. p0_split_tile=3,sigma_split_tile=3")
    ap.add_argument("--capacity", action="store_true",
                    help="one split-precision evaluation at nk=64, nao=64, naux=300 (device memory held, time)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 2 GiB device-to-device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "sharding over q with one Sigma all-reduce is designed, not run.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        return copy_calibration()
    if args.vjp:
        return bench_vjp(args)
    if args.batch:
        return bench_batch(args)
    if args.hf:
        return bench_hf(args)
    if args.star:
        return bench_star(args)
    if args.dyson:
        return bench_dyson(args)
    if args.capacity:
        return bench_capacity(args)

    import numpy as np
    import torch

    from gwhip import inputs, irgrid
    from gwhip import lib as gl

    w = WORKLOAD
    T = irgrid.Transformer.from_grid(irgrid.load_grid(w["lam"]), w["beta"])
    print(f"[bench] workload nk={w['nk']} nao={w['nao']} NQ={w['NQ']} nts={T.nts} nw_b={T.nw_b}", file=sys.stderr,
          flush=True)
    G = inputs.generate_G(w["seed"], w["nk"], w["nao"], w["ns"], T.tau, w["beta"])
    k = gl.GwHipKernel(w["nk"], w["nao"], w["nao"], w["ns"], w["NQ"], T.nts, T.nw_b, w["beta"], mesh=w["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.generate_V(w["seed"], w["sigma"])
    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G.view(np.float64)).to(dev)
    S_d = torch.empty_like(G_d)
    prec = gl.PREC_FP64 if args.precision == "fp64" else gl.PREC_SPLIT
    if prec == gl.PREC_SPLIT:
        k.set_split_slices(args.slices)
    torch.cuda.synchronize()
    t_tune = time.perf_counter()
    if not args.no_autotune:
        k.autotune(G_d.data_ptr(), prec)
    t_tune = time.perf_counter() - t_tune
    if args.tuning:
        k.set_tuning(**{kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.tuning.split(",")})
    print(f"[bench] autotuned in {t_tune:.2f} s: {k.tuning()}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        k.solve_device(G_d.data_ptr(), S_d.data_ptr(), prec)
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    stats = {name: [0, 0.0] for name in gl.KSTATS}
    t0 = time.perf_counter()
    for i in range(args.steps):
        k.solve_device(G_d.data_ptr(), S_d.data_ptr(), prec)
        for name, (n, ms) in k.kernel_stats().items():
            stats[name][0] += n
            stats[name][1] += ms
        print(f"[bench] step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    regions = k.timings()
    fl = algorithmic_flops(w["nk"], w["nao"], w["NQ"], T.nts, T.nw_b)
    S_host = S_d.cpu().numpy().view(np.complex128).copy()
    roof = roofline(stats, args.precision, w, T, args.slices, k.tuning(), args.steps)

    check = {}
    if prec == gl.PREC_SPLIT and not args.no_check:
        # the plain FP64 kernel on the same inputs, tuned for FP64, one untimed evaluation, then timed
        S64_d = torch.empty_like(S_d)
        if not args.no_autotune:
            k.autotune(G_d.data_ptr(), gl.PREC_FP64)
        k.solve_device(G_d.data_ptr(), S64_d.data_ptr(), gl.PREC_FP64)
        torch.cuda.synchronize()
        n64 = max(1, min(args.steps, 3))
        t1 = time.perf_counter()
        for i in range(n64):
            k.solve_device(G_d.data_ptr(), S64_d.data_ptr(), gl.PREC_FP64)
            print(f"[bench] fp64 step {i} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        fp64_ms = (time.perf_counter() - t1) / n64 * 1e3
        S_ref = S64_d.cpu().numpy().view(np.complex128)
        err = float(np.abs(S_host - S_ref).max() / np.abs(S_ref).max())
        E_ref = device_gm_energy(G_d, S64_d, T, w["nk"], w["nao"], w["ns"])
        E_split = device_gm_energy(G_d, S_d, T, w["nk"], w["nao"], w["ns"])
        check = {"fp64_ms_per_step": fp64_ms, "speedup_vs_fp64": fp64_ms / ms_per_step,
                 "max_rel_sigma_err_vs_fp64": err, "gm_energy_fp64": E_ref, "gm_energy_split": E_split,
                 "gm_energy_rel_err": abs(E_split - E_ref) / abs(E_ref), "slices": args.slices,
                 "fp64_steps": n64, "fp64_kernel_ms": {k_: v_[1] for k_, v_ in k.kernel_stats().items()},
                 "fp64_tuning": k.tuning()}

    out = {
        "metric": METRIC, "value": 1e3 / ms_per_step, "unit": "Σ-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == gl.PREC_FP64 else "int8-slices/i32/f64",
        "data": "synthetic (seed 0: Haar-random Fock eigenbasis, V ~ N(0, 0.1^2) generated on device)",
        "config": {"workload": "GW Sigma(tau,k), 4x4x4 k mesh, nao 26, naux 300, ns 1, beta 100, IR 1e4",
                   "nk": w["nk"], "nao": w["nao"], "naux": w["NQ"], "nts": T.nts, "nw_b": T.nw_b,
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roof,
        "precision_check": check,
        "algorithmic_tflops_per_step": fl["total"] / 1e12,
        "whole_step_tflops": fl["total"] / (ms_per_step * 1e-3) / 1e12,
        "regions_ms": regions,
        "kernel_ms": {k_: v_[1] / args.steps for k_, v_ in stats.items()},
        "tuning": k.tuning(), "autotune_s": t_tune,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, T)
    k.close()
    print(json.dumps(out), flush=True)


VJP_WORKLOAD = dict(nk=27, mesh=(3, 3, 3), nao=26, NQ=250, ns=1, beta=100.0, lam=1e4, seed=0, sigma=0.1)


def bench_vjp(args):
    """VJP of Sigma[G] with a Hermitian-per-tau N(0,1) cotangent, P-tilde recomputed vs stored."""
    import numpy as np
    import torch

    from gwhip import inputs, irgrid
    from gwhip import lib as gl

    w = VJP_WORKLOAD
    T = irgrid.Transformer.from_grid(irgrid.load_grid(w["lam"]), w["beta"])
    G = inputs.generate_G(w["seed"], w["nk"], w["nao"], w["ns"], T.tau, w["beta"])
    rng = np.random.default_rng(1)
    Sb = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    Sb = np.ascontiguousarray(0.5 * (Sb + np.conj(np.swapaxes(Sb, -1, -2))))
    k = gl.GwHipKernel(w["nk"], w["nao"], w["nao"], w["ns"], w["NQ"], T.nts, T.nw_b, w["beta"], mesh=w["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.generate_V(w["seed"], w["sigma"])
    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G.view(np.float64)).to(dev)
    Sb_d = torch.from_numpy(Sb.view(np.float64)).to(dev)
    S_d = torch.empty_like(G_d)
    Gb_d = torch.empty_like(G_d)
    torch.cuda.synchronize()
    out = {}
    base_bytes = k.device_bytes()
    for mode, name in ((gl.VJP_RECOMPUTE, "recompute"), (gl.VJP_STORED, "stored")):
        k.set_vjp_store(mode == gl.VJP_STORED)
        k.solve_device(G_d.data_ptr(), S_d.data_ptr(), gl.PREC_FP64)
        for _ in range(args.warmup):
            k.vjp_device(G_d.data_ptr(), Sb_d.data_ptr(), Gb_d.data_ptr(), mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            k.vjp_device(G_d.data_ptr(), Sb_d.data_ptr(), Gb_d.data_ptr(), mode)
            print(f"[bench] vjp {name} step {i} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        out[name] = {"ms_per_vjp": (time.perf_counter() - t0) / args.steps * 1e3,
                     "device_gb": k.device_bytes() / 1e9}
    t0 = time.perf_counter()
    k.solve_device(G_d.data_ptr(), S_d.data_ptr(), gl.PREC_FP64)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3
    # one central finite difference along a Hermitian-per-tau direction, step 1e-6
    Gbar = Gb_d.cpu().numpy().view(np.complex128)
    dG = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    dG = 0.5 * (dG + np.conj(np.swapaxes(dG, -1, -2)))
    h = 1e-6
    k.set_vjp_store(False)
    Sp = k.solve(G + h * dG)
    Sm = k.solve(G - h * dG)
    fd = float(np.real(np.vdot(Sb, (Sp - Sm) / (2 * h))))
    an = float(np.real(np.vdot(Gbar, dG)))
    k.close()
    ms = out["stored"]["ms_per_vjp"]
    print(json.dumps({
        "metric": "Sigma[G] VJPs/s @nk27 nao26 naux250 (FP64, P-tilde stored)", "value": 1e3 / ms, "unit": "VJP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seed 0 G and V, Hermitian-per-tau N(0,1) cotangent)",
        "config": {"workload": "VJP of GW Sigma(tau,k) wrt G, 3x3x3 k mesh, nao 26, naux 250, ns 1, beta 100, IR 1e4",
                   "nts": T.nts, "nw_b": T.nw_b, "parallelism": "single GPU"},
        "modes": out, "forward_ms": fwd_ms, "base_device_gb": base_bytes / 1e9,
        "fd_check": {"step": h, "fd": fd, "vjp": an, "rel_diff": abs(fd - an) / abs(an)},
    }), flush=True)


def hf_cpu_baseline(w, budget_s=15.0):
    """Time the oracle's exchange unit (one (k, k') pair: Y = V dm, Y V^H over Q) on the host and
    scale to the nk^2 pairs of one Fock build; the direct and Madelung terms are O(nk) and left out."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    nao, NQ, nk, ns = w["nao"], w["NQ"], w["nk"], w["ns"]
    rng = np.random.default_rng(0)
    cores = host_cores()
    with threadpool_limits(limits=cores):
        v = (rng.standard_normal((NQ, nao, nao)) + 1j * rng.standard_normal((NQ, nao, nao))) * 0.1
        dm = rng.standard_normal((nao, nao)) + 1j * rng.standard_normal((nao, nao))
        n, t0 = 0, time.perf_counter()
        while True:
            Y = v @ dm
            _ = np.einsum("Qia,Qja->ij", Y, np.conj(v))
            n += 1
            dt = time.perf_counter() - t0
            if dt > budget_s:
                break
    t_unit = dt / n
    return dict(value=1.0 / (t_unit * nk * nk * ns), unit="Fock builds/s", cores=cores, kind="port",
                sample=(f"oracle exchange unit (hf_oracle.hf_scalar inner step) at nao={nao} naux={NQ}: {n} units "
                        f"({t_unit * 1e3:.2f} ms each), scaled by nk^2 * ns = {nk * nk * ns}; numpy, {cores} threads"))


def bench_hf(args):
    """Hartree-Fock Fock build (direct + exchange + Madelung) from the resident integrals."""
    import numpy as np
    import torch

    from gwhip import lib as gl

    w = WORKLOAD
    nk, nao, NQ, ns = w["nk"], w["nao"], w["NQ"], w["ns"]
    # the HF entry point reads only the integrals and k + q maps: a minimal tau grid is enough
    k = gl.GwHipKernel(nk, nao, nao, ns, NQ, 4, 3, w["beta"], mesh=w["mesh"])
    k.generate_V(w["seed"], w["sigma"])
    rng = np.random.default_rng(2)
    x = rng.standard_normal((ns, nk, nao, nao)) + 1j * rng.standard_normal((ns, nk, nao, nao))
    dm = np.ascontiguousarray(0.5 * (x + np.conj(np.swapaxes(x, -1, -2))))
    S = np.ascontiguousarray(np.broadcast_to(np.eye(nao, dtype=np.complex128), dm.shape))
    dev = torch.device("cuda:0")
    dm_d = torch.from_numpy(dm.view(np.float64)).to(dev)
    S_d = torch.from_numpy(S.view(np.float64)).to(dev)
    F_d = torch.empty_like(dm_d)
    mad = 0.37
    torch.cuda.synchronize()
    for i in range(args.warmup):
        k.hf_device(dm_d.data_ptr(), S_d.data_ptr(), mad, F_d.data_ptr())
        print(f"[bench] hf warmup {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ex_ms, ex_n = 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        k.hf_device(dm_d.data_ptr(), S_d.data_ptr(), mad, F_d.data_ptr())
        m, n = k.hf_exchange_stats()
        ex_ms += m
        ex_n += n
        print(f"[bench] hf step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    F = F_d.cpu().numpy().view(np.complex128)
    herm = float(np.abs(F - np.conj(np.swapaxes(F, -1, -2))).max() / np.abs(F).max())
    k.close()
    avg_s = ex_ms / max(ex_n, 1) * 1e-3
    flops = 8.0 * 2 * nk * nk * NQ * nao ** 3         # per launch: two nao^3 complex products per (k, k', Q)
    vbytes = 16.0 * nk * nk * NQ * nao * nao           # V(k, k') read once per launch
    roof = {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "kernel": "orb_reduce_kernel<2,true> (exchange sum_Q V dm V^H over all (k, q), FP64 MFMA)",
            "launches": ex_n, "avg_launch_ms": avg_s * 1e3, "algorithmic_bytes": vbytes,
            "hbm_GBps_algorithmic": vbytes / avg_s / 1e9, "traffic": None}
    roof["frac"] = roof["achieved"] / roof["peak"]
    out = {
        "metric": "Fock builds/s @nk64 nao26 naux300 (FP64)", "value": 1e3 / ms, "unit": "Fock builds/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (V ~ N(0, 0.1^2) generated on device, Hermitian N(0,1) density matrix, S = I)",
        "config": {"workload": "Hartree-Fock Fock build, 4x4x4 k mesh, nao 26, naux 300, ns 1, madelung 0.37",
                   "parallelism": "single GPU"},
        "roofline": roof, "fock_hermiticity": herm,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = hf_cpu_baseline(w)
    print(json.dumps(out), flush=True)


def bench_star(args):
    """Sigma on a symmetry-reduced mesh by the star-of-q scheme: the nk = 64 workload with a made-up
    wedge of 8 irreducible k and q points (k -> k mod 8; random unitary U_k and U_q elsewhere).  The
    tables are not a real space group, so only the time is meaningful: P0 and P-tilde for 8 q, Sigma
    for 8 k over all 64 q, against 64 and 64 on the full mesh."""
    import numpy as np
    import torch

    from gwhip import inputs, irgrid
    from gwhip import lib as gl

    w = WORKLOAD
    nk, nao, NQ, ns, nir = w["nk"], w["nao"], w["NQ"], w["ns"], 8
    T = irgrid.Transformer.from_grid(irgrid.load_grid(w["lam"]), w["beta"])
    G = inputs.generate_G(w["seed"], nk, nao, ns, T.tau, w["beta"])
    G_ibz = np.ascontiguousarray(G[:, :, :nir])
    rng = np.random.default_rng(3)

    def unitaries(n):
        U = np.empty((nk, n, n), np.complex128)
        for i in range(nk):
            q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
            U[i] = np.eye(n) if i < nir else q
        return U

    b2i = np.arange(nk, dtype=np.int64) % nir
    i2b = np.arange(nir, dtype=np.int64)
    zeros = np.zeros(nk, np.int64)
    k = gl.GwHipKernel(nk, nao, nao, ns, NQ, T.nts, T.nw_b, w["beta"], mesh=w["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.generate_V(w["seed"], w["sigma"])
    k.set_k_symmetry(nir, b2i, i2b, unitaries(nao), zeros)
    k.set_q_symmetry(nir, b2i, i2b, unitaries(NQ), zeros)
    prec = gl.PREC_FP64 if args.precision == "fp64" else gl.PREC_SPLIT
    if prec == gl.PREC_SPLIT:
        k.set_split_slices(args.slices)
    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G_ibz.view(np.float64)).to(dev)
    S_d = torch.empty_like(G_d)
    torch.cuda.synchronize()
    for i in range(args.warmup):
        k.solve_ibz_device(G_d.data_ptr(), S_d.data_ptr(), prec)
        print(f"[bench] star warmup {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    stats = {name: [0, 0.0] for name in gl.KSTATS}
    t0 = time.perf_counter()
    for i in range(args.steps):
        k.solve_ibz_device(G_d.data_ptr(), S_d.data_ptr(), prec)
        for name, (n, ms) in k.kernel_stats().items():
            stats[name][0] += n
            stats[name][1] += ms
        print(f"[bench] star step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    k.close()
    out = {
        "metric": "GW Sigma-evals/s @nk64 nao26 naux300, star of q over 8 irreducible k and q", "value": 1e3 / ms,
        "unit": "Sigma-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == gl.PREC_FP64 else "int8-slices/i32/f64",
        "data": "synthetic (seed 0 G and V as the main workload; made-up wedge k -> k mod 8, random unitary "
                "U_k, U_q; timing only)",
        "config": {"workload": "GW Sigma(tau, k_ir), 4x4x4 k mesh, 8 irreducible k and q, nao 26, naux 300, ns 1, "
                               "beta 100, IR 1e4", "precision": args.precision, "parallelism": "single GPU"},
        "kernel_ms": {name: v[1] / args.steps for name, v in stats.items()},
    }
    print(json.dumps(out), flush=True)


def bench_dyson(args):
    """Dyson G step (Sigma(i w_n), batched inverse, G(tau)) and the chemical-potential search."""
    import numpy as np
    import torch

    from gwhip import irgrid
    from gwhip import lib as gl

    w = WORKLOAD
    nk, n, ns, beta = w["nk"], w["nao"], w["ns"], w["beta"]
    T = irgrid.Transformer.from_grid(irgrid.load_grid(w["lam"]), beta)
    wn = (2 * T.n_fermi + 1) * np.pi / beta
    nts, nw = T.nts, len(wn)
    rng = np.random.default_rng(3)
    c = lambda *sh: rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    herm = lambda x: 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))
    U = np.linalg.qr(c(ns, nk, n, n))[0]
    eps = rng.uniform(-1.0, 1.0, (ns, nk, n))
    H = np.einsum("skij,skj,sklj->skil", U, eps, np.conj(U))
    S = np.eye(n) + 0.05 * herm(c(ns, nk, n, n)) / np.sqrt(n)
    S1 = 0.05 * herm(c(ns, nk, n, n))
    St = 0.002 * herm(c(nts, ns, nk, n, n))
    nel = float(2 * np.sum(eps < 0.0) / nk)      # filled to mu ~ 0
    d = gl.Dyson(nts, nw, ns, nk, nk, n, n)
    d.set_grid(T.T_ft, T.T_tf, wn)
    d.set_system(H, S)
    dev = torch.device("cuda:0")
    s1_d = torch.from_numpy(np.ascontiguousarray(S1).view(np.float64)).to(dev)
    st_d = torch.from_numpy(np.ascontiguousarray(St).view(np.float64)).to(dev)
    g_d = torch.empty_like(st_d)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        d.compute_G_device(s1_d.data_ptr(), st_d.data_ptr(), 0.0, g_d.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        d.compute_G_device(s1_d.data_ptr(), st_d.data_ptr(), 0.0, g_d.data_ptr())
        print(f"[bench] dyson step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms_g = (time.perf_counter() - t0) / args.steps * 1e3
    t0 = time.perf_counter()
    nel_found, mu = d.find_mu(S1, St, 0.0, nel, 1e-9)
    ms_mu = (time.perf_counter() - t0) * 1e3
    d.close()
    # CPU: the oracle's batched inverse (numpy / LAPACK) on the same shapes
    from threadpoolctl import threadpool_limits

    from oracle import dyson as od
    with threadpool_limits(limits=host_cores()):
        t0 = time.perf_counter()
        od.compute_G(H, S, S1, St, 0.0, T.T_ft, T.T_tf, wn)
        ms_cpu = (time.perf_counter() - t0) * 1e3
    nmat = nw * ns * nk
    flops = 8.0 * n ** 3 * nmat                     # Gauss-Jordan inverse: n^3 complex multiply-adds
    print(json.dumps({
        "metric": "Dyson G steps/s @nk64 nso26 (FP64)", "value": 1e3 / ms_g, "unit": "Dyson steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_g, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (random Hermitian H with eigenvalues in [-1, 1], S near I, small Hermitian Sigma)",
        "config": {"workload": "Dyson G(tau) at nk 64, nso 26, ns 1, beta 100, IR 1e4", "nts": nts, "nw": nw,
                   "parallelism": "single GPU"},
        "inverse_flops_per_step": flops, "inverse_matrices": nmat,
        "find_mu": {"ms": ms_mu, "mu": mu, "nel": nel_found, "target": nel},
        "cpu_baseline": {"value": 1e3 / ms_cpu, "unit": "Dyson steps/s", "cores": host_cores(), "kind": "port",
                         "sample": "the whole step with the oracle (numpy batched inverse, LAPACK)"},
    }), flush=True)


def copy_calibration():
    """Three launches of torch's copy kernel over 2 GiB: 2 GiB read + 2 GiB written each."""
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "ms_per_copy": dt * 1e3,
                      "GB_per_s": 16 * n / dt / 1e9}), flush=True)


BATCH_SYSTEMS = [(8, 8, 60), (8, 14, 120), (27, 18, 160), (27, 26, 250), (27, 40, 400), (64, 12, 100),
                 (64, 26, 300), (64, 32, 350)]


def bench_batch(args):
    """8 independent ns=2 systems (spin-split Fock +-0.05 Ha), beta 100, IR 1e4, one graph per eval."""
    import numpy as np
    import torch

    from gwhip import inputs, irgrid
    from gwhip import lib as gl

    beta, lam = 100.0, 1e4
    T = irgrid.Transformer.from_grid(irgrid.load_grid(lam), beta)
    dims, Gs = [], []
    for i, (nk, nao, NQ) in enumerate(BATCH_SYSTEMS):
        n = round(nk ** (1 / 3))
        dims.append(dict(nk=nk, nao=nao, nso=nao, ns=2, NQ=NQ, nts=T.nts, nw_b=T.nw_b, beta=beta, mesh=(n, n, n)))
        Gs.append(inputs.generate_G(i, nk, nao, 2, T.tau, beta, spin_split=0.05))
    b = gl.Batch(dims)
    for i, s in enumerate(b.systems):
        s.set_grid(T.T_bt, T.T_tb)
        s.generate_V(i, 0.1)
    dev = torch.device("cuda:0")
    G_d = [torch.from_numpy(G.view(np.float64)).to(dev) for G in Gs]
    S_d = [torch.empty_like(g) for g in G_d]
    gp, sp = [g.data_ptr() for g in G_d], [x.data_ptr() for x in S_d]
    torch.cuda.synchronize()
    flops = sum(algorithmic_flops(nk, nao, NQ, T.nts, T.nw_b, pseudo_ns=2)["total"] for nk, nao, NQ in BATCH_SYSTEMS)
    res, host, energies = {}, {}, {}
    for prec, name in ((gl.PREC_SPLIT, "split"), (gl.PREC_FP64, "fp64")):
        t0 = time.perf_counter()
        b.autotune(gp, prec)
        tune_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(max(args.warmup, 1)):  # the first call captures the graph
            b.eval_device(gp, sp, prec)
        first_s = (time.perf_counter() - t0) / max(args.warmup, 1)
        print(f"[bench] batch {name} warm", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for i in range(args.steps):
            b.eval_device(gp, sp, prec)
            print(f"[bench] batch {name} step {i} done", file=sys.stderr, flush=True)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        st = b.stats()
        res[name] = {"ms_per_batch_eval": ms, "graph_launch_ms": st["last_ms"], "graph_nodes": st["graph_nodes"],
                     "workspace_gb": st["workspace_bytes"] / 1e9, "autotune_s": tune_s, "first_call_s": first_s,
                     "tuning": [s.tuning() for s in b.systems], "algorithmic_tflops": flops / (ms * 1e-3) / 1e12}
        host[name] = [x.cpu().numpy().view(np.complex128).copy() for x in S_d]
        energies[name] = [device_gm_energy(G_d[i], S_d[i], T, nk, nao, 2)
                          for i, (nk, nao, NQ) in enumerate(BATCH_SYSTEMS)]
    per_sys = []
    for i, (nk, nao, NQ) in enumerate(BATCH_SYSTEMS):
        Sf, Ss = host["fp64"][i], host["split"][i]
        Ef, Es = energies["fp64"][i], energies["split"][i]
        per_sys.append({"nk": nk, "nao": nao, "naux": NQ,
                        "max_rel_sigma_err_split_vs_fp64": float(np.abs(Ss - Sf).max() / np.abs(Sf).max()),
                        "gm_energy_fp64": Ef, "gm_energy_rel_err": abs(Es - Ef) / abs(Ef)})
    b.close()
    ms = res["split"]["ms_per_batch_eval"]
    print(json.dumps({
        "metric": "GW Sigma batch-evals/s, 8 uneven ns=2 systems in one captured graph (split precision)",
        "value": 1e3 / ms, "unit": "batch-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-slices/i32/f64", "data": "synthetic (seeds 0-7: spin-split Fock +-0.05 Ha, V ~ N(0, 0.1^2))",
        "config": {"workload": "8 GW systems ns=2, (nk,nao,naux) " + str(BATCH_SYSTEMS) + ", beta 100, IR 1e4",
                   "nts": T.nts, "nw_b": T.nw_b, "parallelism": "single GPU"},
        "modes": res, "systems": per_sys, "algorithmic_flops_per_batch_eval": flops,
        "speedup_split_vs_fp64": res["fp64"]["ms_per_batch_eval"] / ms,
    }), flush=True)


if __name__ == "__main__":
    main()
