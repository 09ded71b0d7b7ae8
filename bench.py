"""Headline benchmark: semi-implicit gradient-flow steps/s on one 256^3 grid
(get_min_u_F, LPsrc.cpp:40-117) in FP64 and in mixed precision (FP32 fields
and FFTs, FP64 sums), with the HBM roofline of the step and the CPU oracle
timed beside it.

A "step" is one iteration of get_min_u_F on the whole field: the pre-step
energy E and ||grad F||, the update u_F' = (u_F + dt f_hat)/(1 + dt G), the
residual max|Re(u_F' - u_F)|, and the new field phi' = IFFT(u_F').

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--N 256]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def algorithmic_bytes_per_step(N: int, s: int) -> int:
    """B_step = 2 R s + 5 C 2s: r2c of f (phi in, f_hat out), the spectral
    update (u_F and f_hat in, u_F' out), c2r (u_F' in, phi' out); R = N^3,
    C = N^2 (N/2 + 1)."""
    R = N ** 3
    C = N * N * (N // 2 + 1)
    return 2 * R * s + 5 * C * 2 * s


# the kernels of one step: two passes (N = 64, 128, 256) or four
STEP_KERNELS_2 = (("k_sj_cols", "SOpGradFlowStep"), ("k_sj_rows", "ProSJFinalize"))
STEP_KERNELS_4 = (("k_cols", "SOpGradFlowStep"), ("k_cols", "ProGradFlowReduce"), ("k_rows", "ProGradFlowFinalize"),
                  ("k_cols", "true, false, true, SOpNone, ProNone"))


def measured_traffic(N, dtype, passes):
    """HBM bytes of one step, summed over its kernels, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/gradflow_step_<N>.md);
    None if there is none for this size and step."""
    path = os.path.join(ROOT, "profiles", "gradflow_step_%d.md" % N)
    if not os.path.exists(path):
        return None
    kernels = STEP_KERNELS_2 if passes == 2 else STEP_KERNELS_4
    total, seen = 0.0, set()
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) < 7 or not cells[1].startswith("`k_"):
            continue
        name = cells[1]
        if ("<%s, %d," % (dtype, N)) not in name:
            continue
        for k in kernels:
            if name.startswith("`" + k[0] + "<") and k[1] in name and k not in seen:
                seen.add(k)
                total += (float(cells[4]) + float(cells[5])) * 1e6
    return total if len(seen) == len(kernels) else None


def seed_field(seeds, N):
    # icosahedral 30-vector star, |q| = 1, in a sphere of radius 6 (reference metric k = n/L)
    return seeds.icosahedral_nucleus(N, L=26.0, kscale=1.0, amp=0.1, seed=0)


class DeviceWarm:
    """Untimed gradient-flow steps on a separate context, run right before a
    timed run's W warm-up steps: the GPU leaves its idle clocks only over a few
    ms of sustained load (a 20-step window started cold runs about 10 % slower
    than the same steps after it), and the metric is the sustained step rate."""

    def __init__(self, spm, torch, N, prec, phi0, passes=0):
        self.torch = torch
        self.g = spm.Grid(N, prec, step_passes=passes)
        phi = torch.as_tensor(phi0).to(self.g.dtype).cuda()
        self.g.gf_begin(phi, 1 << 30, res_tol=-1.0, gn_tol=0.0)
        torch.cuda.synchronize()

    def run(self, seconds):
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < seconds:
            self.g.gf_steps(64)
            self.torch.cuda.synchronize()
            n += 64
        return n

    def close(self):
        self.g.gf_end(None, 0)
        self.g.close()


def time_gpu(spm, torch, N, prec, steps, warmup, phi0, shape=0, passes=0, warm_s=0.0):
    warm = DeviceWarm(spm, torch, N, prec, phi0, passes) if warm_s > 0 else None
    g = spm.Grid(N, prec, shape=shape, step_passes=passes)
    phi = torch.as_tensor(phi0).to(g.dtype).cuda()
    total = warmup + steps
    g.gf_begin(phi, total + 1, res_tol=-1.0, gn_tol=0.0)  # no early stop inside the timed region
    torch.cuda.synchronize()
    warm_steps = warm.run(warm_s) if warm else 0
    g.gf_steps(warmup)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    g.gf_steps(steps)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    it, hist = g.gf_end(phi, total + 1)
    assert it == total, (it, total)
    assert np.all(np.isfinite(hist)), "non-finite energy in the timed steps"
    E_after = g.energy(phi)  # get_Ene of the field after the last step
    field = phi.double().cpu().numpy()
    g.close()
    if warm:
        warm.close()
    return dict(wall_s=wall, event_ms=ev_ms, hist=hist, ctx_bytes=g.ctx_bytes, E_after=E_after, field=field,
                warm_steps=warm_steps)


# One cell of the {110} shell in the reference's metric (k = n / L with L = sqrt 2):
# its 12 wave vectors sit at |k| = 1, the minimum of G, and a BCC-seeded
# relaxation reaches the ordered minimum by its tolerance, not by a step cap.
CELL = dict(L=2 ** 0.5, kscale=1.0, Tau=-0.0061, Gamma=0.30, dt=0.3)


def converged_relax(N, gn_tol=1e-8, switch_tol=1e-6, max_iter=100000):
    """The mixed pipeline's accuracy at a converged state: relaxation of the BCC
    {110} seed in the cell to ||grad F|| <= gn_tol, in FP64 and mixed (FP32 steps
    to switch_tol, then FP64 steps to gn_tol).  Returns both fields (host) and
    the runs' figures."""
    pipeline = importlib.import_module("spm-iqc_amd.pipeline")
    seeds = importlib.import_module("spm-iqc_amd.seeds")
    phi0 = seeds.crystal(N, "bcc", L=CELL["L"], kscale=CELL["kscale"], amp=0.1, seed=0)
    runs, fields = {}, {}
    for mode in ("fp64", "mixed"):
        phi, info = pipeline.relax(N, phi0, mode, gn_tol=gn_tol, max_iter=max_iter, polish=0, params=CELL,
                                   switch_tol=switch_tol if mode == "mixed" else None)
        fields[mode] = phi.cpu().numpy()
        runs[mode] = {"E": info["E"], "grad_norm": info["gn"], "steps_fp32": info["iters_fp32"],
                      "steps_fp64": info["iters_fp64"], "seconds": round(info["t"], 4)}
    return fields, runs


def cpu_baseline(po, N, phi0, budget_s=20.0, check_fields=None):
    """The oracle (the reference's FP64 algorithm, OpenMP threads) on a bounded
    sample of the same workload: gradient-flow steps at the same N.  As the
    checker it also evaluates get_Ene (LPsrc.cpp:8-38) at the GPU runs' final
    fields (check_fields: name -> real field); those evaluations are not timed."""
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = po.Oracle(N, nthreads=nthreads)
    energies = {name: o.energy_of(f) for name, f in (check_fields or {}).items()}
    uR = phi0.astype(np.complex128)
    uF = o.my_fft(uR)
    o.get_min_u_F(uR, uF, 1, -1.0)  # warm-up (page-in of the work arrays)
    n, t = 0, 0.0
    while t < budget_s and n < 50:
        t0 = time.perf_counter()
        o.get_min_u_F(uR, uF, 1, -1.0)
        t += time.perf_counter() - t0
        n += 1
    o.close()
    return dict(value=n / t, unit="steps/s", cores=nthreads, kind="port",
                sample="%d FP64 gradient-flow steps of the oracle at %d^3 (%.1f s)" % (n, N, t)), energies


SERVE_WORKLOAD = ("reduced serving stream: %d Poisson requests at %.0f/s, 60%% relax (N 64/128/256, to "
                  "|grad F| 1e-6 or 2000 steps), 25%% 4-mode LOBPCG (N 64/128), 15%% SPM 100 iterations (N 128); "
                  "configuration 5 as written (500 requests at 8/s, relaxations to |grad F| 1e-8 within 100000 "
                  "steps, 10 LOBPCG calls, Spring_Pairing(30000, 1e-9)) is served by serve.py --as-written")


def serve_stream(n, rate):
    """p50/p99 latency of a Poisson stream of relax / eigenmode / saddle requests
    (64^3-256^3, mixed precision, continuous batching) and of the same stream
    served one request at a time in FP64."""
    serve = importlib.import_module("spm-iqc_amd.serve")
    reqs = serve.make_requests(n, rate, seed=0)
    states = serve.prepare_states(reqs)
    recs, st = serve.serve(reqs, states, mixed=True)
    base, sb = serve.serve_one_at_a_time(reqs, states)
    keep = ("requests", "wall_s", "throughput_rps", "p50_s", "p99_s", "errors")
    return {"requests": n, "rate_rps": rate, "workload": SERVE_WORKLOAD % (n, rate),
            "p99_s": round(st["p99_s"], 4), "p50_s": round(st["p50_s"], 4),
            "mixed_batched": {k: st[k] for k in keep},
            "one_at_a_time_fp64": {k: sb[k] for k in keep},
            "mixed_vs_fp64": serve.compare(recs, base)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--device-warm", type=float, default=0.3,
                    help="seconds of untimed steps on a separate context before each run's warm-up "
                         "(brings the GPU to its sustained clocks); 0 disables")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 512 MiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--shape", default="-1",
                    help="FFT pass shape index (threads per pencil of the three passes), or FP64,FP32 indices; "
                         "-1: measure them all first and use the fastest (the tuner; only for the four-pass step)")
    ap.add_argument("--step-passes", type=int, default=0, choices=(0, 2, 4),
                    help="array passes per step: 0 = the library default (2 at N = 64, 128, 256, else 4)")
    ap.add_argument("--serve-requests", type=int, default=120,
                    help="requests of the serving stream (p50/p99 latency); 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=12.0)
    ap.add_argument("--converged-N", type=int, default=256,
                    help="grid of the mixed-vs-FP64 relaxation to a converged state (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the workload is one 256^3 field); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    spm = importlib.import_module("spm-iqc_amd")
    seeds = importlib.import_module("spm-iqc_amd.seeds")
    N = args.N
    phi0 = seed_field(seeds, N)

    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    shapes = spm.shapes(N)
    passes = args.step_passes or (2 if N in (64, 128, 256) else 4)
    tuned = {}
    want = [int(x) for x in str(args.shape).split(",")]
    want = want * 2 if len(want) == 1 else want
    for prec in (spm.FP64, spm.FP32):
        if want[prec] >= 0 or passes == 2:  # the two-pass step has no shape to tune
            tuned[prec] = (max(want[prec], 0), None)
        else:  # the tuner: every shape timed on 30 graph-replayed steps of a noise field
            best, us = spm.tune(N, prec, nmem=1, steps=30)
            tuned[prec] = (best, us)
    r64 = time_gpu(spm, torch, N, spm.FP64, args.steps, args.warmup, phi0, tuned[spm.FP64][0], passes,
                   args.device_warm)
    r32 = time_gpu(spm, torch, N, spm.FP32, args.steps, args.warmup, phi0, tuned[spm.FP32][0], passes,
                   args.device_warm)

    sps64 = args.steps / (r64["event_ms"] / 1e3)
    sps32 = args.steps / (r32["event_ms"] / 1e3)
    B64 = algorithmic_bytes_per_step(N, 8)
    B32 = algorithmic_bytes_per_step(N, 4)
    ach64 = B64 * sps64
    ach32 = B32 * sps32

    conv_fields, conv = None, None
    if args.converged_N > 0:
        conv_fields, conv = converged_relax(args.converged_N)

    serving = None
    if args.serve_requests > 0:
        serving = serve_stream(args.serve_requests, args.serve_rate)

    cpu, E_oracle = None, {}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po

        cpu, E_oracle = cpu_baseline(po, N, phi0, args.cpu_budget,
                                     check_fields={"fp64": r64["field"], "fp32": r32["field"]})
    # energy after the last step: the mixed run against the FP64 run, both
    # evaluated by the oracle's get_Ene at the final fields (FP32 steps only;
    # the mixed pipeline ends with an FP64 polish, spm-iqc_amd/pipeline.py)
    mixed_err = {"E_fp64": r64["E_after"], "E_mixed": r32["E_after"],
                 "E_rel_mixed_vs_fp64": abs(r32["E_after"] - r64["E_after"]) / abs(r64["E_after"])}
    if conv is not None:
        e64, emx = conv["fp64"]["E"], conv["mixed"]["E"]
        conv.update({
            "workload": "BCC {110} seed (amplitude 0.1, seeded phases) in one {110} cell, L = sqrt 2, k = n/L, "
                        "%d^3, relaxed to |grad F| <= 1e-8; mixed: FP32 steps to |grad F| 1e-6, then FP64 steps"
                        % args.converged_N,
            "E_rel_mixed_vs_fp64": abs(emx - e64) / abs(e64),
            "field_max_diff": float(np.abs(conv_fields["mixed"] - conv_fields["fp64"]).max()),
            "time_ratio_fp64_vs_mixed": round(conv["fp64"]["seconds"] / conv["mixed"]["seconds"], 3)})
        if not args.no_cpu:
            o = po.Oracle(args.converged_N, L=CELL["L"], kscale=CELL["kscale"])
            eo = {m: o.energy_of(conv_fields[m]) for m in ("fp64", "mixed")}
            o.close()
            conv.update({"E_oracle_at_fp64_field": eo["fp64"], "E_oracle_at_mixed_field": eo["mixed"],
                         "E_rel_mixed_vs_fp64_oracle": abs(eo["mixed"] - eo["fp64"]) / abs(eo["fp64"])})
    if E_oracle:
        mixed_err.update({
            "E_oracle_at_fp64_field": E_oracle["fp64"], "E_oracle_at_mixed_field": E_oracle["fp32"],
            "E_rel_mixed_vs_fp64_oracle": abs(E_oracle["fp32"] - E_oracle["fp64"]) / abs(E_oracle["fp64"]),
            "E_rel_fp64_gpu_vs_oracle": abs(r64["E_after"] - E_oracle["fp64"]) / abs(E_oracle["fp64"]),
        })

    out = {
        "metric": "gradient-flow steps/s at 256^3 FP64 vs mixed, % HBM roofline; p99 solve latency",
        "value": round(sps64, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["event_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded icosahedral nucleus, 30-vector star |q|=1, sphere r=6, tanh edge 1)",
        "config": {"workload": "semi-implicit gradient flow (get_min_u_F), one %d^3 field, L=26, k=n/L, "
                               "Tau=-0.0061, Gamma=0.30, dt=0.3" % N, "N": N, "batch": 1,
                   "step_passes": passes,
                   "fft_shape_fp64": shapes[tuned[spm.FP64][0]] if passes == 4 else None,
                   "fft_shape_fp32": shapes[tuned[spm.FP32][0]] if passes == 4 else None,
                   "serving_workload": (SERVE_WORKLOAD % (args.serve_requests, args.serve_rate)
                                        if args.serve_requests > 0 else None)},
        "mixed": {"dtype": "f32 fields/FFTs, f64 sums", "value": round(sps32, 3),
                  "ms_per_step": round(r32["event_ms"] / args.steps, 4),
                  "speedup_vs_fp64": round(sps32 / sps64, 3),
                  "roofline_frac": round(ach32 / HBM_PEAK, 4),
                  "energy_error": mixed_err,
                  "converged": conv},
        "roofline": {"bound": "hbm",
                     "kernel": ("gradient-flow step (2 launches: j stage B + i-axis FFT + update + inverse + B^-1; "
                                "j stage A^-1 + k-axis c2r + f(phi) + r2c + A)" if passes == 2 else
                                "gradient-flow step (4 launches: i-axis fwd FFT + update + inv FFT; "
                                "j-axis inv FFT; k-axis c2r + f(phi) + r2c; j-axis fwd FFT)"),
                     "achieved": round(ach64 / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(ach64 / HBM_PEAK, 4), "traffic": measured_traffic(N, "double", passes),
                     "algorithmic_bytes_per_step": B64},
        "cpu_baseline": cpu,
        "serving": serving,
        "wall_s": {"fp64": round(r64["wall_s"], 4), "fp32": round(r32["wall_s"], 4)},
        "device_warm": {"seconds": args.device_warm, "fp64_steps": r64["warm_steps"], "fp32_steps": r32["warm_steps"],
                        "note": "untimed steps on a separate context before each run's W warm-up steps, so the "
                                "timed K steps run at the GPU's sustained clocks"},
        "energy_last": {"fp64": float(r64["hist"][-1, 0]), "fp32": float(r32["hist"][-1, 0])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
