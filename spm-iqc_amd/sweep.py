"""Batched parameter sweep of gradient-flow relaxations (get_min_u_F,
LPsrc.cpp:40-117) over (Tau, Gamma), with uneven grid sizes.

Jobs of one grid size share a batch (one launch per pass for all members,
steps replayed from HIP graphs, FFT pass shapes from the tuner).  Mixed
precision: every job relaxes in an FP32 batch until ||grad F|| <= fp32_tol
(or its FP32 budget), then moves to an FP64 batch of its size for the polish
to gn_tol; a job moves as soon as it stops, so the FP32 and FP64 batches of
all sizes run side by side on their own streams.

    python spm-iqc_amd/sweep.py [--jobs 64] [--max-iter 20000] [--baseline]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
spm = importlib.import_module("spm-iqc_amd")
seeds = importlib.import_module("spm-iqc_amd.seeds")

KINDS = ("bcc", "fcc", "lam", "ico")  # round-robin seed families
KSCALE = 2.0 * np.pi                  # k = 2 pi n / L: the |q| = 1 stars are resolved from 64^3 up


def make_jobs(n_jobs=64, sizes=(64, 96, 128), seed=0, L=26.0):
    """(Tau, Gamma) on the 8x8 grid linspace(-0.02, 0.01, 8) x linspace(0.20,
    0.40, 8) in row-major order, N drawn from `sizes` with a seeded RNG, seed
    families round-robin, amplitude 0.1 with phases seeded by the job index."""
    rng = np.random.default_rng(seed)
    taus = np.linspace(-0.02, 0.01, 8)
    gammas = np.linspace(0.20, 0.40, 8)
    grid = [(t, g) for t in taus for g in gammas]
    jobs = []
    for j in range(n_jobs):
        Tau, Gamma = grid[j % len(grid)]
        jobs.append(dict(id=j, N=int(rng.choice(sizes)), Tau=float(Tau), Gamma=float(Gamma),
                         kind=KINDS[j % len(KINDS)], seed=j, L=L))
    return jobs


def seed_field(job, dtype=None):
    """The job's seed as a device tensor (FP64 unless dtype is given)."""
    torch = spm._torch()
    kind = "nucleus" if job["kind"] == "ico" else job["kind"]
    return seeds.device_field(torch, job["N"], kind, job["L"], KSCALE, 0.1, job["seed"], dtype=dtype)


class _Stage:
    """One batch (size, precision) with a queue of jobs waiting for a slot."""

    def __init__(self, N, prec, capacity, hist_cap, shape, dt=0.3, L=26.0):
        self.b = spm.Batch(N, prec, capacity=capacity, hist_cap=hist_cap, L=L, kscale=KSCALE, dt=dt, shape=shape)
        self.free = list(range(capacity))
        self.slot_job = {}
        self.queue = []

    def admit(self, load):
        while self.free and self.queue:
            job, phi, max_iter, tol = self.queue.pop(0)
            slot = self.free.pop(0)
            self.b.load(slot, phi, job["Tau"], job["Gamma"], max_iter, -1.0, tol)
            self.slot_job[slot] = job
            load(job)

    def busy(self):
        return bool(self.slot_job) or bool(self.queue)


def run(jobs, mixed=True, gn_tol=1e-8, fp32_tol=1e-7, max_iter=20000, chunk=64, shape=-1, capacity=None,
        verbose=False):
    """Relax every job; returns (per-job results, stats)."""
    torch = spm._torch()
    sizes = sorted({j["N"] for j in jobs})
    by_n = {N: [j for j in jobs if j["N"] == N] for N in sizes}
    stages = {}
    for N in sizes:
        cap = capacity or len(by_n[N])
        if mixed:
            stages[(N, spm.FP32)] = _Stage(N, spm.FP32, cap, max_iter, shape)
        stages[(N, spm.FP64)] = _Stage(N, spm.FP64, cap, max_iter, shape)
    results = {}
    first = spm.FP32 if mixed else spm.FP64
    fields = [seed_field(j, torch.float32 if mixed else torch.float64) for j in jobs]  # inputs, before the clock
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, phi in zip(jobs, fields):
        stages[(j["N"], first)].queue.append((j, phi, max_iter, fp32_tol if mixed else gn_tol))
        results[j["id"]] = dict(id=j["id"], N=j["N"], Tau=j["Tau"], Gamma=j["Gamma"], kind=j["kind"],
                                iters_fp32=0, iters_fp64=0)
    member_steps = 0
    launches = 0
    while any(s.busy() for s in stages.values()):
        for s in stages.values():
            s.admit(lambda job: None)
        active = [s for s in stages.values() if s.slot_job]
        for s in active:
            s.b.steps(chunk)
            launches += 1
        for (N, prec), s in stages.items():
            if not s.slot_job:
                continue
            running, iters, stopped = s.b.poll()
            for slot in [k for k in s.slot_job if stopped[k]]:
                job = s.slot_job.pop(slot)
                phi, hist = s.b.read(slot)
                s.b.release(slot)
                s.free.append(slot)
                r = results[job["id"]]
                n = len(hist)
                member_steps += n
                if prec == spm.FP32:
                    r["iters_fp32"] = n
                    r["gn_fp32"] = float(hist[-1, 1]) if n else float("nan")
                    rest = max(max_iter - n, 1)
                    stages[(N, spm.FP64)].queue.append((job, phi.double(), rest, gn_tol))
                else:
                    r["iters_fp64"] = n
                    r["E"] = float(hist[-1, 0]) if n else float("nan")
                    r["gn"] = float(hist[-1, 1]) if n else float("nan")
                    r["phi_absmax"] = float(phi.abs().max().item())
                    r["t_done"] = time.perf_counter() - t0
                    if verbose:
                        print("  job %2d N %3d done: %5d + %5d its, E %.12e |grad F| %.2e" % (
                            job["id"], N, r["iters_fp32"], n, r["E"], r["gn"]), flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for s in stages.values():
        s.b.close()
    stats = dict(jobs=len(jobs), wall_s=wall, member_steps=member_steps, member_steps_per_s=member_steps / wall,
                 graph_launch_rounds=launches, mixed=mixed, chunk=chunk)
    return [results[j["id"]] for j in jobs], stats


def run_one_at_a_time(jobs, gn_tol=1e-8, max_iter=20000, verbose=False):
    """The same relaxations in FP64, one single-field context at a time."""
    torch = spm._torch()
    out = []
    fields = [seed_field(j) for j in jobs]  # inputs, before the clock
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = 0
    for j, phi in zip(jobs, fields):
        with spm.Grid(j["N"], spm.FP64, L=j["L"], kscale=KSCALE, Tau=j["Tau"], Gamma=j["Gamma"]) as g:
            it, hist = g.gradflow(phi, max_iter, -1.0, gn_tol, check_every=256)
            steps += it
            out.append(dict(id=j["id"], iters=it, E=float(hist[-1, 0]), gn=float(hist[-1, 1])))
            if verbose:
                print("  single job %2d N %3d: %5d its, E %.12e" % (j["id"], j["N"], it, hist[-1, 0]), flush=True)
    wall = time.perf_counter() - t0
    return out, dict(wall_s=wall, member_steps=steps, member_steps_per_s=steps / wall)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--gn-tol", type=float, default=1e-8)
    ap.add_argument("--fp32-tol", type=float, default=1e-7)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--baseline", action="store_true", help="also run the jobs one at a time in FP64")
    ap.add_argument("--out", default=None)
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    jobs = make_jobs(a.jobs)
    res, st = run(jobs, mixed=not a.fp64_only, gn_tol=a.gn_tol, fp32_tol=a.fp32_tol, max_iter=a.max_iter,
                  chunk=a.chunk, verbose=not a.quiet)
    out = {"batched": st, "results": res}
    if a.baseline:
        if not a.fp64_only:  # the batching gain alone: the same sweep batched in FP64
            _, s64 = run(jobs, mixed=False, gn_tol=a.gn_tol, max_iter=a.max_iter, chunk=a.chunk)
            out["batched_fp64"] = s64
        ref, sb = run_one_at_a_time(jobs, a.gn_tol, a.max_iter, verbose=not a.quiet)
        out["one_at_a_time_fp64"] = sb
        out["one_at_a_time_results"] = ref
        out["speedup_wall"] = sb["wall_s"] / st["wall_s"]
        # energies: relative where the minimum is ordered (|E| > 1e-10), absolute at the disordered one
        ordered = [(r, q) for r, q in zip(res, ref) if abs(q["E"]) > 1e-10]
        flat = [(r, q) for r, q in zip(res, ref) if abs(q["E"]) <= 1e-10]
        dE = [abs(r["E"] - q["E"]) / abs(q["E"]) for r, q in ordered]
        out["ordered_minima"] = len(ordered)
        out["E_rel_vs_fp64_ordered_median"] = float(np.median(dE)) if dE else None
        out["E_rel_vs_fp64_ordered_max"] = max(dE) if dE else None
        out["E_abs_vs_fp64_disordered_max"] = max([abs(r["E"] - q["E"]) for r, q in flat] or [0.0])
    txt = json.dumps(out, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k not in ("results", "one_at_a_time_results")}, indent=1),
          flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    return out


if __name__ == "__main__":
    main()
