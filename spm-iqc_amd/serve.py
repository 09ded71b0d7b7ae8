"""Serving loop for relax-to-minimum, eigenmode and saddle-search requests of
different grid sizes, with continuous batching.

Requests arrive as a Poisson stream.  Relaxations (get_min_u_F,
LPsrc.cpp:40-117) join a device batch of their grid size as soon as a slot
is free, in FP32, and move to an FP64 batch for the polish when they stop;
the batches of every size and precision advance together, each on its own
stream, a chunk of graph-replayed iterations at a time.  Eigenmode solves
(hiosd_lobpcg_initialization + hiosd_lobpcg_inip, LOBPCG.cpp:10-487) and
saddle searches (Spring_Pairing, SPM.cpp:170-263) run on worker threads, each
with its own context and stream, admitted while the estimated device memory
of the running work stays under the budget.  Fields stay on the device from
admission to completion.

Latency of a request = completion time - arrival time.  The reference
point is the same stream served one request at a time in FP64, in arrival
order.

    python spm-iqc_amd/serve.py [--requests 500] [--rate 8] [--baseline] [--as-written]
    python spm-iqc_amd/serve.py --as-written --baseline-only --out fp64.json
    python spm-iqc_amd/serve.py --compare mixed.json fp64.json

The default stream is reduced (relaxations to ||grad F|| 1e-6 or 2000 steps, 2
LOBPCG calls, 100 Spring_Pairing iterations).  --as-written serves the
configuration as stated: relaxations to ||grad F|| <= 1e-8 within the
reference's 100000 steps (LPsrc.cpp:52; mixed: FP32 to 1e-6, then FP64), the
4 lowest modes from 10 LOBPCG calls of 21 iterations (main_SPM.cpp:88-91),
and Spring_Pairing(30000, 1e-9) (main_SPM.cpp:122).  Its one-at-a-time FP64
reference can run as a separate process (--baseline-only) and be compared
afterwards (--compare), without a GPU.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import queue
import sys
import threading
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
spm = importlib.import_module("spm-iqc_amd")
seeds = importlib.import_module("spm-iqc_amd.seeds")
sweep = importlib.import_module("spm-iqc_amd.sweep")
pipeline = importlib.import_module("spm-iqc_amd.pipeline")

KSCALE = sweep.KSCALE
L = 26.0
PARAMS = dict(L=L, kscale=KSCALE, Tau=-0.0061, Gamma=0.30, dt=0.3)


# ---------------------------------------------------------------------------
# workload
# ---------------------------------------------------------------------------
# the stream as the configuration states it (see the module docstring)
AS_WRITTEN = dict(relax_tol=1e-8, relax_iters=100000, switch_tol=1e-6, modes_calls=10, spm_iters=30000,
                  spm_tol=1e-9)


def make_requests(n=500, rate=8.0, seed=0, relax_sizes=(64, 128, 256), modes_sizes=(64, 128), spm_size=128,
                  relax_tol=1e-6, relax_iters=2000, modes_calls=2, spm_iters=100, switch_tol=None, spm_tol=1e-9):
    """Poisson arrivals (rate per second) with a seeded mix: 60% relax (seeds
    round-robin over the four families, (Tau, Gamma) from the sweep grid), 25%
    lowest-4-mode LOBPCG on a relaxed state, 15% SPM from a relaxed
    icosahedral nucleus offset along its first mode.  switch_tol: the
    ||grad F|| at which a mixed relaxation hands over from FP32 to FP64 (None:
    at its own tolerance)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    u = rng.uniform(size=n)
    grid = [(a, b) for a in np.linspace(-0.02, 0.01, 8) for b in np.linspace(0.20, 0.40, 8)]
    reqs = []
    for i in range(n):
        if u[i] < 0.60:
            Tau, Gamma = grid[int(rng.integers(len(grid)))]
            reqs.append(dict(id=i, t=float(t[i]), kind="relax", N=int(rng.choice(relax_sizes)),
                             Tau=float(Tau), Gamma=float(Gamma), seed_kind=sweep.KINDS[i % 4], seed=i,
                             tol=relax_tol, max_iter=relax_iters, switch_tol=switch_tol))
        elif u[i] < 0.85:
            reqs.append(dict(id=i, t=float(t[i]), kind="modes", N=int(rng.choice(modes_sizes)), seed=i,
                             calls=modes_calls))
        else:
            reqs.append(dict(id=i, t=float(t[i]), kind="saddle", N=spm_size, seed=i, iters=spm_iters, tol=spm_tol))
    return reqs


def relax_seed(r, dtype=None):
    """The request's seed field, generated on the device."""
    return sweep.seed_field(dict(N=r["N"], L=L, kind=r["seed_kind"], seed=r["seed"]), dtype)


def prepare_states(reqs, verbose=False):
    """Relaxed states (and their first mode) that the eigenmode and saddle
    requests start from: relaxed icosahedral nuclei per size, computed once
    before the stream starts (FP64)."""
    torch = spm._torch()
    sizes = sorted({r["N"] for r in reqs if r["kind"] in ("modes", "saddle")})
    states = {}
    for N in sizes:
        phi0 = seeds.device_field(torch, N, "nucleus", L, KSCALE, 0.1, 0)
        phi, info = pipeline.relax(N, phi0, "fp64", gn_tol=1e-8, max_iter=4000, params=PARAMS)
        ritz, V, _ = pipeline.lowest_modes(N, phi, "fp64", k=8, nwant=4, calls=3, params=PARAMS)
        states[N] = dict(phi=phi, v1=V[0].clone())
        del V
        if verbose:
            print("  state N %d: E %.12e |grad F| %.2e, lowest %s" % (N, info["E"], info["gn"], ritz), flush=True)
    torch.cuda.synchronize()
    return states


# per context: the Gram partials and block coefficients of the block operations
# (2048 blocks x a 32 x 32 tile of doubles, plus 4 x 4096 doubles)
CTX_SCRATCH = 8 * (2048 * 32 * 32 + 4 * 4096)


def footprint(r):
    """Device bytes a worker request holds at its peak: the context's two work
    spectra and scratch, its solver workspace, and the fields the request keeps
    (peak over the stages: the mixed requests' FP64 polish holds FP64 copies)."""
    N = r["N"]
    R = N ** 3

    def ctx(b):
        return 2 * N * N * (N // 2 + 1 + 16) * 2 * b + CTX_SCRATCH

    if r["kind"] == "modes":
        k = 8  # workspace U, W, Y, Vp, Up, T (6k fields), the block V (FP64 outside), phi; at N <= 128
        # the block's operators run as batch members with k spectra of their own

        def spectra(b):
            return k * N * N * (N // 2 + 16) * 2 * b if N <= 128 else 0

        fp32 = ctx(4) + (6 * k + 2) * R * 4 + k * R * 8 + spectra(4)
        fp64 = ctx(8) + (6 * k + 1) * R * 8 + k * R * 8 + spectra(8)
        return max(fp32, fp64)
    # saddle: two replicas, d and the kept spectrum (workspace), ngrad scratch; the FP64
    # replicas made from the FP32 ones, and phi, v1
    fp32 = ctx(4) + 4 * R * 4 + N * N * (N // 2 + 17) * 2 * 4 + 4 * R * 8
    fp64 = ctx(8) + 4 * R * 8 + N * N * (N // 2 + 17) * 2 * 8 + 2 * R * 8
    return max(fp32, fp64)


# ---------------------------------------------------------------------------
# worker requests (one context per request, on the worker's stream)
# ---------------------------------------------------------------------------
def run_modes(r, states, mixed):
    torch = spm._torch()
    N, k = r["N"], 8
    phi = states[N]["phi"]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(r["seed"])
    start = torch.rand((k, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 2.0 - 1.0
    ritz, V, info = pipeline.lowest_modes(N, phi, "mixed" if mixed else "fp64", k=k, nwant=4,
                                          calls=r["calls"], seed=r["seed"], modes_polish=1, params=PARAMS,
                                          start=start)
    del V
    return dict(eigenvalues=[float(x) for x in ritz])


def run_saddle(r, states, mixed):
    st = states[r["N"]]
    n = r["iters"]
    p0, p1, info = pipeline.saddle(r["N"], st["phi"], st["v1"], "mixed" if mixed else "fp64", offset=0.01,
                                   maxiter=n, tol=r.get("tol", 1e-9), spm_polish=max(n // 5, 1), params=PARAMS)
    return dict(E=info["E"], res=info["res"], iters=info["iters"])


def run_relax_single(r):
    torch = spm._torch()
    with spm.Grid(r["N"], spm.FP64, L=L, kscale=KSCALE, Tau=r["Tau"], Gamma=r["Gamma"]) as g:
        phi = relax_seed(r)
        it, hist = g.gradflow(phi, r["max_iter"], -1.0, r["tol"], check_every=128)
    return dict(iters=it, E=float(hist[-1, 0]), gn=float(hist[-1, 1]))


# Worker requests start shortest expected service first: an eigenmode solve
# (a few hundred LOBPCG iterations) before a saddle search (thousands of
# Spring_Pairing iterations), each kind in arrival order.
PRIORITY = {"modes": 0, "saddle": 1}


class _Workers:
    def __init__(self, n, states, mixed):
        self.q = queue.PriorityQueue()  # (priority, arrival, id, request); None-request entries stop a thread
        self.done = queue.Queue()
        self.states, self.mixed = states, mixed
        self.threads = [threading.Thread(target=self._loop, daemon=True) for _ in range(n)]
        for t in self.threads:
            t.start()

    def _loop(self):
        torch = spm._torch()
        s = torch.cuda.Stream()
        while True:
            r = self.q.get()[3]
            if r is None:
                return
            try:
                with torch.cuda.stream(s):
                    out = run_modes(r, self.states, self.mixed) if r["kind"] == "modes" else \
                        run_saddle(r, self.states, self.mixed)
                    s.synchronize()
                self.done.put((r, out, None))
            except Exception as e:  # reported with the request
                self.done.put((r, None, repr(e)))

    def put(self, r):
        self.q.put((PRIORITY[r["kind"]], r["t"], r["id"], r))

    def close(self):
        for i, _ in enumerate(self.threads):
            self.q.put((len(PRIORITY), float("inf"), i, None))
        for t in self.threads:
            t.join()


# ---------------------------------------------------------------------------
# the server
# ---------------------------------------------------------------------------
def serve(reqs, states, mixed=True, capacity=None, chunk=32, workers=4, budget_bytes=None, verbose=False):
    """Serve the stream in real time; returns (per-request records, stats)."""
    torch = spm._torch()
    capacity = capacity or {64: 16, 128: 8, 256: 4}
    if budget_bytes is None:
        free, total = torch.cuda.mem_get_info()
        budget_bytes = int(0.8 * free)
    sizes = sorted({r["N"] for r in reqs if r["kind"] == "relax"})
    max_iter = max([r["max_iter"] for r in reqs if r["kind"] == "relax"] or [1])
    stages = {}
    for N in sizes:
        if mixed:
            stages[(N, spm.FP32)] = sweep._Stage(N, spm.FP32, capacity.get(N, 4), max_iter, -1, L=L)
        stages[(N, spm.FP64)] = sweep._Stage(N, spm.FP64, capacity.get(N, 4), max_iter, -1, L=L)
    reserved = sum(s.b.dev_bytes for s in stages.values())
    pool = _Workers(workers, states, mixed)
    in_flight = {}  # worker requests: id -> footprint
    waiting = []    # worker requests not yet admitted
    recs = {}
    pending = sorted(reqs, key=lambda r: r["t"])
    first = spm.FP32 if mixed else spm.FP64
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nxt = 0
    shown = 0
    while len(recs) < len(reqs):
        now = time.perf_counter() - t0
        # arrivals
        while nxt < len(pending) and pending[nxt]["t"] <= now:
            r = pending[nxt]
            nxt += 1
            if r["kind"] == "relax":
                phi = relax_seed(r, torch.float32 if mixed else torch.float64)
                tol = max(r["tol"], r.get("switch_tol") or 0.0) if mixed else r["tol"]
                stages[(r["N"], first)].queue.append((r, phi, r["max_iter"], tol))
            else:
                waiting.append(r)
        # admission of worker requests by device-memory footprint
        while waiting:
            fp = footprint(waiting[0])
            if in_flight and reserved + sum(in_flight.values()) + fp > budget_bytes:
                break
            r = waiting.pop(0)
            in_flight[r["id"]] = fp
            r["t_admit"] = time.perf_counter() - t0
            pool.put(r)
        # batches: admit, advance, retire
        for s in stages.values():
            s.admit(lambda req: req.__setitem__("t_admit", time.perf_counter() - t0))
        active = [s for s in stages.values() if s.slot_job]
        for s in active:
            s.b.steps(chunk)
        for (N, prec), s in stages.items():
            if not s.slot_job:
                continue
            running, iters, stopped = s.b.poll()
            for slot in [k for k in s.slot_job if stopped[k]]:
                r = s.slot_job.pop(slot)
                phi, hist = s.b.read(slot)
                s.b.release(slot)
                s.free.append(slot)
                n = len(hist)
                if prec == spm.FP32:
                    r["iters_fp32"] = n
                    rest = max(r["max_iter"] - n, 1)
                    stages[(N, spm.FP64)].queue.append((r, phi.double(), rest, r["tol"]))
                else:
                    t_done = time.perf_counter() - t0
                    recs[r["id"]] = dict(id=r["id"], kind="relax", N=N, t=r["t"], t_done=t_done,
                                         latency=t_done - r["t"], iters_fp32=r.get("iters_fp32", 0),
                                         iters_fp64=n, E=float(hist[-1, 0]), gn=float(hist[-1, 1]))
        # worker completions
        while True:
            try:
                r, out, err = pool.done.get_nowait()
            except queue.Empty:
                break
            in_flight.pop(r["id"], None)
            t_done = time.perf_counter() - t0
            rec = dict(id=r["id"], kind=r["kind"], N=r["N"], t=r["t"], t_done=t_done, latency=t_done - r["t"])
            if err:
                rec["error"] = err
            else:
                rec.update(out)
            recs[r["id"]] = rec
        if not active and len(recs) < len(reqs):
            # nothing on the batches: wait for the next arrival or a worker
            wait = pending[nxt]["t"] - (time.perf_counter() - t0) if nxt < len(pending) else 0.005
            time.sleep(min(max(wait, 0.0), 0.005))
        if verbose and time.perf_counter() - t0 >= shown:
            shown += 2.0
            occ = " ".join("%d%s:%d/%d+%d" % (N, "s" if p == spm.FP32 else "d", len(s.slot_job), s.b.capacity,
                                                len(s.queue)) for (N, p), s in stages.items())
            print("  t %5.1f s: arrived %d, served %d, workers %d running %d waiting; batches %s" % (
                time.perf_counter() - t0, nxt, len(recs), len(in_flight), len(waiting), occ), flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pool.close()
    for s in stages.values():
        s.b.close()
    out = [recs[r["id"]] for r in reqs]
    return out, summarize(out, wall)


def serve_one_at_a_time(reqs, states, verbose=False):
    """The same stream served one request at a time in FP64, in arrival order."""
    torch = spm._torch()
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    shown = 0.0
    for r in sorted(reqs, key=lambda r: r["t"]):
        now = time.perf_counter() - t0
        if r["t"] > now:
            time.sleep(r["t"] - now)
        if r["kind"] == "relax":
            out = run_relax_single(r)
        elif r["kind"] == "modes":
            out = run_modes(r, states, False)
        else:
            out = run_saddle(r, states, False)
        torch.cuda.synchronize()
        t_done = time.perf_counter() - t0
        rec = dict(id=r["id"], kind=r["kind"], N=r["N"], t=r["t"], t_done=t_done, latency=t_done - r["t"])
        rec.update(out)
        recs.append(rec)
        if verbose and (len(recs) % 50 == 0 or t_done - shown >= 10.0):
            shown = t_done
            print("  one at a time: %d / %d at t = %.1f s" % (len(recs), len(reqs), t_done), flush=True)
    wall = time.perf_counter() - t0
    recs.sort(key=lambda x: x["id"])
    return recs, summarize(recs, wall)


def summarize(recs, wall):
    lat = np.array([r["latency"] for r in recs])
    per = {}
    for k in ("relax", "modes", "saddle"):
        lk = np.array([r["latency"] for r in recs if r["kind"] == k])
        if len(lk):
            per[k] = dict(n=int(len(lk)), p50_s=float(np.percentile(lk, 50)), p99_s=float(np.percentile(lk, 99)))
    return dict(requests=len(recs), wall_s=wall, throughput_rps=len(recs) / wall,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), errors=sum(1 for r in recs if "error" in r), per_kind=per)


def compare(a, b, tol=None):
    """Differences between two served copies of the same stream (a against b).
    Relaxations: where both reached their tolerance, energies compared
    absolutely at the disordered minimum (|E| <= 1e-10) and relatively at
    ordered ones; pairs whose energies differ by more than 1e-6 relative (1e-12
    absolute) ended in different minima and are counted apart; where a run
    stopped at its step budget instead, the relative difference of the energy
    reached after the same budget."""
    conv_abs, conv_rel, capped_rel, n_conv, n_cap = [], [], [], 0, 0
    eig, sad, other = [], [], []
    for x, y in zip(a, b):
        if x["kind"] == "relax":
            t = tol if tol is not None else 1e-6
            if x["gn"] <= t and y["gn"] <= t:
                n_conv += 1
                d = abs(x["E"] - y["E"])
                if d > max(1e-12, 1e-6 * abs(y["E"])):
                    # a different minimum: the trajectories parted near a saddle
                    other.append(x["id"])
                elif abs(y["E"]) <= 1e-10:
                    conv_abs.append(d)
                else:
                    conv_rel.append(d / abs(y["E"]))
            else:
                n_cap += 1
                capped_rel.append(abs(x["E"] - y["E"]) / max(abs(y["E"]), 1e-300))
        elif x["kind"] == "modes":
            eig.append(float(np.max(np.abs(np.subtract(x["eigenvalues"], y["eigenvalues"])))))
        elif "E" in x and "E" in y:
            sad.append(abs(x["E"] - y["E"]) / max(abs(y["E"]), 1e-300))

    def mx(v):
        return float(max(v)) if v else None

    def med(v):
        return float(np.median(v)) if v else None

    return {"relax_converged": n_conv, "relax_converged_other_minimum": len(other),
            "relax_converged_other_minimum_ids": other, "relax_converged_E_abs_max_disordered": mx(conv_abs),
            "relax_converged_E_rel_max_ordered": mx(conv_rel), "relax_budget_capped": n_cap,
            "relax_capped_E_rel_median": med(capped_rel), "relax_capped_E_rel_max": mx(capped_rel),
            "eig_abs_max": mx(eig), "eig_abs_median": med(eig), "saddle_E_rel_max": mx(sad),
            "saddle_E_rel_median": med(sad)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--requests", type=int, default=500)
    ap.add_argument("--rate", type=float, default=8.0, help="arrivals per second")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--baseline", action="store_true", help="also serve one request at a time in FP64")
    ap.add_argument("--baseline-only", action="store_true", help="serve one request at a time in FP64 only")
    ap.add_argument("--as-written", action="store_true",
                    help="relaxations to 1e-8 within 100000 steps, 10 LOBPCG calls, Spring_Pairing(30000, 1e-9)")
    ap.add_argument("--compare", nargs=2, metavar=("MIXED_JSON", "FP64_JSON"),
                    help="compare the records of two earlier runs (no GPU) and print the differences")
    ap.add_argument("--out", default=None)
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    if a.compare:
        runs = []
        for fn in a.compare:
            with open(fn) as f:
                runs.append(json.load(f))
        recs = runs[0]["records"]
        base = runs[1]["records_fp64"] or runs[1]["records"]
        tol = runs[0].get("relax_tol", 1e-6)
        out = {"mixed_vs_fp64": compare(recs, base, tol)}
        print(json.dumps(out), flush=True)
        return out
    kw = dict(AS_WRITTEN) if a.as_written else {}
    reqs = make_requests(a.requests, a.rate, a.seed, **kw)
    states = prepare_states(reqs, verbose=not a.quiet)
    out = {"rate_rps": a.rate, "requests": a.requests, "seed": a.seed, "as_written": a.as_written,
           "relax_tol": kw.get("relax_tol", 1e-6)}
    recs, base = None, None
    if not a.baseline_only:
        recs, st = serve(reqs, states, mixed=True, chunk=a.chunk, workers=a.workers, verbose=not a.quiet)
        out["mixed_batched"] = st
        if not a.quiet:
            print(json.dumps({"mixed_batched": st}), flush=True)
    if a.baseline or a.baseline_only:
        base, sb = serve_one_at_a_time(reqs, states, verbose=not a.quiet)
        out["one_at_a_time_fp64"] = sb
        if recs is not None:
            out["mixed_vs_fp64"] = compare(recs, base, out["relax_tol"])
    print(json.dumps({k: v for k, v in out.items()}), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(dict(out, records=recs, records_fp64=base), f, indent=1)
    return out


if __name__ == "__main__":
    main()
