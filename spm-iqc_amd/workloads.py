"""Fixed 256^3 workloads of the path, for profilers and quick timing: one
command per kernel family, printing its own timing.

    python spm-iqc_amd/workloads.py gradflow [--precision fp64|fp32] [--steps 50] [--shape -1] [--step-passes 0]
    python spm-iqc_amd/workloads.py lobpcg   [--precision ...] [--iters 5]
    python spm-iqc_amd/workloads.py spm      [--precision ...] [--iters 20]
    python spm-iqc_amd/workloads.py hisd     [--precision ...] [--iters 20]
    python spm-iqc_amd/workloads.py variants [--precision ...] [--steps 50]

gradflow: get_min_u_F iterations on the icosahedral nucleus seed (LPsrc.cpp:40-117).
lobpcg:   hiosd_lobpcg_inip iterations, block 8, on the same field (LOBPCG.cpp:207-487).
spm:      Spring_Pairing iterations from phi -/+ 0.01 v (SPM.cpp:170-263).
hisd:     hiosd_lobpcg_sieh steps, k = 1, from phi + 0.01 v (HISD.cpp:13-30, 86-485).
variants: the two-pass step's variants timed by spm_tune_step (graph-replayed steps of a
          noise spectrum), the fastest recorded in the tuning cache.
"""
from __future__ import annotations

import argparse
import importlib
import os
import sys
import time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
spm = importlib.import_module("spm-iqc_amd")
seeds = importlib.import_module("spm-iqc_amd.seeds")


def _field(torch, g, N):
    # icosahedral 30-vector star, |q| = 1 under k = n / L, amplitude 0.1, sphere of radius 6
    return seeds.device_field(torch, N, "nucleus", 26.0, 1.0, 0.1, 0, dtype=g.dtype)


def _shape(N, prec, shape):
    if shape >= 0:
        return shape
    best, _ = spm.tune(N, prec, nmem=1, steps=30)
    return best


def gradflow(N, prec, steps, shape, passes=0):
    torch = spm._torch()
    shape = _shape(N, prec, shape)
    with spm.Grid(N, prec, shape=shape, step_passes=passes) as g:
        phi = _field(torch, g, N)
        g.gf_begin(phi, steps + 33, res_tol=-1.0, gn_tol=0.0)
        g.gf_steps(32)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.gf_steps(steps)
        e1.record()
        torch.cuda.synchronize()
        it, _ = g.gf_end(phi, steps + 33)
        us = e0.elapsed_time(e1) / steps * 1e3
    print("gradflow N %d prec %d shape %s: %d steps, %.1f us/step" % (N, prec, spm.shapes(N)[shape], steps, us))


def lobpcg(N, prec, iters, shape, passes=0):
    torch = spm._torch()
    shape = _shape(N, prec, shape)
    with spm.Grid(N, prec, shape=shape, step_passes=passes) as g:
        phi = _field(torch, g, N)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(0)
        start = (torch.rand((8, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 2 - 1).to(g.dtype)
        V = g.lobpcg_init(start)
        del start
        g.lobpcg(phi, V, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ritz = g.lobpcg(phi, V, iters)
        dt = time.perf_counter() - t0
    print("lobpcg N %d prec %d: %d iterations, %.2f ms/iteration, lowest %s" % (N, prec, iters, dt / iters * 1e3,
                                                                              ritz[:4]))


def saddle(N, prec, iters, shape, passes=0):
    torch = spm._torch()
    shape = _shape(N, prec, shape)
    with spm.Grid(N, prec, shape=shape, step_passes=passes) as g:
        phi = _field(torch, g, N)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(1)
        v = torch.rand((N, N, N), generator=gen, dtype=torch.float64, device="cuda").to(g.dtype) - 0.5
        v /= g.nrmp(v)
        p0, p1 = phi - 0.01 * v, phi + 0.01 * v
        g.spring_pairing(p0, p1, 1, 0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, res, _ = g.spring_pairing(p0, p1, iters, 0.0)
        dt = time.perf_counter() - t0
    print("spm N %d prec %d: %d iterations, %.2f ms/iteration, res %.3e" % (N, prec, it, dt / max(it, 1) * 1e3, res))


def hisd(N, prec, iters, shape, passes=0):
    torch = spm._torch()
    shape = _shape(N, prec, shape)
    with spm.Grid(N, prec, shape=shape, step_passes=passes) as g:
        phi = _field(torch, g, N)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(2)
        start = (torch.rand((1, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 2 - 1).to(g.dtype)
        V = g.lobpcg_init(start)
        x = phi + 0.01 * V[0]
        g.hisd(x, V, epsf=0.0, maxiter=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, res = g.hisd(x, V, epsf=0.0, maxiter=iters)
        dt = time.perf_counter() - t0
    print("hisd N %d prec %d: %d steps, %.2f ms/step, res %.3e" % (N, prec, it, dt / max(it, 1) * 1e3, res))


def variants(N, prec, steps):
    best, us = spm.tune_step(N, prec, nmem=1, steps=steps)
    names = []
    for f in range(spm.STEP_VARIANTS):
        u = {0: "U hint by size", spm.STEP_U_CACHED: "U cached", spm.STEP_U_STREAM: "U streamed"}[f & 6]
        names.append(("U gathered" if f & spm.STEP_GATHER_U else "U preloaded") + ", " + u)
    for f in range(spm.STEP_VARIANTS):
        print("variants N %d prec %d flags %d (%s): %.1f us/step%s" % (N, prec, f, names[f], us[f],
                                                                     "  <- fastest" if f == best else ""))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("workload", choices=("gradflow", "lobpcg", "spm", "hisd", "variants"))
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--precision", choices=("fp64", "fp32", "both"), default="both")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--shape", default="-1", help="FFT pass shape index, or FP64,FP32 indices; -1: the tuner's choice")
    ap.add_argument("--step-passes", type=int, default=0, choices=(0, 2, 4),
                    help="array passes per step and transform round trip: 0 = the library default")
    a = ap.parse_args(argv)
    precs = {"fp64": [spm.FP64], "fp32": [spm.FP32], "both": [spm.FP64, spm.FP32]}[a.precision]
    want = [int(x) for x in a.shape.split(",")]
    want = want * 2 if len(want) == 1 else want
    for p in precs:
        if a.workload == "gradflow":
            gradflow(a.N, p, a.steps, want[p], a.step_passes)
        elif a.workload == "lobpcg":
            lobpcg(a.N, p, a.iters or 5, want[p], a.step_passes)
        elif a.workload == "variants":
            variants(a.N, p, a.steps)
        elif a.workload == "spm":
            saddle(a.N, p, a.iters or 20, want[p], a.step_passes)
        else:
            hisd(a.N, p, a.iters or 20, want[p], a.step_passes)


if __name__ == "__main__":
    main()
