"""Batched gradient flows (members with their own Tau, Gamma and limits, one
launch per pass for all members, graph-replayed steps) and the FFT pass
shapes / tuner.

A member runs the same kernels on the same data as a single-field
get_min_u_F run (LPsrc.cpp:40-117), so its history and field must equal the
single run's exactly, whether it runs alone or beside others, and whenever
it was loaded.  Every shape of a grid size is checked against the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _single(spm, torch, N, prec, phi, Tau, Gamma, max_iter, gn_tol=0.0, res_tol=1e-16, shape=0):
    with spm.Grid(N, prec, kscale=2 * np.pi, Tau=Tau, Gamma=Gamma, shape=shape) as g:
        t = torch.as_tensor(phi).to(g.dtype).cuda()
        it, hist = g.gradflow(t, max_iter, res_tol, gn_tol, check_every=7)
        return it, hist, t.cpu().numpy()


MEMBERS = [  # (seed kind, Tau, Gamma, max_iter, gn_tol)
    ("bcc", -0.02, 0.20, 23, 0.0),
    ("fcc", 0.01, 0.40, 40, 0.0),
    ("ico", -0.0061, 0.30, 35, 0.0),
]


def _seed(seeds, N, kind, s):
    if kind == "ico":
        return seeds.icosahedral_nucleus(N, 26.0, 2 * np.pi, 0.1, s, radius=8.0)
    return seeds.crystal(N, kind, 26.0, 2 * np.pi, 0.1, s)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_members_equal_single_runs(spm, seeds, torch, prec):
    N = 32
    phis = [_seed(seeds, N, k, i) for i, (k, *_) in enumerate(MEMBERS)]
    with spm.Batch(N, prec, capacity=4, hist_cap=64, kscale=2 * np.pi) as b:
        ts = [torch.as_tensor(p).to(b.dtype).cuda() for p in phis]
        for m, ((_, Tau, Gamma, mi, gt), t) in enumerate(zip(MEMBERS, ts)):
            b.load(m, t, Tau, Gamma, mi, 1e-16, gt)
        b.steps(37)  # two graph replays of 16 steps and 5 single steps
        running, it, st = b.poll()
        assert running == 1 and list(st[:3]) == [1, 0, 1] and st[3] == 1
        b.steps(16)
        running, it, st = b.poll()
        assert running == 0 and list(it[:3]) == [23, 40, 35]
        got = [b.read(m) for m in range(3)]
    for m, (_, Tau, Gamma, mi, gt) in enumerate(MEMBERS):
        it1, h1, f1 = _single(spm, torch, N, prec, phis[m], Tau, Gamma, mi, gt)
        phi_b, h_b = got[m]
        assert it1 == mi and len(h_b) == mi
        assert np.array_equal(h_b, h1), "history of member %d differs from the single run" % m
        assert np.array_equal(phi_b.cpu().numpy(), f1), "field of member %d differs" % m


def test_batch_continuous_admission(spm, seeds, torch):
    """Members loaded into free slots while others run, and a slot reused,
    give the same results as each run alone."""
    N, prec = 32, 1
    jobs = [("bcc", -0.01, 0.25, 30), ("lam", 0.005, 0.35, 20), ("fcc", 0.0, 0.30, 25), ("ico", -0.0061, 0.3, 18)]
    phis = [_seed(seeds, N, k, 10 + i) for i, (k, *_) in enumerate(jobs)]
    results = {}
    with spm.Batch(N, prec, capacity=2, hist_cap=64, kscale=2 * np.pi) as b:
        ts = [torch.as_tensor(p).float().cuda() for p in phis]
        b.load(0, ts[0], *jobs[0][1:3], jobs[0][3])
        b.steps(5)
        b.load(1, ts[1], *jobs[1][1:3], jobs[1][3])  # joins a running batch
        slot_job = {0: 0, 1: 1}
        pending = [2, 3]
        while slot_job:
            b.steps(3)
            running, it, st = b.poll()
            for slot in list(slot_job):
                if st[slot]:
                    j = slot_job.pop(slot)
                    results[j] = b.read(slot)
                    b.release(slot)
                    if pending:
                        jn = pending.pop(0)
                        b.load(slot, ts[jn], *jobs[jn][1:3], jobs[jn][3])
                        slot_job[slot] = jn
    assert sorted(results) == [0, 1, 2, 3]
    for j, (_, Tau, Gamma, mi) in enumerate(jobs):
        it1, h1, f1 = _single(spm, torch, N, prec, phis[j], Tau, Gamma, mi)
        phi_b, h_b = results[j]
        assert np.array_equal(h_b, h1) and np.array_equal(phi_b.cpu().numpy(), f1), "job %d" % j


def test_batch_errors(spm, torch):
    with spm.Batch(16, 0, capacity=2, hist_cap=8) as b:
        t = torch.zeros((16,) * 3, dtype=torch.float64, device="cuda")
        with pytest.raises(spm.SpmError):
            b.load(2, t, -0.0061, 0.3, 5)  # slot out of range
        with pytest.raises(spm.SpmError):
            b.load(0, t, -0.0061, 0.3, 9)  # beyond hist_cap
        b.load(0, t, -0.0061, 0.3, 8, -1.0)
        with pytest.raises(spm.SpmError):
            b.read(0)  # still running
        b.steps(8)
        assert b.poll()[0] == 0
        _, h = b.read(0)
        assert len(h) == 8


@pytest.mark.parametrize("N", [64, 96, 128])
def test_every_shape_matches_oracle(spm, seeds, po, torch, N):
    """Each FFT pass shape of N (different radices and tiles) against the oracle,
    in the axis-by-axis step the shapes define (step_passes = 4; at N = 64 and
    128 the default step is the two-pass one, tested in test_gpu_two_pass.py)."""
    steps = 6
    phi = seeds.crystal(N, "bcc", 26.0, 2 * np.pi, 0.1, 3)
    o = po.Oracle(N, kscale=2 * np.pi)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    it, E, _, _ = o.get_min_u_F(uR, uF, steps, -1.0)
    assert it == steps
    shapes = spm.shapes(N)
    assert len(shapes) >= 2
    for k in range(len(shapes)):
        with spm.Grid(N, 0, kscale=2 * np.pi, shape=k, step_passes=4) as g:
            t = torch.as_tensor(phi).cuda()
            itg, hist = g.gradflow(t, steps, -1.0, 0.0, check_every=steps)
            out = t.cpu().numpy()
        assert itg == steps
        assert np.all(np.abs(hist[:, 0] - E) <= 1e-12 * np.abs(E)), "shape %s energy" % (shapes[k],)
        assert np.abs(out - uR.real).max() <= 1e-10, "shape %s field" % (shapes[k],)


def test_tuner_records_and_resolves(spm, torch, tmp_path, monkeypatch):
    cache = tmp_path / "shapes.txt"
    monkeypatch.setenv("SPM_TUNE_CACHE", str(cache))
    best, us = spm.tune(64, spm.FP32, nmem=3, steps=4)
    assert 0 <= best < len(spm.shapes(64)) and len(us) == len(spm.shapes(64)) and min(us) > 0
    assert us[best] == min(us)
    line = cache.read_text().strip().splitlines()[-1].split()
    assert line[1:5] == ["64", "1", "4", str(best)]  # N, precision, members (bucketed), shape
    # shape -1 resolves from the cache (the recorded key is read back)
    with spm.Batch(64, spm.FP32, capacity=3, hist_cap=4, shape=-1) as b:
        assert b.capacity == 3
    assert len(cache.read_text().strip().splitlines()) == 1  # no second measurement


def test_step_tuner_records_and_resolves(spm, torch, tmp_path, monkeypatch):
    """spm_tune_step times every two-pass step variant, records the fastest
    under <arch>/sj, and step_variant -1 resolves from that record."""
    cache = tmp_path / "shapes.txt"
    monkeypatch.setenv("SPM_TUNE_CACHE", str(cache))
    best, us = spm.tune_step(64, spm.FP32, nmem=2, steps=4)
    assert 0 <= best < spm.STEP_VARIANTS and len(us) == spm.STEP_VARIANTS and min(us) > 0
    assert us[best] == min(us)
    line = cache.read_text().strip().splitlines()[-1].split()
    assert line[0].endswith("/sj")
    assert line[1:5] == ["64", "1", "2", str(best)]  # N, precision, members (bucketed), flags
    with spm.Batch(64, spm.FP32, capacity=2, hist_cap=4, step_variant=-1) as b:
        assert b.capacity == 2
    assert len(cache.read_text().strip().splitlines()) == 1  # no second measurement


def test_step_variant_rejected(spm, torch):
    for bad in (3, 8 | 6, 8 | 7, 16, -2):
        with pytest.raises(spm.SpmError):
            spm.Grid(64, spm.FP64, step_variant=bad)


@pytest.fixture(scope="module")
def sweep():
    import importlib

    return importlib.import_module("spm-iqc_amd.sweep")


def test_sweep_fp64_batched_equals_one_at_a_time(sweep):
    jobs = sweep.make_jobs(6, sizes=(32, 48), seed=1)
    res, st = sweep.run(jobs, mixed=False, gn_tol=1e-8, max_iter=300, chunk=32, shape=0)
    ref, _ = sweep.run_one_at_a_time(jobs, gn_tol=1e-8, max_iter=300)
    assert st["member_steps"] == sum(r["iters"] for r in ref)
    for r, q in zip(res, ref):
        assert r["iters_fp64"] == q["iters"] and r["E"] == q["E"] and r["gn"] == q["gn"], (r, q)


def test_sweep_mixed_converges_to_fp64_minimum(sweep):
    """FP32 batch to ||grad F|| <= 1e-7, FP64 polish batch to 1e-8: the same
    minima as FP64 alone (disordered-phase parameters, where the relaxation
    converges within the budget)."""
    jobs = [j for j in sweep.make_jobs(16, sizes=(32,), seed=2) if j["Tau"] <= -0.015][:4]
    assert len(jobs) == 4
    res, st = sweep.run(jobs, mixed=True, gn_tol=1e-8, fp32_tol=1e-7, max_iter=6000, chunk=64, shape=0)
    ref, _ = sweep.run_one_at_a_time(jobs, gn_tol=1e-8, max_iter=6000)
    for r, q in zip(res, ref):
        assert r["iters_fp32"] > 0 and r["gn"] <= 1e-8 and q["gn"] <= 1e-8, (r, q)
        assert abs(r["E"] - q["E"]) <= 1e-12, (r, q)


@pytest.mark.parametrize("kind", ["bcc", "fcc", "lam", "nucleus"])
def test_device_seeds_match_host_seeds(seeds, torch, kind):
    N, ks = 32, 2 * np.pi
    dev = seeds.device_field(torch, N, kind, 26.0, ks, 0.1, 5).cpu().numpy()
    host = seeds.icosahedral_nucleus(N, 26.0, ks, 0.1, 5) if kind == "nucleus" else \
        seeds.crystal(N, kind, 26.0, ks, 0.1, 5)
    assert np.abs(dev - host).max() <= 1e-12
