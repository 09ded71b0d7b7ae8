"""The two-pass gradient-flow step and operator round trips of 64^3, 128^3 and 256^3
grids (the j transform split between the row pass and the i pass; the
default at N = 64, 128 and 256) against the oracle, against the axis-by-axis passes, in
batches, and through their own state layout.  Tolerances as for every
get_min_u_F check: FP64 energy 1e-12 relative; FP32 against FP64 1e-6
relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NS = [64, 128, 256]
TWO = dict(step_passes=2)


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev_seed(seeds, torch, N, kind, s, dtype, kscale=1.0):
    return seeds.device_field(torch, N, kind, 26.0, kscale, 0.1, s, dtype=dtype)


@pytest.mark.parametrize("N", NS)
def test_two_pass_gradflow_matches_oracle(spm, po, seeds, torch, N):
    """get_min_u_F (LPsrc.cpp:40-117), FP64: per-step E, ||grad F||, residual
    and the final field against the oracle."""
    print(f"[case] two-pass gradflow {N} vs oracle", flush=True)
    steps = 4
    phi = seeds.icosahedral_nucleus(N, seed=0)
    o = po.Oracle(N)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    it, E, res, gn = o.get_min_u_F(uR, uF, steps, 1e-16, gnorm=True)
    with spm.Grid(N, 0, **TWO) as g:
        pt = torch.as_tensor(phi).cuda()
        itg, hist = g.gradflow(pt, steps, 1e-16, 0.0, check_every=2)
        out = pt.cpu().numpy()
    assert itg == it == steps
    assert np.all(np.abs(hist[:, 0] - E) <= 1e-12 * np.abs(E))
    assert np.all(np.abs(hist[:, 1] - gn) <= 1e-9 * gn + 1e-14)
    assert np.all(np.abs(hist[:, 2] - res) <= 1e-8 * res + 1e-15)
    assert np.abs(out - uR.real).max() <= 1e-10


def _variants(spm):
    # every flag set spm_tune_step chooses from
    return [spm.STEP_EXPLICIT | f for f in range(spm.STEP_VARIANTS)]


@pytest.mark.parametrize("N", [64, 128])
def test_step_variants_match_oracle(spm, po, seeds, torch, N):
    """Every two-pass step variant (state preloaded or gathered, cache hints by
    size, off or on), FP64: 40 iterations through graph replays, per-step E,
    ||grad F||, residual and the final field against the oracle (get_min_u_F,
    LPsrc.cpp:40-117)."""
    steps = 40
    phi = seeds.icosahedral_nucleus(N, seed=0)
    o = po.Oracle(N)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    it, E, res, gn = o.get_min_u_F(uR, uF, steps, 1e-16, gnorm=True)
    for v in _variants(spm):
        print(f"[case] step variant {v} N={N} vs oracle", flush=True)
        with spm.Grid(N, 0, step_variant=v, **TWO) as g:
            pt = torch.as_tensor(phi).cuda()
            itg, hist = g.gradflow(pt, steps, 1e-16, 0.0, check_every=32)
            out = pt.cpu().numpy()
        assert itg == it == steps, v
        assert np.all(np.abs(hist[:, 0] - E) <= 1e-12 * np.abs(E)), v
        assert np.all(np.abs(hist[:, 1] - gn) <= 1e-9 * gn + 1e-14), v
        assert np.all(np.abs(hist[:, 2] - res) <= 1e-8 * res + 1e-15), v
        assert np.abs(out - uR.real).max() <= 1e-10, v


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_step_variants_compute_the_same(spm, seeds, torch, N, prec):
    """The variants differ only in when the state is loaded and in cache hints:
    the history and the field equal the built-in variant's bit for bit."""
    steps = 24
    dt = torch.float64 if prec == 0 else torch.float32
    phi = _dev_seed(seeds, torch, N, "bcc", 5, dt)
    ref = None
    for v in [0] + _variants(spm):
        print(f"[case] step variant {v} N={N} prec={prec}", flush=True)
        with spm.Grid(N, prec, step_variant=v, **TWO) as g:
            t = phi.clone()
            it, hist = g.gradflow(t, steps, -1.0, 0.0, check_every=16)
            got = (it, hist.copy(), t.cpu().numpy())
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0] == steps, v
        assert np.array_equal(got[1], ref[1]), v
        assert np.array_equal(got[2], ref[2]), v


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_two_pass_equals_four_pass(spm, seeds, torch, N, prec):
    """Same iterations as the axis-by-axis step: FP64 to rounding (1e-13),
    FP32 to its own rounding (1e-6 relative energy, 1e-5 field)."""
    print(f"[case] two vs four passes N={N} prec={prec}", flush=True)
    steps = 40
    out = {}
    for passes in (2, 4):
        with spm.Grid(N, prec, kscale=2 * np.pi, Tau=0.005, step_passes=passes) as g:
            t = _dev_seed(seeds, torch, N, "bcc", 3, g.dtype, 2 * np.pi)
            it, hist = g.gradflow(t, steps, -1.0, 0.0, check_every=16)
            out[passes] = (it, hist.copy(), t.double().cpu().numpy())
    (i2, h2, f2), (i4, h4, f4) = out[2], out[4]
    assert i2 == i4 == steps
    tol_e, tol_f = (1e-13, 1e-13) if prec == 0 else (1e-6, 1e-5)
    assert np.all(np.abs(h2[:, 0] - h4[:, 0]) <= tol_e * np.abs(h4[:, 0]))
    assert np.all(np.abs(h2[:, 1] - h4[:, 1]) <= tol_e * 100 * h4[:, 1])
    assert np.abs(f2 - f4).max() <= tol_f * np.abs(f4).max()


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_two_pass_layout_round_trip(spm, seeds, torch, N, prec):
    """gf_begin then gf_end with no step: phi -> (row pass, i pass) -> state
    -> (i pass, row pass) -> phi, through the two-pass state layout."""
    print(f"[case] two-pass round trip N={N} prec={prec}", flush=True)
    with spm.Grid(N, prec, **TWO) as g:
        phi = _dev_seed(seeds, torch, N, "nucleus", 1, g.dtype)
        g.gf_begin(phi, 10, 1e-16, 0.0)
        back = torch.empty_like(phi)
        it, _ = g.gf_end(back, 10)
    assert it == 0
    err = (back - phi).abs().max().item() / phi.abs().max().item()
    assert err < (64 * 2e-15 if prec == 0 else 3e-6)


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_two_pass_batch_equals_single_runs(spm, seeds, torch, N, prec):
    """Batch members (own Tau, Gamma, limits) give exactly the single runs."""
    print(f"[case] two-pass batch N={N} prec={prec}", flush=True)
    members = [("bcc", -0.02, 0.20, 21), ("fcc", 0.01, 0.40, 30)]
    kw = dict(kscale=2 * np.pi, **TWO)
    with spm.Batch(N, prec, capacity=2, hist_cap=64, **kw) as b:
        phis = [_dev_seed(seeds, torch, N, k, i, b.dtype, 2 * np.pi) for i, (k, *_) in enumerate(members)]
        for m, ((_, Tau, Gamma, mi), t) in enumerate(zip(members, phis)):
            b.load(m, t, Tau, Gamma, mi, 1e-16, 0.0)
        b.steps(32)
        running, it, _ = b.poll()
        assert running == 0 and list(it[:2]) == [21, 30]
        got = [b.read(m) for m in range(2)]
    for m, (_, Tau, Gamma, mi) in enumerate(members):
        with spm.Grid(N, prec, Tau=Tau, Gamma=Gamma, **kw) as g:
            t = phis[m].clone()
            it1, h1 = g.gradflow(t, mi, 1e-16, 0.0, check_every=7)
        phi_b, h_b = got[m]
        assert it1 == mi and len(h_b) == mi
        assert np.array_equal(h_b, h1), "history of member %d differs from the single run" % m
        assert torch.equal(phi_b, t), "field of member %d differs" % m


def _rel(a, b):
    return (a - b).abs().max().item() / max(b.abs().max().item(), 1e-300)


@pytest.mark.parametrize("N", NS)
def test_two_pass_ngrad_matches_oracle(spm, po, seeds, torch, N):
    """ngrad_cam (LPsrc.cpp:119-173) through the two-pass round trip, FP64,
    against the oracle (1e-10 max norm, as at the other sizes)."""
    print(f"[case] two-pass ngrad {N} vs oracle", flush=True)
    phi = seeds.icosahedral_nucleus(N, seed=2)
    o = po.Oracle(N)
    g_ref = o.ngrad_cam(phi + 0j).real
    with spm.Grid(N, 0, **TWO) as g:
        gr, nrm = g.ngrad(torch.as_tensor(phi).cuda())
        gr = gr.cpu().numpy()
    assert np.abs(gr - g_ref).max() <= 1e-10
    assert abs(nrm - o.nrmp(g_ref + 0j)) <= 1e-10 * nrm


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_two_pass_operators_equal_four_pass(spm, seeds, torch, N, prec):
    """hv_cam, precond_camnew, ngrad_cam (LOBPCG.cpp:113-199, LPsrc.cpp:119-173)
    and the SPM forces (SPM.cpp:34-168) through the two-pass round trip equal
    the axis-by-axis passes to rounding (FP64 1e-13, FP32 2e-6 relative)."""
    print(f"[case] two-pass operators N={N} prec={prec}", flush=True)
    tol = 1e-13 if prec == 0 else 2e-6
    out = {}
    for passes in (2, 4):
        with spm.Grid(N, prec, step_passes=passes) as g:
            phi = _dev_seed(seeds, torch, N, "nucleus", 4, g.dtype)
            gen = torch.Generator(device="cuda").manual_seed(5)
            v = (torch.rand((N, N, N), generator=gen, dtype=torch.float64, device="cuda") - 0.5).to(g.dtype)
            r = {"hv": g.hv(phi, v).clone(), "precond": g.precond(v.clone())}
            gr, nrm = g.ngrad(phi)
            r["ngrad"], r["nrm"] = gr.clone(), nrm
            x = phi + 0.01 * v
            r["natural"] = g.natural_force(x.clone(), v / g.nrmp(v)).clone()
            r["minmax"] = g.minmax_force(x.clone(), v / g.nrmp(v)).clone()
            out[passes] = r
    for k in ("hv", "precond", "ngrad", "natural", "minmax"):
        assert _rel(out[2][k].double(), out[4][k].double()) < tol, k
    assert abs(out[2]["nrm"] - out[4]["nrm"]) <= tol * out[4]["nrm"]


@pytest.mark.parametrize("N", NS)
def test_two_pass_lobpcg_and_spm_equal_four_pass(spm, seeds, torch, N):
    """hiosd_lobpcg_inip and Spring_Pairing over the two-pass round
    trips follow the axis-by-axis runs (Ritz values 1e-9, SPM residual 1e-9
    relative after 6 iterations)."""
    print(f"[case] two-pass lobpcg/spm N={N}", flush=True)
    out = {}
    for passes in (2, 4):
        with spm.Grid(N, 0, step_passes=passes) as g:
            phi = _dev_seed(seeds, torch, N, "nucleus", 0, g.dtype)
            gen = torch.Generator(device="cuda").manual_seed(0)
            start = torch.rand((4, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 2 - 1
            V = g.lobpcg_init(start)
            del start
            ritz = g.lobpcg(phi, V, 4)
            del V
            v = (torch.rand((N, N, N), generator=gen, dtype=torch.float64, device="cuda") - 0.5)
            v /= g.nrmp(v)
            p0, p1 = phi - 0.01 * v, phi + 0.01 * v
            it, res, _ = g.spring_pairing(p0, p1, 6, 0.0)
            out[passes] = (ritz.copy(), it, res, p0.cpu())
    (r2, i2, s2, f2), (r4, i4, s4, f4) = out[2], out[4]
    assert np.all(np.abs(r2 - r4) <= 1e-9)
    assert i2 == i4 == 6 and abs(s2 - s4) <= 1e-9 * s4
    assert (f2 - f4).abs().max().item() <= 1e-12


@pytest.mark.parametrize("k", [3, 8])
def test_two_pass_lobpcg_block_operators(spm, seeds, torch, k):
    """The two-pass LOBPCG sweep applies H and the preconditioner to a block as
    batch members (the whole block of up to 8 fields per launch of each pass):
    Ritz values within 1e-9 of the axis-by-axis sweep, which goes field by field,
    after 6 iterations at 64^3."""
    N = 64
    print(f"[case] two-pass lobpcg block k={k}", flush=True)
    out = {}
    for passes in (2, 4):
        with spm.Grid(N, 0, step_passes=passes) as g:
            phi = _dev_seed(seeds, torch, N, "nucleus", 0, g.dtype)
            gen = torch.Generator(device="cuda").manual_seed(1)
            start = torch.rand((k, N, N, N), generator=gen, dtype=torch.float64, device="cuda") * 2 - 1
            V = g.lobpcg_init(start)
            out[passes] = g.lobpcg(phi, V, 6).copy()
    r2, r4 = out[2], out[4]
    assert np.all(np.isfinite(r2)) and np.all(np.abs(r2 - r4) <= 1e-9), (r2, r4)
