"""HIP path vs the oracle and the reference's golden vectors, through the
C ABI (marked gpu). Tolerances: BEM field 1e-11 relative to the largest
component at each point (FP64 sums of thousands of ring terms, with strong
cancellation next to the surface, in a different order and with ~1 ulp
K/E that differ from SciPy's); pair forces
1e-13; integrator runs: identical attempt/accept counts per group,
identical screen-event identities and order, final state 1e-9 relative."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib

    _lib.load()  # fails loudly if the HIP library is missing
    return torch


def _rel_rows(got, ref):
    scale = np.abs(ref).max(axis=-1, keepdims=True)
    scale = np.where(scale > 0, scale, 1.0)
    return np.max(np.abs(got - ref) / scale)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("name", ["bem_sphere.npz", "bem_post80.npz"])
def test_bem_field_matches_golden(torch_cuda, name):
    from bem_amd.bem.solver import AxisymmetricBEMSolution

    _say("bem field", name)
    d = golden(name)
    sol = AxisymmetricBEMSolution(d["profile"], d["sigma"])
    f = sol.field(d["points"])
    assert _rel_rows(f, d["pert"]) < 1e-11


@pytest.mark.parametrize("shape", [None, (256, 64), (192, 64), (256, 8), (192, 1)])
def test_bem_field_post400_against_oracle(torch_cuda, shape):
    """Every gather layout: the automatic shape, 64 points per block (blocks
    of 64, 64 and 22 points: wave-major and slice-major side by side), 8
    points per block and one point per block (sub-panel striding)."""
    from bem_amd.bem.solver import AxisymmetricBEMSolution
    from oracle import bem as ob

    _say("bem field post400", shape)
    d = golden("bem_post400.npz")
    rng = np.random.default_rng(7)
    # points near and on the cap, along the wall and far away
    th = np.deg2rad(rng.uniform(0, 90, 146))
    r = 1e-6 * (1 + 10 ** rng.uniform(-5, 0.5, 146))
    phi = rng.uniform(0, 2 * np.pi, 146)
    pts = np.stack([r * np.sin(th) * np.cos(phi), r * np.sin(th) * np.sin(phi), 10e-6 + r * np.cos(th)], -1)
    pts = np.concatenate([pts, [[1e-6, 0, 5e-6], [1.00001e-6, 0, 3e-6], [0, 0, 11e-6], [2e-6, 1e-6, 30e-6]]])
    sol = AxisymmetricBEMSolution(d["profile"], d["sigma"])
    ctx = sol._upload(8)
    if shape:
        ctx.check(ctx.lib.spt_set_field_shape(ctx.h, *shape), "spt_set_field_shape")
    try:
        got = sol.field(pts)
    finally:
        ctx.lib.spt_set_field_shape(ctx.h, 0, 0)
    ref = ob.evaluate_field(d["profile"], d["sigma"], pts)
    assert _rel_rows(got, ref) < 1e-11


def test_wave_major_trace_matches_oracle(torch_cuda):
    """A lockstep BEM trace whose field blocks hold 48 particles (64 points
    per block: the wave-major gather, scalar node loads) against the oracle
    run of the same 48 cap electrons in 3 groups of 16 on the 80-panel post:
    identical attempt/accept counts per group, state 1e-9."""
    from bem_amd import SpecificParticleTracer
    from bem_amd.bem.geometry import PostTipBEMGeometry
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup
    from oracle import tracker as otr

    _say("wave-major trace")
    p = golden("bem_post80.npz")
    R, H = float(p["R"]), float(p["H"])
    src = cap_source(48, R, H, seed=5)
    pg = ParticleGroup(dict(src, weight=np.full(48, 1.602176634e-19), status=np.ones(48), species="electron"))
    geo = PostTipBEMGeometry(-1e8, R, H, n_segments=80, sigma=p["sigma"])
    screens, t_max = [12e-6, 20e-6], 1.5e-12
    tr = SpecificParticleTracer(pg, 16, geo, screens, rtol=1e-9, atol=1e-12, t_max=t_max, mode="lockstep")
    tr.field_shape = (256, 64)
    tr.run()
    out = tr.last_run
    st = otr.setup(*(src[k] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), 16)
    o = otr.run(st, dict(kind="post_bem", E_gun=-1e8, R=R, H=H, profile=p["profile"], sigma=p["sigma"],
                         z_handoff=20 * (R + H), kill_z_below=0.0), screens, t_max, rtol=1e-9, atol=1e-12)
    _say(f"wave-major trace: attempts {out['attempts'].tolist()} vs oracle {o['attempts'].tolist()}")
    np.testing.assert_array_equal(out["attempts"], o["attempts"])
    np.testing.assert_array_equal(out["accepted"], o["accepted"])
    pos = out["pos"].reshape(3, 16, 3)
    np.testing.assert_allclose(pos, o["pos"], rtol=1e-9, atol=1e-9 * np.abs(o["pos"]).max())


def test_geometry_fields(torch_cuda):
    from bem_amd import FlatCathode, HemisphericalTip
    from bem_amd.bem.geometry import HemisphericalTipBEMGeometry
    from oracle import bem as ob
    from oracle import fields as of

    _say("geometry fields")
    rng = np.random.default_rng(0)
    pts = rng.normal(0, 1e-6, (200, 3))
    np.testing.assert_array_equal(FlatCathode(-1e8).field(pts), of.gun_field(pts, -1e8))
    np.testing.assert_allclose(HemisphericalTip(-1e8, 0.5e-6).field(pts), of.tip_field(pts, -1e8, 0.5e-6),
                               rtol=1e-15, atol=1e-6)
    d = golden("bem_sphere.npz")
    g = HemisphericalTipBEMGeometry(float(d["Ez"]), float(d["R"]), sigma=d["sigma"])
    got = g.field(d["gpoints"])
    assert _rel_rows(got, d["gfield"]) < 1e-11


def test_kill_masks(torch_cuda):
    from bem_amd import FlatCathode, HemisphericalTip
    from oracle import tracker as otr

    _say("kill masks")
    rng = np.random.default_rng(1)
    pts = rng.normal(0, 1e-6, (500, 3))
    g = FlatCathode(-1e8)
    np.testing.assert_array_equal(g.kill_mask(pts), otr.kill_mask(dict(kind="flat", kill_z_below=0.0), pts))
    t = HemisphericalTip(-1e8, 1e-6)
    np.testing.assert_array_equal(t.kill_mask(pts), otr.kill_mask(dict(kind="tip", R=1e-6, kill_z_below=0.0), pts))


def test_pair_forces(torch_cuda):
    from bem_amd import FlatCathode, HemisphericalTip
    from bem_amd.geometry import coulomb_force
    from oracle import forces as ofo

    _say("pair forces")
    rng = np.random.default_rng(2)
    G, E = 5, 24
    pos = rng.normal(0, 50e-9, (G, E, 3))
    pos[..., 2] = np.abs(pos[..., 2]) + 60e-9
    q = np.full((G, E), -1.602176634e-19)
    act = rng.random((G, E)) > 0.2
    np.testing.assert_allclose(coulomb_force(pos, q, act, 1e-12), ofo.coulomb(pos, q, act, 1e-12), rtol=1e-13,
                               atol=1e-13 * np.abs(ofo.coulomb(pos, q, act, 1e-12)).max())
    ref = ofo.flat_image(pos, q, act, 3e-9, 1e-12)
    np.testing.assert_allclose(FlatCathode(-1e8).image_force(pos, q, act, 1e-12), ref, rtol=1e-13,
                               atol=1e-13 * np.abs(ref).max())
    ref = ofo.tip_image(pos, q, act, 50e-9 - 3e-9, 1e-12, 3e-9)
    np.testing.assert_allclose(HemisphericalTip(-1e8, 50e-9).image_force(pos, q, act, 1e-12), ref, rtol=1e-12,
                               atol=1e-13 * np.abs(ref).max())


def test_bem_assembly_and_solve(torch_cuda):
    from bem_amd.bem.solver import assemble, solve_sigma
    from oracle import bem as ob

    _say("bem assembly and solve")
    d = golden("bem_post80.npz")
    A = assemble(d["profile"])
    Ar = ob.assemble(d["profile"])
    np.testing.assert_allclose(A, Ar, rtol=1e-12, atol=1e-13 * np.abs(Ar).max())
    s = solve_sigma(d["profile"], float(d["Ez"]))
    np.testing.assert_allclose(s, d["sigma"], rtol=1e-10, atol=1e-11 * np.abs(d["sigma"]).max())
    d = golden("bem_post400.npz")
    s = solve_sigma(d["profile"], float(d["Ez"]))
    np.testing.assert_allclose(s, d["sigma"], rtol=1e-10, atol=1e-11 * np.abs(d["sigma"]).max())


def _tracker_for(name, block_threads=0, mode="auto"):
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.bem.geometry import HemisphericalTipBEMGeometry, PostTipBEMGeometry
    from bem_amd.particles import ParticleGroup
    from oracle import bem as ob

    d = golden(f"run_{name}.npz")
    src = {k[3:]: d[k] for k in d.files if k.startswith("in_")}
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    if name.startswith("c1"):
        geo = FlatCathode(-1e7 if name == "c1" else -1e8, z0=3e-9)
    elif name == "tipbem":
        R = 1e-6
        prof = ob.sphere_profile(R, ob.n_for_length(R * np.pi, R * np.pi / 39))
        geo = HemisphericalTipBEMGeometry(-1e8, R, sigma=ob.solve(prof, lambda r, z: -1e8 * z))
    elif name == "well":
        from bem_amd.bem.geometry import CylindricalWellBEMGeometry

        w = golden("bem_well.npz")
        geo = CylindricalWellBEMGeometry(float(w["Ez"]), float(w["R"]), float(w["H"]),
                                         field_max_length=float(w["sharp_max_length"]), sigma=w["sharp_sigma"])
    else:
        p = golden("bem_post80.npz")
        geo = PostTipBEMGeometry(-1e8, float(p["R"]), float(p["H"]), n_segments=80, sigma=p["sigma"])
    tr = SpecificParticleTracer(pg, int(d["n_emit"]), geo, list(d["screens"]), rtol=float(d["rtol"]),
                                atol=float(d["atol"]), t_max=float(d["t_max"]), block_threads=block_threads,
                                mode=mode)
    return d, tr


@pytest.mark.parametrize("mode", ["persistent", "lockstep"])
@pytest.mark.parametrize("name", ["c1small", "tipbem", "post80", "c1", "well"])
def test_trace_matches_reference_runs(torch_cuda, name, mode):
    _say("trace", name, mode)
    d, tr = _tracker_for(name, mode=mode)
    screens, _ = tr.run()
    out = tr.last_run
    G = tr.n_groups
    np.testing.assert_array_equal(out["attempts"], d["attempts"])
    np.testing.assert_array_equal(out["accepted"], d["accepted"])
    pos = out["pos"].reshape(G, -1, 3)
    vel = out["vel"].reshape(G, -1, 3)
    np.testing.assert_allclose(pos, d["pos"], rtol=1e-9, atol=1e-9 * np.abs(d["pos"]).max())
    np.testing.assert_allclose(vel, d["vel"], rtol=1e-9, atol=1e-9 * np.abs(d["vel"]).max())
    for si, s in enumerate(screens):
        np.testing.assert_array_equal(s.id, d[f"s{si}_id"])
        np.testing.assert_allclose(s.t, d[f"s{si}_t"], rtol=1e-11)
        for f in ("x", "y", "z", "px", "py", "pz"):
            ref = d[f"s{si}_{f}"]
            np.testing.assert_allclose(getattr(s, f), ref, rtol=1e-9, atol=1e-9 * max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("mode", ["persistent", "lockstep"])
def test_trace_output_times_match_reference_run(torch_cuda, mode):
    """t_out snapshots (reference tracker.py t_out, trajectories.py): the
    steps land on every output time, also one equal to a group's start
    time; records in the reference's order, identical step counts."""
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    _say("trace output times", mode)
    d = golden("run_c1out.npz")
    src = {k[3:]: d[k] for k in d.files if k.startswith("in_")}
    n = len(src["x"])
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, int(d["n_emit"]), FlatCathode(-1e8, z0=3e-9), list(d["screens"]),
                                t_out=list(d["t_out"]), rtol=float(d["rtol"]), atol=float(d["atol"]),
                                t_max=float(d["t_max"]), mode=mode)
    screens, traj = tr.run()
    np.testing.assert_array_equal(tr.last_run["attempts"], d["attempts"])
    np.testing.assert_array_equal(tr.last_run["accepted"], d["accepted"])
    assert len(traj) == len(d["t_out"])
    for k, s in enumerate(traj):
        np.testing.assert_array_equal(s.id, d[f"o{k}_id"])
        if len(s.id):
            np.testing.assert_allclose(s.t, d[f"o{k}_t"], rtol=1e-15)
            for f in ("x", "y", "z", "px", "py", "pz"):
                ref = d[f"o{k}_{f}"]
                np.testing.assert_allclose(getattr(s, f), ref, rtol=1e-9, atol=1e-9 * max(np.abs(ref).max(), 1e-300))


def test_trace_block_sizes_agree(torch_cuda):
    """Workgroup size changes only the field partial-sum order: results stay
    within 1e-12 and step counts are unchanged."""
    _say("trace block sizes")
    outs = []
    for T in (64, 256, 1024):
        d, tr = _tracker_for("post80", block_threads=T, mode="persistent")
        tr.run()
        outs.append(tr.last_run)
    for o in outs[1:]:
        np.testing.assert_array_equal(o["attempts"], outs[0]["attempts"])
        np.testing.assert_allclose(o["pos"], outs[0]["pos"], rtol=1e-11, atol=1e-11 * np.abs(outs[0]["pos"]).max())


@pytest.mark.parametrize("mode", ["persistent", "lockstep"])
def test_uneven_groups_match_per_group_oracle(torch_cuda, mode):
    """CSR groups of sizes 1..7: each group equals an oracle run of that
    group alone (groups never interact; RMS normalised by 6 * size)."""
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.emission import flat_gaussian_source
    from bem_amd.particles import ParticleGroup
    from oracle import tracker as otr

    _say("uneven groups")
    sizes = [3, 1, 7, 2, 5]
    n = sum(sizes)
    src = flat_gaussian_source(n, seed=11)
    src["t"] = np.linspace(0, 4e-14, n)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, None, FlatCathode(-1e8, z0=3e-9), [0.3e-6, 1e-6], rtol=1e-8, t_max=1.2e-12,
                                group_sizes=sizes, mode=mode)
    screens, _ = tr.run()
    out = tr.last_run
    start = 0
    for g, s in enumerate(sizes):
        sl = slice(start, start + s)
        st = otr.setup(*(src[k][sl] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), s)
        o = otr.run(st, dict(kind="flat", E_gun=-1e8, z0=3e-9, kill_z_below=0.0), [0.3e-6, 1e-6], 1.2e-12,
                    rtol=1e-8)
        assert out["attempts"][g] == o["attempts"][0]
        assert out["accepted"][g] == o["accepted"][0]
        np.testing.assert_allclose(out["pos"][sl], o["pos"][0], rtol=1e-9, atol=1e-15)
        start += s
