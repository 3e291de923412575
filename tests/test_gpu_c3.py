"""C3's shape under test on the GPU: the cylindrical well (R 50 nm, H 30 nm)
with uneven CSR groups whose sizes span 1..256 and per-group emission
offsets, in the 64/128/256-thread instantiations of both trace kernels,
with and without the image-charge BEM (z0 3 nm, both kernels); and the
2000-panel well of the benchmark's C3 sample.

Each compared group equals an oracle run of that group alone: the
reference's groups never interact (reference batched_rk45.py:4-22), and
its RMS error norm runs over 6 x the group's size (reference
batched_rk45.py:387), so a group's run does not depend on the batch it
sits in. Identical attempt/accept counts, state within 1e-9 relative at
the device's end time (a run stopped after a fixed number of attempts ends
at a time that moves with roundoff-level step-size differences:
tests/trajectory.py).
Runs are capped at a few attempts (reference image_force is "not chunked"
and the oracle takes ~30 s per 6 attempts of a 256-member group with the
image BEM: reference bem/image_charge.py:942-958)."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import _lib

    _lib.load()
    return torch


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _well_image():
    """Device geometry and the oracle's dict for the reference's well with
    image charges (golden bem_well_image.npz: its sigma and operators)."""
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry

    w = golden("bem_well_image.npz")
    geo = CylindricalWellBEMGeometry(float(w["Ez"]), float(w["R"]), float(w["H"]), z0=float(w["z0"]),
                                     field_max_length=3e-9, image_max_length=float(w["image_max_length"]),
                                     image_n_max=int(w["n_max"]), sigma=w["sigma"])
    og = dict(kind="well_bem", E_gun=float(w["Ez"]), R=float(w["R"]), H=float(w["H"]), profile=w["real_profile"],
              sigma=w["sigma"], z_handoff=float(w["z_handoff"]), bottom_fillet=float(w["bottom_fillet"]),
              rim_fillet=float(w["rim_fillet"]), z0=float(w["z0"]), image_profile=w["image_profile"],
              image_A_inv=np.linalg.inv(w["A"]), n_max=int(w["n_max"]), d_lo=float(w["d_lo"]), d_hi=float(w["d_hi"]))
    return geo, og


def _well_plain():
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry

    w = golden("bem_well.npz")
    geo = CylindricalWellBEMGeometry(float(w["Ez"]), float(w["R"]), float(w["H"]),
                                     field_max_length=float(w["sharp_max_length"]), sigma=w["sharp_sigma"])
    og = dict(kind="well_bem", E_gun=float(w["Ez"]), R=float(w["R"]), H=float(w["H"]), profile=w["sharp_profile"],
              sigma=w["sharp_sigma"], z_handoff=float(w["sharp_z_handoff"]))
    return geo, og


def _compare_groups(geo, og, sizes, compare, screens, t_max, max_steps, seed, **kw):
    """Trace all groups on the device, capped at max_steps attempts; each
    group in `compare` against the oracle run of that group alone."""
    from bem_amd import SpecificParticleTracer
    from bem_amd.emission import well_bottom_source
    from bem_amd.particles import ParticleGroup
    from oracle import tracker as otr
    from trajectory import assert_same_state

    sizes = np.asarray(sizes)
    src = well_bottom_source(sizes, geo.R, geo.H, t_offset_max=5e-15, radius_fraction=0.5, seed=seed)
    n = int(sizes.sum())
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, None, geo, screens, rtol=1e-8, atol=1e-12, t_max=t_max, group_sizes=sizes,
                                max_steps=max_steps, **kw)
    out = tr.trace()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for g in compare:
        sl = slice(int(off[g]), int(off[g + 1]))
        st = otr.setup(*(src[k][sl] for k in ("x", "y", "z", "px", "py", "pz", "t", "id")), int(sizes[g]))
        o = otr.run(st, og, screens, t_max, rtol=1e-8, atol=1e-12, max_steps=max_steps, raise_at_cap=False)
        _say(f"group {g} (size {sizes[g]}): attempts {out['attempts'][g]} / oracle {o['attempts'][0]}, "
             f"accepted {out['accepted'][g]} / {o['accepted'][0]}")
        assert out["attempts"][g] == o["attempts"][0]
        assert out["accepted"][g] == o["accepted"][0]
        # the state at the device's end time (tests/trajectory.py): the
        # oracle's state carried by its own velocity and acceleration
        td, td_ref = out["t_death"][sl], o["t_death"][0]
        np.testing.assert_array_equal(np.isfinite(td), np.isfinite(td_ref))
        dead = np.isfinite(td_ref)
        dt = np.where(dead, td - np.where(dead, td_ref, 0.0), out["t_end"][g] - o["t"][0])
        born = st["t_birth"][0] <= o["t"][0]
        accel = otr.make_accel(og, st["charge"], st["mass_p"], st["q_over_m"], 1e-12)
        acc = accel(o["pos"], o["vel"], born[None], np.array([0]))[0]
        assert_same_state(out["pos"][sl], out["vel"][sl], o["pos"][0], o["vel"][0], acc, dt,
                          np.where(dead, td_ref, o["t"][0]))
        _say(f"  end-time shift {abs(out['t_end'][g] - o['t'][0]) / o['t'][0]:.1e} relative")
    return out


@pytest.mark.parametrize("mode", ["lockstep", "persistent"])
def test_image_bem_groups_up_to_256_match_oracle(gpu, mode):
    """The image-charge BEM, largest group 256 (lockstep: the 256-thread
    instantiation; persistent: 256-thread workgroups that solve their own
    group): groups of 256 and 180 members, and a single electron."""
    geo, og = _well_image()
    sizes = [7, 256, 1, 180, 33, 97, 140, 2]
    _say(f"image BEM [{mode}], sizes", sizes)
    out = _compare_groups(geo, og, sizes, [1, 3, 2], [-25e-9, -15e-9], 8e-14, 6, seed=31, mode=mode)
    assert np.all(out["attempts"] == 6)


def test_image_bem_groups_up_to_128_match_oracle(gpu):
    """The 128-thread instantiation: groups of 128, 97 and 5 members."""
    geo, og = _well_image()
    sizes = [97, 5, 128, 64, 70]
    _say("image BEM, sizes", sizes)
    _compare_groups(geo, og, sizes, [0, 2, 1], [-25e-9, -15e-9], 8e-14, 6, seed=32)


@pytest.mark.parametrize("mode,threads", [("persistent", 256), ("persistent", 128), ("lockstep", 0)])
def test_static_field_groups_match_oracle(gpu, mode, threads):
    """The well's static BEM field and in-group Coulomb forces: the
    persistent kernel at 128 and 256 threads (one group per workgroup) and
    lockstep, on groups of up to 128 / 256 members, 20 attempts."""
    geo, og = _well_plain()
    sizes = [128, 3, 77, 1, 120] if threads == 128 else [256, 3, 150, 1, 90]
    _say(f"static field [{mode}, {threads}], sizes {sizes}")
    _compare_groups(geo, og, sizes, [0, 2, 3], [20e-9, 100e-9], 2.5e-13, 20, seed=33, mode=mode,
                    block_threads=threads)


def test_well_2000_panels_sigma_and_field_match_oracle(gpu):
    """The C3 sample's real surface: 2000 panels (field_max_length 0.04223
    nm, the corners rounded as z0 = 3 nm rounds them). The device solve's
    sigma reproduces the Dirichlet data E_gun z at sampled nodes through
    the oracle's adaptive potential (1e-9 of the largest), and the device
    perturbation field equals the oracle's evaluate_field on that sigma
    (reference bem/axisymmetric.py:501-605) at far, near and on-surface
    points (1e-11 of each point's largest component)."""
    from bem_amd.bem.geometry import CylindricalWellBEMGeometry
    from oracle import bem as ob

    R, H, z0 = 50e-9, 30e-9, 3e-9
    r = min(0.5 * z0, 0.2 * (H - z0))
    geo = CylindricalWellBEMGeometry(-1e8, R, H, field_max_length=0.04223e-9, bottom_fillet_radius=r,
                                     rim_fillet_radius=z0 + r)
    prof, sig = geo.solution.profile, geo.solution.sigma
    assert len(prof) - 1 == 2000
    rng = np.random.default_rng(41)
    k = np.sort(rng.choice(len(prof), 40, replace=False))
    nodes = np.stack([prof[k, 0], np.zeros(len(k)), prof[k, 1]], -1)
    phi = ob.potential(prof, sig, nodes)
    b = -1e8 * prof[k, 1]
    scale = np.abs(-1e8 * prof[:, 1]).max()
    _say(f"2000-panel well: Dirichlet residual {np.abs(phi - b).max() / scale:.2e}")
    assert np.abs(phi - b).max() / scale < 1e-9
    # points: above the plane, inside the well, near the wall and corners, on nodes
    n = 240
    rho = np.concatenate([rng.uniform(0, 3 * R, n // 3), rng.uniform(0, 0.95 * R, n // 3),
                          R - 10 ** rng.uniform(-11, -8.5, n // 3)])
    z = np.concatenate([10 ** rng.uniform(-10, -6.5, n // 3), -H * rng.uniform(0.02, 0.98, n // 3),
                        -H * rng.uniform(0.05, 0.95, n // 3)])
    ph = rng.uniform(0, 2 * np.pi, n)
    pts = np.stack([rho * np.cos(ph), rho * np.sin(ph), z], -1)
    on = prof[rng.integers(1, len(prof) - 1, 16)]
    pts = np.concatenate([pts, np.stack([on[:, 0], np.zeros(16), on[:, 1]], -1)])
    got = geo.solution.field(pts)
    ref = ob.evaluate_field(prof, sig, pts)
    scale = np.abs(ref).max(axis=-1, keepdims=True)
    err = np.max(np.abs(got - ref) / scale)
    _say(f"2000-panel well: field max rel {err:.2e} over {len(pts)} points")
    assert err < 1e-11
