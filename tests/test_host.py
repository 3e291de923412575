"""Host-side checks that need no GPU: the C ABI library and its exports,
the device K/E coefficients, emission ordering and samplers, profiles,
tracker set-up."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy import special

from conftest import ROOT
from bem_amd import emission
from bem_amd.bem.mesh import n_for_length, post_profile, sphere_profile

HEADER = os.path.join(ROOT, "include", "spt.h")
LIB = os.path.join(ROOT, "bem_amd", "libspt_hip.so")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(spt_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    from bem_amd import _lib

    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libspt_hip.so not built")
    lib = ctypes.CDLL(LIB)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts():
    from bem_amd import _lib

    assert ctypes.sizeof(_lib.Geometry_t) == 88
    assert ctypes.sizeof(_lib.Event_t) == 80
    assert _lib.TraceArgs_t.n_groups.offset == 88
    assert _lib.TraceArgs_t.screens.offset % 8 == 0


def test_struct_layouts_match_the_c_header(tmp_path):
    """sizeof/offsetof of every ABI struct, compiled from include/spt.h by
    the host C compiler, against the ctypes mirrors."""
    import shutil
    import subprocess

    from bem_amd import _lib

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    structs = {"spt_geometry": _lib.Geometry_t, "spt_event": _lib.Event_t, "spt_mlp_cfg": _lib.MlpCfg_t,
               "spt_trace_args": _lib.TraceArgs_t}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "spt.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, ct in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(ct), cname
        for fname, _ in ct._fields_:
            assert got[(cname, fname)] == getattr(ct, fname).offset, (cname, fname)


def _coeffs(name):
    src = open(os.path.join(ROOT, "bem_amd", "csrc", "ellip_coeffs.h")).read()
    return [float(x) for x in re.search(r"SPT_ELLIP_%s \{ (.*) \}" % name, src).group(1).split(",")]


def _fma(a, b, c):
    """fma in 64-bit-mantissa long double, rounded once more to double
    (within 2^-11 ulp of the exactly rounded fma)."""
    L = np.longdouble
    return (L(a) * L(b) + L(c)).astype(np.float64)


def _device_log(x, drop_lo=False, n_terms=7):
    """The device logarithm (spt_device.h log_pos) restated in NumPy, the
    hardware reciprocal estimate modelled by a float32-rounded 1/d.
    drop_lo / n_terms plant mistakes for the test's own check."""
    src = open(os.path.join(ROOT, "bem_amd", "csrc", "log_coeffs.h")).read()
    T = [float(v) for v in re.search(r"SPT_LOG_T \{ (.*) \}", src).group(1).split(",")][:n_terms]
    hi = float(re.search(r"SPT_LOG_LN2_HI (\S+)", src).group(1))
    lo = 0.0 if drop_lo else float(re.search(r"SPT_LOG_LN2_LO (\S+)", src).group(1))
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    low = m < 0.70710678118654752440
    m = np.where(low, m + m, m)
    e = np.where(low, e - 1, e).astype(np.float64)
    f = m - 1.0
    d = 2.0 + f
    with np.errstate(divide="ignore"):
        r = (1.0 / d).astype(np.float32).astype(np.float64)
    r = _fma(r, _fma(-d, r, 1.0), r)
    s = f * r
    s = _fma(r, _fma(-d, s, f), s)
    z = s * s
    t = np.full_like(z, T[-1])
    for c in reversed(T[:-1]):
        t = _fma(t, z, c)
    R = z * t
    hf = 0.5 * (f * f)
    v = _fma(e, hi, -((hf - _fma(s, hf + R, e * lo)) - f))
    return np.where(x == 0.0, -np.inf, v)


def test_device_log_against_numpy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.random(20000), 10 ** rng.uniform(-300, 0, 20000), 1 - 10 ** rng.uniform(-16, 0, 20000),
                        [5e-324, 2.0**-1022, 0.5, 0.70710678118654752440, np.nextafter(1.0, 0.0)]])
    got, ref = _device_log(x), np.log(x)
    nz = ref != 0
    assert np.max(np.abs(got[nz] / ref[nz] - 1)) < 2 * 2.0**-52
    assert _device_log(np.array([1.0]))[0] == 0.0
    assert _device_log(np.array([0.0]))[0] == -np.inf
    # planted slips fail the same bound: ln 2's low part dropped, the
    # polynomial one term short
    for kw in (dict(drop_lo=True), dict(n_terms=6)):
        bad = _device_log(x, **kw)
        assert np.max(np.abs(bad[nz] / ref[nz] - 1)) > 2 * 2.0**-52, kw


def _log_table():
    src = open(os.path.join(ROOT, "bem_amd", "csrc", "log_coeffs.h")).read()

    def arr(name):
        return np.array([float(v) for v in re.search(r"%s \{ (.*) \}" % name, src).group(1).split(",")])

    return arr("SPT_LOGT_INVC"), arr("SPT_LOGT_LOGC"), arr("SPT_LOGT_Q"), float(
        re.search(r"SPT_LOG_LN2 (\S+)", src).group(1))


def _device_log_tab(x, index_shift=0, n_q=4):
    """spt_device.h log_tab restated in NumPy (x > 0). index_shift / n_q
    plant mistakes for the test's own check."""
    inv, logc, Q, ln2 = _log_table()
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    i = ((m.view(np.uint64) >> np.uint64(44)).astype(np.int64) + index_shift) & (len(inv) - 1)
    r = _fma(m, inv[i], -1.0)
    Q = Q[:n_q]
    q = np.full_like(r, Q[-1])
    for c in reversed(Q[:-1]):
        q = _fma(q, r, c)
    return _fma(e.astype(np.float64), ln2, logc[i] + _fma(r * r, q, r))


def test_device_table_log_against_numpy():
    """log_tab's contract: |error| <= 2^-52 max(1, |ln x|), ln 1 = 0, the
    table spanning [1/2, 1) in 256 intervals with |m/c_i - 1| <= 2^-9."""
    inv, logc, Q, ln2 = _log_table()
    assert len(inv) == 256 and ln2 == np.log(2.0)
    lo = 0.5 + np.arange(256) / 512.0
    assert np.all(np.abs(lo * inv - 1) <= 2.0**-9) and np.all(np.abs((lo + 1 / 512.0) * inv - 1) <= 2.0**-9)
    np.testing.assert_allclose(logc, -np.log(inv), rtol=2e-16, atol=1e-19)
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.random(20000), 10 ** rng.uniform(-300, 0, 20000), 1 - 10 ** rng.uniform(-16, 0, 20000),
                        [2.0**-1022, 0.5, np.nextafter(1.0, 0.0)]])
    ref = np.log(x)
    scale = np.maximum(1.0, np.abs(ref))
    assert np.max(np.abs(_device_log_tab(x) - ref) / scale) < 2.0**-52
    assert _device_log_tab(np.array([1.0]))[0] == 0.0
    # planted slips fail the same bound: the neighbouring interval's entry,
    # log1p one term short
    for kw in (dict(index_shift=1), dict(n_q=3)):
        assert np.max(np.abs(_device_log_tab(x, **kw) - ref) / scale) > 2.0**-52, kw


def _device_ke(p, log=None):
    """The device formula (spt_device.h ellip_ke_lp) restated in NumPy, ln p
    by log_pos (default) or the given logarithm."""
    PK, QK = _coeffs("PK"), _coeffs("QK")

    def hd(c, x):
        s = np.full_like(x, c[10])
        d = np.zeros_like(x)
        for a in reversed(c[:10]):
            d = d * x + s
            s = s * x + a
        return s, d

    lp = (log or _device_log)(p)
    P, dP = hd(PK, p)
    Q, dQ = hd(QK, p)
    K = P - lp * Q
    pkp = p * dP - Q - (p * lp) * dQ
    return K, p * K - 2.0 * (1.0 - p) * pkp


@pytest.mark.parametrize("log", ["log_pos", "log_tab"])
def test_device_ellip_coefficients_against_scipy_and_identities(log):
    """K and E with either device logarithm (log_pos: toroidal seeds and
    potentials; log_tab: the field gather's ring kernel)."""
    lg = _device_log if log == "log_pos" else _device_log_tab
    rng = np.random.default_rng(0)
    p = np.concatenate([10 ** rng.uniform(-300, 0, 4000), rng.random(4000), 1 - 10 ** rng.uniform(-16, 0, 2000),
                        [1.0]])
    K, E = _device_ke(p, lg)
    np.testing.assert_allclose(K, special.ellipkm1(p), rtol=4e-16 * 4)
    np.testing.assert_allclose(E, special.ellipe(1 - p), rtol=4e-16 * 4)
    # K(0) = E(0) = pi/2; E(1) = 1 (DLMF 19.6.1)
    K1, E1 = _device_ke(np.array([1.0]), lg)
    assert abs(K1[0] - np.pi / 2) < 4e-16 and abs(E1[0] - np.pi / 2) < 4e-16
    # Legendre relation at m = 1/2 (DLMF 19.7.1): 2 E K - K^2 = pi/2
    Kh, Eh = _device_ke(np.array([0.5]), lg)
    assert abs(2 * Eh[0] * Kh[0] - Kh[0] ** 2 - np.pi / 2) < 2e-15


def test_group_indices_sort_and_errors():
    ids = np.array([5, 3, 9, 1])
    np.testing.assert_array_equal(emission.group_indices(ids, 2), [[3, 1], [0, 2]])
    with pytest.raises(ValueError):
        emission.group_indices(ids, 3)
    with pytest.raises(ValueError):
        emission.group_indices(np.array([1, 1, 2, 3]), 2)
    order, off = emission.group_offsets(ids, [1, 3])
    np.testing.assert_array_equal(order, [3, 1, 0, 2])
    np.testing.assert_array_equal(off, [0, 1, 4])
    with pytest.raises(ValueError):
        emission.group_offsets(ids, [2, 3])


def test_samplers_are_seeded():
    a = emission.flat_gaussian_source(16, seed=0)
    b = emission.flat_gaussian_source(16, seed=0)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    c = emission.cap_source(32, 1e-6, 10e-6, seed=0)
    r = np.sqrt(c["x"] ** 2 + c["y"] ** 2 + (c["z"] - 10e-6) ** 2)
    np.testing.assert_allclose(r, 1e-6, rtol=1e-12)
    th = np.degrees(np.arccos((c["z"] - 10e-6) / 1e-6))
    assert th.max() <= 60.0
    # momentum points outward (cos-weighted about the normal)
    n = np.stack([c["x"], c["y"], c["z"] - 10e-6], -1)
    p = np.stack([c["px"], c["py"], c["pz"]], -1)
    assert np.all(np.sum(n * p, -1) >= 0)


def test_profiles():
    s = sphere_profile(1.0, n_for_length(np.pi, np.pi / 39))
    assert s.shape == (40, 2)
    p = post_profile(1e-6, 10e-6, 400)
    assert p.shape == (401, 2)
    assert p[0, 0] == 0.0 and p[-1, 0] == 0.0
    L = np.linalg.norm(np.diff(p, axis=0), axis=1)
    assert L.max() / L.min() < 1.02
    np.testing.assert_allclose(p[:, 1], -p[::-1, 1], atol=1e-20)   # mirror symmetric
    np.testing.assert_allclose(p[:, 0], p[::-1, 0], atol=1e-20)


def test_tracker_setup_matches_oracle():
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.particles import ParticleGroup
    from oracle import tracker as otr

    src = emission.flat_gaussian_source(12, seed=1)
    src["id"] = np.arange(12)[::-1] + 100
    pg = ParticleGroup(dict(src, weight=np.full(12, 1.602176634e-19), status=np.ones(12), species="electron"))
    tr = SpecificParticleTracer(pg, 4, FlatCathode(-1e8), [1e-6])
    st = otr.setup(src["x"], src["y"], src["z"], src["px"], src["py"], src["pz"], src["t"], src["id"], 4)
    np.testing.assert_array_equal(tr.pos0.reshape(3, 4, 3), st["pos"])
    np.testing.assert_array_equal(tr.vel0.reshape(3, 4, 3), st["vel"])
    np.testing.assert_array_equal(tr.ids.reshape(3, 4), st["ids"])
    assert tr.t_max == otr.default_t_max(st, -1e8, [1e-6])
    with pytest.raises(ValueError):
        SpecificParticleTracer(pg, 4, FlatCathode(-1e8), [1e-6], n_workers=2)
    with pytest.raises(ValueError):
        SpecificParticleTracer(pg, 5, FlatCathode(-1e8), [1e-6])
    with pytest.raises(ValueError):
        SpecificParticleTracer(pg, 4, FlatCathode(-1e8), [1e-6], backend="cpu")
    with pytest.warns(UserWarning):
        SpecificParticleTracer(pg, 4, FlatCathode(-1e8), [2e-6], z_max=1e-6)


def test_weight_override_warning():
    from bem_amd import FlatCathode, SpecificParticleTracer
    from bem_amd.particles import ParticleGroup

    src = emission.flat_gaussian_source(4, seed=1)
    pg = ParticleGroup(dict(src, weight=np.full(4, 2e-19), status=np.ones(4), species="electron"))
    with pytest.warns(UserWarning):
        tr = SpecificParticleTracer(pg, 2, FlatCathode(-1e8), [1e-6])
    np.testing.assert_array_equal(tr.weight, 1.602176634e-19)


def test_surrogate_band_rule():
    """The band is the smallest validated distance beyond which every
    validated distance's worst error is within the tolerance; none -> inf,
    and a surrogate with an infinite band is inert (never installed)."""
    from bem_amd.surrogate import FieldSurrogate, band_for

    cands, worst = [1.0, 2.0, 4.0, 8.0], [3e-4, 2e-5, 8e-7, 4e-7]
    assert band_for(cands, worst, 1e-6) == 4.0
    assert band_for(cands, worst, 1e-4) == 2.0
    assert band_for(cands, worst, 1e-3) == 1.0
    assert band_for(cands, [3e-4, 2e-5, 8e-7, 2e-6], 1e-6) == np.inf
    assert FieldSurrogate(None, dict(band=np.inf), None).inert
    assert not FieldSurrogate(None, dict(band=4.0), None).inert
