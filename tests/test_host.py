"""CPU-side tests: the C-ABI library, host logic, configs and IC builders.

No GPU compute is called here; the library is only loaded and inspected.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import lfm_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lfm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lfm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from lfm_hip import _abi, _build
    path = _build.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), f"missing export {s}"
    bound = {name for name, _, _ in _abi.SIGNATURES}
    assert set(syms) == bound, set(syms) ^ bound


def declared_arity():
    """Parameter count of every function declared in the header."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(lfm_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = args.strip()
        out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_binding_arity_matches_header():
    from lfm_hip import _abi
    ar = declared_arity()
    for name, _, argtypes in _abi.SIGNATURES:
        assert len(argtypes) == ar[name], (name, len(argtypes), ar[name])


def test_abi_version_and_error_string():
    from lfm_hip import _abi
    lib = _abi.load()
    assert lib.lfm_abi_version() == _abi.LFM_ABI_VERSION == 6
    # argument validation happens before any device work: N < 8 is refused
    p = _abi.LfmParams(4, 0, 4e-4, 1 / 63, 0, 19, 0, 0.1, 0, 0, 0, 0)
    dummy = [1 << 20] * 4
    rc = lib.lfm_step_real_f32(*dummy, None, *[(2 << 20) + (q << 16) for q in range(4)],
                               ctypes.byref(p), None)
    assert rc == -1
    assert b"out of range" in lib.lfm_last_error()
    # and above the largest supported lattice (N = 1290, the largest N with N^3 < 2^31)
    p.N = 1291
    rc = lib.lfm_step_real_f32(*dummy, None, *[(2 << 20) + (q << 16) for q in range(4)],
                               ctypes.byref(p), None)
    assert rc == -1 and b"out of range" in lib.lfm_last_error()


def test_abi_rejects_in_place_and_missing_inputs():
    from lfm_hip import _abi
    lib = _abi.load()
    N = 16
    p = _abi.LfmParams(N, 0, 4e-4, 1 / 63, 0, 19, 0, 0.1, 0, 0, 0, 0)
    base = 1 << 24
    E, Ep, chi, chip = base, base + 4 * N ** 3, base + 8 * N ** 3, base + 12 * N ** 3
    # output aliases an input
    rc = lib.lfm_step_real_f32(E, Ep, chi, chip, None, E, None, base + 16 * N ** 3, None,
                               ctypes.byref(p), None)
    assert rc == -1 and b"overlaps" in lib.lfm_last_error()
    rc = lib.lfm_step_real_f32(None, Ep, chi, chip, None, base + 16 * N ** 3, None,
                               base + 20 * N ** 3, None, ctypes.byref(p), None)
    assert rc == -1 and b"NULL" in lib.lfm_last_error()
    # mask flag without a mask
    p.flags = _abi.LFM_F_MASK
    rc = lib.lfm_step_real_f32(E, Ep, chi, chip, None, base + 16 * N ** 3, None,
                               base + 20 * N ** 3, None, ctypes.byref(p), None)
    assert rc == -1 and b"mask" in lib.lfm_last_error()
    # keep flag without the factors; keep overlapping an output
    p.flags = _abi.LFM_F_KEEP_PSI
    rc = lib.lfm_step_real_f32(E, Ep, chi, chip, None, base + 16 * N ** 3, None,
                               base + 20 * N ** 3, None, ctypes.byref(p), None)
    assert rc == -1 and b"keep" in lib.lfm_last_error()
    p.keep = base + 16 * N ** 3
    rc = lib.lfm_step_real_f32(E, Ep, chi, chip, None, base + 16 * N ** 3, None,
                               base + 20 * N ** 3, None, ctypes.byref(p), None)
    assert rc == -1 and b"overlaps" in lib.lfm_last_error()
    # not supported by the two-step kernel
    rc = lib.lfm_step2_real_f32(E, Ep, chi, chip, None, base + 24 * N ** 3, base + 28 * N ** 3,
                                base + 32 * N ** 3, base + 36 * N ** 3, ctypes.byref(p), None)
    assert rc == -1 and b"KEEP" in lib.lfm_last_error()
    p.keep = None
    # COLOR with kappa_tube but no S_a
    p.flags, p.kappa_tube = 0, 0.5
    n3 = N ** 3 * 4
    args = [base + q * 3 * n3 for q in range(4)] + [base + 12 * n3, base + 13 * n3, None]
    outs = [base + (20 + 3 * q) * n3 for q in range(4)] + [base + 40 * n3, None]
    rc = lib.lfm_step_color_f32(*args, *outs, None, None, None, ctypes.byref(p), None)
    assert rc == -1 and b"S_a" in lib.lfm_last_error()


def test_slab_step_validation():
    """lfm_step_slab_f32 refuses bad shapes and flags before any device work."""
    from lfm_hip import _abi
    lib = _abi.load()
    N = 16
    p = _abi.LfmParams(N, 0, 4e-4, 1 / 63, 0, 19, 0, 0.1, 0, 0, 0, 0)
    fld = 4 * 3 * N * N * N
    ptr = [(1 << 24) + q * fld for q in range(16)]
    ins = [ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], None, None]
    outs = [ptr[6], ptr[7], ptr[8], None]
    for level, planes, flags, what in ((1, 2, 0, b"planes"), (1, N + 1, 0, b"planes"),
                                       (3, 8, 0, b"level"), (1, 8, _abi.LFM_F_DIRECT, b"DIRECT"),
                                       (1, 8, _abi.LFM_F_MASK, b"mask"),
                                       (0, 8, _abi.LFM_F_SA_SCV, b"SA_SCV")):
        p.flags = flags
        rc = lib.lfm_step_slab_f32(level, planes, *ins, *outs, ctypes.byref(p), None)
        assert rc == -1 and what in lib.lfm_last_error(), (level, planes, flags)
    p.flags = 0
    # an output aliasing an input
    rc = lib.lfm_step_slab_f32(1, 8, *ins, ptr[0], ptr[7], ptr[8], None, ctypes.byref(p), None)
    assert rc == -1 and b"overlaps" in lib.lfm_last_error()
    # COLOR with kappa_tube and no S_a input
    p.kappa_tube = 0.5
    rc = lib.lfm_step_slab_f32(2, 8, *ins, *outs, ctypes.byref(p), None)
    assert rc == -1 and b"S_a" in lib.lfm_last_error()


def test_plane_source_and_detector_validation():
    """LFM_F_SOURCE and LFM_F_DETECTOR are checked before any device work:
    their pointers, axes, planes, lengths and interval, an output over the
    envelope, the amounts or the counter, frames over an input, and the entry
    points that do not take them (COLOR, two-step, slab share)."""
    from lfm_hip import _abi
    lib = _abi.load()
    N = 16
    n3 = N ** 3
    ptr = [(1 << 24) + q * 4 * 3 * n3 for q in range(16)]
    env, amp, cnt, frames = ptr[10], ptr[11], ptr[12], ptr[13]

    def params(flags=_abi.LFM_F_SOURCE, **kw):
        p = _abi.LfmParams(N, flags, 4e-4, 1 / 63, 0, 19, 0, 0.1, 0, 0, 0, 0)
        p.src_env, p.src_amp, p.step_counter, p.src_len = env, amp, cnt, 100
        p.src_axis, p.src_plane, p.step_off = 2, 4, 1
        p.det_frames, p.det_len, p.det_axis, p.det_plane, p.det_every = frames, 10, 1, 5, 1
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def real(p, out=ptr[6]):
        return lib.lfm_step_real_f32(ptr[0], ptr[1], ptr[2], ptr[3], None, out, None, ptr[7], None,
                                     ctypes.byref(p), None)

    def cplx(p):
        return lib.lfm_step_complex_f32(ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], None,
                                        ptr[6], None, ptr[7], None, ptr[8], None, ctypes.byref(p),
                                        None)
    S, D = _abi.LFM_F_SOURCE, _abi.LFM_F_DETECTOR
    for flags, kw, what in ((S, {"src_env": None}, b"src_env"),
                            (S, {"step_counter": None}, b"step_counter"),
                            (D, {"step_counter": None}, b"step_counter"),
                            (S, {"src_axis": 3}, b"axis"), (S, {"src_plane": N}, b"plane"),
                            (S, {"src_plane": -1}, b"plane"), (S, {"src_len": -1}, b"src_len"),
                            (D, {"det_frames": None}, b"det_frames"),
                            (D, {"det_axis": -1}, b"axis"), (D, {"det_plane": N}, b"plane"),
                            (D, {"det_every": 0}, b"interval"), (D, {"det_field": 2}, b"field"),
                            (D, {"det_len": -1}, b"frames")):
        for fn in (real, cplx):
            rc = fn(params(flags, **kw))
            assert rc == -1 and what in lib.lfm_last_error(), (flags, kw, fn)
    # an output over the envelope, the amounts or the counter; frames over an input
    for out in (env, amp + 64, cnt):
        rc = real(params(), out=out)
        assert rc == -1 and b"overlaps" in lib.lfm_last_error(), out
    rc = real(params(D, det_frames=ptr[1] + 64))
    assert rc == -1 and b"overlaps" in lib.lfm_last_error()
    for flags in (S, D):
        p = params(flags)
        rc = lib.lfm_step_color_f32(ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], None, ptr[6],
                                    None, ptr[7], None, ptr[8], None, None, None, None,
                                    ctypes.byref(p), None)
        assert rc == -1 and b"COMPLEX levels" in lib.lfm_last_error()
        rc = lib.lfm_step2_real_f32(ptr[0], ptr[1], ptr[2], ptr[3], None, ptr[6], ptr[7], ptr[8],
                                    ptr[9], ctypes.byref(p), None)
        assert rc == -1 and b"not supported" in lib.lfm_last_error()
        rc = lib.lfm_step_slab_f32(1, 8, ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], None,
                                   None, ptr[6], ptr[7], ptr[8], None, ctypes.byref(p), None)
        assert rc == -1 and b"not supported" in lib.lfm_last_error()


def test_continuous_source_amounts():
    """ContinuousSource's amounts and envelope follow the reference's formulas
    (lfm/experiment/source.py:131-149,196-228): injection
    dt^2 A sin(omega t) boost rounded to float32, skipped below 1e-30, the
    imaginary amount scaled by sin(phase) only at COMPLEX with a phase; the
    envelope a float32 Gaussian about N/2 (peak 1 at the centre cell)."""
    import math
    import types
    import lfm_hip as L

    class FakeSim:
        def __init__(self, level, N):
            self.config = types.SimpleNamespace(grid_size=N, dt=0.02)
            self.evolver = types.SimpleNamespace(level=level)

        def _to_device(self, arr):
            return np.asarray(arr, np.float32).ravel()
    N = 32
    s = L.ContinuousSource(FakeSim(1, N), axis=0, position=0.25, omega=21.0, amplitude=2.0,
                           envelope_sigma=0.25, phase=1.0, boost=10.0)
    assert s.position == 8
    env = s.envelope
    assert env.dtype == np.float32 and env.shape == (N, N) and env[N // 2, N // 2] == 1.0
    # symmetric about N/2 and Gaussian of width N/4
    np.testing.assert_array_equal(env[N // 2 + 3], env[N // 2 - 3])
    assert abs(float(env[N // 2, N // 2 + 8]) - math.exp(-0.5)) < 1e-6
    for step in (1, 7, 250):
        inj = 0.02 ** 2 * 2.0 * math.sin(21.0 * step * 0.02) * 10.0
        a_r, a_i = s.amounts(step)
        assert a_r == float(np.float32(inj)) and a_i == float(np.float32(inj * math.sin(1.0)))
    assert s.amounts(0) == (0.0, 0.0)          # sin(0): the reference skips
    r = L.ContinuousSource(FakeSim(0, N), position=5, phase=1.0)
    assert r.position == 5 and r.amounts(3)[1] == 0.0     # REAL: no imaginary part
    c = L.ContinuousSource(FakeSim(1, N), phase=1e-13)
    assert c.amounts(3)[1] == 0.0                          # phase below 1e-12
    with pytest.raises(ValueError):
        L.ContinuousSource(FakeSim(1, N), axis=3)


def test_integration_stub_matches_params_layout():
    """The ctypes stub that INTEGRATION.md shows a maintainer lists the
    fields of lfm_params_t in the order the library's own binding does."""
    from lfm_hip import _abi
    text = open(os.path.join(os.path.dirname(HEADER), "..", "INTEGRATION.md")).read()
    block = text[text.index("class LfmParams(ctypes.Structure)"):]
    block = block[:block.index("lib.lfm_step_real_f32.argtypes")]
    names = re.findall(r'\("([A-Za-z_0-9]+)", (?:i32|i64|vp)\)', block)
    floats = re.search(r'for n in \(([^)]*)\)', block).group(1)
    floats = re.findall(r'"([a-z_0-9]+)"', floats)
    stub = names[:2] + floats + names[2:]
    assert stub == [n for n, _ in _abi.LfmParams._fields_], stub


def test_barrier_geometry_and_factors():
    """Barrier (lfm/experiment/barrier.py:98-427) on a host-array simulation:
    the default two slits, the slab, the chi template, the absorber and the
    which-path detector factor 1 - gamma, gamma = 1 - (1 - s)^(1/transit),
    apply() on the live buffers, and fold() refusing overlapping factors."""
    import types
    import lfm_hip as L
    N = 64

    class HostSim:
        def __init__(self):
            self.config = types.SimpleNamespace(grid_size=N, chi0=19.0)
            self.chi = [np.full((N, N, N), 20.0, np.float32) for _ in range(4)]
            self.pr = np.ones((N, N, N), np.float32)
            self.evolver = types.SimpleNamespace(level=0)

        def _to_device(self, arr):
            return np.asarray(arr, np.float32).ravel()

        def _native_chi_pair(self):
            return tuple(self.chi)

        def _native_psi_real(self):
            return self.pr

        def _native_psi_imag(self):
            return None
    sim = HostSim()
    b = L.Barrier(sim, slits=[L.Slit(center=24, width=2),
                              L.Slit(center=40, width=2, detector=True, detector_strength=0.95)])
    assert (b.position, b.axis, b.height) == (32, 2, 69.0)
    # slab k in [31, 33); slits along axis 0 at [23, 25) and [39, 41)
    solid = np.zeros((N, N, N), bool)
    solid[:, :, 31:33] = True
    solid[23:25, :, 31:33] = False
    solid[39:41, :, 31:33] = False
    np.testing.assert_array_equal(b.mask, solid)
    chi = b.chi_template()
    assert set(np.unique(chi)) == {19.0, 69.0} and bool((chi[solid] == 69.0).all())
    gamma = 1.0 - (1.0 - 0.95) ** (1.0 / 10)
    k = b.keep_factors()
    assert bool((k[solid] == 0.0).all()) and bool((k[39:41, :, 31:33] == np.float32(1 - gamma)).all())
    assert bool((k[23:25, :, 31:33] == 1.0).all()) and float(k[:, :, :31].min()) == 1.0
    # construction applied the barrier to the live buffers
    for c in sim.chi:
        np.testing.assert_array_equal(c, chi)
    np.testing.assert_array_equal(sim.pr, k)
    # default geometry: two slits max(4, N/8) either side of the centre
    d = L.Barrier(HostSim())
    assert [(s.center, s.width) for s in d.slits] == [(24, 2), (40, 2)]
    # overlapping detector slits cannot fold into one factor per site
    o = L.Barrier(HostSim(), slits=[L.Slit(center=20, width=4, detector=True, detector_strength=0.5),
                                    L.Slit(center=21, width=4, detector=True, detector_strength=0.5)])
    with pytest.raises(ValueError, match="overlap"):
        o.fold()


def test_host_mask_bit_identical_to_oracle():
    from lfm_hip.fields import boundary_mask_host
    for N, bf in ((8, 0.3), (16, 0.15), (48, 0.3), (64, 0.3)):
        np.testing.assert_array_equal(boundary_mask_host(N, bf), O.boundary_mask(N, bf))


def test_config_defaults_and_validation():
    import lfm_hip as L
    c = L.SimulationConfig()
    assert (c.grid_size, c.chi0, c.dt, c.boundary_fraction, c.report_interval) == (128, 19.0, 0.02, 0.3, 5000)
    assert abs(c.kappa - 1 / 63) < 1e-15 and c.epsilon_w == pytest.approx(0.1)
    assert c.e_amplitude == 3.6 and c.sigma == pytest.approx(128 / 12)
    with pytest.raises(ValueError, match="CFL"):
        L.SimulationConfig(dt=0.06)
    with pytest.raises(ValueError):
        L.SimulationConfig(grid_size=4)
    with pytest.raises(ValueError):
        L.SimulationConfig(boundary_fraction=1.0)
    fp = L.full_physics(32)
    assert fp.field_level == L.FieldLevel.COLOR and fp.sa_enabled
    assert fp.kappa_tube == pytest.approx(30 / 63) and fp.lambda_self == pytest.approx(4 / 31)
    assert fp.epsilon_cc == pytest.approx(2 / 17) and fp.kappa_string == pytest.approx(1 / 189)
    ge = L.gravity_em(16)
    assert ge.field_level == L.FieldLevel.COMPLEX and ge.lambda_self == pytest.approx(4 / 31)
    sc = L.SimulationConfig.for_scale(L.PhysicsScale.NUCLEAR, grid_size=16)
    assert sc.field_level == L.FieldLevel.COLOR and sc.kappa_c == pytest.approx(1 / 189)


def test_get_backend_refuses_cpu_and_unknown():
    import lfm_hip as L
    with pytest.raises(ValueError, match="no CPU backend"):
        L.get_backend("cpu")
    with pytest.raises(ValueError, match="Unknown"):
        L.get_backend("tpu")


def test_backend_needs_gpu_when_absent():
    import torch
    import lfm_hip as L
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="GPU"):
        L.HipBackend()


def test_ic_builders_match_formulas():
    from lfm_hip import fields
    N = 16
    r, i = fields.gaussian_soliton(N, (8, 7, 9), 3.0, 2.0, phase=0.5)
    assert r.dtype == np.float32 and r.shape == (N, N, N)
    assert r[8, 7, 9] == pytest.approx(3.0 * np.cos(0.5), rel=1e-6)
    assert i[8, 7, 9] == pytest.approx(3.0 * np.sin(0.5), rel=1e-6)
    assert r[8, 7, 11] == pytest.approx(3.0 * np.cos(0.5) * np.exp(-4 / 8), rel=1e-6)
    pr, pi = fields.place_solitons(N, [(4, 4, 4), (10, 10, 10)], 2.0, 1.5, colors=[0, 2])
    assert pr.shape == (3, N, N, N) and np.all(pr[1] == 0)
    assert pr[0, 4, 4, 4] == pytest.approx(2.0, rel=1e-5)
    a = fields.seed_noise(N, 1e-3, rng=np.random.default_rng(42))[0]
    b = (np.random.default_rng(42).standard_normal((N, N, N)) * 1e-3).astype(np.float32)
    np.testing.assert_array_equal(a, b)


def test_python_constants_match_header_defines():
    """Every LFM_* constant the Python binding defines equals the header's."""
    from lfm_hip import _abi
    text = open(HEADER).read()
    defs = {k: int(v, 0) for k, v in re.findall(r"#define\s+(LFM_[A-Z0-9_]+)\s+\(?(-?(?:0x)?[0-9a-fA-F]+)\)?",
                                                 text)}
    shared = [k for k in dir(_abi) if k.startswith("LFM_") and k in defs]
    assert {"LFM_F_MASK", "LFM_F_CHI_DELTA", "LFM_RELAX_SLOT_WORDS"} <= set(shared)
    for k in shared:
        assert getattr(_abi, k) == defs[k], k


def test_params_struct_layout(tmp_path):
    """LfmParams mirrors lfm_params_t: the same fields in declaration order at
    the offsets a C compiler gives them (gcc compiles the header and prints
    offsetof of every field and the struct size)."""
    import shutil
    import subprocess
    from lfm_hip import _abi
    names = [n for n, _ in _abi.LfmParams._fields_]
    assert names[:2] == ["N", "flags"] and names[12:15] == ["tile", "chunk", "keep"]
    text = open(HEADER).read()
    body = re.search(r"typedef struct lfm_params \{(.*?)\} lfm_params_t;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"[A-Za-z_0-9]+(?=[,;])", body)
    assert fields == names, (fields, names)
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"lfm_hip.h\"\nint main(void) {\n"
                   + "".join(f'    printf("%zu\\n", offsetof(lfm_params_t, {n}));\n' for n in names)
                   + '    printf("%zu\\n", sizeof(lfm_params_t));\n    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]
    want = [getattr(_abi.LfmParams, n).offset for n in names] + [ctypes.sizeof(_abi.LfmParams)]
    assert got == want, (got, want)


def test_relax_fused_step_bookkeeping():
    """The relaxation loop's launch sequence (host logic only, launches mocked):
    step n writes max slot n % 3; it reads the previous step's slot as its
    pending rescale unless E was rescaled in between; every convergence check
    rescales and folds max|dE| at a multiple of the check interval; E is
    rescaled on return.  Graph-replayed blocks of six steps follow the same
    sequence as eager steps."""
    from lfm_hip import eigenmode as em

    class Mock(em._Relaxer):
        def __init__(self):
            self.log, self.E, self.E_alt, self.sw, self._graphs = [], "A", "B", 1, {}

            class Slots:
                def __getitem__(self, i):
                    return self

                def zero_(self):
                    pass
            self.slots = Slots()

        def _step(self, k, scale_k, amplitude, dt):
            self.log.append(("step", k, scale_k))
            self.E, self.E_alt = self.E_alt, self.E

        def _rescale(self, k, amplitude, E_prev=None):
            self.log.append(("rescale", k, E_prev is not None))

        def slot_max(self, i):
            return 1.0

        def _block_graph(self, k0, amplitude, dt):
            me = self

            class G:
                def replay(self):
                    for u in range(me.GRAPH_STEPS):
                        k = (k0 + u) % 3
                        me._step(k, (k - 1) % 3, amplitude, dt)
            return G()

    for steps in (1, 5, 6, 7, 12, 13, 199, 200, 201, 405):
        for ci in (1, 2, 7, 200):
            for graphs in (True, False):
                R = Mock()
                R.relax(14.0, 0.01, steps, ci, -1.0, graphs=graphs)
                n, pending, checks = 0, None, []
                for ev in R.log:
                    if ev[0] == "step":
                        n += 1
                        assert ev[1] == n % 3 and ev[2] == pending, (steps, ci, n, ev, pending)
                        pending = ev[1]
                    else:
                        assert ev[1] == pending, (steps, ci, ev, pending)
                        if ev[2]:
                            checks.append(n)
                        pending = None
                assert n == steps and pending is None
                assert checks == [m for m in range(1, steps + 1) if m % ci == 0]
                assert R.E == ("A" if steps % 2 == 0 else "B")


# ------------------------------------------------------ drop-in array protocol
def test_device_array_protocol():
    """DeviceArray (what HipBackend hands out) answers the NumPy/CuPy array
    calls the reference makes on backend buffers: copy() is a new buffer,
    fill() and slice assignment write in place, reshape() views share storage
    and stay DeviceArrays (lfm/core/evolver.py:161-176,355-429;
    lfm/experiment/barrier.py:321-334).  Checked on CPU tensors; the class
    adds no device logic."""
    import torch
    from lfm_hip.backend import DeviceArray, as_device_array
    a = as_device_array(torch.arange(27, dtype=torch.float32))
    assert isinstance(a, DeviceArray)
    v = a.reshape(3, 3, 3)
    assert isinstance(v, DeviceArray) and v.data_ptr() == a.data_ptr()
    snap = a.copy()
    assert isinstance(snap, DeviceArray) and snap.data_ptr() != a.data_ptr()
    v.fill(19.0)
    assert float(a.sum()) == 19.0 * 27 and float(snap.sum()) == sum(range(27))
    v[1] = 22.0
    assert float(a[9:18].min()) == 22.0
    a[:] = snap
    assert torch.equal(a, torch.arange(27, dtype=torch.float32))
    v *= torch.zeros(3, 3)
    assert float(a.abs().sum()) == 0.0
    d = a.astype(np.float64)
    assert d.dtype == torch.float64 and d.data_ptr() != a.data_ptr()


def test_hip_backend_raises_import_error_without_gpu():
    """lfm/core/backends/__init__.py:57-59 raises ImportError when the GPU
    stack is missing; HipBackend raises an ImportError subclass likewise."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from lfm_hip.backend import GpuUnavailableError, HipBackend
    from lfm_hip.evolver import get_backend
    assert issubclass(GpuUnavailableError, ImportError)
    with pytest.raises(ImportError):
        HipBackend()
    with pytest.raises(ImportError):
        get_backend("gpu")
    with pytest.raises(ValueError):
        get_backend("nonsense")
