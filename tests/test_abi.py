"""CPU-side checks of the C ABI: the HIP library builds, loads and exports
every entry point include/gravis.h declares. No compute call is made (there
is no GPU here); GPU parity lives in tests/test_gpu_*.py."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(gravis_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for n in ("gravis_create", "gravis_accel", "gravis_kdk", "gravis_diagnostics"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from gravis_amd import _lib
    L = _lib.load()
    missing = [n for n in sorted(_declared()) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= _declared()


def test_version_and_status_strings():
    from gravis_amd import _lib
    L = _lib.load()
    assert b"gfx950" in L.gravis_version()
    for s in range(6):
        assert L.gravis_status_string(s)


def test_library_is_built_for_gfx950_only():
    from gravis_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"sm_"):
        assert b"amdgcn-amd-amdhsa--" + other not in data


def test_null_context_is_rejected_without_touching_the_gpu():
    from gravis_amd import _lib
    L = _lib.load()
    s = _lib.Solver(0, 0, 0.05, 0.5, 16, 64)
    assert L.gravis_accel(None, C.byref(s), 10, *([None] * 7)) == _lib.INVALID_ARG
    assert L.gravis_kdk(None, C.byref(s), 10, *([None] * 10), 0.01, 1) == _lib.INVALID_ARG
    assert L.gravis_create(0, None) == _lib.INVALID_ARG
    assert L.gravis_last_error(None) == b"no context"


def test_package_has_no_oracle_import():
    """The product never imports the oracle (no CPU fallback)."""
    pkg = os.path.join(ROOT, "gravis_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "liboracle" not in txt, f


def test_scenario_plummer_bit_exact_with_oracle():
    """The product's IC generator (host C++) reproduces the oracle's
    ChaCha20/Plummer particles bit for bit (the oracle is pinned by the
    reference's m1_plummer_1k_diag.csv)."""
    import numpy as np
    import gravis_amd as g
    import oracle as O
    for n, seed in ((1000, 42), (4097, 7)):
        p = g.scenarios.plummer(n, seed)
        s = O.plummer(n, seed)
        for f, of in (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("vy", "vy"),
                      ("vz", "vz"), ("mass", "m")):
            assert np.array_equal(getattr(p, f), getattr(s, of)), (n, seed, f)
    assert g.scenarios.plummer_suggested_softening(1000) == O.plummer_suggested_softening(1000)
    assert g.scenarios.plummer_suggested_dt() == O.plummer_suggested_dt()
    t = g.scenarios.two_body(0.7)
    o = O.two_body(0.7)
    assert np.array_equal(t.x, o.x) and np.array_equal(t.vy, o.vy)
    assert g.scenarios.two_body_period() == O.two_body_period()


def test_ensemble_layout_matches_oracle():
    import numpy as np
    import gravis_amd.ensemble as E
    import oracle as O
    assert (E.counts(42, 256) == O.ens_counts(42, 256)).all()
    t = E.targets(10007, 999)
    o = O.ens_targets(10007, 999)
    for a, b in zip(t, o):
        assert np.array_equal(a, b)


def test_scenario_evrard_bit_exact_with_oracle():
    """gravis_scenario_evrard (host code in the library) reproduces the
    oracle's EvrardCollapse::generate bit for bit (same ChaCha20 stream)."""
    import numpy as np
    import gravis_amd as g
    import oracle as O
    for n, seed in ((3000, 11), (257, 42)):
        p = g.scenarios.evrard(n, seed=seed)
        s = O.evrard(n, seed=seed)
        for a, b in (("x", "x"), ("y", "y"), ("z", "z"), ("mass", "m"),
                     ("smoothing_length", "h"), ("internal_energy", "u"), ("alpha_visc", "alpha")):
            assert np.array_equal(getattr(p, a), getattr(s, b)), (n, a)
        assert np.array_equal(p.particle_type, s.type)
        assert not p.vx.any() and not p.density.any()
        assert g.scenarios.evrard_softening(n) == O.evrard_softening(n)


def test_scenario_evrard_1r_bit_exact_with_oracle():
    """The rho ~ 1/r generator in the library equals the oracle's bit for bit."""
    import numpy as np
    import gravis_amd as g
    import oracle as O
    for n, seed in ((2000, 42), (513, 5)):
        p = g.scenarios.evrard_1r(n, seed=seed)
        s = O.evrard_1r(n, seed=seed)
        for a, b in (("x", "x"), ("y", "y"), ("z", "z"), ("mass", "m"),
                     ("smoothing_length", "h"), ("internal_energy", "u"), ("alpha_visc", "alpha")):
            assert np.array_equal(getattr(p, a), getattr(s, b)), (n, a)
        assert np.array_equal(p.particle_type, s.type)
