"""CPU: the C-ABI library loads, exports every function include/hadros.h
declares, parses names like the reference, and fails loudly (no CPU path)
when no GPU is present."""
import os
import re

import numpy as np
import pytest

from hadros_amd import abi, ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hadros.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hadros_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = ffi.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(ffi.EXPORTS) <= set(names) | {"hadros_last_error"}


def test_struct_sizes_match_header_layout():
    import ctypes as C

    assert C.sizeof(abi.Camera) == 8 * 4 + 4 * 2 + 8 * 3 + 4 * 2
    assert C.sizeof(abi.Medium) == 8 + 8 * 12
    assert C.sizeof(abi.Source) == 8 + 8 * 14
    assert C.sizeof(abi.Spectrum) == 8 + 8 * 4
    assert C.sizeof(abi.ImageOut) == 8 * 11
    assert C.sizeof(abi.Stats) == 8 * 4 + 4 * 4
    assert C.sizeof(abi.Mev) == 4 * 8 + 8 * 16
    assert C.sizeof(abi.MevOut) == 8 * 5


def test_mev_names_and_defaults_match_reference():
    """parse_mev_* names and aliases (src/mev_neutrino_physics.cpp:36-233) and the
    MeVThermalParams defaults (include/radiative_transfer.hpp:66-106) through the C-ABI."""
    import ctypes as C

    L = ffi.lib()
    for fn, table in (("model", abi.MEV_MODELS), ("flavor", abi.MEV_FLAVORS), ("thermal", abi.MEV_THERMAL),
                      ("ye", abi.MEV_YE), ("spectral", abi.MEV_SPECTRAL)):
        for name, k in table.items():
            v = C.c_int32(-1)
            assert getattr(L, "hadros_parse_mev_" + fn)(name.encode(), C.byref(v)) == 0 and v.value == k
        v = C.c_int32(-1)
        assert getattr(L, "hadros_parse_mev_" + fn)(b"Physical_", C.byref(v)) == abi.ERR_KIND
    m = abi.Mev()
    L.hadros_mev_defaults(C.byref(m))
    ref = abi.mev()
    for f, _ in abi.Mev._fields_:
        assert getattr(m, f) == getattr(ref, f), f


def test_name_parsing_matches_reference_aliases():
    for name, k in abi.PROFILE_KINDS.items():
        assert ffi.parse_profile(name) == k
    for name, k in abi.SOURCE_KINDS.items():
        assert ffi.parse_source(name) == k
    for name, k in abi.SPECTRUM_KINDS.items():
        assert ffi.parse_spectrum(name) == k
    for bad, fn in (("Gaussian", ffi.parse_profile), ("ringx", ffi.parse_source), ("flat", ffi.parse_spectrum)):
        with pytest.raises(ffi.HadrosError) as ei:
            fn(bad)
        assert ei.value.code == abi.ERR_KIND
    L = ffi.lib()
    assert L.hadros_profile_name(6) == b"powerlaw_funnel_envelope"
    assert L.hadros_source_name(3) == b"shock_layer"
    assert L.hadros_spectrum_name(2) == b"powerlaw_cutoff"


def test_sigma_loader_follows_reference_parser(tmp_path):
    p = tmp_path / "t.dat"
    p.write_text("# model test\n\n1e3 1 2\nbad line\n2e3 -1 3\n3e3 2 4\n4e3 3 5 extra\n")
    e, g, s = ffi.sigma_load(str(p))
    assert list(e) == [1e3, 3e3, 4e3] and list(s) == [2.0, 4.0, 5.0]
    p.write_text("1e3 1 2\n")
    with pytest.raises(ffi.HadrosError) as ei:
        ffi.sigma_load(str(p))
    assert ei.value.code == abi.ERR_TABLE
    p.write_text("2e3 1 2\n1e3 1 2\n")
    with pytest.raises(ffi.HadrosError) as ei:
        ffi.sigma_load(str(p))
    assert ei.value.code == abi.ERR_TABLE
    with pytest.raises(ffi.HadrosError) as ei:
        ffi.sigma_load(str(tmp_path / "missing.dat"))
    assert ei.value.code == abi.ERR_IO


def test_no_cpu_fallback_without_gpu():
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("GPU present")
    from hadros_amd import workloads

    e, _, s = workloads.synthetic_sigma()
    with pytest.raises(ffi.HadrosError) as ei:
        ffi.trace_rt(workloads.c1_camera(4, 4), workloads.c1_scene(), e, s)
    assert ei.value.code == abi.ERR_DEVICE


def test_c3_manifest_is_the_seeded_draw():
    import json

    from hadros_amd import workloads

    committed = json.load(open(os.path.join(ROOT, "hadros_amd", "manifests", "c3_cameras.json")))
    assert committed == workloads.c3_manifest()
    assert len(committed) == 256
    res = np.array([m["res"] for m in committed])
    assert set(res) <= {64, 128, 256, 512}
