"""The C-ABI library loads on any host and exports every entry point
include/dic_hip.h declares; struct layouts agree with the header."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "dic_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dic_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_are_exported(pkg):
    lib = pkg.load_library()
    names = _declared()
    assert len(names) >= 29
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg._lib.EXPORTED_SYMBOLS)


def test_struct_layouts(pkg):
    assert pkg.RESULT_DTYPE.itemsize == 136
    assert pkg.INITIAL_DTYPE.itemsize == 56
    assert ctypes.sizeof(pkg.IcgnParams) == 48
    assert pkg.RESULT_DTYPE.fields["status"][1] == 15 * 8


def test_version_string(pkg):
    assert b"gfx950" in pkg.load_library().dic_version()


def test_config_from_dict_mirrors_binding(pkg):
    cfg = pkg.config_from_dict({
        "subset": {"radius": 41}, "shape_function": {"order": 2},
        "optimization": {"method": "icgn", "max_iterations": 50, "convergence_threshold": 1e-6},
        "correlation": {"criterion": "ssd"}, "interpolation": {"degree": 3},
        "reliability_propagation": {"spacing": 3},
    })
    assert (cfg.subset_radius, cfg.shape_order, cfg.max_iterations, cfg.objective) == (41, 2, 50, "ssd")
    assert cfg.image_precompute.degree == 3 and cfg.propagation_spacing == 3
    # degree 5 is not mapped by the binding: the quintic default stays
    assert pkg.config_from_dict({"interpolation": {"degree": 5}}).image_precompute.degree == 5
    import pytest
    with pytest.raises(RuntimeError, match="criterion"):
        pkg.config_from_dict({"correlation": {"criterion": "zncc"}})
    with pytest.raises(RuntimeError, match="order"):
        pkg.config_from_dict({"shape_function": {"order": 3}})


def test_cpp_host_layer_builds():
    """include/dic_hip.hpp compiles with plain g++ into the C++ example."""
    exe = os.path.join(ROOT, "traditional-dic_amd", "solve_pair")
    assert os.path.exists(exe) and os.access(exe, os.X_OK)
