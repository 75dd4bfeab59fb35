"""The C-ABI library loads on any host and exports every entry point
include/dic_hip.h declares; struct layouts agree with the header."""
import ctypes
import os
import re
import subprocess

import pytest

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
    assert ctypes.sizeof(pkg._lib.IntegerSearchParams) == 32
    assert ctypes.sizeof(pkg._lib.SeedParams) == 144
    assert pkg._lib.ESTIMATE_DTYPE.itemsize == 80
    assert pkg._lib.SEED_EVAL_DTYPE.itemsize == 176


def test_ctypes_mirrors_match_the_header(pkg, tmp_path):
    """Every struct of include/dic_hip.h, compiled as plain C by gcc: its size
    (and the offset of each field the Python layer fills) equals the ctypes or
    numpy mirror's."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    lib = pkg._lib
    mirrors = {
        "dic_icgn_params": lib.IcgnParams, "dic_batch_item": lib.BatchItem, "dic_batch_tuning": lib.BatchTuning,
        "dic_server_config": lib.ServerConfig, "dic_request": lib.Request, "dic_server_stats": lib.ServerStats,
        "dic_propagation_params": lib.PropagationParams, "dic_integer_search_params": lib.IntegerSearchParams,
        "dic_seed_params": lib.SeedParams,
    }
    dtypes = {"dic_initial": lib.INITIAL_DTYPE, "dic_result": lib.RESULT_DTYPE, "dic_seed": lib.SEED_DTYPE,
              "dic_estimate": lib.ESTIMATE_DTYPE, "dic_seed_evaluation": lib.SEED_EVAL_DTYPE}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dic_hip.h"', "int main(void) {"]
    for name in list(mirrors) + list(dtypes):
        lines.append(f'    printf("{name} %zu\\n", sizeof({name}));')
    for name, st in mirrors.items():
        for field, _ in st._fields_:
            lines.append(f'    printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    for name, st in mirrors.items():
        assert int(got[name]) == ctypes.sizeof(st), name
        for field, _ in st._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(st, field).offset, (name, field)
    for name, dt in dtypes.items():
        assert int(got[name]) == dt.itemsize, name


def test_no_fallback_without_library_or_device():
    """The product path fails loudly: without libdic_hip.so the loader raises,
    and on a host without a GPU creating a context raises (no CPU path)."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import __graft_entry__ as g\n"
        "pkg = g.load_package()\n"
        "try:\n"
        "    pkg._lib.load_library('/nonexistent/libdic_hip.so')\n"
        "    print('loaded')\n"
        "except ImportError as e:\n"
        "    print('missing:', 'no CPU fallback' in str(e))\n"
        "pkg._lib._lib = None\n"
        "import torch\n"
        "if not torch.cuda.is_available():\n"
        "    try:\n"
        "        pkg.Context(0)\n"
        "        print('context')\n"
        "    except pkg.DicError as e:\n"
        "        print('nodevice:', e.code)\n"
    ) % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300).stdout
    assert "missing: True" in out, out
    assert "loaded" not in out and "context" not in out, out
    import torch
    if not torch.cuda.is_available():
        assert "nodevice: 5" in out, out  # DIC_ERR_NO_DEVICE


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


def test_seed_config_from_dict_mirrors_binding(pkg):
    """bind_subset.cpp:201-228: initialization and seed_selection keys; the
    seed objective follows correlation.criterion; keys the binding does not
    read keep their defaults."""
    import pytest
    cfg = pkg.config_from_dict({
        "subset": {"radius": 41, "truncate_roi_subsets": True},
        "correlation": {"criterion": "ssd"},
        "initialization": {
            "method": "integer_search",
            "integer_search": {"subset_radius": 10, "search_radius": 30, "sift_enabled": True,
                               "pyramid_enabled": False, "pyramid_scale": 4, "pyramid_refinement_radius": 4},
            "subpixel_refinement": {"enabled": True, "shape_function": "second_order", "optimizer": "fgn",
                                    "subset_radius": 41, "max_iterations": 50, "convergence_threshold": 1e-6}},
        "seed_selection": {"seed_count": 16, "max_znssd": 2.0, "min_zncc": 0.9, "quality_metric": "zncc",
                           "min_texture_std": 0.02, "kmeans_iterations": 20, "kmeans_sample_limit": 20000},
    })
    si, ss = cfg.seed_initialization, cfg.seed_selection
    assert (si.integer_search.subset_radius, si.integer_search.search_radius) == (10, 30)
    assert si.integer_search.sift_enabled and not si.integer_search.pyramid_enabled
    assert (si.subpixel.subset_radius, si.subpixel.max_iterations, si.subpixel.convergence_threshold) == (41, 50, 1e-6)
    assert (si.subpixel.shape_function, si.subpixel.optimizer) == ("first_order", "icgn")  # not read
    assert si.subpixel.objective == "ssd"
    assert (ss.seed_count, ss.max_znssd, ss.min_texture_std) == (16, 2.0, 0.02)
    assert (ss.min_zncc, ss.quality_metric) == (0.7, "znssd")  # not read
    assert cfg.truncate_roi_subsets is False  # not read
    with pytest.raises(RuntimeError, match="integer_search"):
        pkg.config_from_dict({"initialization": {"method": "sift"}})
    p = pkg.seed_params(cfg)
    assert (p.search.subset_radius, p.search.search_radius, p.search.pyramid_enabled) == (10, 30, 0)
    assert (p.subpixel_radius, p.subpixel_objective, p.seed_count) == (41, 1, 16)
    # padding.cpp:28-36: search + max radius + border
    cfg.truncate_roi_subsets = True
    assert pkg.recommended_subset_padding(cfg) == 30 + 41 + 3


def test_cpp_pair_guard(tmp_path):
    """dic_hip.hpp refuses a PreparedPair handed over with other images or
    settings (std::invalid_argument) and accepts the matching one.  Compiled
    with plain g++; the guard runs before any device call."""
    src = tmp_path / "guard.cpp"
    src.write_text(r'''
#include <cstdio>
#include "dic_hip.hpp"
using dic_hip::detail::PairKey;
static int refused(const PairKey &have, const PairKey &want)
{
    try { dic_hip::detail::check_pair(have, want); } catch (const std::invalid_argument &) { return 1; }
    return 0;
}
int main()
{
    int a, b, c;
    const PairKey k{&a, &b, 5, 3, 1, 0, 0};
    PairKey other_def = k, other_ref = k, swapped = k, degree = k, border = k, exact = k, prec = k, grad = k;
    other_def.deformed = &c; other_ref.reference = &c; swapped.reference = &b; swapped.deformed = &a;
    degree.degree = 3; border.border = 4; exact.exact = 0; prec.precision = 1; grad.gradient_mode = 1;
    std::printf("%d %d %d %d %d %d %d %d %d\n", refused(k, k), refused(k, other_def), refused(k, other_ref),
                refused(k, swapped), refused(k, degree), refused(k, border), refused(k, exact), refused(k, prec),
                refused(k, grad));
    return 0;
}
''')
    exe = tmp_path / "guard"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), str(src)], check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.split()
    assert out == ["0"] + ["1"] * 8, out
