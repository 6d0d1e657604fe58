"""The C-ABI: the library loads without a GPU, exports every symbol include/zombie_amd.h
declares, and the ctypes structures match the C layouts (checked with gcc)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from zombie_amd import _abi as A

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "zombie_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:zg_status|void|int32_t|const char\*)\s+(zg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in A.SIGNATURES, f"{n} missing from the ctypes signature table"
    assert lib.zg_version().decode() == "0.1.0"


def test_defaults_match_walk_settings_constructor():
    ws = A.zg_walk_settings()
    A.lib().zg_walk_settings_default(C.byref(ws), 1e-3, 2e-3, 512, 1)
    assert ws.silhouette_precision == pytest.approx(1e-3)
    assert ws.russian_roulette_threshold == 0.0
    assert ws.splitting_threshold == pytest.approx(A.FLT_MAX)
    assert ws.steps_before_using_maximal_spheres == 512 and ws.steps_before_applying_tikhonov == 0
    assert ws.solve_double_sided == 1 and ws.use_gradient_control_variates == 1
    assert ws.ignore_source_contribution == 0


def test_invalid_arguments_are_reported_not_fatal():
    h = C.c_void_p()
    st = A.lib().zg_scene_create(4, None, 0, None, 0, None, None, None, 0, 1e-3, None, C.byref(h))
    assert st == A.ZG_ERR_INVALID_ARGUMENT
    assert "dim" in A.last_error()


STRUCTS = ["zg_walk_settings", "zg_function", "zg_pde", "zg_build_options", "zg_scene_info", "zg_sample_points",
           "zg_statistics", "zg_solve_options", "zg_solve_report", "zg_grid_field", "zg_mlp_config",
           "zg_bvc_samples", "zg_bvc_eval_points", "zg_bvc_statistics"]


def test_ctypes_layouts_match_c(tmp_path):
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){"]
    for s in STRUCTS:
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(A, s)._fields_:
            src.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for s in STRUCTS:
        cls = getattr(A, s)
        assert int(got[s]) == C.sizeof(cls), s
        for f, _ in cls._fields_:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def test_points_dimension_check():
    """the solve wrappers hand plain pointers to the C-ABI: points, normals and derivative
    directions of the wrong width are refused on the host before any call"""
    import torch
    from zombie_amd import engine as E
    n = 5
    ok = E.DevicePoints(pt=torch.zeros(n, 3), dist_to_absorbing=torch.ones(n), n_walks=torch.ones(n, dtype=torch.int32),
                        n_walks_max=1, normal=torch.zeros(n, 3))
    ok.check_dim(3)
    with pytest.raises(ValueError):
        ok.check_dim(2)
    bad = E.DevicePoints(pt=torch.zeros(n, 3), dist_to_absorbing=torch.ones(n), n_walks=torch.ones(n, dtype=torch.int32),
                         n_walks_max=1, direction_for_derivative=torch.zeros(n, 2))
    with pytest.raises(ValueError):
        bad.check_dim(3)


def _points(n, dim=3, **kw):
    import torch
    from zombie_amd import engine as E
    a = dict(pt=torch.zeros(n, dim), dist_to_absorbing=torch.ones(n), n_walks=torch.ones(n, dtype=torch.int32),
             n_walks_max=1)
    a.update(kw)
    return E.DevicePoints(**a)


@pytest.mark.parametrize("field,value,err", [
    ("n_walks", "int64", TypeError),             # torch.full's default integer type: wrong walk counts
    ("pt", "float64", TypeError),                # 8-byte coordinates misread as floats
    ("pt", "strided", ValueError),               # a sliced view: the kernel reads the wrong values
    ("dist_to_absorbing", "short", ValueError),  # shorter than n: reads past its end
    ("n_walks", "short", ValueError),
    ("type", "int32", TypeError),
    ("first_sphere_radius", "short", ValueError),
    ("n_walks", "too_many", ValueError),         # a point asks for more walks than n_walks_max
])
def test_points_validation_refuses_mismatched_arrays(field, value, err):
    """Every array the solve wrappers pass to the C-ABI is checked on the host: element type,
    length, contiguity (these checks come before the device check, so they run on the CPU)."""
    import torch
    n = 6
    bad = {
        "int64": torch.ones(n, dtype=torch.int64),
        "float64": torch.zeros(n, 3, dtype=torch.float64),
        "strided": torch.zeros(n, 6)[:, ::2],
        "short": torch.ones(n - 1, dtype=torch.int32 if field == "n_walks" else torch.float32),
        "int32": torch.zeros(n, dtype=torch.int32),
        "too_many": torch.full((n,), 2, dtype=torch.int32),
    }[value]
    p = _points(n, **{field: bad})
    with pytest.raises(err):
        p.validate(3)


def test_points_validation_requires_a_gpu_device():
    p = _points(4)
    with pytest.raises(ValueError, match="GPU"):
        p.validate(3)


@pytest.mark.parametrize("field,kind", [("n", "float"), ("mean", "short"), ("m2", "float32"),
                                        ("walk_length", "int32"), ("grad_mean", "shape")])
def test_stats_validation_refuses_mismatched_arrays(field, kind):
    import torch
    from zombie_amd import engine as E
    n = 5
    st = E.DeviceStats.zeros(n, device="cpu", dim=3)
    st.validate(n, 3, torch.device("cpu"), need_gradient=True)
    cur = getattr(st, field)
    bad = {"float": cur.to(torch.float64), "short": cur[:-1].clone(), "float32": cur.to(torch.float32),
           "int32": cur.to(torch.int32), "shape": torch.zeros(n, 2, dtype=torch.float64)}[kind]
    setattr(st, field, bad)
    with pytest.raises((TypeError, ValueError)):
        st.validate(n, 3, torch.device("cpu"), need_gradient=True)


def test_stats_validation_needs_gradient_members_for_gradient_points():
    import torch
    from zombie_amd import engine as E
    st = E.DeviceStats.zeros(4, device="cpu")
    st.validate(4, 3, torch.device("cpu"))
    with pytest.raises(ValueError, match="gradient"):
        st.validate(4, 3, torch.device("cpu"), need_gradient=True)
    with pytest.raises(ValueError, match="on"):
        E.DeviceStats.zeros(4, device="cpu").validate(4, 3, torch.device("meta"))


def test_cpp_adapter_compiles_and_links(tmp_path):
    """include/zombie_amd.hpp (the reference-signature C++ adapter) builds with plain g++ -Wall
    -Wextra against the C-ABI: tests/adapter_check.cpp, the caller the GPU test runs, compiles
    and links against libzombie_amd.so (no GPU needed to link)."""
    from zombie_amd import _build
    out = _build.build_adapter_check(out=tmp_path / "adapter_check")
    assert out.exists()
