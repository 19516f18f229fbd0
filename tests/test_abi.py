"""The C ABI library loads and exports every function include/*.h declares
(no compute calls: runs without a GPU), and the host-side plan plumbing."""
import ctypes
import glob
import os
import re

import pytest

from gigalens_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b(gl_[a-z0-9_]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    assert {"gl_plan_create", "gl_plan_destroy", "gl_plan_query", "gl_simulate", "gl_simulate_vjp", "gl_log_prob",
            "gl_map_adam_step", "gl_last_error"} <= names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(nat.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(nat.EXPORTED_SYMBOLS) <= declared_functions()


def test_struct_layout_matches_header():
    """ctypes mirrors of the ABI structs have the C sizes and field offsets."""
    src = os.path.join(ROOT, "_abi_probe.c")
    out = os.path.join(ROOT, "_abi_probe")
    import subprocess
    import tempfile
    d = tempfile.mkdtemp()
    src = os.path.join(d, "probe.c")
    out = os.path.join(d, "probe")
    open(src, "w").write('#include <stdio.h>\n#include <stddef.h>\n#include "gigalens_amd.h"\n'
                         'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(gl_component_desc), '
                         'sizeof(gl_plan_desc), sizeof(gl_plan_info), sizeof(gl_batch_opts), sizeof(gl_batch_info), '
                         'offsetof(gl_batch_opts, band_lds_bytes), offsetof(gl_batch_info, workspace_bytes));'
                         'return 0;}\n')
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                        src, "-o", out], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no C compiler for the layout probe: " + r.stderr[-200:])
    sizes = [int(v) for v in subprocess.run([out], capture_output=True, text=True).stdout.split()]
    assert sizes == [ctypes.sizeof(nat.ComponentDesc), ctypes.sizeof(nat.PlanDesc), ctypes.sizeof(nat.PlanInfo),
                     ctypes.sizeof(nat.BatchOpts), ctypes.sizeof(nat.BatchInfo), nat.BatchOpts.band_lds_bytes.offset,
                     nat.BatchInfo.workspace_bytes.offset]


def test_layout_and_prior_table():
    from gigalens_amd import workloads as wl
    from gigalens_amd import distributions as tfd
    from gigalens_amd.plan import layout_from_model, layout_from_paths
    prior = wl.epl_shear_prior()
    phys = wl.epl_shear_model()
    lay_p = layout_from_paths(phys, [p for p, _ in tfd.leaves(prior)])
    lay_m = layout_from_model(phys)
    assert lay_p == lay_m  # prior flatten order == params flatten order (tf.nest)
    assert [k for _, _, k in lay_m[:8]] == ["center_x", "center_y", "e1", "e2", "gamma", "theta_E", "gamma1", "gamma2"]
    assert [k for _, _, k in lay_m[8:15]] == ["Ie", "R_sersic", "center_x", "center_y", "e1", "e2", "n_sersic"]
    kinds, args = tfd.prior_table(prior)
    assert len(kinds) == 22 and len(args) == 88


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", "/nonexistent/libgigalens_amd.so")
    with pytest.raises(nat.NativeLibraryMissing):
        nat.lib()
