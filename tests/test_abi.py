"""The C-ABI library loads without a GPU and exports every declared symbol;
argument errors come back as status codes (no device work is started)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w\s\*]*?\b(dugma_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from dugma_amd import _lib
    return _lib.load()


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dugma_init", "dugma_update_sig", "dugma_compi", "dugma_fr_m", "dugma_energy",
                 "dugma_energy_grad", "dugma_knn"):
        assert must in names


def test_every_declared_symbol_is_exported(lib):
    from dugma_amd import _lib
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    unbound = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not unbound, unbound


def test_argument_errors_are_status_codes(lib):
    d = ctypes.POINTER(ctypes.c_double)
    out = ctypes.c_void_p()
    x = (ctypes.c_double * 9)()
    # D = 4 is rejected before any device call
    rc = lib.dugma_init(ctypes.cast(x, d), ctypes.cast(x, d), ctypes.cast(x, d), ctypes.cast(x, d),
                        1, 1, 4, ctypes.byref(out), None)
    assert rc == -1 and b"D must be 2 or 3" in lib.dugma_last_error()
    rc = lib.dugma_init(None, None, None, None, 1, 1, 3, ctypes.byref(out), None)
    assert rc == -1
    rc = lib.dugma_init(ctypes.cast(x, d), ctypes.cast(x, d), ctypes.cast(x, d), ctypes.cast(x, d),
                        20000000, 1, 3, ctypes.byref(out), None)
    assert rc == -4
    assert lib.dugma_energy(None, None, None, 0, None) == -1
    assert lib.dugma_energy_grad_hess(None, None, None, 0, None, None, None) == -1
    assert lib.dugma_update_sig(None, None, None, 1.0, 1.0, 3.0, None) == -1
    assert lib.dugma_compi(None, None, None, 1.0, 1.0, 3.0, 1.0, None) == -1
    assert lib.dugma_knn(None, 0, 8, None, None, None, None) == -1
    assert lib.dugma_fr_m(None, None) == -1
    assert lib.dugma_version().startswith(b"dugma")


def test_monomial_table_matches_reference_files():
    """The generated 637/45 monomial tables equal energy3D.m / energy2D.m."""
    ref = "/root/reference"
    if not os.path.exists(os.path.join(ref, "energy3D.m")):
        pytest.skip("reference sources not present on this machine")
    import importlib.util
    spec = importlib.util.spec_from_file_location("gm", os.path.join(ROOT, "dugma_amd", "csrc", "gen_monomials.py"))
    gm = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gm)
    m3, m2 = gm.tables(ref)
    text = open(os.path.join(ROOT, "dugma_amd", "csrc", "dugma_monomials.inc")).read()
    assert gm.emit(m3, m2) == text


def test_batch_argument_errors_before_any_device_call():
    """dugma_batch_create* validate every trial before touching the device:
    empty trials and covariances that are not s I are status -1 with a
    message; the other batch entry points reject null handles."""
    import numpy as np
    import dugma_amd as dg
    from dugma_amd import _lib
    from dugma_amd import workloads as W
    X = np.random.default_rng(0).normal(size=(3, 20))
    C = np.tile(np.eye(3), (20, 1, 1))
    C[5, 0, 1] = C[5, 1, 0] = 0.1  # not isotropic
    with pytest.raises(_lib.DugmaError) as e:
        dg.Batch([dict(X=X, X_C=W.blocks(C), Y=X, Y_C=W.blocks(np.tile(np.eye(3), (20, 1, 1))))])
    assert e.value.code == -1 and "isotropic" in str(e.value)
    with pytest.raises(_lib.DugmaError) as e:
        dg.Batch([dict(X=X[:, :0], sX=np.zeros(0), Y=X, sY=np.ones(20))])
    assert e.value.code == -1 and "empty trial" in str(e.value)
    L = _lib.load()
    assert L.dugma_batch_register(None, None, None, None, None, None, None, None, None) == -1
    assert L.dugma_batch_autotune(None, None, None, 1, None, None, None, None, None) == -1
    assert L.dugma_batch_free(None) == -1
    ch = (ctypes.c_int * 4)(1, 1, 1, 1)
    assert L.dugma_batch_set_shape(None, 2, ch, 8192) == -1
    assert L.dugma_batch_get_shape(None, 0, None, None, None) == -1
    assert L.dugma_batch_set_level_shape(None, 0, 2, ch, 8192) == -1


def test_option_structs_match_the_header(tmp_path):
    """The ctypes mirrors of the option structs have the C layout: field
    offsets and sizes from a C program compiled against include/dugma.h."""
    import shutil
    import subprocess
    from dugma_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"dugma_batch_opts": _lib.BatchOpts, "dugma_register_opts": _lib.RegisterOpts,
               "dugma_register_energy_opts": _lib.RegisterEnergyOpts}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "dugma.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c = {}
    for line in out:
        if line:
            name, field, value = line.split()
            c[(name, field)] = int(value)
    for cname, py in structs.items():
        assert c[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert c[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_pose_buffers_check_shapes():
    """The per-context pose buffers of the Python mirror take R (D x D) and t
    (D entries) only, and copy them (column-major R as the C ABI expects)."""
    import numpy as np
    import dugma_amd as dg
    b = dg._PoseBuffers(3)
    R = np.arange(9.0).reshape(3, 3)
    b.pose(R, [1.0, 2.0, 3.0])
    assert b.R.flags.f_contiguous
    np.testing.assert_array_equal(b.R, R)
    np.testing.assert_array_equal(np.frombuffer(b.R.tobytes(order="A")), R.T.ravel())
    np.testing.assert_array_equal(b.t, [1.0, 2.0, 3.0])
    with pytest.raises(ValueError):
        b.pose(np.eye(2), [0.0, 0.0, 0.0])
    with pytest.raises(ValueError):
        b.pose(np.eye(3), [0.0, 0.0])
    b2 = dg._PoseBuffers(2)
    b2.pose(np.eye(2), np.array([[1.0], [2.0]]))
    np.testing.assert_array_equal(b2.t, [1.0, 2.0])
