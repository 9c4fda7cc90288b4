"""The C-ABI library without a GPU: it loads, exports every symbol that
include/tcp_gpu.h declares, and rejects bad parameters before any device call."""
import ctypes
import os
import re
import subprocess

import pytest

from tcp_allpair_amd import gpu
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import CParams, FP64, MIXED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tcp_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tcp_\w+)\s*\(", text, re.M)))


def test_header_declares_the_bound_functions():
    names = declared_functions()
    assert len(names) >= 15
    assert set(names) == set(gpu.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = gpu.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_params_struct_layout_matches_header(tmp_path):
    """ctypes' CParams against the C compiler's layout of tcp_params and tcpo_params."""
    fields = [f[0] for f in CParams._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%zu %zu\\n", offsetof(tcp_params, {f}), offsetof(tcpo_params, {f}));\n'
                   for f in fields)
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"tcp_gpu.h\"\n"
                   "#include \"tcp_oracle.h\"\nint main(void){\n"
                   '  printf("%zu %zu\\n", sizeof(tcp_params), sizeof(tcpo_params));\n'
                   + body + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"),
                    str(src), "-o", str(exe)], check=True)
    rows = [tuple(int(v) for v in line.split())
            for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
            if line]
    assert rows[0] == (ctypes.sizeof(CParams), ctypes.sizeof(CParams))
    for f, (a, b) in zip(fields, rows[1:]):
        assert a == b == getattr(CParams, f).offset, f


def _create(p, precision=FP64):
    lib = gpu.load()
    h = ctypes.c_void_p()
    c = CParams.from_params(p)
    rc = lib.tcp_create(ctypes.byref(h), ctypes.byref(c), precision)
    return rc, lib.tcp_last_error(None).decode(), h


def test_cutoff_error_before_device_use():
    p, _, _ = make_plasma(10, 10)
    rc, msg, h = _create(p.replace(rc=0.51 * p.box[0]))
    assert rc == -2 and "cut-off radius" in msg and not h.value


@pytest.mark.parametrize("change,code", [
    (dict(K=(24, 32, 32)), -4),      # not a power of two
    (dict(order=5), -4),             # odd order
    (dict(order=14), -4),
    (dict(recip=2, kmax=0), -1),
    (dict(ne=1, ni=0), -1),
    (dict(rs=0.0), -1),
    (dict(flags=4), -1),             # unknown regularisation flag
    (dict(xm=(1.0, 0.0)), -1),       # masses divide the kicks
    (dict(Te=-1.0), -1),             # negative VDF bin width (4 Te / 500)
    (dict(ne=1 << 22, ni=1), -1),    # beyond the per-context particle limit
])
def test_parameter_errors(change, code):
    p, _, _ = make_plasma(10, 10)
    rc, msg, h = _create(p.replace(**change))
    assert rc == code and msg and not h.value


@pytest.mark.parametrize("arc,precision,code", [
    (4.5, MIXED, 0), (4.6, MIXED, -4), (6.0, FP64, 0), (float("inf"), FP64, -1),
    (float("nan"), MIXED, -1)])
def test_free_alpha_range(arc, precision, code):
    """alpha > 0 overrides 3.1/rc (force.f90:23); a mixed context's alpha*rc must stay in
    [0, 4.5], the range of its FP32 erfcx table (FP64 has no such bound)."""
    p, _, _ = make_plasma(10, 10)
    rc, msg, h = _create(p.replace(alpha=arc / p.rc), precision)
    if code == 0:  # past validation: created, or stopped at the first device call
        assert rc in (0, -3), (rc, msg)
        if h.value:
            gpu.load().tcp_destroy(h)
        return
    assert rc == code and msg and not h.value


def test_bad_precision():
    p, _, _ = make_plasma(10, 10)
    rc, msg, _ = _create(p, precision=7)
    assert rc == -1


def _batch_create(ps, precision=FP64, nrep=None):
    lib = gpu.load()
    h = ctypes.c_void_p()
    arr = (CParams * len(ps))(*[CParams.from_params(p) for p in ps])
    rc = lib.tcp_batch_create(ctypes.byref(h), len(ps) if nrep is None else nrep, arr, precision)
    return rc, lib.tcp_batch_last_error(None).decode(), h


def test_batch_argument_errors_before_device_use():
    p, _, _ = make_plasma(10, 10, recip=2, kmax=3)
    rc, msg, h = _batch_create([p], nrep=0)
    assert rc == -1 and msg and not h.value
    # the replica count is checked before any parameter set is read
    rc, msg, h = _batch_create([p], nrep=70000)
    assert rc == -1 and "65535" in msg and not h.value
    rc, msg, h = _batch_create([p, p.replace(kmax=4)])
    assert rc == -1 and "share" in msg and not h.value
    rc, msg, h = _batch_create([p, p.replace(flags=1)])
    assert rc == -1 and "share" in msg and not h.value
    # SPME replicas: one spline order, brick-binned grids (every K a multiple of 8)
    s = p.replace(recip=1, K=(16, 16, 16))
    rc, msg, h = _batch_create([s, s.replace(order=6)])
    assert rc == -1 and "share" in msg and not h.value
    rc, msg, h = _batch_create([s, s.replace(K=(16, 12, 16))])
    assert rc == -4 and "batches" in msg and not h.value
    # a replica's own parameter error is reported with its index
    rc, msg, h = _batch_create([p, p.replace(rc=0.51 * p.box[0])])
    assert rc == -2 and "replica 1" in msg and not h.value
