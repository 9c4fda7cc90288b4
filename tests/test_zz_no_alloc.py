"""tcp_force_direct_noalloc (no device allocation, scratch in a static device
array) against tcp_force_direct on the same positions, through the C ABI.

Tolerances: bit for bit in both precisions (same kernels, block shape and
summation order); RDF counts identical.
Two calls in a row share the scratch array, the second on a smaller system.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

from tcp_allpair_amd import gpu
from tcp_allpair_amd.gpu import Context
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import FP64, KELBG, MIXED, RECIP_NONE, CParams

HIST_STRIDE = 512  # device RDF row length


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(p, precision, x, with_rdf=True):
    """Device SoA inputs and outputs (torch tensors), one call, then a sync."""
    lib = gpu.load()
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
    f = torch.full_like(xd, float("nan"))
    ur = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    rdf = torch.zeros(3 * HIST_STRIDE, dtype=torch.int64, device=dev) if with_rdf else None
    stream = torch.cuda.current_stream()
    c = CParams.from_params(p)
    ptrs = [t.data_ptr() for t in (xd[0], xd[1], xd[2], f[0], f[1], f[2], ur)]
    rc = lib.tcp_force_direct_noalloc(ctypes.byref(c), int(precision), *ptrs,
                                      rdf.data_ptr() if with_rdf else None,
                                      ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, lib.tcp_last_error(None).decode()
    torch.cuda.synchronize()
    h = rdf.cpu().numpy().reshape(3, HIST_STRIDE) if with_rdf else None
    return f.cpu().numpy().T.copy(), ur.item(), (None if h is None else h[:, :501].astype(float))


def context(p, precision, x):
    with Context(p, precision) as c:
        c.set_state(x=x)
        ur = c.force_direct()
        f = c.get_state()[2]
        rdf, _ = c.histograms()
    return f, ur, rdf


@pytest.mark.gpu
@pytest.mark.parametrize("precision,flags", [(FP64, 0), (MIXED, 0), (MIXED, KELBG)])
def test_noalloc_matches_force_direct(precision, flags):
    """Two calls in a row on one GPU (8000 particles, then 2500): each equals
    tcp_force_direct on a context of the same system."""
    for (ne, ni), seed in (((4000, 4000), 301), ((1300, 1200), 302)):
        say("case noalloc", precision, flags, ne + ni)
        p, x, _ = make_plasma(ne, ni, seed=seed, recip=RECIP_NONE)
        p = p.replace(flags=flags)
        f, ur, rdf = noalloc(p, precision, x)
        fc, urc, rdfc = context(p, precision, x)
        assert np.all(np.isfinite(f))
        np.testing.assert_array_equal(f, fc)
        assert ur == urc
        np.testing.assert_array_equal(rdf, rdfc)
    # without an RDF buffer: the same forces
    f2, ur2, _ = noalloc(p, precision, x, with_rdf=False)
    np.testing.assert_array_equal(f2, fc)
    assert ur2 == urc


def test_noalloc_rejects_large_systems_before_device_use():
    lib = gpu.load()
    p, _, _ = make_plasma(4097, 4096, recip=RECIP_NONE)
    c = CParams.from_params(p)
    one = ctypes.c_void_p(8)  # never dereferenced: the size check comes first
    rc = lib.tcp_force_direct_noalloc(ctypes.byref(c), FP64, one, one, one, one, one, one, one,
                                      None, None)
    assert rc == -1 and "8192" in lib.tcp_last_error(None).decode()
    rc = lib.tcp_force_direct_noalloc(ctypes.byref(c), FP64, None, one, one, one, one, one, one,
                                      None, None)
    assert rc == -1
