"""The allocation-free cell sort (dk_sort_particles_noalloc) against dk_sort_particles, through
the C-ABI.  Two contexts hold the same particles; one sorts with the allocating entry point,
the other with the allocation-free one, and every column must come out identical, bit for bit
(integer and index work), on two calls in a row, the second with fewer particles.  A count
beyond DK_SORT_NOALLOC_MAX is refused before any kernel runs."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 262144   # DK_SORT_NOALLOC_MAX of include/dustk.h


def say(msg):
    print("[no-alloc] " + msg, flush=True)
    sys.stdout.flush()


def _contexts(w, cap):
    from athenak_amd.lib import DustContext
    a = DustContext(w.mesh, w.blocks, w.phys, cap, form=abi.DK_FORM_FP64)
    b = DustContext(w.mesh, w.blocks, w.phys, cap, form=abi.DK_FORM_FP64)
    return a, b


def _load_same(a, b, rdata, idata):
    """The same particle columns (device tensors, n columns) into both contexts, bound."""
    import torch
    n = rdata.shape[1]
    for g in (a, b):
        g.rdata.zero_()
        g.idata.zero_()
        g.rdata[:, :n] = rdata
        g.idata[:, :n] = idata
    torch.cuda.synchronize()
    a.bind_particles(n)
    b.bind_particles(n)


def _sort_noalloc(g):
    from athenak_amd.lib import lib
    g._check(lib().dk_sort_particles_noalloc(g.h), "dk_sort_particles_noalloc")


def _columns(g):
    g.stream.synchronize()
    n = g.npart
    return g.rdata[:, :n].cpu().numpy(), g.idata[:, :n].cpu().numpy()


def test_noalloc_sort_equals_the_allocating_sort_on_two_calls():
    import torch
    # 16^3 cells in 8 blocks, 3 particles per cell (12,288), moved for 3 steps so that they
    # cross cells and blocks and the order is no longer the set-up's
    w = workloads.c1(n=16, nb=2, ppc=3)
    a, b = _contexts(w, w.npart + 16)
    a.setup(w.setup)
    for _ in range(3):
        a.step(w.dt, w.src)
    a.check_device_error()
    a.stream.synchronize()
    n1 = a.npart
    r1, i1 = a.rdata[:, :n1].clone(), a.idata[:, :n1].clone()
    _load_same(a, b, r1, i1)
    say("call 1: %d particles" % n1)
    a.sort()
    _sort_noalloc(b)
    ra, ia = _columns(a)
    rb, ib = _columns(b)
    assert not np.array_equal(ia, i1.cpu().numpy()), "the input was already in cell order"
    np.testing.assert_array_equal(ia, ib)
    np.testing.assert_array_equal(ra.view(np.uint64), rb.view(np.uint64))

    # call 2: fewer particles (a third), shuffled, into the same contexts
    n2 = n1 // 3
    gen = torch.Generator().manual_seed(7)
    pick = torch.randperm(n1, generator=gen)[:n2].to(r1.device)
    r2, i2 = r1[:, pick].contiguous(), i1[:, pick].contiguous()
    _load_same(a, b, r2, i2)
    say("call 2: %d particles" % n2)
    a.sort()
    _sort_noalloc(b)
    ra, ia = _columns(a)
    rb, ib = _columns(b)
    assert ia.shape[1] == n2
    np.testing.assert_array_equal(ia, ib)
    np.testing.assert_array_equal(ra.view(np.uint64), rb.view(np.uint64))
    a.check_device_error()
    b.check_device_error()
    a.close()
    b.close()


def test_noalloc_sort_refuses_counts_beyond_its_maximum():
    from athenak_amd.lib import DustError
    w = workloads.c1(n=8, nb=1, ppc=1)
    _, b = _contexts(w, NOALLOC_MAX + 1)
    b.rdata.fill_(0.25)
    b.idata.zero_()
    b.bind_particles(NOALLOC_MAX + 1)
    say("call beyond the maximum")
    with pytest.raises(DustError, match="capacity|maximum|size"):
        _sort_noalloc(b)
    b.stream.synchronize()
    # nothing ran: the columns are as they were
    assert float(b.rdata[:, :NOALLOC_MAX + 1].min()) == 0.25
    b.close()
