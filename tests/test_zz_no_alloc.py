"""GPU: the allocation-free ASP filter entry point
(sk_asp_simplify_filter_device_noalloc) gives exactly what the allocating
device entry point gives, on two calls in a row (the second smaller), leaves
the library's device pool untouched, refuses inputs beyond its static
scratch, and recovers after a malformed input."""
import ctypes as C

import numpy as np
import pytest

import strata_k as sk

pytestmark = pytest.mark.gpu


class Lcg:
    def __init__(self, s):
        self.s = s

    def upto(self, n):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return (self.s >> 11) % n


def ground_program(seed, n_rules, n_atoms, max_body):
    rng = Lcg(seed)
    head, bs, bl = [], [0], []
    for _ in range(n_rules):
        head.append(-1 if rng.upto(8) == 0 else rng.upto(n_atoms))
        for _ in range(rng.upto(max_body + 1)):
            a, positive = rng.upto(n_atoms), rng.upto(2) == 1
            bl.append(a + 1 if positive else -(a + 1))
        bs.append(len(bl))
    poss = [1 if rng.upto(4) != 0 else 0 for _ in range(n_atoms)]
    cert = [1 if poss[i] == 1 and rng.upto(3) == 0 else 0 for i in range(n_atoms)]
    return head, bs, bl, poss, cert


def filter_on_device(ctx, prog, noalloc):
    import torch
    head, bs, bl, poss, cert = prog
    dev = "cuda:0"
    t = [torch.tensor(np.asarray(a, dt), device=dev) for a, dt in
         ((head, np.int32), (np.asarray(bs, np.uint32).view(np.int32), np.int32), (bl, np.int32),
          (poss, np.uint8), (cert, np.uint8))]
    oh = torch.full((max(1, len(head)),), -7, dtype=torch.int32, device=dev)
    ob = torch.full((len(head) + 1,), -7, dtype=torch.int32, device=dev)
    ol = torch.full((max(1, len(bl)),), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    nr, nl = ctx.asp_simplify_filter_device(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), len(head), len(bl),
                                            t[3].data_ptr(), t[4].data_ptr(), len(poss), oh.data_ptr(), ob.data_ptr(),
                                            ol.data_ptr(), noalloc=noalloc)
    return (nr, nl, oh[:nr].cpu().numpy().tolist(), ob[:nr + 1].cpu().numpy().view(np.uint32).tolist(),
            ol[:nl].cpu().numpy().tolist())


def test_noalloc_equals_allocating_entry(ctx):
    L = ctx.L
    progs = [ground_program(11, 300_000, 4096, 6), ground_program(12, 20_000, 512, 5), ground_program(13, 0, 8, 2)]
    for i, prog in enumerate(progs):  # each call smaller than the one before
        want = filter_on_device(ctx, prog, noalloc=False)
        held = L.sk_ctx_device_bytes(ctx.h)
        got = filter_on_device(ctx, prog, noalloc=True)
        assert L.sk_ctx_device_bytes(ctx.h) == held, "the pool grew during the allocation-free call"
        assert got == want, i
    assert want[:2] == (0, 0) and want[3] == [0]


def test_noalloc_limits_and_malformed_input(ctx):
    import torch
    # beyond the static scratch: refused before any device work
    buf = torch.zeros(16, dtype=torch.int32, device="cuda:0")
    p = buf.data_ptr()
    nr, nl = C.c_uint64(), C.c_uint64()
    for n_rules, n_atoms in ((1 << 27) + 1, 4), (4, (1 << 22) + 1):
        code = ctx.L.sk_asp_simplify_filter_device_noalloc(ctx.h, p, p, p, n_rules, 0, p, p, n_atoms, p, p, p,
                                                          C.byref(nr), C.byref(nl))
        assert code == 104  # SK_ERR_INVALID_ARGUMENT
        info = sk.engine.SkErrorInfo()
        ctx.L.sk_ctx_error(ctx.h, C.byref(info))
        assert b"allocation-free" in info.message
    # a malformed literal raises the flag; the next call starts clean
    bad = ([1], [0, 1], [9], [1, 1], [0, 0])
    with pytest.raises(sk.EngineError):
        filter_on_device(ctx, bad, noalloc=True)
    good = ground_program(14, 5000, 64, 4)
    assert filter_on_device(ctx, good, noalloc=True) == filter_on_device(ctx, good, noalloc=False)
