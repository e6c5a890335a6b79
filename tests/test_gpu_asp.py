"""GPU: the grounding simplification filter against the oracle's restatement
of the reference's CPU filter, on the reference's own random corpus
(crates/strata-gpu/tests/asp_diff.rs:24-70: 200 seeds of 64 rules over 24
atoms, bodies up to 5, cert within poss, 1/8 constraints) and on a large
instantiated rule set; malformed literals are refused."""
import numpy as np
import pytest

import strata_k as sk
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class Lcg:
    def __init__(self, s):
        self.s = s

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return self.s

    def upto(self, n):
        return (self.next() >> 11) % n


def random_program(seed, n_rules, n_atoms, max_body):
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


def same(got, want):
    gh, gb, gl = got
    wh, wb, wl = want
    return gh.tolist() == wh and gb.tolist() == wb and gl.tolist() == wl


def test_reference_corpus(ctx):
    for seed in range(200):
        prog = random_program((seed * 0x9E3779B9 + 1) & 0xFFFFFFFFFFFFFFFF, 64, 24, 5)
        assert same(sk.asp_simplify_filter(*prog, ctx=ctx), O.asp_simplify_filter(*prog)), seed


def test_large_rule_set(ctx):
    prog = random_program(0xA5A5, 300_000, 4096, 6)
    got = sk.asp_simplify_filter(*prog, ctx=ctx)
    assert same(got, O.asp_simplify_filter(*prog))
    assert 0 < len(got[0]) < 300_000


@pytest.mark.parametrize("n_atoms", [16384, 16385, 100_000])
def test_atom_table_sizes(ctx, n_atoms):
    """Atom states staged in LDS (up to 16384 atoms) and gathered from the
    global tables beyond."""
    prog = random_program(n_atoms, 50_000, n_atoms, 6)
    assert same(sk.asp_simplify_filter(*prog, ctx=ctx), O.asp_simplify_filter(*prog))


def test_long_bodies(ctx):
    """Bodies up to 40 literals (read back in several chunks), with mostly
    possible atoms so that long rules survive."""
    for seed in range(6):
        head, bs, bl, poss, cert = random_program(seed + 17, 6000, 300, 40)
        poss = [1 if i % 50 else 0 for i in range(300)]
        cert = [1 if i % 7 == 0 and poss[i] else 0 for i in range(300)]
        want = O.asp_simplify_filter(head, bs, bl, poss, cert)
        assert max(want[1][i + 1] - want[1][i] for i in range(len(want[0]))) > 8
        assert same(sk.asp_simplify_filter(head, bs, bl, poss, cert, ctx=ctx), want), seed


def test_empty_and_bodiless(ctx):
    assert same(sk.asp_simplify_filter([], [0], [], [1], [0], ctx=ctx), ([], [0], []))
    assert same(sk.asp_simplify_filter([3, -1], [0, 0, 0], [], [1], [1], ctx=ctx), ([3, -1], [0, 0, 0], []))


@pytest.mark.parametrize("lits", [[5], [0], [-7]])
def test_malformed_literals_refused(ctx, lits):
    with pytest.raises(sk.EngineError):
        sk.asp_simplify_filter([1], [0, len(lits)], lits, [1, 1], [0, 0], ctx=ctx)


@pytest.mark.parametrize("n_rules", [1, 0])
def test_certain_atom_must_be_possible(ctx, n_rules):
    """An atom certain but not possible is refused, with or without rules:
    the reference's CPU filter (asp.rs:53-66) and its device filter decide
    such an atom differently."""
    head, bs, bl = ([1], [0, 1], [2]) if n_rules else ([], [0], [])
    with pytest.raises(sk.EngineError):
        sk.asp_simplify_filter(head, bs, bl, [1, 0], [0, 1], ctx=ctx)
    # the same atoms with cert within poss are accepted
    assert same(sk.asp_simplify_filter(head, bs, bl, [1, 1], [0, 1], ctx=ctx),
                O.asp_simplify_filter(head, bs, bl, [1, 1], [0, 1]))


def test_device_resident_entry_and_kernel_records(ctx):
    """The device-buffer entry point equals the host-buffer one; the context's
    kernel records name the count and emit passes."""
    import torch
    head, bs, bl, poss, cert = random_program(99, 20_000, 512, 5)
    want = sk.asp_simplify_filter(head, bs, bl, poss, cert, ctx=ctx)
    dev = "cuda:0"
    t = [torch.tensor(np.asarray(a, dt), device=dev) for a, dt in
         ((head, np.int32), (np.asarray(bs, np.uint32).view(np.int32), np.int32), (bl, np.int32),
          (poss, np.uint8), (cert, np.uint8))]
    oh = torch.empty(len(head), dtype=torch.int32, device=dev)
    ob = torch.empty(len(head) + 1, dtype=torch.int32, device=dev)
    ol = torch.empty(len(bl), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx.kernel_reset()
    ctx.set_profiling(True)
    for _ in range(2):
        nr, nl = ctx.asp_simplify_filter_device(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), len(head), len(bl),
                                                t[3].data_ptr(), t[4].data_ptr(), len(poss), oh.data_ptr(),
                                                ob.data_ptr(), ol.data_ptr())
    ctx.set_profiling(False)
    assert (nr, nl) == (len(want[0]), len(want[2]))
    assert oh[:nr].cpu().numpy().tolist() == want[0].tolist()
    assert ob[:nr + 1].cpu().numpy().view(np.uint32).tolist() == want[1].tolist()
    assert ol[:nl].cpu().numpy().tolist() == want[2].tolist()
    recs = {r["name"]: r for r in ctx.kernel_records()}
    assert recs["asp_count"]["launches"] == 2 and recs["asp_emit"]["launches"] == 2
    # per rule its end offset, per literal the literal, per block of rules two totals
    rest = recs["asp_count"]["bytes"] - 2 * (4.0 * len(head) + 4.0 * len(bl))
    assert rest > 0 and rest % 32 == 0
    assert recs["asp_emit"]["bytes"] == 2 * (4.0 * len(head) + 4.0 * len(bl) + 12.0 * nr + 4.0 * nl)
