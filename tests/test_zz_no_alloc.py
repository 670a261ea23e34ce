"""md_fm_loss_grad_noalloc (the library's fixed device scratch, no device
allocation) against md_fm_loss_grad on the same engine: both run the same
kernels in the same order, so loss and gradient must be bit-identical.  Two
calls in a row, the second on fewer frames; two engines sharing the one
scratch array; and the capacity check."""
import sys

import numpy as np
import pytest

from md_gpu_amd import systems as S

pytestmark = pytest.mark.gpu

MD_E_CAPACITY = -4


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def frames_of(base, nf, seed):
    rng = np.random.default_rng(seed)
    fr = np.stack([S.wrap_centered(base.pos + rng.normal(0, 0.2, base.pos.shape), base.box)
                   for _ in range(nf)]).astype(np.float32)
    tg = rng.normal(0, 40.0, (nf, base.n, 3)).astype(np.float32)
    return fr, tg


def same(a, b):
    la, g1a, g2a, gqa = a
    lb, g1b, g2b, gqb = b
    assert la == lb, (la, lb)
    assert gqa == gqb, (gqa, gqb)
    np.testing.assert_array_equal(g1a, g1b)
    np.testing.assert_array_equal(g2a, g2b)


@pytest.mark.parametrize("fused", [1, 0])
def test_noalloc_matches_allocating_entry_point(fused):
    from md_gpu_amd import engine as E
    base = S.nacl(nsites=16)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=4)
    e = E.Engine.from_system(base)
    e.set_launch(fm_fused=fused)
    # first call: 6 frames, the no-allocation entry point before the allocating one has run
    fr, tg = frames_of(base, 6, seed=1)
    say("noalloc fused", fused, "6 frames")
    e.fm_set_frames(fr, tg)
    a = e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0, noalloc=True)
    b = e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0)
    assert a[0] > 0
    same(a, b)
    # second call: fewer frames (smaller input), other parameters
    fr, tg = frames_of(base, 2, seed=2)
    say("noalloc fused", fused, "2 frames")
    e.fm_set_frames(fr, tg)
    a = e.fm_loss_grad(tgt_ff.lj1, tgt_ff.lj2, tgt_q, noalloc=True)
    b = e.fm_loss_grad(tgt_ff.lj1, tgt_ff.lj2, tgt_q)
    same(a, b)
    e.close()


def test_two_engines_share_the_scratch():
    """Sequential calls from two engines on one GPU: each call writes every
    slot it reads, so what the other engine left behind does not matter."""
    from md_gpu_amd import engine as E
    b1 = S.nacl(nsites=16)
    b2 = S.nacl(nsites=12, seed=7)
    e1 = E.Engine.from_system(b1)
    e2 = E.Engine.from_system(b2)
    e1.fm_set_frames(*frames_of(b1, 4, seed=3))
    e2.fm_set_frames(*frames_of(b2, 3, seed=4))
    say("noalloc two engines")
    r1 = e1.fm_loss_grad(b1.ff.lj1, b1.ff.lj2, 1.0, noalloc=True)
    r2 = e2.fm_loss_grad(b2.ff.lj1, b2.ff.lj2, 1.0, noalloc=True)
    r1b = e1.fm_loss_grad(b1.ff.lj1, b1.ff.lj2, 1.0, noalloc=True)
    same(r1, r1b)
    same(r1, e1.fm_loss_grad(b1.ff.lj1, b1.ff.lj2, 1.0))
    same(r2, e2.fm_loss_grad(b2.ff.lj1, b2.ff.lj2, 1.0))
    e1.close()
    e2.close()


def test_noalloc_rejects_more_tiles_than_the_scratch_holds():
    """Two types: 80 bytes of per-tile sums per 64-atom tile; 1 MiB less 2 KiB
    holds 13,081 tiles, so 208 frames of 4,096 ions (13,312 tiles) are refused
    and 204 frames (13,056 tiles) are not."""
    from md_gpu_amd import engine as E
    base = S.nacl(nsites=16)
    e = E.Engine.from_system(base)
    rng = np.random.default_rng(9)
    fr = np.repeat(base.pos[None], 208, axis=0).astype(np.float32)
    tg = rng.normal(0, 40.0, fr.shape).astype(np.float32)
    say("noalloc capacity 208 frames")
    e.fm_set_frames(fr, tg)
    with pytest.raises(E.MDError) as x:
        e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0, noalloc=True)
    assert x.value.code == MD_E_CAPACITY
    say("noalloc capacity 204 frames")
    e.fm_set_frames(fr[:204], tg[:204])
    same(e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0, noalloc=True), e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0))
    e.close()
