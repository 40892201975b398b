"""Serving form on the GPU: requests of mixed sizes, boundary conditions and precisions,
admitted by HBM footprint and run as concurrent batches.

The bar: every result is the same alone and in the mix (bitwise, since a member's
arithmetic does not depend on its batch), FP64 and mixed results meet 1e-12 against the
oracle, admission under a small pool still completes every request, and requests that can
never fit are rejected at submission.
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

# (shape (nz, ny, nx), bc, precision)
REQUESTS = [
    ((17, 48, 32), "PPPPNN", 0),
    ((40, 32, 64), "NNNNDD", 1),
    ((33, 48, 48), "DDDDDD", 0),
    ((16, 64, 96), "DDNNNN", 1),
    ((100, 32, 32), "PPPPDD", 0),
    ((64, 96, 64), "DDNDDD", 1),
    ((5, 32, 48), "NNNNNN", 0),
    ((128, 64, 32), "NNDDND", 1),
    ((40, 32, 64), "PPPPNN", 1),
    ((17, 48, 32), "DDDDNN", 0),
]


def rel_maxdiff(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_server_alone_equals_mix():
    print("server alone vs mix", flush=True)
    rhs = [prng.uniform_pm1(shape, seed=500 + i) for i, (shape, _, _) in enumerate(REQUESTS)]
    with pa.Server(hbm_bytes=1 << 30, staging_bytes=256 << 20, streams=4) as S:
        alone = []
        for (shape, bc, prec), f in zip(REQUESTS, rhs):
            u, lat = S.wait(S.submit(shape, bc, precision=prec, rhs=f))
            assert lat > 0
            alone.append(u)
        tickets = [S.submit(shape, bc, precision=prec, rhs=f) for (shape, bc, prec), f in zip(REQUESTS, rhs)]
        mixed = [S.wait(t)[0] for t in tickets]
        st = S.stats()
        assert st["completed"] == 2 * len(REQUESTS)
    for i, ((shape, bc, prec), f) in enumerate(zip(REQUESTS, rhs)):
        assert np.array_equal(alone[i], mixed[i]), f"request {i} {bc} p{prec}: alone != in mix"
        ref = oracle.solve(f, bc, zero_mode=1)
        err = rel_maxdiff(mixed[i], ref)
        assert err <= 1e-12, f"request {i} {bc} {shape} p{prec}: {err:.3e}"


def test_server_admission_small_pool():
    """a pool that holds about two requests at a time still completes a queue of many"""
    print("server small pool", flush=True)
    shape = (32, 64, 64)
    with pa.Server(hbm_bytes=6 << 20, staging_bytes=64 << 20, streams=2) as S:
        tickets = [S.submit(shape, "NNNNDD", precision=i % 2, seed=900 + i) for i in range(12)]
        outs = [S.wait(t)[0] for t in tickets]
        st = S.stats()
        assert st["completed"] == 12 and st["batches"] >= 6
    for i in (0, 1):
        ref = oracle.solve(prng.uniform_pm1(shape, seed=900 + i), "NNNNDD", zero_mode=1)
        assert rel_maxdiff(outs[i], ref) <= 1e-12


def test_server_rejects_oversized():
    print("server reject", flush=True)
    with pa.Server(hbm_bytes=1 << 20, staging_bytes=1 << 20, streams=1) as S:
        with pytest.raises(pa.PittPackError) as e:
            S.submit((64, 64, 64), "NNNNDD")
        assert e.value.code == pa.ALLOCATION_FAIL
        # sizes nothing can solve are refused at submission (the request is planned there)
        with pytest.raises(pa.PittPackError) as e:
            S.submit((16, 2000, 32), "NNNNDD", seed=1)
        assert e.value.code == pa.UNSUPPORTED_SIZE


def test_server_exact_footprint_padded_shape():
    """A shape whose z tiles are padded (nx = 20 -> 32 columns, nz = 17 -> 32 rows) takes more
    than 24 B per point; a pool one byte short of its planned footprint rejects it at
    submission, and a pool of exactly that footprint runs it (no admission loop that can
    never place a request)."""
    print("server exact footprint", flush=True)
    shape, bc = (17, 32, 20), "NNNNDD"
    with pa.Server(hbm_bytes=64 << 20, staging_bytes=1 << 20, streams=1) as S:
        need = S.footprint(shape, bc)
    pts = 20 * 32 * 17
    assert need > 24 * pts + 65536  # the padded tiles
    with pa.Server(hbm_bytes=need - 256, staging_bytes=1 << 20, streams=1) as S:
        with pytest.raises(pa.PittPackError) as e:
            S.submit(shape, bc, seed=3)
        assert e.value.code == pa.ALLOCATION_FAIL
    f = prng.uniform_pm1(shape, seed=3)
    with pa.Server(hbm_bytes=need, staging_bytes=1 << 20, streams=1) as S:
        u, _ = S.wait(S.submit(shape, bc, rhs=f))
    assert rel_maxdiff(u, oracle.solve(f, bc, zero_mode=1)) <= 1e-12


def test_server_large_request_not_starved():
    """A large request that fits only the empty pool, queued behind small ones in flight and
    ahead of many more small requests: it may be overtaken by at most PP_OPT_MAX_OVERTAKE admitted
    batches, so it runs within (limit + streams + 1) launches of its submission instead of
    after the whole stream of small requests."""
    print("server aging", flush=True)
    big, small, bc = (96, 128, 128), (16, 32, 32), "NNNNDD"
    limit, streams = 2, 2
    with pa.Server(hbm_bytes=64 << 20, staging_bytes=1 << 20, streams=1) as S:
        nb, ns = S.footprint(big, bc), S.footprint(small, bc)
    # the big request fits only an empty pool; small ones fit beside each other
    with pa.Server(hbm_bytes=nb + ns // 2, staging_bytes=256 << 20, streams=streams) as S:
        S.set_option(pa.OPT_BATCH_MEMBERS, 2)
        S.set_option(pa.OPT_MAX_OVERTAKE, limit)
        head = [S.submit(small, bc, seed=10 + i) for i in range(4)]
        tb = S.submit(big, bc, seed=99)
        tail = [S.submit(small, bc, seed=100 + i) for i in range(60)]
        seqs = [S.wait(t, with_batch=True)[2] for t in head]
        u, _, seq_big = S.wait(tb, with_batch=True)
        tail_seqs = [S.wait(t, with_batch=True)[2] for t in tail]
        st = S.stats()
    assert st["completed"] == 65
    before = sum(1 for q in tail_seqs if q < seq_big)
    print(f"big request ran in batch {seq_big}; {before} of 60 later requests ran before it", flush=True)
    assert seq_big <= max(seqs) + limit + streams + 1
    assert before <= 2 * (limit + streams)
    ref = oracle.solve(prng.uniform_pm1(big, seed=99), bc, zero_mode=1)
    assert rel_maxdiff(u, ref) <= 1e-12
