"""Serving: continuous batching under an HBM budget.  A request's result is
the same alone and in the mix; a request larger than the budget is rejected
while the others finish; queued requests can be cancelled."""
import numpy as np
import pytest

from petakit5d_amd import synth

pytestmark = pytest.mark.gpu


def _vol(shape, seed):
    return synth.rng(seed, 0).integers(80, 400, size=shape).astype(np.uint16)


def test_mixed_stream_equals_single_calls(oracle):
    import petakit5d_amd as pk
    print("[serve] mixed stream", flush=True)
    psf1 = synth.gaussian_psf((9, 9, 11), 1.0, 2.5)
    psf2 = synth.gaussian_psf((9, 9, 13), 1.3, 3.2)
    ctx = pk.Context(0, 0)
    srv = pk.Server(ctx, max_active=3)  # forces queueing: later requests join as others finish
    jobs = []
    for i, (shape, meth, it, psf) in enumerate([((64, 48, 32), "omw", 2, psf1), ((50, 40, 24), "rl", 5, psf2),
                                                ((96, 64, 30), "omw", 2, psf2), ((48, 48, 40), "rl", 3, psf1),
                                                ((40, 36, 20), "omw", 3, psf1), ((72, 60, 27), "rl", 4, psf2)]):
        I = _vol(shape, 30 + i)
        pb = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
        jobs.append((srv.submit(I, psf, pb, it, method=meth, Background=100, save16bit=False), I, psf, pb, meth, it))
    for req, I, psf, pb, meth, it in jobs:
        J = req.result(timeout_s=60)
        inf = req.info()
        assert inf["latency_ms"] > 0 and inf["iterations_done"] == it
        if meth == "omw":
            R, _ = pk.decon_lucy_omw_function(I, psf, pb, it, Background=100, save16bit=False, pad=True, ctx=ctx)
        else:
            R, _, _ = pk.decon_lucy_function(I, psf, it, Background=100, save16bit=False, pad=True, ctx=ctx)
        np.testing.assert_array_equal(J, R)
    st = srv.stats()
    assert st["completed"] == 6 and st["failed"] == 0
    # served results against the oracle (one request per method)
    import pk_oracle
    from conftest import rel_l2
    for req, I, psf, pb, meth, it in (jobs[0], jobs[1]):
        J = req.result(timeout_s=60)
        if meth == "omw":
            R, _ = pk_oracle.decon_lucy_omw_function(I, psf, pb, it, Background=100, save16bit=False)
        else:
            R, _, _ = pk_oracle.decon_lucy_function(I, psf, it, Background=100, save16bit=False)
        e = rel_l2(J, R)
        print(f"  served {meth}-{it} {I.shape}: rel-L2 vs oracle FP32 {e:.3e}", flush=True)
        assert e <= 1e-4
    srv.stop()
    ctx.close()


def test_budget_rejects_oversized_request():
    import petakit5d_amd as pk
    print("[serve] budget", flush=True)
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    ctx = pk.Context(0, 64 << 20)  # 64 MiB budget
    srv = pk.Server(ctx, max_active=4)
    small = [srv.submit(_vol((32, 32, 16), 50 + i), psf, psf, 2, method="omw", Background=100) for i in range(3)]
    big = srv.submit(_vol((256, 256, 128), 60), psf, psf, 2, method="omw", Background=100)
    for r in small:
        r.result(timeout_s=60)
    with pytest.raises(pk.PkDeconError) as e:
        big.result(timeout_s=60)
    assert e.value.status == "PK_ERR_BUDGET"
    srv.stop()
    ctx.close()


def test_cancel_queued_request():
    import petakit5d_amd as pk
    psf = synth.gaussian_psf((7, 7, 9), 1.0, 2.5)
    ctx = pk.Context(0, 0)
    srv = pk.Server(ctx, max_active=1)
    reqs = [srv.submit(_vol((64, 64, 32), 70 + i), psf, psf, 3, method="omw", Background=100) for i in range(4)]
    last = reqs[-1]
    try:
        last.cancel()
        cancelled = True
    except pk.PkDeconError as err:  # it may already be running on a fast device
        cancelled = False
        assert err.status == "PK_ERR_BUSY"
    for r in reqs[:-1]:
        r.result(timeout_s=60)
    if cancelled:
        assert last.info()["state"] == 4
    srv.stop()
    ctx.close()


def test_injected_failure_fails_only_that_request(monkeypatch):
    """A request that fails mid-flight (forced from the host: the scheduler
    fails request 3 before its third iteration, PK_SERVE_INJECT_FAILURE) fails
    alone: its error reaches the caller, its memory and OTF pins are released
    after its queued steps have run, the requests beside it and one submitted
    after the failure equal their standalone calls bitwise, and the server's
    counters record one failure."""
    import petakit5d_amd as pk
    print("[serve] injected failure of request 3 at iteration 2", flush=True)
    monkeypatch.setenv("PK_SERVE_INJECT_FAILURE", "3:2")
    psf1 = synth.gaussian_psf((9, 9, 11), 1.0, 2.5)
    psf2 = synth.gaussian_psf((9, 9, 13), 1.3, 3.2)
    ctx = pk.Context(0, 0)
    srv = pk.Server(ctx, max_active=8)
    spec = [((64, 48, 32), "omw", 2, psf1), ((50, 40, 24), "rl", 5, psf2), ((56, 44, 28), "rl", 6, psf1),
            ((96, 64, 30), "omw", 3, psf2), ((48, 48, 40), "rl", 4, psf1)]
    jobs = []
    for i, (shape, meth, it, psf) in enumerate(spec):
        I = _vol(shape, 90 + i)
        pb = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
        jobs.append((srv.submit(I, psf, pb, it, method=meth, Background=100, save16bit=False), I, psf, pb, meth, it))
    assert [j[0].id for j in jobs] == [1, 2, 3, 4, 5]

    def standalone(I, psf, pb, meth, it):
        if meth == "omw":
            return pk.decon_lucy_omw_function(I, psf, pb, it, Background=100, save16bit=False, pad=True, ctx=ctx)[0]
        return pk.decon_lucy_function(I, psf, it, Background=100, save16bit=False, pad=True, ctx=ctx)[0]

    for n, (req, I, psf, pb, meth, it) in enumerate(jobs):
        if n == 2:
            with pytest.raises(pk.PkDeconError) as e:
                req.result(timeout_s=60)
            assert e.value.status == "PK_ERR_INVALID_ARG" and "injected failure" in str(e.value)
            inf = req.info()
            assert inf["state"] == 3 and inf["iterations_done"] == 2  # FAILED after two iterations
            continue
        np.testing.assert_array_equal(req.result(timeout_s=60), standalone(I, psf, pb, meth, it))
    # the server keeps serving after the failure
    I6 = _vol((40, 36, 20), 99)
    r6 = srv.submit(I6, psf1, psf1[::-1, ::-1, ::-1].copy(), 3, method="omw", Background=100, save16bit=False)
    np.testing.assert_array_equal(r6.result(timeout_s=60),
                                  standalone(I6, psf1, psf1[::-1, ::-1, ::-1].copy(), "omw", 3))
    st = srv.stats()
    assert st["completed"] == 5 and st["failed"] == 1 and st["active"] == 0
    srv.stop()
    ctx.close()


def test_reset_between_calls_rebuilds_otfs_identically():
    """pk_reset (the reference's reset(gpuDevice) between runs) drops the
    cached OTFs and the workspace; the next calls rebuild them and return the
    same bits, for single calls, a batch and a served request on the same
    context."""
    import petakit5d_amd as pk
    print("[serve] reset between calls", flush=True)
    psf = synth.gaussian_psf((9, 9, 11), 1.1, 2.8)
    pb = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    I = _vol((72, 60, 36), 120)
    chunks = [_vol((48, 40, 24), 121), _vol((60, 52, 30), 122)]
    ctx = pk.Context(0, 0)

    def run():
        a = pk.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False, pad=True, ctx=ctx)[0]
        b = pk.decon_lucy_function(I, psf, 4, Background=100, save16bit=False, pad=True, ctx=ctx)[0]
        c = pk.decon_batch(chunks, psf, pb, 2, method="omw", Background=100, save16bit=False, ctx=ctx)
        return [a, b] + list(c)

    first = run()
    assert ctx.device_bytes() > 0
    ctx.reset()
    assert ctx.device_bytes() == 0
    again = run()
    for x, y in zip(first, again):
        np.testing.assert_array_equal(x, y)
    ctx.reset()
    srv = pk.Server(ctx, max_active=2)
    J = srv.submit(I, psf, pb, 2, method="omw", Background=100, save16bit=False).result(timeout_s=60)
    np.testing.assert_array_equal(J, first[0])
    srv.stop()
    ctx.close()
