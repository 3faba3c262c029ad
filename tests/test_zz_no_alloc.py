"""ipt_hip_mlp_gradient_noalloc (the gradient entry point that allocates
nothing: its partial gradients live in a fixed device array) against
ipt_hip_mlp_gradient, through the C ABI: bit for bit, on two calls in a row
(the second with fewer samples, so fewer partial slices than the first call
left in the shared array), on the null stream and on a handle's stream."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402

P = api.MLP_NPARAM


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    p[0:192] = rng.standard_normal(192) * 0.8
    p[192:256] = rng.standard_normal(64) * 0.3
    p[256:4352] = rng.standard_normal(4096) / 8.0
    p[4352:4416] = rng.standard_normal(64) * 0.3
    p[4416:4480] = rng.standard_normal(64) / 8.0
    p[4480] = 0.1
    return p


def _samples(count, seed):
    rng = np.random.default_rng(seed)
    feat = rng.standard_normal((count, 3)).astype(np.float32)
    q = rng.standard_normal(count).astype(np.float32)
    t = rng.standard_normal(count).astype(np.float32)
    w = rng.uniform(0.1, 2.0, count).astype(np.float32)
    return feat, q, t, w


def test_noalloc_gradient_equals_gradient():
    hip = api.hip()
    handle = api.create_handle()
    ptrs = []

    def put(a):
        p = hip.to_device(np.ascontiguousarray(a))
        ptrs.append(p)
        return p

    try:
        p = _params(11)
        dp = put(p)
        dg, dl = put(np.zeros(P, np.float32)), put(np.zeros(1, np.float32))
        dg2, dl2 = put(np.zeros(P, np.float32)), put(np.zeros(1, np.float32))
        # 70000 samples: all 1024 partial slices; then 5003: 79 slices
        for count, seed, h in [(70000, 1, None), (5003, 2, None), (130, 3, handle), (70000, 4, handle)]:
            print("no-alloc gradient: count %d, %s stream" % (count, "handle" if h else "null"), flush=True)
            feat, q, t, w = _samples(count, seed)
            dfe, dq, dt, dw = put(feat.ravel()), put(q), put(t), put(w)
            api.mlp_gradient(dp, dfe, dq, dt, dw, count, dg, dl, h)
            api.mlp_gradient_noalloc(dp, dfe, dq, dt, dw, count, dg2, dl2, h)
            g, loss = hip.to_host(dg, (P,), np.float32), hip.to_host(dl, (1,), np.float32)[0]
            g2, loss2 = hip.to_host(dg2, (P,), np.float32), hip.to_host(dl2, (1,), np.float32)[0]
            assert np.array_equal(g, g2) and loss == loss2, (count, np.abs(g - g2).max(), loss, loss2)
            if count <= 5003:
                # and both are the gradient (FP32 against the FP64 oracle)
                gref, lref = orc.mlp_grad(*(a.astype(np.float64) for a in (p, feat.ravel(), q, t, w)))
                assert abs(loss2 - lref) <= 1e-5 * lref
                assert np.abs(g2 - gref).max() <= 1e-4 * np.abs(gref).max()
    finally:
        for ptr in ptrs:
            hip.free(ptr)
        api.destroy_handle(handle)


def test_noalloc_gradient_rejects_bad_arguments():
    with pytest.raises(api.IptError) as e:
        api.mlp_gradient_noalloc(0, 0, 0, 0, 0, 10, 0, 0)
    assert e.value.status == 1
