"""pbf_mlp_loss_grad_noalloc (the entry point for callers that cannot allocate) against
pbf_mlp_loss_grad on the same inputs: bit-identical gradients and loss, on two calls in a row,
the second with fewer samples (the shared scratch keeps the first call's rows beyond the second
call's workgroups, which the reduction must not read)."""
import numpy as np
import pytest

from pbf_amd import DeviceBuffer, Mlp

pytestmark = pytest.mark.gpu


def test_noalloc_gradient_equals_the_allocating_entry_point():
    rng = np.random.default_rng(21)
    m = Mlp(seed=4)
    p = (rng.standard_normal(m.nparams) * 0.2).astype(np.float32)
    m.set_params(p)
    m.set_normalization(rng.standard_normal(16).astype(np.float32) * 0.1, np.full(16, 0.8, np.float32))
    nmax = 20000
    x = rng.standard_normal((nmax, 16)).astype(np.float32)
    t = rng.standard_normal(nmax).astype(np.float32)
    dx = DeviceBuffer(x.nbytes)
    dt = DeviceBuffer(t.nbytes)
    dg = DeviceBuffer((m.nparams + 1) * 4)
    dx.from_host(x)
    dt.from_host(t)
    for n in (nmax, 700):  # 256 workgroups (the scratch's maximum), then 11
        print(f"[case] noalloc gradient n={n}", flush=True)
        loss, g = m.loss_grad(x[:n], t[:n])
        m.loss_grad_noalloc(dx.ptr, dt.ptr, n, dg.ptr)
        out = dg.to_host(np.float32, m.nparams + 1)
        assert np.array_equal(out[:-1].view(np.uint32), g.view(np.uint32)), f"gradient n={n}"
        assert np.float32(out[-1] / np.float32(n)) == np.float32(loss), f"loss n={n}"
        assert np.abs(g).max() > 0
    for b in (dx, dt, dg):
        b.close()
    m.close()
