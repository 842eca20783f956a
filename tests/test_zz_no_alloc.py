"""The allocation-free sketch entry point (mpisdf_kr_sketch_device_noalloc)
against mpisdf_kr_sketch_device on the same inputs: bit-identical sketch,
norms and non-finite count, on two calls in a row whose second is smaller
(the shared device scratch keeps the first call's partial sums, which the
second must not see), and a clear error past the scratch's size."""
import numpy as np
import pytest

from mp_isdf_amd import capi

pytestmark = pytest.mark.gpu


def _case(torch, rng, n, d, ell, form):
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), device="cuda")
    dp = (d + 15) // 16 * 16
    seg = {2: 1, 3: 2}[form]
    esz = torch.uint8 if form == 3 else torch.bfloat16
    u, v = rng.normal(size=(d, ell)), rng.normal(size=(d, ell))
    ut = torch.empty((ell, seg * dp), dtype=esz, device="cuda")
    vt = torch.empty((ell, seg * dp), dtype=esz, device="cuda")
    capi.call("mpisdf_sketch_prepare_omega", dev(u), dev(v), d, ell, dp, form, ut, vt, None)
    phi = rng.normal(size=(n, d)) * np.exp(-rng.uniform(0, 6, size=(n, 1)))
    sw = dev(rng.uniform(0.1, 1.0, size=n))
    scale = torch.empty(n, dtype=torch.float32, device="cuda")
    if form == 2:
        p = torch.zeros((n, dp), dtype=torch.bfloat16, device="cuda")
        p[:, :d] = dev(phi).to(torch.bfloat16)
        capi.call("mpisdf_sketch_scales_bf16", sw, n, ell, scale, None)
    else:
        p = torch.zeros((n, 2 * dp), dtype=torch.uint8, device="cuda")
        capi.call("mpisdf_sketch_quantize_fp8", dev(phi).float().contiguous(), d, n, d, sw, ell, p, dp, scale, None)
    return ut, vt, p, scale, seg * dp


def _run(torch, entry, form, ut, vt, p, scale, ell, k, n):
    y = torch.full((n, ell), float("nan"), dtype=torch.float32, device="cuda")
    yn = torch.zeros(n, dtype=torch.float64, device="cuda")
    bad = torch.zeros(1, dtype=torch.int64, device="cuda")
    capi.call(entry, 3 if form == 3 else 2, ut, vt, ell, k, p, k, n, scale, y, ell, yn, bad, None)
    torch.cuda.synchronize()
    return y, yn, bad


def test_noalloc_sketch_equals_allocating_entry(mp_isdf):
    import torch
    rng = np.random.default_rng(29)
    for n, d, ell, form in ((20000, 200, 700, 2), (700, 75, 300, 3)):
        ut, vt, p, scale, k = _case(torch, rng, n, d, ell, form)
        ya, na, ba = _run(torch, "mpisdf_kr_sketch_device", form, ut, vt, p, scale, ell, k, n)
        yb, nb, bb = _run(torch, "mpisdf_kr_sketch_device_noalloc", form, ut, vt, p, scale, ell, k, n)
        assert torch.equal(ya, yb)
        assert torch.equal(na, nb)
        assert int(ba.item()) == int(bb.item()) == 0
        assert bool(torch.isfinite(yb).all())


def test_noalloc_sketch_rejects_sizes_past_its_scratch(mp_isdf):
    import torch
    ell, n = 1000, (1 << 22) // 4 + 1  # ceil(1000 / 256) = 4 row blocks per point
    with pytest.raises(ValueError, match="allocation-free"):
        capi.call("mpisdf_kr_sketch_device_noalloc", 2, None, None, ell, 64, None, 64, n, None,
                  torch.empty(1, device="cuda"), ell, torch.empty(1, device="cuda"), torch.empty(1, device="cuda"),
                  None)
