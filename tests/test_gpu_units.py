"""Device pieces against the oracle, bit for bit: the draw transforms and the wave
heap -> postfix decode."""
import numpy as np
import pytest

from gpu_util import dev, say

pytestmark = pytest.mark.gpu


def test_draw_transforms_bit_exact(oracle, gpr):
    import torch
    n = 1 << 16
    rng = np.random.default_rng(1)
    words = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    words[:4] = [0, 0xffffffff, 0x80000000, 1]
    L = oracle.lib()
    with gpr.Context() as ctx:
        out = torch.zeros(n * 10, dtype=torch.float32, device="cuda")
        d_words = dev(words)            # keep a reference: a temporary's block is reused at once
        say("draw transforms")
        ctx.test_draws_dev(d_words.data_ptr(), n, 42, out.data_ptr())
        g = out.cpu().numpy().reshape(n, 10)
    sub = slice(0, 4096)
    ref = np.zeros((4096, 5), dtype=np.float32)
    for i, w in enumerate(words[sub]):
        u = L.gpo_u01(int(w))
        ref[i] = [u, L.gpo_normal_icdf(u), L.gpo_logf(float(np.uint32(w & 0x7fffffff).view(np.float32))),
                  L.gpo_expf(float(np.float32(np.int32(np.uint32(w).view(np.int32))) * np.float32(2 ** -24))),
                  L.gpo_sin2pi(u)]
    for j in range(5):
        a, b = g[sub, j], ref[:, j]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), (j, np.flatnonzero(~same)[:5], a[~same][:5], b[~same][:5])
    # reference transforms drawn in sequence from substream (42, chain=i): random_float,
    # random_normal, random_geometric, host random_normal, random_int(27)
    from ctypes import byref
    for i in range(512):
        st = oracle.GpoStream()
        L.gpo_stream_init(byref(st), 42, i, 0, 0, 0)
        seq = [L.gpo_stream_next(byref(st)) >> 1 for _ in range(40)]
        it = iter(seq)

        def rint(nn):
            div = 0x7fffffff // (nn + 1)
            while True:
                r = next(it) // div
                if r < nn:
                    return r

        def rfloat():
            return np.float32(np.float64(np.float32(rint(1000000))) / 1000000.0)

        rf = rfloat()
        u, v = rfloat(), rfloat()
        rn = np.float32(np.sqrt(np.float32(-2.0) * np.float32(L.gpo_logf(float(u))))) * np.float32(L.gpo_sin2pi(float(v)))
        rg = np.ceil(np.float32(L.gpo_logf(float(rfloat()))) / np.float32(float.fromhex("-0x1.26bb1cp+1")))
        hu = np.float32(next(it)) / np.float32(0x7fffffff)
        hv = np.float32(next(it)) / np.float32(0x7fffffff)
        hn = np.float32(np.sqrt(np.float32(-2.0) * np.float32(L.gpo_logf(float(hu))))) * np.float32(L.gpo_sin2pi(float(hv)))
        ri = rint(27)
        assert g[i, 5] == rf and g[i, 6] == rn and g[i, 7] == rg and g[i, 8] == hn and g[i, 9] == ri, i


def test_wave_heap_decode_matches_host(oracle, gpr):
    import torch
    rng = np.random.default_rng(2)
    nh = 4000
    hop = rng.integers(0, 27, (nh, 31)).astype(np.uint8)
    hop[:100] = rng.choice([3, 4], (100, 31))            # A/B-heavy heaps: long pass-through chains
    hc = rng.normal(0, 3, (nh, 31)).astype(np.float32)
    with gpr.Context() as ctx:
        d_ops = torch.zeros(nh * 64, dtype=torch.uint8, device="cuda")
        d_cs = torch.zeros(nh * 64, dtype=torch.float32, device="cuda")
        d_meta = torch.zeros(nh * 3, dtype=torch.int32, device="cuda")
        d_hop, d_hc = dev(hop), dev(hc)  # keep references: a temporary's block is reused at once
        say("wave heap decode")
        ctx.test_ref_decode_dev(d_hop.data_ptr(), d_hc.data_ptr(), nh, d_ops.data_ptr(), d_cs.data_ptr(),
                                d_meta.data_ptr())
        g_ops = d_ops.cpu().numpy().reshape(nh, 64)
        g_cs = d_cs.cpu().numpy().reshape(nh, 64)
        meta = d_meta.cpu().numpy().reshape(nh, 3)
    L = oracle.lib()
    for h in range(nh):
        ops, cs, _ = oracle.ref_decode(hop[h], hc[h])
        n = len(ops)
        assert meta[h, 0] == n, h
        assert np.array_equal(g_ops[h, :n], ops), h
        assert np.array_equal(g_cs[h, :n].view(np.uint32), cs.view(np.uint32)), h
        assert meta[h, 1] == L.gpo_ref_fold_count(oracle.u8(hop[h].copy()), 1)
        assert meta[h, 2] == L.gpo_ref_fold_length(oracle.u8(hop[h].copy()), 1)

