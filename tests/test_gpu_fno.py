"""FNO warm-start kernels: the BF16 MFMA GEMM against numpy on BF16-rounded
operands (strided and transposed views), and the network's forward pass and
gradients against the float64 autograd reference (oracle/fno_oracle.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _record(what, err):
    import json
    import os

    os.makedirs("gpurun_out", exist_ok=True)
    name = os.environ.get("PYTEST_CURRENT_TEST", "?").split(" ")[0]
    with open("gpurun_out/parity_errors.jsonl", "a") as f:
        f.write(json.dumps({"test": name, "what": what, "dtype": "bf16xbf16->f32", "rel_err": err}) + "\n")


def bf16_round(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("m,n,k,batch,ta,tb", [(16, 16, 16, 1, 0, 0), (64, 64, 64, 1, 0, 0), (37, 70, 45, 3, 0, 1),
                                              (100, 33, 129, 2, 1, 0), (5, 200, 1024, 4, 1, 1)])
def test_gemm_matches_numpy(ggp, m, n, k, batch, ta, tb, fp32):
    import torch
    from gpe_hip_amd.fno import _Ops

    rng = np.random.default_rng(m * n + k)
    A = rng.standard_normal((batch, k, m) if ta else (batch, m, k)).astype(np.float32)
    B = rng.standard_normal((batch, n, k) if tb else (batch, k, n)).astype(np.float32)
    C0 = rng.standard_normal((batch, m, n)).astype(np.float32)
    Ar, Br = (A, B) if fp32 else (bf16_round(A), bf16_round(B))
    opA = np.transpose(Ar, (0, 2, 1)) if ta else Ar
    opB = np.transpose(Br, (0, 2, 1)) if tb else Br
    want = 0.5 * np.einsum("bik,bkj->bij", opA.astype(np.float64), opB.astype(np.float64)) - 2.0 * C0
    dA, dB, dC = (torch.tensor(x, device="cuda:0") for x in (A, B, C0))
    a = (1, m, k * m) if ta else (k, 1, m * k)
    b = (1, k, n * k) if tb else (n, 1, k * n)
    _Ops().mm(dA, dB, dC, m, n, k, a, b, (n, 1, m * n), batch=batch, alpha=0.5, beta=-2.0, fp32=fp32)
    got = dC.cpu().numpy()
    _record("gemm_f32" if fp32 else "gemm", float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    # FP32 accumulation of exactly representable products (BF16) or FP32 products
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want)) + 1e-5


def test_fno_forward_and_gradients_match_reference(ggp):
    torch = pytest.importorskip("torch")
    from gpe_hip_amd.fno import FNO
    from oracle import fno_oracle

    N, m, W, layers, B = 16, 4, 8, 2, 2
    net = FNO(N, in_ch=4, out_ch=4, width=W, modes=m, layers=layers, seed=3)
    for k in ("bc", "b0", "bp"):  # nonzero biases so their gradients are exercised
        net.p[k].copy_(torch.tensor(np.random.default_rng(1).standard_normal(tuple(net.p[k].shape)) * 0.1,
                                    dtype=torch.float32))
    rng = np.random.default_rng(7)
    x = rng.standard_normal((B, 4, N, N)).astype(np.float32)
    t = rng.standard_normal((B, 4, N, N)).astype(np.float32)
    params = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in net.p.items()}
    out_ref, loss_ref, g_ref = fno_oracle.loss_and_grads(params, x, t, N, m, layers)
    net.zero_grad()
    out, _ = net.loss_and_grad(torch.tensor(x, device="cuda:0"), torch.tensor(t, device="cuda:0"))
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    # BF16 operands: ~3 significant digits per product
    _record("forward", float(np.linalg.norm(out - out_ref) / np.linalg.norm(out_ref)))
    assert np.linalg.norm(out - out_ref) <= 2e-2 * np.linalg.norm(out_ref)
    for k, gr in g_ref.items():
        g = net.g[k].cpu().numpy()
        _record("grad_" + k, float(np.linalg.norm(g - gr) / np.linalg.norm(gr)))
        assert np.linalg.norm(g - gr) <= 5e-2 * np.linalg.norm(gr) + 1e-7, (k, np.linalg.norm(g - gr),
                                                                             np.linalg.norm(gr))


def test_band_limit_on_device_matches_host(ggp):
    """FP32 band limit on the device against the float64 truncated-mode sum."""
    torch = pytest.importorskip("torch")
    from gpe_hip_amd.fno import FNO

    N, m = 64, 8
    net = FNO(N, in_ch=4, out_ch=4, width=4, modes=m, layers=1, seed=0)
    rng = np.random.default_rng(11)
    y = rng.standard_normal((3, 4, N, N)).astype(np.float32)
    got = net.band_limit(torch.tensor(y, device="cuda:0")).cpu().numpy()
    F = np.fft.fft2(y.astype(np.float64), axes=(-2, -1))
    kx = np.r_[np.arange(m), np.arange(N - m, N)]
    G = np.zeros_like(F)
    c = np.where(np.arange(m) == 0, 1.0, 2.0)
    G[..., kx, :m] = F[..., kx, :m] * c
    want = np.fft.ifft2(G, axes=(-2, -1)).real
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    _record("band_limit_f32", float(err))
    assert err <= 1e-5
