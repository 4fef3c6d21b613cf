"""GPU tests of the learned field surrogate: the BF16 MFMA GEMM, the MLP's forward pass and
gradient against the numpy restatement (oracle/mlp_oracle.py, BF16 rounding points
emulated), training, and the surrogate-corrected tracing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def _bf16_tensor(a):
    import torch
    from oracle.mlp_oracle import bf16
    r = bf16(a).astype(np.float32)
    u = (r.view(np.uint32) >> 16).astype(np.uint16)
    return torch.tensor(u.view(np.int16), device="cuda"), r.astype(np.float64)


@pytest.mark.parametrize("ta,tb,ksplit", [(False, False, 1), (False, True, 1), (True, False, 1), (True, False, 7), (False, False, 3)])
def test_bf16_gemm_matches_numpy(gpu, ta, tb, ksplit):
    """Every operand layout the MLP uses (A or B transposed through strides) and split-K:
    products of BF16 values are exact, so only the FP32 summation order differs."""
    import torch
    from traceon_amd import backend
    from traceon_amd.device import ptr, stream_ptr
    rng = np.random.default_rng(3)
    M, N, K = 150, 97, 300
    a = rng.standard_normal((K, M) if ta else (M, K))
    b = rng.standard_normal((N, K) if tb else (K, N))
    at, ar = _bf16_tensor(a)
    bt, br = _bf16_tensor(b)
    A = ar.T if ta else ar
    Bm = br.T if tb else br
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    backend.check(backend.backend_lib.tb_gemm_bf16(M, N, K, ptr(at), sam, sak, ptr(bt), sbk, sbn, ptr(C), N, 0, ksplit, stream_ptr()))
    ref = A @ Bm
    err = np.max(np.abs(C.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print("bf16 gemm", ta, tb, ksplit, err, flush=True)
    assert err < 1e-5


def test_mlp_forward_and_gradient_vs_oracle(gpu):
    """Forward outputs and every parameter block of the gradient against the oracle with the
    kernels' BF16 rounding points (a transposed or dropped term would be an O(1) error)."""
    import torch
    from oracle import mlp_oracle as mo
    from traceon_amd import backend
    from traceon_amd.device import ptr, stream_ptr
    L = backend.backend_lib
    assert L.tb_mlp_nparams() == mo.NPARAM
    rng = np.random.default_rng(5)
    B = 700
    P = mo.init_params(2, scale=0.7)
    X = rng.uniform(-1, 1, (B, 2))
    T = rng.standard_normal((B, 5))
    Pt = torch.tensor(P, dtype=torch.float32, device="cuda")
    P32 = Pt.double().cpu().numpy()          # the FP32 master values the kernels use
    Wbf = torch.empty(3 * 128 * 128, dtype=torch.int16, device="cuda")
    backend.check(L.tb_mlp_to_bf16(ptr(Pt), ptr(Wbf), stream_ptr()))
    Xt = torch.tensor(X, dtype=torch.float32, device="cuda")
    Tt = torch.tensor(T, dtype=torch.float32, device="cuda")
    X32 = Xt.double().cpu().numpy()
    ws = torch.empty(L.tb_mlp_workspace_bytes(B), dtype=torch.uint8, device="cuda")
    Y = torch.empty((B, 5), dtype=torch.float32, device="cuda")
    backend.check(L.tb_mlp_forward(ptr(Pt), ptr(Wbf), ptr(Xt), B, ptr(Y), ptr(ws), stream_ptr()))
    Yo, _ = mo.forward(P32, X32, rounding=True)
    ey = np.max(np.abs(Y.cpu().numpy() - Yo)) / np.max(np.abs(Yo))
    G = torch.zeros(mo.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    backend.check(L.tb_mlp_grad(ptr(Pt), ptr(Wbf), ptr(Xt), ptr(Tt), B, ptr(G), ptr(loss), ptr(ws), stream_ptr()))
    Go, lo = mo.grad(P32, X32, Tt.double().cpu().numpy(), rounding=True)
    g = G.cpu().numpy()
    print("mlp forward rel", ey, "loss", float(loss.item()), lo, flush=True)
    assert ey < 1e-3
    assert abs(float(loss.item()) - lo) <= 1e-3 * lo
    for name, (o, n) in mo.OFF.items():
        e = np.linalg.norm(g[o:o + n] - Go[o:o + n]) / np.linalg.norm(Go[o:o + n])
        print("grad", name, e, flush=True)
        assert e < 1e-2, name
