"""Device-resident building blocks over the Tier B entry points of libtraceon_hip.so.

PyTorch (ROCm) provides device memory and the current HIP stream; every numeric step
runs in this project's HIP kernels.  These are the pieces the solver, field and tracer
modules assemble:

* ``Elements``       — a line mesh on the device: Gauss-point buffers, SoA sources, row data
* ``assemble``       — dense matrix fill + singular diagonal (FP64, optional FP32 copy)
* ``lu_factor`` / ``lu_solve`` / ``solve_system`` — LU with partial pivoting, FP64 or
                       FP32 with FP64 iterative refinement
* ``Sources``        — flattened ring sources of a charge set, field sums at point batches,
                       axial derivatives
* ``axial_coefficients`` — FieldRadialAxial interpolation coefficients
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import backend as B

_L = B.backend_lib
F64 = torch.float64


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("traceon_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None):
    return None if t is None else t.data_ptr()


def to_dev(a, dtype=F64) -> torch.Tensor:
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device())


def check(status: int, name: str) -> None:
    B.check(status, name)


class Elements:
    """Line elements (N,4,3) with excitation types/values on the device."""

    def __init__(self, lines, types=None, values=None):
        lines = np.ascontiguousarray(lines, dtype=np.float64)
        assert lines.ndim == 3 and lines.shape[1:] == (4, 3)
        N = len(lines)
        self.N = N
        if types is None:
            types = np.ones(N, dtype=np.uint8)
        if values is None:
            values = np.zeros(N)
        types = np.ascontiguousarray(types, dtype=np.uint8)
        bad = ~np.isin(types, [1, 2, 3, 5, 6])
        if np.any(bad):
            # radial.c:305-308 exits the process; this raises instead
            raise ValueError(f"ExcitationType unknown at element {int(np.argmax(bad))}")
        self.lines = to_dev(lines)
        self.types = to_dev(types, torch.uint8)
        self.values = to_dev(values)
        dev = device()
        self.jac = torch.empty((N, 16), dtype=F64, device=dev)
        self.pos = torch.empty((N, 16, 2), dtype=F64, device=dev)
        self.src = torch.empty((3 * 16 * max(N, 1),), dtype=F64, device=dev)
        self.rows = torch.empty((max(N, 1) * _L.tb_rowinfo_bytes(),), dtype=torch.uint8, device=dev)
        self.bad = torch.zeros(1, dtype=torch.int32, device=dev)
        if N:
            check(_L.tb_prepare_elements(ptr(self.lines), ptr(self.types), ptr(self.values), N, ptr(self.jac), ptr(self.pos),
                                         ptr(self.src), ptr(self.rows), ptr(self.bad), stream_ptr()), "prepare_elements")


def assemble(E: Elements, fp64: bool = True, fp32: bool = False, diagonal: bool = True):
    """Dense matrix of the BEM system (solver.py:206-253): off-diagonal fill + singular
    diagonal.  Returns (M64 or None, M32 or None)."""
    N = E.N
    dev = device()
    M = torch.empty((N, N), dtype=F64, device=dev) if fp64 else None
    M32 = torch.empty((N, N), dtype=torch.float32, device=dev) if fp32 else None
    if M is None and M32 is None:
        raise ValueError("assemble: nothing to produce")
    if M is None:
        # the fill accumulates in FP64 registers and stores both; without an FP64 target use a row-block scratch
        blk = max(1, min(N, (256 << 20) // (8 * max(N, 1))))
        tmp = torch.empty((blk, N), dtype=F64, device=dev)
        for r0 in range(0, N, blk):
            r1 = min(N, r0 + blk) - 1
            view = tmp.data_ptr() - r0 * N * 8
            check(_L.tb_fill_matrix(view, N, ptr(M32), N, ptr(E.rows), ptr(E.src), N, r0, r1, 0, stream_ptr()), "fill")
        if diagonal:
            check(_L.tb_fill_diagonal(None, 0, ptr(M32), N, ptr(E.lines), ptr(E.rows), N, 0, N - 1, None, stream_ptr()), "diag")
        return None, M32
    check(_L.tb_fill_matrix(ptr(M), N, ptr(M32), N, ptr(E.rows), ptr(E.src), N, 0, N - 1, 0, stream_ptr()), "fill")
    if diagonal:
        check(_L.tb_fill_diagonal(ptr(M), N, ptr(M32), N, ptr(E.lines), ptr(E.rows), N, 0, N - 1, None, stream_ptr()), "diag")
    return M, M32


def self_integrals(E: Elements) -> torch.Tensor:
    out = torch.empty(E.N, dtype=F64, device=device())
    check(_L.tb_fill_diagonal(None, 0, None, 0, ptr(E.lines), ptr(E.rows), E.N, 0, E.N - 1, ptr(out), stream_ptr()), "diag")
    return out


@dataclass
class LU:
    A: torch.Tensor
    ipiv: torch.Tensor
    info: int
    fp32: bool
    ws: torch.Tensor   # block inverses and permutation used by the solves


class KernelTiming(C.Structure):
    """tb_kernel_timing_t: launches, summed HIP-event ms, algorithmic FLOP and HBM bytes of a kernel."""
    _fields_ = [("launches", C.c_longlong), ("ms", C.c_double), ("flops", C.c_double), ("bytes", C.c_double)]


def lu_factor(A: torch.Tensor, overwrite: bool = True, gemm_timing: "KernelTiming | None" = None) -> LU:
    """In-place (by default) LU with partial pivoting of a square row-major matrix.
    gemm_timing: accumulates the MFMA GEMM launches of this factorisation (measurement)."""
    assert A.dim() == 2 and A.shape[0] == A.shape[1] and A.is_contiguous()
    N = A.shape[0]
    fp32 = A.dtype == torch.float32
    if not overwrite:
        A = A.clone()
    dev = A.device
    ipiv = torch.empty(N, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(_L.tb_lu_workspace_bytes(N, int(fp32)), dtype=torch.uint8, device=dev)
    if gemm_timing is not None:
        check(_L.tb_lu_factor_timed(ptr(A), N, N, ptr(ipiv), ptr(ws), ptr(info), int(fp32), C.byref(gemm_timing),
                                    stream_ptr()), "lu_factor")
    else:
        f = _L.tb_lu_factor_f32 if fp32 else _L.tb_lu_factor
        check(f(ptr(A), N, N, ptr(ipiv), ptr(ws), ptr(info), stream_ptr()), "lu_factor")
    lu = LU(A, ipiv, int(info.item()), fp32, ws)
    if lu.info:
        # numpy.linalg.solve (solver.py:173) raises LinAlgError on an exactly singular matrix
        raise np.linalg.LinAlgError(f"Singular matrix (zero or NaN pivot in column {lu.info - 1})")
    return lu


def check_solve_status(err: int) -> None:
    """Raise if a triangular solve reported a dependency wait past its spin limit (its solution
    is then NaN): a failure of the device, not of the matrix."""
    if err:
        raise RuntimeError("lu_solve: a triangular-solve dependency wait passed its spin limit; the solution is invalid")


def lu_solve(lu: LU, X: torch.Tensor, status: torch.Tensor | None = None) -> torch.Tensor:
    """In-place X (nrhs, N) FP64 <- A^{-1} X.  Without `status` the call waits and raises on a
    failed solve; with it (a device int32), the solve status is written there (nonzero on
    failure) for the caller to read with other results."""
    assert X.dtype == F64 and X.is_contiguous()
    N = lu.A.shape[0]
    X2 = X.view(-1, N)
    f = _L.tb_lu_solve_f32 if lu.fp32 else _L.tb_lu_solve
    st = status if status is not None else torch.zeros(1, dtype=torch.int32, device=X.device)
    for r0 in range(0, X2.shape[0], 64):
        blk = X2[r0:r0 + 64]
        check(f(ptr(lu.A), N, N, ptr(lu.ipiv), ptr(blk), blk.shape[0], ptr(lu.ws), stream_ptr()), "lu_solve")
        # each call clears the flag: fold every block's status into st
        tmp = st if r0 == 0 else torch.zeros(1, dtype=torch.int32, device=X.device)
        check(_L.tb_lu_solve_status(ptr(lu.ws), N, int(lu.fp32), ptr(tmp), stream_ptr()), "lu_solve_status")
        if r0:
            st.bitwise_or_(tmp)
    if status is None:
        check_solve_status(int(st.item()))
    return X


def absmax_into(x: torch.Tensor, out: torch.Tensor) -> None:
    """out (one int64 slot) <- bit pattern of max |x| (library kernel; NaN gives a NaN pattern)."""
    check(_L.tb_absmax(ptr(x), x.numel(), ptr(out), stream_ptr()), "absmax")


def residual(A: torch.Tensor, B_: torch.Tensor, X: torch.Tensor):
    N = A.shape[0]
    nrhs = X.shape[0]
    R = torch.empty_like(X)
    rmax = torch.zeros(nrhs, dtype=torch.int64, device=A.device)
    check(_L.tb_residual(ptr(A), N, N, ptr(B_), ptr(X), ptr(R), nrhs, ptr(rmax), stream_ptr()), "residual")
    return R, rmax.view(torch.float64)


@dataclass
class SolveStats:
    """mode, refinement steps, max |b - A x| (with refinement: at the start of the last step,
    before its correction, which is at the rounding level of x), LAPACK-style info."""
    mode: str
    iterations: int
    residual: float
    info: int


def solve_system(M64: torch.Tensor | None, F: np.ndarray | torch.Tensor, M32: torch.Tensor | None = None,
                 mode: str = "fp64", max_iter: int = 30, refine: bool = True, lu: LU | None = None,
                 gemm_timing: KernelTiming | None = None, fallback: bool = True):
    """Charges for right-hand sides F (nrhs, N) of the assembled system M64 x = F.

    mode "fp64":  LU of the FP64 matrix (the dgesv of solver.py:173), then, if ``refine``,
                  iterative refinement with double-double residuals against M64.
    mode "mixed": LU of the FP32 copy M32 (MFMA; FP32 accuracy), FP64 solves, refinement with
                  double-double residuals against M64.  Refinement converges when
                  cond(M) * eps32 < 1; when it does not reach the rounding level of x (it
                  stalls or diverges) and ``fallback`` is set, the system is solved again
                  with an FP64 LU, and the stats report mode "mixed->fp64".
    Refinement stops when the update is at the rounding level of x (2 eps |x|), stops
    shrinking, or shrinks by a factor rho per step such that the rest of the geometric series,
    dmax rho / (1 - rho), is below that level.  gemm_timing: see lu_factor.  Returns (X (nrhs,N) FP64 tensor, SolveStats, LU)."""
    Bt = F if isinstance(F, torch.Tensor) else to_dev(np.atleast_2d(F))
    Bt = Bt.reshape(-1, Bt.shape[-1]).contiguous()
    if mode == "fp64":
        if lu is None:
            lu = lu_factor(M64, overwrite=False, gemm_timing=gemm_timing)
    elif mode == "mixed":
        if lu is None:
            try:
                lu = lu_factor(M32, overwrite=True, gemm_timing=gemm_timing)
            except np.linalg.LinAlgError:
                # singular in FP32 only (a zero or NaN FP32 pivot): the FP64 LU decides
                if not (fallback and M64 is not None):
                    raise
                X, st, lu = solve_system(M64, Bt, mode="fp64", max_iter=max_iter, refine=True, fallback=False)
                return X, SolveStats("mixed->fp64", st.iterations, st.residual, lu.info), lu
    else:
        raise ValueError(mode)
    X = Bt.clone()
    it = 0
    converged = True
    rlast = None
    if refine or mode == "mixed":
        lu_solve(lu, X)
        last = float("inf")
        eps = 2.0 ** -52
        converged = False
        # per step, one readback: max |update|, max |x| (bit patterns) and the solve status
        stat = torch.zeros(3, dtype=torch.int64, device=X.device)
        for it in range(1, max_iter + 1):
            R, rmax = residual(M64, Bt, X)
            rlast = rmax
            stat.zero_()
            lu_solve(lu, R, status=stat[2:].view(torch.int32)[:1])
            absmax_into(R, stat[0:1])
            absmax_into(X, stat[1:2])
            h = stat.cpu()
            check_solve_status(int(h[2]))
            dmax, xmax = (float(v) for v in h[:2].view(torch.float64))
            if not math.isfinite(dmax) or (dmax >= 0.5 * last and dmax > 64.0 * eps * xmax):
                break                      # diverging or stalled above the rounding level
            check(_L.tb_axpy(ptr(X), ptr(R), X.numel(), stream_ptr()), "axpy")
            # converged: the update is at the rounding level, or it stopped shrinking, or the
            # updates contract by rho = dmax / last so that the next would be below the
            # rounding level (dmax rho / (1 - rho) bounds the rest of the series)
            rho = dmax / last if math.isfinite(last) and last > 0.0 else 1.0
            tail_below = rho < 1.0 and dmax * rho / (1.0 - rho) <= 2.0 * eps * xmax
            if dmax <= 2.0 * eps * xmax or dmax >= 0.5 * last or tail_below:
                converged = True
                break
            last = dmax
    else:
        lu_solve(lu, X)
    if mode == "mixed" and not converged and fallback and M64 is not None and lu.info == 0:
        X, st, lu = solve_system(M64, Bt, mode="fp64", max_iter=max_iter, refine=True, fallback=False)
        return X, SolveStats("mixed->fp64", it + st.iterations, st.residual, lu.info), lu
    if rlast is None:
        _, rlast = residual(M64, Bt, X)
    rstat = torch.zeros(1, dtype=torch.int64, device=X.device)
    absmax_into(rlast, rstat)   # the largest residual over the right-hand sides (library kernel)
    return X, SolveStats(mode, it, float(rstat.cpu().view(torch.float64)[0]), lu.info), lu


class Sources:
    """Flattened ring sources (w = q*jac, r, z) of one charge set on the device."""

    def __init__(self, charges, jac, pos, kind: str = "line"):
        nq, pstride, zcol = (16, 2, 1) if kind == "line" else (12, 3, 2)
        q = charges if isinstance(charges, torch.Tensor) else to_dev(np.asarray(charges, dtype=np.float64))
        j = jac if isinstance(jac, torch.Tensor) else to_dev(jac)
        p = pos if isinstance(pos, torch.Tensor) else to_dev(pos)
        N = q.numel()
        self.kind = kind
        self.n = N * nq
        dev = device()
        self.w = torch.empty(max(self.n, 1), dtype=F64, device=dev)
        self.r = torch.empty_like(self.w)
        self.z = torch.empty_like(self.w)
        self._keep = (q, j, p)
        if N:
            check(_L.tb_sources(ptr(q), ptr(j), ptr(p), N, nq, pstride, zcol, ptr(self.w), ptr(self.r), ptr(self.z), stream_ptr()),
                  "sources")

    def field_points(self, pts: torch.Tensor, pot: bool = True, field: bool = True, out_pot=None, out_field=None,
                     accumulate: bool = False):
        M = pts.shape[0]
        dev = pts.device
        P = (out_pot if out_pot is not None else torch.zeros(M, dtype=F64, device=dev)) if pot else None
        Fd = (out_field if out_field is not None else torch.zeros((M, 3), dtype=F64, device=dev)) if field else None
        if self.n == 0 or M == 0:
            return P, Fd
        if self.kind == "line":
            check(_L.tb_field_points(ptr(pts), M, ptr(self.w), ptr(self.r), ptr(self.z), self.n, ptr(P), ptr(Fd), int(accumulate),
                                     stream_ptr()), "field_points")
        else:
            if field:
                check(_L.tb_current_points(ptr(pts), M, ptr(self.w), ptr(self.r), ptr(self.z), self.n, ptr(Fd), int(accumulate),
                                           stream_ptr()), "current_points")
        return P, Fd

    def axial_derivatives(self, z: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        Nz = z.numel()
        out = out if out is not None else torch.zeros((Nz, 9), dtype=F64, device=z.device)
        if self.n:
            check(_L.tb_axial_derivatives(ptr(out), ptr(z), Nz, ptr(self.w), ptr(self.r), ptr(self.z), self.n,
                                          int(self.kind != "line"), stream_ptr()), "axial_derivatives")
        return out

    def current_axial_potential(self, z: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(z.numel(), dtype=F64, device=z.device)
        if self.n:
            check(_L.tb_current_axial_potential(ptr(z), z.numel(), ptr(self.w), ptr(self.r), ptr(self.z), self.n, ptr(out),
                                                stream_ptr()), "current_axial_potential")
        return out


def axial_coefficients(derivs: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    Nz = z.numel()
    C_ = torch.zeros((Nz - 1, 9, 6), dtype=F64, device=z.device)
    scratch = torch.empty(6 * Nz, dtype=F64, device=z.device)
    check(_L.tb_axial_coefficients(ptr(derivs.contiguous()), ptr(z), Nz, ptr(C_), ptr(scratch), stream_ptr()), "axial_coefficients")
    return C_


def axial_points(pts: torch.Tensor, z: torch.Tensor, coeffs: torch.Tensor, pot: bool = True, field: bool = True):
    M = pts.shape[0]
    P = torch.zeros(M, dtype=F64, device=pts.device) if pot else None
    Fd = torch.zeros((M, 3), dtype=F64, device=pts.device) if field else None
    check(_L.tb_axial_points(ptr(pts), M, ptr(z), z.numel(), ptr(coeffs), ptr(P), ptr(Fd), stream_ptr()), "axial_points")
    return P, Fd


def jacobian_3d(tri) -> tuple[torch.Tensor, torch.Tensor]:
    t = to_dev(np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3))
    N = t.shape[0]
    jac = torch.empty((N, 12), dtype=F64, device=t.device)
    pos = torch.empty((N, 12, 3), dtype=F64, device=t.device)
    if N:
        check(_L.tb_jacobian_3d(ptr(t), N, ptr(jac), ptr(pos), stream_ptr()), "jacobian_3d")
    return jac, pos


def gemm(A: torch.Tensor, Bm: torch.Tensor, C_: torch.Tensor | None = None, alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
    M, K = A.shape
    K2, N = Bm.shape
    assert K == K2 and A.dtype == Bm.dtype
    if C_ is None:
        C_ = torch.zeros((M, N), dtype=A.dtype, device=A.device)
    f = _L.tb_gemm_f64 if A.dtype == F64 else _L.tb_gemm_f32
    check(f(M, N, K, alpha, ptr(A), A.stride(0), ptr(Bm), Bm.stride(0), beta, ptr(C_), C_.stride(0), stream_ptr()), "gemm")
    return C_
