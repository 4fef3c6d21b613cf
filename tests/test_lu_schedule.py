"""CPU model of the speculative pass of the blocked LU (lu_factor's `speculate` lambda in
traceon_amd/csrc/lu.hip): the same launch sequence on the two streams, with the same offsets,
leading dimensions, group sizes, buffer rotation, device-side gating and event waits,
replayed in numpy with a small panel width (NB = 8) on bounds-checked flat buffers.

It checks what the GPU tests cannot show directly:
- every buffer access stays inside its allocation (the workspace layout of LuWorkspace);
- the matrix after the pass, for a forced no-pivot failure at every panel in turn (and the
  driver's recovery product for a failure inside a group), is unpivoted elimination of the
  panels before the failure, under both extreme orders of the two streams;
- every pair of launches on different streams that neither stream order nor an event wait
  orders touches disjoint memory (no write of one meets a read or write of the other).

The model mirrors the C++ by transcription: a change to the launch sequence in lu.hip
changes this file with it.  (The diagonal-block kernel's own in-block arithmetic is modelled
as unpivoted elimination; its kernels are covered by the GPU tests.)"""
import numpy as np
import pytest

NB = 8          # the model's panel width (LU_NB = 128 on the GPU)
GMAX = 4        # the widest panel group (LU_GROUP_MAX)


class Buf:
    def __init__(self, name, n):
        self.name, self.a = name, np.full(n, np.nan)


def view(buf, off, ld, rows, cols):
    """rows x cols at flat offset `off` with leading dimension `ld`, bounds-checked."""
    if rows <= 0 or cols <= 0:
        return np.zeros((max(rows, 0), max(cols, 0)))
    assert off >= 0 and cols <= ld, (buf.name, off, ld, cols)
    assert off + (rows - 1) * ld + cols <= len(buf.a), (buf.name, off, ld, rows, cols, len(buf.a))
    return np.lib.stride_tricks.as_strided(buf.a[off:], shape=(rows, cols), strides=(ld * 8, 8))


def region(buf, off, ld, rows, cols):
    """A rows x cols block of a buffer, materialised as a set of cells only by the race check."""
    return (buf.name, (off, ld, max(rows, 0), max(cols, 0)))


def cells(reg):
    off, ld, rows, cols = reg
    return frozenset(off + r * ld + c for r in range(rows) for c in range(cols))


class Pass:
    def __init__(self, A0, lda, pair_min, quad_min, fail_at=(), nb=NB):
        self.NB, self.GMAX = nb, GMAX
        self.dry = False   # bounds only: no arithmetic (set by Driver)
        N = A0.shape[0]
        self.N, self.lda, self.nblk = N, lda, (N + self.NB - 1) // self.NB
        self.ldu = (N + 31) // 32 * 32
        self.A = Buf("A", N * lda)
        view(self.A, 0, lda, N, N)[...] = A0
        self.Lp = Buf("Lp", 2 * N * self.GMAX * self.NB)
        self.Up = Buf("Up", 2 * self.GMAX * self.NB * self.ldu)
        self.Ltmp = Buf("Ltmp", 2 * N * self.NB)
        self.U12 = Buf("U12", self.NB * N)
        self.U12b = Buf("U12b", self.NB * N)
        self.Dbak = np.zeros((self.NB, self.NB))
        self.Y, self.Z, self.lmax = {}, {}, 0.0
        self.fail_at = set(fail_at)
        self.pair_min, self.quad_min = pair_min, quad_min
        self.q = {"sa": [], "st": []}
        self.ver = {}
        self.groups = []

    # ---- enqueue ----
    def enq(self, s, fn, rd=(), wr=()):
        self.q[s].append(("op", fn, rd, wr))

    def record(self, s, ev):
        self.ver[ev] = self.ver.get(ev, 0) + 1
        self.q[s].append(("rec", (ev, self.ver[ev])))

    def wait(self, s, ev):
        self.q[s].append(("wait", (ev, self.ver.get(ev, 0))))

    def set_identity_pivots(self, k0, nb):
        """k_diag_block with factor = 1 writes ipiv[k0 + i] = k0 + i for its panel."""
        if hasattr(self, "ipiv"):
            self.ipiv[k0:k0 + nb] = np.arange(k0, k0 + nb)

    def gated(self, gp):
        return gp is not None and self.first_fail <= gp

    def gemm(self, s, gp, M, Nn, K, alpha, Ab, ao, la, Bb, bo, lb, beta, Cb, co, lc):
        if M <= 0 or Nn <= 0:
            return

        def fn():
            if self.gated(gp):
                return
            a, b, c = view(Ab, ao, la, M, K), view(Bb, bo, lb, K, Nn), view(Cb, co, lc, M, Nn)
            if self.dry:
                return
            a, b = a.copy(), b.copy()
            assert not np.isnan(a).any() and not np.isnan(b).any(), ("unwritten operand", Ab.name, Bb.name)
            c[...] = alpha * (a @ b) if beta == 0 else beta * c + alpha * (a @ b)
        rc = region(Cb, co, lc, M, Nn)
        self.enq(s, fn, rd=(region(Ab, ao, la, M, K), region(Bb, bo, lb, K, Nn), rc), wr=(rc,))

    def diag_l21(self, q, Cb, co, ldc):
        N, lda = self.N, self.lda
        k0 = q * self.NB
        nb = min(self.NB, N - k0)
        k1, m2 = k0 + nb, N - k0 - nb

        def diag():
            if self.gated(q):
                return
            D = view(self.A, k0 * lda + k0, lda, nb, nb)
            self.set_identity_pivots(k0, nb)
            if self.dry:
                return
            self.Dbak[:nb, :nb] = D
            L, U = np.eye(nb), D.copy()
            for j in range(nb):
                for i in range(j + 1, nb):
                    L[i, j] = U[i, j] / U[j, j]
                    U[i, :] -= L[i, j] * U[j, :]
            D[...] = np.tril(L, -1) + np.triu(U)
            self.Y[q], self.Z[q] = np.linalg.inv(L), np.linalg.inv(U)
            self.lmax = np.abs(np.tril(L, -1)).max() if nb > 1 else 0.0
        rd = region(self.A, k0 * lda + k0, lda, nb, nb)
        self.enq("sa", diag, rd=(rd,), wr=(rd,))
        if m2 > 0:
            def l21():
                if self.gated(q):
                    return
                c = view(Cb, co, ldc, m2, nb)
                a = view(self.A, k1 * lda + k0, lda, m2, nb)
                if self.dry:
                    return
                c[...] = a @ self.Z[q]
                self.lmax = max(self.lmax, np.abs(c).max())
            self.enq("sa", l21, rd=(region(self.A, k1 * lda + k0, lda, m2, nb),), wr=(region(Cb, co, ldc, m2, nb),))

    def check(self, q):
        def fn():
            if not self.gated(q) and (self.lmax > 1 or q in self.fail_at):
                self.first_fail = min(self.first_fail, q)
        self.enq("sa", fn)

    def u12(self, q, M, Nn, Bo, Cb, co, ldc):
        self.check(q)   # gemm_checked: the panel's test, then U12 = inv(L11) A12

        def fn():
            if not self.gated(q):
                c, b = view(Cb, co, ldc, M, Nn), view(self.A, Bo, self.lda, M, Nn)
                if not self.dry:
                    c[...] = self.Y[q][:M, :M] @ b
        self.enq("sa", fn, rd=(region(self.A, Bo, self.lda, M, Nn),), wr=(region(Cb, co, ldc, M, Nn),))

    def copy_panel(self, gp, Lb, lo, ldl, a21, Ub, uo, ldu, a12, rows, cols, nb):
        lda = self.lda

        def fn():
            if not self.gated(gp):
                view(self.A, a21, lda, rows, nb)[...] = view(Lb, lo, ldl, rows, nb)
                view(self.A, a12, lda, nb, cols)[...] = view(Ub, uo, ldu, nb, cols)
        self.enq("st", fn, rd=(region(Lb, lo, ldl, rows, nb), region(Ub, uo, ldu, nb, cols)),
                 wr=(region(self.A, a21, lda, rows, nb), region(self.A, a12, lda, nb, cols)))

    # ---- the pass (lu.hip, lu_factor: speculate) ----
    def speculate(self, s0=0):
        N, lda, nblk, ldu = self.N, self.lda, self.nblk, self.ldu
        self.first_fail = nblk
        LDL = self.GMAX * self.NB
        Lgr = lambda gi: (gi & 1) * N * LDL
        Ugr = lambda gi: (gi & 1) * LDL * ldu
        q, gi = s0, 0
        while True:
            G = 4 if N - q * self.NB >= self.quad_min else 2 if N - q * self.NB >= self.pair_min else 1
            while G > 1 and q + G >= nblk:
                G //= 2
            if G == 1:
                break
            b0, bG = q * self.NB, q * self.NB + G * self.NB
            b2G = min(bG + G * self.NB, N)
            Lq, Uq = Lgr(gi), Ugr(gi)
            for j in range(G):
                bj, bj1 = b0 + j * self.NB, b0 + j * self.NB + self.NB
                self.diag_l21(q + j, self.Lp, Lq + bj1 * LDL + j * self.NB, LDL)
                self.u12(q + j, self.NB, N - bj1, bj * lda + bj1, self.Up, Uq + j * self.NB * ldu + bj1, ldu)
                if j + 1 == G:
                    break
                Lj, Uj = Lq + bj1 * LDL + j * self.NB, Uq + j * self.NB * ldu
                self.gemm("sa", q + j, N - bj1, bG - bj1, self.NB, -1, self.Lp, Lj, LDL, self.Up, Uj + bj1, ldu, 1, self.A,
                          bj1 * lda + bj1, lda)
                self.gemm("sa", q + j, bG - bj1, N - bG, self.NB, -1, self.Lp, Lj, LDL, self.Up, Uj + bG, ldu, 1, self.A,
                          bj1 * lda + bG, lda)
            self.groups.append((q, G, gi))
            gl, KG = q + G - 1, G * self.NB
            if gi > 0:
                self.wait("sa", "ev_r")
            self.gemm("sa", gl, N - bG, b2G - bG, KG, -1, self.Lp, Lq + bG * LDL, LDL, self.Up, Uq + bG, ldu, 1, self.A,
                      bG * lda + bG, lda)
            self.gemm("sa", gl, b2G - bG, N - b2G, KG, -1, self.Lp, Lq + bG * LDL, LDL, self.Up, Uq + b2G, ldu, 1, self.A,
                      bG * lda + b2G, lda)
            self.record("sa", "ev_u")
            self.wait("st", "ev_u")
            self.gemm("st", gl, N - b2G, N - b2G, KG, -1, self.Lp, Lq + b2G * LDL, LDL, self.Up, Uq + b2G, ldu, 1, self.A,
                      b2G * lda + b2G, lda)
            for j in range(G):
                bj, bj1 = b0 + j * self.NB, b0 + j * self.NB + self.NB
                self.copy_panel(q + j, self.Lp, Lq + bj1 * LDL + j * self.NB, LDL, bj1 * lda + bj, self.Up, Uq + j * self.NB * ldu + bj1,
                                ldu, bj * lda + bj1, N - bj1, N - bj1, self.NB)
            self.record("st", "ev_r")
            q += G
            gi += 1
        Lbuf = lambda q: (q & 1) * N * self.NB
        Ubuf = lambda q: self.U12b if q & 1 else self.U12
        while q < nblk:
            k0 = q * self.NB
            nb = min(self.NB, N - k0)
            m2 = n2 = N - k0 - nb
            k1 = k0 + nb
            self.diag_l21(q, self.Ltmp, Lbuf(q), nb)
            if n2 <= 0:
                self.check(q)
                break
            nb1 = min(self.NB, n2)
            if q > s0:
                self.wait("sa", "ev_r")
            self.u12(q, nb, n2, k0 * lda + k1, Ubuf(q), 0, n2)
            self.record("sa", "ev_u")
            self.gemm("sa", q, n2, nb1, nb, -1, self.Ltmp, Lbuf(q), nb, Ubuf(q), 0, n2, 1, self.A, k1 * lda + k1, lda)
            self.wait("st", "ev_u")
            self.gemm("st", q, n2, n2 - nb1, nb, -1, self.Ltmp, Lbuf(q), nb, Ubuf(q), nb1, n2, 1, self.A, k1 * lda + k1 + nb1, lda)
            self.copy_panel(q, self.Ltmp, Lbuf(q), nb, k1 * lda + k0, Ubuf(q), 0, n2, k0 * lda + k1, m2, n2, nb)
            self.record("st", "ev_r")
            q += 1

    def run(self, first):
        """Execute both queues; `first` runs as far as it can before the other stream moves."""
        pos, reached = {"sa": 0, "st": 0}, {}
        order = [first, "st" if first == "sa" else "sa"]
        while pos["sa"] < len(self.q["sa"]) or pos["st"] < len(self.q["st"]):
            moved = False
            for s in order:
                while pos[s] < len(self.q[s]):
                    it = self.q[s][pos[s]]
                    if it[0] == "wait" and it[1][1] > 0 and reached.get(it[1][0], 0) < it[1][1]:
                        break
                    if it[0] == "rec":
                        reached[it[1][0]] = max(reached.get(it[1][0], 0), it[1][1])
                    elif it[0] == "op":
                        it[1]()
                    pos[s] += 1
                    moved = True
                if moved:
                    break
            assert moved, "the streams wait on each other"

    def unordered_conflicts(self):
        """Pairs of launches on different streams with no order between them that touch the same memory."""
        recpos = {it[1]: i for s in ("sa", "st") for i, it in enumerate(self.q[s]) if it[0] == "rec"}
        knows = {}
        for s in ("sa", "st"):   # entries of the other stream ordered before each entry of s
            k, arr = 0, []
            for it in self.q[s]:
                if it[0] == "wait" and it[1][1] > 0:
                    k = max(k, recpos[it[1]])
                arr.append(k)
            knows[s] = arr
        bad = 0
        cache = {}

        def memo(reg):
            if reg not in cache:
                cache[reg] = cells(reg)
            return cache[reg]
        ops = {s: [(i, it) for i, it in enumerate(self.q[s]) if it[0] == "op"] for s in ("sa", "st")}
        for i, a in ops["sa"]:
            for j, b in ops["st"]:
                if i < knows["st"][j] or j < knows["sa"][i]:
                    continue
                for na, ra in a[3]:
                    bad += sum(1 for nb_, rb in tuple(b[2]) + tuple(b[3]) if na == nb_ and memo(ra) & memo(rb))
                for nb_, rb in b[3]:
                    bad += sum(1 for na, ra in a[2] if na == nb_ and memo(ra) & memo(rb))
        return bad

    def driver_after_failure(self):
        """lu_factor's driver after a failed pass: the recovery product of a failure at panel
        j > 0 of a group, then the failed panel's diagonal block restored from its copy."""
        N, lda, kf = self.N, self.lda, self.first_fail
        for gq, G, gi in self.groups:
            if gq < kf < gq + G:
                j, bG, LDL = kf - gq, (gq + G) * self.NB, self.GMAX * self.NB
                M = N - bG
                if M > 0:
                    a = view(self.Lp, (gi & 1) * N * LDL + bG * LDL, LDL, M, j * self.NB)
                    b = view(self.Up, (gi & 1) * LDL * self.ldu + bG, self.ldu, j * self.NB, M)
                    c = view(self.A, bG * lda + bG, lda, M, M)
                    c[...] = c - a @ b
        k0 = kf * self.NB
        nb = min(self.NB, N - k0)
        view(self.A, k0 * lda + k0, lda, nb, nb)[...] = self.Dbak[:nb, :nb]


def eliminated(A0, panels):
    """Unpivoted elimination of the first `panels` * NB columns, the factors stored in place."""
    A = A0.copy()
    for j in range(min(panels * NB, A.shape[0])):
        A[j + 1:, j] /= A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A


MODES = {"singles": (10 ** 9, 10 ** 9), "pairs": (0, 10 ** 9), "quads": (0, 0), "quads-pairs-singles": (NB * 4, NB * 11)}


@pytest.mark.parametrize("N", [NB * 9 + 3, NB * 16, NB * 21 + 1])
@pytest.mark.parametrize("mode", list(MODES))
def test_pass_state_for_every_failing_panel(N, mode):
    pair_min, quad_min = MODES[mode]
    rng = np.random.default_rng(N)
    A0 = rng.normal(size=(N, N)) + N * np.eye(N)
    nblk = (N + NB - 1) // NB
    for lda in (N, N + 5):
        for fail in [None] + list(range(nblk)):
            for first in ("sa", "st"):
                P = Pass(A0, lda, pair_min, quad_min, () if fail is None else (fail,))
                P.speculate()
                P.run(first)
                kf = P.first_fail
                assert kf == (nblk if fail is None else fail)
                if kf < nblk:
                    P.driver_after_failure()
                ref = eliminated(A0, kf)
                got = view(P.A, 0, lda, N, N)
                assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (N, lda, mode, fail, first)


@pytest.mark.parametrize("N", [NB * 16, NB * 21 + 1])
@pytest.mark.parametrize("mode", list(MODES))
def test_no_unordered_launches_share_memory(N, mode):
    pair_min, quad_min = MODES[mode]
    P = Pass(np.eye(N) * 2.0, N, pair_min, quad_min)
    P.speculate()
    assert P.unordered_conflicts() == 0


def test_model_detects_a_missing_wait():
    """The race check itself: without the side stream's wait for the previous group's bulk
    update, the strips of the next group race with it."""
    N = NB * 21 + 1
    P = Pass(np.eye(N) * 2.0, N, 0, 0)
    P.speculate()
    P.q["sa"] = [it for it in P.q["sa"] if not (it[0] == "wait" and it[1][0] == "ev_r")]
    assert P.unordered_conflicts() > 0


# ---------------------------------------------------------------------------------------
# The driver around the passes (lu_factor's loop in lu.hip): speculative pass from panel q;
# after a failure at kf, the recovery product of a failure inside a group, the restore of
# kf's diagonal block from its copy, the pivoted panel (k_panel, cooperative over G
# workgroups), k_diag_block with factor = 0, k_laswp, U12 = inv(L11) A12 into the U12 buffer,
# its copy into the matrix and the A22 update; then the next pass from kf + 1.
# ---------------------------------------------------------------------------------------
WS_G = 1024   # LuWorkspace::G: pivot-candidate slots of the cooperative panel


class Driver(Pass):
    def __init__(self, A0, lda, pair_min, quad_min, nb=NB, G=5, dry=False, forced=()):
        super().__init__(A0, lda, pair_min, quad_min, fail_at=forced, nb=nb)
        self.dry = dry
        self.G = G
        N, nblk = self.N, self.nblk
        self.ipiv = np.full(N, -1, dtype=np.int64)
        self.cand_row = Buf("cand_row", 2 * WS_G * nb + 2 * nb)
        self.Linv = Buf("Linv", nblk * nb * nb)
        self.Uinv = Buf("Uinv", nblk * nb * nb)
        self.Dbak_ws = Buf("Dbak", nb * nb)
        self.launches = 0

    def pivoted_panel(self, q):
        N, lda, NB, G = self.N, self.lda, self.NB, self.G
        k0 = q * NB
        nb = min(NB, N - k0)
        m = N - k0
        # k_panel: workgroup g holds rows [k0 + g R, +R) of the panel in LDS (R x NB, the dynamic
        # LDS set for Rmax = ceil(N / G) rows) and reads / writes those rows' panel columns
        R, Rmax = -(-m // G), -(-N // G)
        assert R <= Rmax and G <= WS_G
        for g in range(G):
            rbeg = k0 + g * R
            rend = min(rbeg + R, N)
            if rend > rbeg:
                view(self.A, rbeg * lda + k0, lda, rend - rbeg, nb)
        view(self.cand_row, 0, NB, 2 * G + 2, NB)
        Pn = view(self.A, k0 * lda + k0, lda, m, nb)
        for j in range(nb):   # LAPACK's choice: largest |a|, first index on ties
            d = k0 + j
            if self.dry:
                p = N - 1 - (d * 7919) % (N - d)   # any row in [d, N): bounds only
            else:
                p = d + int(np.argmax(np.abs(Pn[j:, j])))
                Pn[[j, p - k0]] = Pn[[p - k0, j]]
                Pn[j + 1:, j] /= Pn[j, j]
                Pn[j + 1:, j + 1:] -= np.outer(Pn[j + 1:, j], Pn[j, j + 1:])
            assert d <= p < N
            self.ipiv[d] = p
        # k_diag_block, factor = 0: the block (all of it through 16-byte segments when
        # nb == NB, else element loads inside nb x nb), inverses into the solve's slots
        blk = NB if nb == NB else nb
        view(self.A, k0 * lda + k0, lda, blk, blk)
        Y, Z = view(self.Linv, q * NB * NB, NB, NB, NB), view(self.Uinv, q * NB * NB, NB, NB, NB)
        if not self.dry:
            D = view(self.A, k0 * lda + k0, lda, nb, nb)
            L = np.tril(D, -1) + np.eye(nb)
            self.Y[q], self.Z[q] = np.linalg.inv(L), np.linalg.inv(np.triu(D))
            Y[...] = 0.0
            Y[:nb, :nb] = self.Y[q]
        # k_laswp: one thread per column outside [k0, k0 + nb), swaps in panel order
        cols = [c for c in (np.arange(0, k0), np.arange(k0 + nb, N)) if len(c)]
        for j in range(nb):
            d, p = k0 + j, int(self.ipiv[k0 + j])
            assert 0 <= p < N
            for c in cols:
                view(self.A, d * lda + int(c[0]), lda, 1, len(c))
                view(self.A, p * lda + int(c[0]), lda, 1, len(c))
            if p != d and not self.dry:
                full = view(self.A, 0, lda, N, N)
                for c in cols:
                    full[[d, p], c[0]:c[-1] + 1] = full[[p, d], c[0]:c[-1] + 1]
        n2 = N - k0 - nb
        if n2 > 0:
            a12 = k0 * lda + k0 + nb
            self.direct_gemm(nb, n2, nb, 1.0, self.Linv, q * NB * NB, NB, self.A, a12, lda, 0.0, self.U12, 0, n2)
            u, a = view(self.U12, 0, n2, nb, n2), view(self.A, a12, lda, nb, n2)
            if not self.dry:
                a[...] = u
            self.direct_gemm(n2, n2, nb, -1.0, self.A, (k0 + nb) * lda + k0, lda, self.U12, 0, n2, 1.0, self.A,
                             (k0 + nb) * lda + k0 + nb, lda)

    def direct_gemm(self, M, Nn, K, alpha, Ab, ao, la, Bb, bo, lb, beta, Cb, co, lc):
        a, b, c = view(Ab, ao, la, M, K), view(Bb, bo, lb, K, Nn), view(Cb, co, lc, M, Nn)
        if not self.dry:
            c[...] = alpha * (a @ b) if beta == 0 else beta * c + alpha * (a @ b)

    def factor(self, first="sa"):
        q = 0
        nblk = self.nblk
        while q < nblk:
            self.q, self.groups = {"sa": [], "st": []}, []
            self.speculate(q)
            self.run(first)
            kf = self.first_fail
            if kf < q or kf >= nblk:
                break
            if not self.dry:
                self.driver_after_failure()
            else:   # the same launches, bounds only
                for gq, G, gi in self.groups:
                    if gq < kf < gq + G:
                        bG, LDL = (gq + G) * self.NB, self.GMAX * self.NB
                        Mr = self.N - bG
                        view(self.Lp, (gi & 1) * self.N * LDL + bG * LDL, LDL, Mr, (kf - gq) * self.NB)
                        view(self.Up, (gi & 1) * LDL * self.ldu + bG, self.ldu, (kf - gq) * self.NB, Mr)
                        view(self.A, bG * self.lda + bG, self.lda, Mr, Mr)
                k0 = kf * self.NB
                nbk = min(self.NB, self.N - k0)
                view(self.A, k0 * self.lda + k0, self.lda, nbk, nbk)
            self.pivoted_panel(kf)
            q = kf + 1
        assert np.all((self.ipiv >= np.arange(self.N)) & (self.ipiv < self.N))


@pytest.mark.parametrize("N", [NB * 9 + 3, NB * 16, NB * 21 + 1, NB * 13 + 4])
@pytest.mark.parametrize("mode", list(MODES))
def test_pivoted_factorisation_matches_lapack(N, mode):
    """The whole driver on matrices that need row interchanges (random, no diagonal dominance):
    natural failures of the no-pivot test, the pivoted panels and the resumed passes give
    LAPACK's getrf factors and pivots, under both stream orders and with a padded leading
    dimension; also with extra forced failures at panels that would have passed."""
    import scipy.linalg as sla
    pair_min, quad_min = MODES[mode]
    rng = np.random.default_rng(N + 7)
    A0 = rng.normal(size=(N, N))
    A0[:, :NB * 2] += 3 * np.eye(N)[:, :NB * 2]     # the first panels pass, later ones pivot
    lu, piv = sla.lu_factor(A0)
    nblk = (N + NB - 1) // NB
    for lda in (N, N + 5):
        for forced in ((), (1, nblk - 1)):
            for first in ("sa", "st"):
                D = Driver(A0, lda, pair_min, quad_min, forced=forced)
                D.factor(first)
                got = view(D.A, 0, lda, N, N)
                assert np.array_equal(D.ipiv, piv), (N, mode, lda, forced, first)
                assert np.abs(got - lu).max() <= 1e-10 * np.abs(lu).max(), (N, mode, lda, forced, first)


@pytest.mark.parametrize("N", [5004, 5000, 4097, 4109, 8204, 8193, 2060])
def test_pivoted_driver_bounds_at_the_gpu_panel_width(N):
    """Bounds of every launch of the driver at the GPU's panel width (NB = 128) and 256
    cooperative workgroups, for sizes whose last panel is 1-15 columns wide (N mod 128 < 16)
    among others, with the no-pivot test failing at every panel, at every third and only at the
    last: the pivoted panel's rows, the diagonal-block loads, the row interchanges (pivot rows
    anywhere below the diagonal), the U12 and A22 products and the group recovery product all
    stay inside their buffers."""
    nb = 128
    nblk = (N + nb - 1) // nb
    A0 = np.zeros((N, N))
    for forced in (tuple(range(nblk)), tuple(range(0, nblk, 3)), (nblk - 1,)):
        for lda in (N, N + 3):
            D = Driver(A0, lda, 4096, 8192, nb=nb, G=256, dry=True, forced=forced)
            D.factor("sa")
