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
    if rows <= 0 or cols <= 0:
        return (buf.name, frozenset())
    return (buf.name, frozenset(off + r * ld + c for r in range(rows) for c in range(cols)))


class Pass:
    def __init__(self, A0, lda, pair_min, quad_min, fail_at=()):
        N = A0.shape[0]
        self.N, self.lda, self.nblk = N, lda, (N + NB - 1) // NB
        self.ldu = (N + 31) // 32 * 32
        self.A = Buf("A", N * lda)
        view(self.A, 0, lda, N, N)[...] = A0
        self.Lp = Buf("Lp", 2 * N * GMAX * NB)
        self.Up = Buf("Up", 2 * GMAX * NB * self.ldu)
        self.Ltmp = Buf("Ltmp", 2 * N * NB)
        self.U12 = Buf("U12", NB * N)
        self.U12b = Buf("U12b", NB * N)
        self.Dbak = np.zeros((NB, NB))
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

    def gated(self, gp):
        return gp is not None and self.first_fail <= gp

    def gemm(self, s, gp, M, Nn, K, alpha, Ab, ao, la, Bb, bo, lb, beta, Cb, co, lc):
        if M <= 0 or Nn <= 0:
            return

        def fn():
            if self.gated(gp):
                return
            a, b, c = view(Ab, ao, la, M, K).copy(), view(Bb, bo, lb, K, Nn).copy(), view(Cb, co, lc, M, Nn)
            assert not np.isnan(a).any() and not np.isnan(b).any(), ("unwritten operand", Ab.name, Bb.name)
            c[...] = alpha * (a @ b) if beta == 0 else beta * c + alpha * (a @ b)
        rc = region(Cb, co, lc, M, Nn)
        self.enq(s, fn, rd=(region(Ab, ao, la, M, K), region(Bb, bo, lb, K, Nn), rc), wr=(rc,))

    def diag_l21(self, q, Cb, co, ldc):
        N, lda = self.N, self.lda
        k0 = q * NB
        nb = min(NB, N - k0)
        k1, m2 = k0 + nb, N - k0 - nb

        def diag():
            if self.gated(q):
                return
            D = view(self.A, k0 * lda + k0, lda, nb, nb)
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
                c[...] = view(self.A, k1 * lda + k0, lda, m2, nb) @ self.Z[q]
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
                view(Cb, co, ldc, M, Nn)[...] = self.Y[q][:M, :M] @ view(self.A, Bo, self.lda, M, Nn)
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
        LDL = GMAX * NB
        Lgr = lambda gi: (gi & 1) * N * LDL
        Ugr = lambda gi: (gi & 1) * LDL * ldu
        q, gi = s0, 0
        while True:
            G = 4 if N - q * NB >= self.quad_min else 2 if N - q * NB >= self.pair_min else 1
            while G > 1 and q + G >= nblk:
                G //= 2
            if G == 1:
                break
            b0, bG = q * NB, q * NB + G * NB
            b2G = min(bG + G * NB, N)
            Lq, Uq = Lgr(gi), Ugr(gi)
            for j in range(G):
                bj, bj1 = b0 + j * NB, b0 + j * NB + NB
                self.diag_l21(q + j, self.Lp, Lq + bj1 * LDL + j * NB, LDL)
                self.u12(q + j, NB, N - bj1, bj * lda + bj1, self.Up, Uq + j * NB * ldu + bj1, ldu)
                if j + 1 == G:
                    break
                Lj, Uj = Lq + bj1 * LDL + j * NB, Uq + j * NB * ldu
                self.gemm("sa", q + j, N - bj1, bG - bj1, NB, -1, self.Lp, Lj, LDL, self.Up, Uj + bj1, ldu, 1, self.A,
                          bj1 * lda + bj1, lda)
                self.gemm("sa", q + j, bG - bj1, N - bG, NB, -1, self.Lp, Lj, LDL, self.Up, Uj + bG, ldu, 1, self.A,
                          bj1 * lda + bG, lda)
            self.groups.append((q, G, gi))
            gl, KG = q + G - 1, G * NB
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
                bj, bj1 = b0 + j * NB, b0 + j * NB + NB
                self.copy_panel(q + j, self.Lp, Lq + bj1 * LDL + j * NB, LDL, bj1 * lda + bj, self.Up, Uq + j * NB * ldu + bj1,
                                ldu, bj * lda + bj1, N - bj1, N - bj1, NB)
            self.record("st", "ev_r")
            q += G
            gi += 1
        Lbuf = lambda q: (q & 1) * N * NB
        Ubuf = lambda q: self.U12b if q & 1 else self.U12
        while q < nblk:
            k0 = q * NB
            nb = min(NB, N - k0)
            m2 = n2 = N - k0 - nb
            k1 = k0 + nb
            self.diag_l21(q, self.Ltmp, Lbuf(q), nb)
            if n2 <= 0:
                self.check(q)
                break
            nb1 = min(NB, n2)
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
        ops = {s: [(i, it) for i, it in enumerate(self.q[s]) if it[0] == "op"] for s in ("sa", "st")}
        for i, a in ops["sa"]:
            for j, b in ops["st"]:
                if i < knows["st"][j] or j < knows["sa"][i]:
                    continue
                for na, ra in a[3]:
                    bad += sum(1 for nb_, rb in tuple(b[2]) + tuple(b[3]) if na == nb_ and ra & rb)
                for nb_, rb in b[3]:
                    bad += sum(1 for na, ra in a[2] if na == nb_ and ra & rb)
        return bad

    def driver_after_failure(self):
        """lu_factor's driver after a failed pass: the recovery product of a failure at panel
        j > 0 of a group, then the failed panel's diagonal block restored from its copy."""
        N, lda, kf = self.N, self.lda, self.first_fail
        for gq, G, gi in self.groups:
            if gq < kf < gq + G:
                j, bG, LDL = kf - gq, (gq + G) * NB, GMAX * NB
                M = N - bG
                if M > 0:
                    a = view(self.Lp, (gi & 1) * N * LDL + bG * LDL, LDL, M, j * NB)
                    b = view(self.Up, (gi & 1) * LDL * self.ldu + bG, self.ldu, j * NB, M)
                    c = view(self.A, bG * lda + bG, lda, M, M)
                    c[...] = c - a @ b
        k0 = kf * NB
        nb = min(NB, N - k0)
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
