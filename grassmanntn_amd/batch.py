"""Batched Levin-Nave TRG: many coupling points advanced together, one HIP-graph launch per step.

A batch of P site tensors of equal shape is one dense tensor T[B, i, j, k, l] whose first leg is
bosonic (statistics (0, 1, 1, -1, -1)).  A bosonic leg carried through every contraction changes no
Grassmann sign, so each step is the reference's TRG step (gauge2d.py:1647-1755) applied to every
point at once: both leg permutations folded into one batched join gather each, the 4P parity blocks
of the two decompositions in one gtn_svd_blocks call (one launch per stage for all of them), every
contraction a batched parity-block GEMM, and the per-point normalisation and FP64 log-norm
accumulation of example.py:191-193 on the device (gtn_normalize_batch).  Nothing is read back during
a run, so a step is recorded once as a HIP graph and replayed.

Truncation: every parity block keeps int(dcut/2) values and the new bond is 2 * 2^ceil(log2
int(dcut/2)), the reference's dense rule (__init__.py:3998-4015) for blocks of full rank.  A block of
lower rank gives zero singular values with zero vectors (the 1e-14 rank rule of :3939-3941 is applied
on the device), i.e. the reference's tensor padded with zero entries, whose norm, traces and ln Z are
the same.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import core
from ._device import DeviceArray
from ._perm import Perm
from ._plan import fresh_char, parse_partition
from .core import dense, einsum, join_legs, split_legs, precision, settings


def _mk(data, stats):
    r = dense.__new__(dense)
    r.data, r.statistics, r.encoder, r.format, r._even = data, tuple(stats), "parity-preserving", "matrix", True
    return r


def _blocks_batch(Mdata, P, m, n, out=None):
    """(P, 2, m/2, n/2) contiguous copy (into `out` when given) of the parity blocks M[b][p::2, p::2]
    (BlockSVD's ME / MO, __init__.py:3990-3996) of every point b."""
    mh, nh = m // 2, n // 2
    G = Perm()
    axes = ((P, m * n), (2, n + 1), (mh, 2 * n), (nh, 2))
    for dim, stride in axes:
        i = G.add_in(dim, stride)
        v = G.add_virt(dim)
        G.add_comp(v, dim, [(i, 1)])
        G.add_out(dim, [(v, 1)])
    return G.run(Mdata, dst=out)


def _interleave_batch(stack, P, rows, cols, row_stride, par_stride, batch_stride):
    """full[b][r][c] = stack[b][r % 2][r // 2][c // 2] where r and c have equal parity, else 0: the
    interleaving of get_full_matrix (__init__.py:4017-4028) for every point b."""
    G = Perm()
    ib = G.add_in(P, batch_stride)
    ir = G.add_in(rows, row_stride, enc=True, split=True, stride_lo=par_stride)
    ic = G.add_in(cols, 1, enc=True, split=True, stride_lo=0)
    vb = G.add_virt(P)
    G.add_comp(vb, P, [(ib, 1)])
    vr = G.add_virt(rows, enc=True)
    G.add_comp(vr, rows, [(ir, 1)], bit=0)
    vc = G.add_virt(cols, enc=True)
    G.add_comp(vc, cols, [(ic, 1)], bit=1)
    G.add_out(P, [(vb, 1)])
    G.add_out(rows, [(vr, 1)])
    G.add_out(cols, [(vc, 1)])
    G.set_table([0, 2, 2, 0])
    return G.run(stack)


def _bond(k):
    return 2 * (2 ** int(math.ceil(math.log2(k))) if k > 1 else 1)


def svd_batch(calls, cutoff):
    """[(T, string), ...] -> [(U, S, V), ...]: the dense svd (__init__.py:4033-4308) of every point of
    batched tensors T[B, ...] (first leg bosonic, the batch) with `string` grouping the other legs,
    all parity blocks of all calls in one device decomposition.  Each block keeps int(cutoff/2)
    values; see the module docstring for blocks of lower rank."""
    lib = L.lib()
    jobs = []
    for T, string in calls:
        if T.statistics[0] != 0:
            raise ValueError("svd_batch: the first leg must be the bosonic batch leg")
        left, right = parse_partition(string, "svd")
        z = fresh_char(left + right)
        P = T.shape[0]
        spec = "(%s)(%s)(%s)" % (z, left, right)
        try:
            # the join gather writes the parity blocks of every point's matrix directly
            G, src, (_, m, n), mstats = core._join_gather(T, spec, "matrix", (0, -1, 1), blocks=True)
            M = None
        except ValueError:  # a group mixing bosonic and fermionic legs: join, then cut the blocks
            G, src = None, None
            M = join_legs(T, spec, "matrix", intermediate_stat=(0, -1, 1))
            _, m, n = M.shape
            mstats = M.statistics
        mh, nh = m // 2, n // 2
        k = min(mh, nh, int(cutoff / 2))
        if k < 1:
            raise ValueError("svd_batch: cutoff %s keeps no values" % cutoff)
        d = _bond(k) // 2
        jobs.append(dict(T=T, left=left, right=right, z=z, P=P, M=M, G=G, src=src, mstats=tuple(mstats), m=m, n=n,
                         mh=mh, nh=nh, k=k, d=d))
    # calls of equal block shape share one buffer each for blocks, U, S and V, so that the device
    # decomposition sees all their blocks at constant strides (one launch per stage)
    shapes = {}
    for j in jobs:
        shapes.setdefault((j["mh"], j["nh"], j["d"]), []).append(j)
    for (mh, nh, d), js in shapes.items():
        tot = sum(j["P"] for j in js)
        blocks = DeviceArray.empty((tot, 2, mh, nh))
        SU, SV, sv = DeviceArray.zeros((tot, 2, mh, d)), DeviceArray.zeros((tot, 2, d, nh)), DeviceArray.zeros((tot, 2, d))
        at = 0
        for j in js:
            P = j["P"]
            j["blocks"] = blocks.flat(at * 2 * mh * nh, P * 2 * mh * nh, (P, 2, mh, nh))
            if j["G"] is not None:
                j["G"].run(j["src"], dst=j["blocks"])
            else:
                _blocks_batch(j["M"].data, P, j["m"], j["n"], out=j["blocks"])
            j["SU"] = SU.flat(at * 2 * mh * d, P * 2 * mh * d, (P, 2, mh, d))
            j["SV"] = SV.flat(at * 2 * d * nh, P * 2 * d * nh, (P, 2, d, nh))
            j["sv"] = sv.flat(at * 2 * d, P * 2 * d, (P, 2, d))
            at += P
    nblk = sum(2 * j["P"] for j in jobs)
    arr = (L.SvdBlock * nblk)()
    t = 0
    for j in jobs:
        mh, nh, d, k = j["mh"], j["nh"], j["d"], j["k"]
        for q in range(2 * j["P"]):
            b = arr[t]
            b.M = j["blocks"].data_ptr() + 8 * q * mh * nh
            b.U = j["SU"].data_ptr() + 8 * q * mh * d
            b.S = j["sv"].data_ptr() + 8 * q * d
            b.Vt = j["SV"].data_ptr() + 8 * q * d * nh
            b.m, b.n, b.lda, b.ldu, b.ldv, b.k = mh, nh, nh, d, nh, k
            t += 1
    opts = L.SvdOpts()
    opts.oversample = settings.svd_oversample
    opts.iters = core.svd_iters_for(max(j["k"] for j in jobs))
    opts.precision, opts.f64_tail = core.svd_spec(settings.svd_precision)
    opts.mode = 0
    opts.seed = settings.seed
    opts.orth_every = settings.svd_orth_every
    ws_bytes = lib.gtn_svd_workspace_bytes(arr, nblk, C.byref(opts))
    ws = DeviceArray.empty(max(ws_bytes // 8, 1), np.float64)
    L.check(lib.gtn_svd_blocks(arr, nblk, C.byref(opts), L.ptr(ws), ws_bytes, L.stream_ptr()), "gtn_svd_blocks")
    svd_batch.last_values = [j["sv"] for j in jobs]
    out = []
    for j in jobs:
        P, m, n, mh, nh, d = j["P"], j["m"], j["n"], j["mh"], j["nh"], j["d"]
        T, left, right, z = j["T"], j["left"], j["right"], j["z"]
        U = _interleave_batch(j["SU"], P, m, 2 * d, row_stride=d, par_stride=mh * d, batch_stride=2 * mh * d)
        V = _interleave_batch(j["SV"], P, 2 * d, n, row_stride=nh, par_stride=d * nh, batch_stride=2 * d * nh)
        S = DeviceArray.empty((P, 2 * d, 2 * d))
        L.check(lib.gtn_diag_interleave(L.ptr(S), P, d, L.ptr(j["sv"]), L.GTN_F64, L.stream_ptr()),
                "gtn_diag_interleave")
        s0, s1 = j["mstats"][1], j["mstats"][2]
        U, S, V = _mk(U, (0, s0, 1)), _mk(S, (0, -1, 1)), _mk(V, (0, -1, s1))
        nl = len(left)
        xstats, xshape = T.statistics[1:], T.shape[1:]
        nc = fresh_char(left + right + z)
        U = split_legs(U, "(%s)(%s)%s" % (z, left, nc), (0,) + tuple(xstats[:nl]) + (1,),
                       (P,) + tuple(xshape[:nl]) + (2 * d,), intermediate_stat=(0, -1, 1))
        S = S.switch_encoder()
        V = split_legs(V, "(%s)%s(%s)" % (z, nc, right), (0, -1) + tuple(xstats[nl:]),
                       (P, 2 * d) + tuple(xshape[nl:]), intermediate_stat=(0, -1, 1))
        res = [U, S, V]
        if T.format == "standard":
            res = [x.switch_format() for x in res]
        if T.encoder == "parity-preserving":
            res = [x.switch_encoder() for x in res]
        for x in res:
            x._even = True
        out.append(tuple(res))
    return out


def trg_batch_step(Tb, dcut):
    """The contractions of one TRG step (gauge2d.py:1660-1734) on a batch Tb[B, i, j, k, l]: returns
    the unnormalised T'[B, i, j, k, l] and the singular values of the first decomposition (T1's
    parity blocks, (B, 2, d): even then odd, each decreasing)."""
    T1 = core.permuted_view(Tb, "Bijkl->Bjkli")
    T2 = core.permuted_view(Tb, "Bijkl->Bklij")
    (U1, S1, V1), (U2, S2, V2) = svd_batch([(T1, "ab|cd"), (T2, "ab|cd")], dcut)
    sv = svd_batch.last_values[0]
    r = core.sqrt(S1)
    U1 = einsum("Babx,Bxc->Babc", U1, r)
    V1 = einsum("Bax,Bxbc->Babc", r, V1)
    r = core.sqrt(S2)
    U2 = einsum("Babx,Bxc->Babc", U2, r)
    V2 = einsum("Bax,Bxbc->Babc", r, V2)
    # VV and UU stay in their GEMMs' parity-block layout: their output gathers run inside the
    # operand gathers of the T' contraction (one pass over each instead of two)
    with core.deferred_outputs():
        VV = einsum("Bkwz,Blxw->Blxzk", V1, V2)
        UU = einsum("Byxi,Bzyj->Bjzxi", U1, U2)
    return einsum("Blxzk,Bjzxi->Bijkl", VV, UU), sv


def normalize_batch(T, nrm2, lognorm=None, hist=None, step=None, out=None):
    """Every point of T[B, ...] divided by its Frobenius norm, in place or into the device array
    `out` (T's element count); nrm2[b] gets the squared norm, lognorm[b] the driver's
    2 lognorm[b] + log ||T_b||, hist[b, step] = log ||T_b||."""
    P = T.shape[0]
    n = T.data.numel() // P
    hl = hist.shape[1] if hist is not None else 0
    if out is not None and out.numel() != T.data.numel():
        raise ValueError("normalize_batch: output of %d values for %d" % (out.numel(), T.data.numel()))
    L.check(L.lib().gtn_normalize_batch(L.ptr(T.data), L.ptr(out) if out is not None else None, P, n,
                                        L.dtype_code(T.data), L.ptr(nrm2),
                                        L.ptr(lognorm) if lognorm is not None else None,
                                        L.ptr(hist) if hist is not None else None, hl,
                                        L.ptr(step) if step is not None else None, L.stream_ptr()),
            "gtn_normalize_batch")
    return T


def logz_batch(T):
    """Per point b: ln of the supertrace einsum('IJIJ', T_b) (gauge2d.py:1599-1615, periodic), the
    complex logarithm where Z_b < 0."""
    Z = einsum("BIJIJ->B", T).data.numpy().astype(np.complex128)
    return np.log(Z)


def wrap_batch(data, P):
    """A batch tensor (statistics (0, 1, 1, -1, -1), parity-preserving) over device data (P, D^4)."""
    D = int(round((data.numel() // P) ** 0.25))
    t = dense.__new__(dense)
    t.data = data.reshape((P, D, D, D, D))
    t.statistics = (0, 1, 1, -1, -1)
    t.encoder, t.format, t._even = "parity-preserving", "standard", True
    return t


class Graph:
    """The device work of one call of fn(), recorded as a HIP graph: allocations made while
    recording come from memory pool `pool`, so every buffer the graph writes stays its own; what fn
    returns refers to buffers that every replay overwrites."""

    def __init__(self, fn, pool):
        lib = L.lib()
        s = L.stream_ptr()
        L.check(lib.gtn_graph_begin(s, pool), "gtn_graph_begin")
        try:
            self.result = fn()
        except BaseException:
            lib.gtn_graph_end(s, None)
            raise
        x = C.c_void_p()
        L.check(lib.gtn_graph_end(s, C.byref(x)), "gtn_graph_end")
        self.exec = x

    def launch(self):
        L.check(L.lib().gtn_graph_launch(self.exec, L.stream_ptr()), "gtn_graph_launch")

    def destroy(self):
        if self.exec is not None:
            L.check(L.lib().gtn_graph_destroy(self.exec), "gtn_graph_destroy")
            self.exec = None


_next_pool = [0]


def _new_pool():
    _next_pool[0] += 1
    return _next_pool[0]


class BatchedTRG:
    """TRG of P coupling points of equal leg dimension together.

    schedule(i) -> (contraction, svd) arithmetic of step i (default FP64 with the library's SVD; see
    core.mixed_schedule).  With graphs=True each distinct (input shape, arithmetic) step is recorded
    once as a HIP graph (after one eager warm-up step that fills the library's plan and table caches)
    and every step of a run is one graph launch; graphs=False runs the same steps eagerly.

    run(T0, steps) -> ln Z per site of every point after `steps` steps (complex where the final
    supertrace is negative), with the per-step log norms in self.log_norms (P x steps) and the last
    step's singular values of T1 (both parity blocks, decreasing, divided by the largest) in
    self.final_sv (P x bond)."""

    def __init__(self, P, D, dcut, steps, schedule=None, graphs=True):
        self.P, self.D, self.dcut, self.steps = P, D, dcut, steps
        self.schedule = schedule or (lambda i: ("f64", settings.svd_precision))
        self.use_graphs = graphs
        self.Dn = _bond(min(D * D // 2, int(dcut / 2)))  # bond after the first step
        self.x0 = DeviceArray.zeros((P, D ** 4))
        self.xs = self.x0 if self.Dn == D else DeviceArray.zeros((P, self.Dn ** 4))
        self.nrm2 = DeviceArray.zeros(P)
        self.lognorm = DeviceArray.zeros(P)
        self.hist = DeviceArray.zeros((P, max(steps, 1)))
        self.stepc = DeviceArray.zeros(1)
        self.svals = DeviceArray.zeros((P, self.Dn))  # singular values of the last step's T1 blocks
        self.pool = _new_pool() if graphs else 0
        self.graphs = {}

    # ---- one step: input buffer -> output buffer -------------------------------------------------
    def _step(self, x_in, x_out, prec, bookkeeping=True):
        from ._device import copy2d
        T = wrap_batch(x_in, self.P)
        with precision(*prec):
            Tn, sv = trg_batch_step(T, self.dcut)
        if bookkeeping:
            normalize_batch(Tn, self.nrm2, self.lognorm, self.hist, self.stepc, out=x_out)
            copy2d(self.svals, self.Dn, sv.reshape((self.P, self.Dn)), self.Dn, self.P, self.Dn)
        else:
            normalize_batch(Tn, DeviceArray.empty(self.P), out=x_out)

    def _key(self, i):
        return (0 if i == 0 else 1, tuple(self.schedule(i)))

    def prepare(self, T0):
        """Warm-up steps and graph captures for every stage of a `steps`-step run from T0 (a host
        array (P, D, D, D, D) or device array of P D^4 values)."""
        if not self.use_graphs:
            return
        self.load(T0)
        w0 = self.x0.clone()
        w1 = DeviceArray.empty(self.xs.shape)
        self._step(w0, w1, self.schedule(0), bookkeeping=False)
        warm = {self._key(0)}
        for i in range(1, self.steps):
            k = self._key(i)
            if k not in warm:
                w2 = DeviceArray.empty(self.xs.shape)
                self._step(w1, w2, self.schedule(i), bookkeeping=False)
                warm.add(k)
        del w0, w1
        for i in range(self.steps):
            k = self._key(i)
            if k in self.graphs:
                continue
            x_in = self.x0 if i == 0 else self.xs
            self.graphs[k] = Graph(lambda: self._step(x_in, self.xs, self.schedule(i)), self.pool)

    def load(self, T0):
        if isinstance(T0, DeviceArray):
            self.x0.copy_from(T0.reshape(self.x0.shape))
        else:
            a = np.ascontiguousarray(np.asarray(T0, dtype=np.float64).reshape(self.x0.shape))
            L.check(L.lib().gtn_memcpy(C.c_void_p(self.x0.data_ptr()), a.ctypes.data_as(C.c_void_p), a.nbytes, 1,
                                       L.stream_ptr()), "gtn_memcpy")

    def start(self, T0):
        """Load T0 and clear the log norms (device work, no synchronisation beyond the upload)."""
        self.load(T0)
        for b in (self.nrm2, self.lognorm, self.hist, self.stepc):
            b.fill_(0.0)

    def advance(self):
        """All `steps` steps (graph launches, or eager steps without graphs); nothing is read back."""
        for i in range(self.steps):
            if self.use_graphs:
                self.graphs[self._key(i)].launch()
            else:
                self._step(self.x0 if i == 0 else self.xs, self.xs, self.schedule(i))

    def result(self):
        """ln Z per site of every point after the run (example.py:186-193)."""
        T = wrap_batch(self.xs, self.P)
        with precision("f64", "f64"):
            lz = logz_batch(T)
        logn = self.lognorm.numpy()
        self.log_norms = self.hist.numpy()
        sv = -np.sort(-self.svals.numpy(), axis=1)
        self.final_sv = sv / np.where(sv[:, :1] > 0, sv[:, :1], 1.0)
        return (lz + logn) / 2.0 ** self.steps

    def run(self, T0):
        self.start(T0)
        self.advance()
        return self.result()

    def close(self):
        for g in self.graphs.values():
            g.destroy()
        self.graphs = {}
        if self.pool:
            L.check(L.lib().gtn_pool_release(self.pool), "gtn_pool_release")
            self.pool = 0
