"""CPU emulation of libgtn_hip.so's kernels -- TEST INFRASTRUCTURE ONLY.

Used by CPU tests to drive the real host layer of grassmanntn_amd (descriptor building, layouts,
sign tables, GEMM grouping) without a GPU: every C-ABI entry point is replaced by a numpy
re-statement of what the kernel computes, with the same index arithmetic, and every memory access
range is checked against the allocation it falls in.  Tensors live on the CPU while it is active.
It is never used by the product: it is installed only by the `emulated` pytest fixture.
"""
import contextlib
import ctypes as C

import numpy as np


def _popc(x):
    x = np.asarray(x, dtype=np.uint64)
    c = np.zeros(x.shape, dtype=np.uint64)
    y = x.copy()
    while np.any(y):
        c += y & np.uint64(1)
        y >>= np.uint64(1)
    return c


def _enc(x):
    x = np.asarray(x, dtype=np.uint64)
    return x ^ (_popc(x >> np.uint64(1)) & np.uint64(1))


def _sigma_bit(x):
    return (_popc(x) >> np.uint64(1)) & np.uint64(1)


class _Registry:
    """Host buffers standing in for device allocations, with their address ranges."""

    def __init__(self):
        self.live = {}

    def alloc(self, nbytes):
        buf = np.zeros(max(int(nbytes), 1), dtype=np.uint8)
        addr = buf.ctypes.data
        self.live[addr] = buf
        return addr

    def free(self, addr):
        if addr not in self.live:
            raise AssertionError("free of an address that was not allocated")
        del self.live[addr]

    def check(self, lo, hi, what):
        """[lo, hi) bytes must lie inside one live allocation."""
        for b, buf in self.live.items():
            if b <= lo and hi <= b + buf.nbytes:
                return
        raise AssertionError("%s: access [%#x, %#x) outside every allocation" % (what, lo, hi))


NP = {0: np.float64, 1: np.float32, 2: np.complex128}
ESZ = {0: 8, 1: 4, 2: 16}


def _view(addr, count, dtype):
    arr = (C.c_byte * (count * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(arr, dtype=dtype)


class FakeLib:
    def __init__(self, reg):
        self.reg = reg
        self.calls = []
        self.err = b""

    # ---- helpers ------------------------------------------------------------------------------
    def _arr(self, p, n, dt, what):
        addr = p.value if isinstance(p, C.c_void_p) else p
        self.reg.check(addr, addr + n * np.dtype(dt).itemsize, what)
        return _view(addr, n, dt)

    def gtn_version(self):
        return 100

    def gtn_last_error(self):
        return self.err

    def gtn_device_info(self, name, n, ncu):
        return 0

    def gtn_set_dry_run(self, n, bases, sizes):
        return 0

    def gtn_mem_info(self, live, cached, free, total):
        for ref, v in ((live, 0), (cached, 0), (free, 64 << 30), (total, 64 << 30)):
            ref._obj.value = v
        return 0

    # memory pools as the runtime keeps them: a block allocated while a capture is open belongs to
    # that capture's pool; freeing it returns it to the pool (it stays allocated, since a recorded
    # graph may address it) for reuse by later allocations of the same pool
    def gtn_malloc(self, nbytes, out):
        pool = getattr(self, "pool_open", 0)
        if not pool:
            out._obj.value = self.reg.alloc(nbytes)
            return 0
        cls = max((int(nbytes) + 255) // 256 * 256, 256)
        self.pool_free = getattr(self, "pool_free", {})
        self.pool_of = getattr(self, "pool_of", {})
        fl = self.pool_free.setdefault(pool, {}).setdefault(cls, [])
        addr = fl.pop() if fl else self.reg.alloc(cls)
        self.pool_of[addr] = (pool, cls)
        out._obj.value = addr
        return 0

    def gtn_free(self, p):
        addr = p.value if isinstance(p, C.c_void_p) else p
        pool_of = getattr(self, "pool_of", {})
        if addr in pool_of:
            pool, cls = pool_of[addr]
            if pool not in getattr(self, "pool_released", set()):
                self.pool_free[pool][cls].append(addr)
                return 0
            del pool_of[addr]
        self.reg.free(addr)
        return 0

    def gtn_memcpy(self, dst, src, nbytes, kind, stream):
        d = dst.value if isinstance(dst, C.c_void_p) else dst
        s_ = src.value if isinstance(src, C.c_void_p) else src
        if kind in (1, 3):
            self.reg.check(d, d + nbytes, "memcpy dst")
        if kind in (2, 3):
            self.reg.check(s_, s_ + nbytes, "memcpy src")
        C.memmove(d, s_, nbytes)
        return 0

    def gtn_fill(self, p, n, dtype, value, stream):
        a = self._arr(p, n, NP[dtype] if dtype in NP else np.int8, "fill")
        a[:] = value
        return 0

    def gtn_copy2d(self, dst, ldd, src, lds, rows, cols, dtype, stream):
        dt = NP[dtype]
        d = self._arr(dst, (rows - 1) * ldd + cols, dt, "copy2d dst")
        s_ = self._arr(src, (rows - 1) * lds + cols, dt, "copy2d src")
        for r in range(rows):
            d[r * ldd:r * ldd + cols] = s_[r * lds:r * lds + cols]
        return 0

    def gtn_synchronize(self):
        return 0

    def gtn_event_create(self, p):
        p._obj.value = 1
        return 0

    def gtn_event_record(self, ev, stream):
        return 0

    def gtn_event_elapsed_ms(self, a, b, ms):
        ms._obj.value = 0.0
        return 0

    def gtn_event_destroy(self, ev):
        return 0

    def gtn_index_tables(self, n, popc, sig, enc, stream):
        i = np.arange(n, dtype=np.uint64)
        if popc:
            self._arr(popc, n, np.int32, "index_tables")[:] = _popc(i)
        if sig:
            self._arr(sig, n, np.int8, "index_tables")[:] = 1 - 2 * _sigma_bit(i).astype(np.int8)
        if enc:
            self._arr(enc, n, np.int32, "index_tables")[:] = _enc(i)
        return 0

    def gtn_sign_tensor(self, ndim, dims, table, out, stream):
        dims = [dims[i] for i in range(ndim)]
        total = int(np.prod(dims)) if dims else 1
        idx = np.arange(total, dtype=np.uint64)
        bits = np.zeros(total, dtype=np.uint64)
        r = idx.copy()
        for ax in range(ndim - 1, -1, -1):
            k = r % np.uint64(dims[ax])
            r //= np.uint64(dims[ax])
            bits |= (_popc(k) & np.uint64(1)) << np.uint64(ax)
        tab = np.array([table[i] for i in range(32)], dtype=np.uint64)
        neg = (tab[(bits >> np.uint64(5)).astype(np.int64)] >> (bits & np.uint64(31))) & np.uint64(1)
        self._arr(out, total, np.int8, "sign_tensor")[:] = np.where(neg == 1, -1, 1)
        return 0

    def gtn_permute(self, dp, src, dst, stream):
        d = dp._obj
        self.calls.append(("permute", d.n_out))
        return self._gather(d, lambda n: self._arr(src, n, NP[d.dtype], "permute src"), dst)

    def gtn_permute2(self, dpo, dpn, src, dst, stream):
        do, dn = dpo._obj, dpn._obj
        self.calls.append(("permute2", do.n_out, dn.n_out))
        if do.dtype != dn.dtype:
            return self._fail("permute2: dtypes differ")
        mid_n = 1
        for a in range(dn.n_out):
            mid_n *= dn.out_dim[a]
        if do.src_elems != mid_n:
            return self._fail("permute2: the outer gather reads %d elements, the inner one makes %d" % (do.src_elems, mid_n))
        mid = np.zeros(mid_n, dtype=NP[dn.dtype])
        self._gather(dn, lambda n: self._arr(src, n, NP[dn.dtype], "permute2 src"), None, out=mid)
        return self._gather(do, lambda n: mid[:n], dst)

    def _gather(self, d, src_of, dst, out=None):
        """The signed gather of descriptor d: src_of(n) gives the first n source elements; the result
        goes to device address dst (or the host array `out`)."""
        n_out, n_virt, n_in = d.n_out, d.n_virt, d.n_in
        comps = []
        for v in range(n_virt):
            comps += [v] * d.virt_ncomp[v]
        if len(comps) != d.n_comp:
            raise AssertionError("components do not add up")
        total = 1
        for a in range(n_out):
            total *= d.out_dim[a]
        dt = NP[d.dtype]
        dst_a = out if out is not None else self._arr(dst, total, dt, "permute dst")
        if total == 0:
            return 0
        lin = np.arange(total, dtype=np.uint64)
        V = [np.zeros(total, dtype=np.uint64) for _ in range(max(n_virt, 1))]
        rest = lin.copy()
        for ax in range(n_out - 1, -1, -1):
            dd = np.uint64(d.out_dim[ax])
            o = rest % dd
            rest //= dd
            for t in range(2):
                v = d.out_virt[ax][t]
                if v >= 0:
                    V[v] += o * np.uint64(d.out_vmul[ax][t])
        F = [np.zeros(total, dtype=np.uint64) for _ in range(max(n_in, 1))]
        bits = np.zeros(total, dtype=np.uint64)
        neg = np.zeros(total, dtype=np.uint64)
        j = d.n_comp
        for v in range(n_virt - 1, -1, -1):
            if np.any(V[v] >= np.uint64(d.virt_dim[v])):
                raise AssertionError("virtual axis %d overflows" % v)
            c = _enc(V[v]) if d.virt_enc[v] else V[v].copy()
            if d.virt_sigma[v]:
                neg ^= _sigma_bit(c)
            for _ in range(d.virt_ncomp[v]):
                j -= 1
                cd = np.uint64(d.comp_dim[j])
                k = c % cd
                c //= cd
                pc = _popc(k)
                if d.comp_bit[j] >= 0:
                    bits ^= (pc & np.uint64(1)) << np.uint64(d.comp_bit[j])
                if d.comp_sigma[j]:
                    neg ^= (pc >> np.uint64(1)) & np.uint64(1)
                for t in range(2):
                    b = d.comp_in[j][t]
                    if b >= 0:
                        F[b] += k * np.uint64(d.comp_mult[j][t])
        off = np.zeros(total, dtype=np.int64)
        for b in range(n_in):
            f = F[b]
            if np.any(f >= np.uint64(d.in_dim[b])):
                raise AssertionError("input axis %d overflows" % b)
            if d.in_sigma[b]:
                neg ^= _sigma_bit(f)
            e = _enc(f) if d.in_enc[b] else f
            if d.in_split[b]:
                off += (e & np.uint64(1)).astype(np.int64) * d.in_stride_lo[b] + (e >> np.uint64(1)).astype(np.int64) * d.in_stride[b]
            else:
                off += e.astype(np.int64) * d.in_stride[b]
        words = np.array([d.table[i] for i in range(len(d.table))], dtype=np.uint64)
        code = (words[(bits >> np.uint64(4)).astype(np.int64)] >> ((bits & np.uint64(15)) * np.uint64(2))) & np.uint64(3)
        zero = code == 2
        sign = np.where((code ^ neg) & np.uint64(1), -1.0, 1.0)
        if off.min() < 0 or off.max() >= d.src_elems:
            raise AssertionError("permute reads outside src_elems")
        src_a = src_of(int(off.max()) + 1)
        alpha = complex(d.alpha_re, d.alpha_im) if d.dtype == 2 else d.alpha_re
        vals = src_a[off] * sign * alpha
        vals[zero] = 0
        dst_a[:] = vals
        return 0

    def gtn_gemm_grouped(self, groups, n, dtype, prec, stream):
        dt = NP[dtype]
        es = ESZ[dtype]
        for g in range(n):
            d = groups[g]
            self.calls.append(("gemm", d.M, d.N, d.K, d.batch))
            if d.M == 0 or d.N == 0 or d.batch == 0:
                continue
            if not (d.sak == 1 or d.sam == 1 or d.M <= 1 or d.K <= 1):
                return self._fail("A needs a unit stride")
            if not (d.sbk == 1 or d.sbn == 1 or d.N <= 1 or d.K <= 1):
                return self._fail("B needs a unit stride")
            for b in range(d.batch):
                a0 = d.A + es * b * d.sab
                b0 = d.B + es * b * d.sbb
                c0 = d.C + es * b * d.scb
                ka = (d.M - 1) * d.sam + (d.K - 1) * d.sak if d.K > 0 else 0
                kb = (d.K - 1) * d.sbk + (d.N - 1) * d.sbn if d.K > 0 else 0
                kc = (d.M - 1) * d.scm + (d.N - 1) * d.scn
                if d.K > 0:
                    self.reg.check(a0, a0 + es * (ka + 1), "gemm A")
                    self.reg.check(b0, b0 + es * (kb + 1), "gemm B")
                self.reg.check(c0, c0 + es * (kc + 1), "gemm C")
                Cv = _view(c0, kc + 1, dt)
                if d.K > 0:
                    Av = _view(a0, ka + 1, dt)
                    Bv = _view(b0, kb + 1, dt)
                    Am = np.lib.stride_tricks.as_strided(Av, (d.M, d.K), (d.sam * es, d.sak * es))
                    Bm = np.lib.stride_tricks.as_strided(Bv, (d.K, d.N), (d.sbk * es, d.sbn * es))
                    if prec in (4, 5):
                        # BF16 operands (mode 5: the first split hi + lo), products accumulated wider
                        from oracle.mlp_oracle import bf16
                        A64 = Am.astype(np.float64)
                        hi = bf16(A64)
                        A2 = hi + bf16(A64 - hi) if prec == 5 else hi
                        prod = (A2 @ bf16(Bm.astype(np.float64))).astype(dt)
                    else:
                        prod = Am @ Bm
                else:
                    prod = np.zeros((d.M, d.N), dtype=dt)
                Cm = np.lib.stride_tricks.as_strided(Cv, (d.M, d.N), (d.scm * es, d.scn * es))
                if d.beta != 0.0:
                    Cm[...] = d.alpha * prod + d.beta * Cm
                else:
                    Cm[...] = d.alpha * prod
        return 0

    def _fail(self, msg):
        self.err = msg.encode()
        return 1

    def gtn_sumsq(self, x, n, dtype, out, stream):
        v = self._arr(x, n, NP[dtype], "sumsq")
        self._arr(out, 1, np.float64, "sumsq out")[0] = float(np.sum(np.abs(v) ** 2))
        return 0

    def gtn_scale_by_norm(self, x, n, dtype, nrm2, stream):
        v = self._arr(x, n, NP[dtype], "scale")
        v /= np.sqrt(self._arr(nrm2, 1, np.float64, "scale nrm")[0])
        return 0

    def gtn_reduce_last(self, src, dst, rows, inner, dtype, stream):
        s = self._arr(src, rows * inner, NP[dtype], "reduce src")
        self._arr(dst, rows, NP[dtype], "reduce dst")[:] = s.reshape(rows, inner).sum(axis=1)
        return 0

    def gtn_lognorm_update(self, ln, nrm2, stream):
        a = self._arr(ln, 1, np.float64, "lognorm")
        a[0] = 2 * a[0] + 0.5 * np.log(self._arr(nrm2, 1, np.float64, "lognorm nrm")[0])
        return 0

    def gtn_power_diag(self, x, n, stride, dtype, p, rcond, stream):
        v = self._arr(x, (n - 1) * stride + 1, NP[dtype], "power")
        w = v[::stride]
        with np.errstate(invalid="ignore"):
            v[::stride] = np.where(np.abs(w) > rcond, np.power(w.astype(np.float64), p), 0.0)
        return 0

    def gtn_elementwise(self, x, n, dtype, op, p, rcond, stream):
        if n == 0:
            return 0
        v = self._arr(x, n, NP[dtype], "elementwise")
        w = v.astype(np.float64)
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            f = 1.0 / w if op == 1 else (np.exp(w) if op == 2 else np.power(w, p))
            v[:] = np.where(np.abs(w) > rcond, f, 0.0)
        return 0

    def gtn_axpby(self, x, y, out, n, dtype, ar, ai, br, bi, stream):
        dt = NP[dtype]
        a = complex(ar, ai) if dtype == 2 else ar
        b = complex(br, bi) if dtype == 2 else br
        xv = self._arr(x, n, dt, "axpby x")
        r = a * xv
        if y:
            r = r + b * self._arr(y, n, dt, "axpby y")
        self._arr(out, n, dt, "axpby out")[:] = r
        return 0

    def gtn_odd_norm(self, x, ndim, dims, ferm, pp, dtype, out, stream):
        dims_ = [dims[i] for i in range(ndim)]
        n = int(np.prod(dims_)) if dims_ else 1
        v = self._arr(x, n, NP[dtype], "odd_norm").reshape(dims_)
        par = np.zeros(dims_, dtype=np.int64)
        for ax in range(ndim):
            if ferm[ax]:
                i = np.arange(dims_[ax])
                p = (i & 1) if pp[ax] else (_popc(i).astype(np.int64) & 1)
                sh = [1] * ndim
                sh[ax] = dims_[ax]
                par = par + p.reshape(sh)
        self._arr(out, 1, np.float64, "odd out")[0] = float(np.sum(np.abs(v[(par % 2) == 1]) ** 2))
        return 0

    def gtn_normalize_batch(self, x, dst, nb, n, dtype, nrm2, lognorm, hist, hist_len, step, stream):
        v = self._arr(x, nb * n, NP[dtype], "normalize x").reshape(nb, n)
        out = self._arr(dst, nb * n, NP[dtype], "normalize dst").reshape(nb, n) if dst else v
        q = (v.astype(np.float64) ** 2).sum(axis=1)
        self._arr(nrm2, nb, np.float64, "normalize nrm2")[:] = q
        ln = 0.5 * np.log(q)
        if lognorm:
            a = self._arr(lognorm, nb, np.float64, "normalize lognorm")
            a[:] = 2 * a + ln
        if hist and step:
            k = int(self._arr(step, 1, np.float64, "normalize step")[0])
            if 0 <= k < hist_len:
                h = self._arr(hist, nb * hist_len, np.float64, "normalize hist").reshape(nb, hist_len)
                h[:, k] = ln
        out[...] = v / np.sqrt(q)[:, None].astype(v.dtype)
        if step:
            self._arr(step, 1, np.float64, "normalize step")[0] += 1
        return 0

    def gtn_diag_interleave(self, dst, nb, d, src, dtype, stream):
        sv = self._arr(src, nb * 2 * d, NP[dtype], "diag src").reshape(nb, 2, d)
        out = self._arr(dst, nb * 4 * d * d, NP[dtype], "diag dst").reshape(nb, 2 * d, 2 * d)
        out[...] = 0
        r = np.arange(2 * d)
        out[:, r, r] = sv[:, r % 2, r // 2]
        return 0

    # ---- MLP stages (the formulas of include/gtn_hip.h; FP32 storage) ---------------------------------
    def gtn_mlp_bias_act(self, y, E, n, c, bias, act, pre, stream):
        from oracle.mlp_oracle import gelu
        Y = self._arr(y, E * n * c, np.float32, "bias_act y").reshape(E, n, c)
        v = Y.astype(np.float64) + self._arr(bias, E * c, np.float32, "bias").reshape(E, 1, c)
        if pre:
            self._arr(pre, E * n * c, np.float32, "bias_act pre")[:] = v.reshape(-1)
        Y[...] = gelu(v) if act == 1 else v
        return 0

    def gtn_mlp_act_backward(self, dy, pre, count, act, stream):
        from oracle.mlp_oracle import gelu_grad
        d = self._arr(dy, count, np.float32, "act_backward dy")
        d[:] = d * gelu_grad(self._arr(pre, count, np.float32, "act_backward pre").astype(np.float64))
        return 0

    def gtn_mlp_colsum_workspace(self, E, c):
        return E * c * 32 * 8

    def gtn_mlp_colsum(self, x, E, n, c, out, part, stream):
        X = self._arr(x, E * n * c, np.float32, "colsum x").reshape(E, n, c).astype(np.float64)
        self._arr(out, E * c, np.float32, "colsum out")[:] = X.sum(axis=1).reshape(-1)
        return 0

    def gtn_mlp_mse(self, out, target, weight, E, n, c, dout, loss, stream):
        O_ = self._arr(out, E * n * c, np.float32, "mse out").reshape(E, n, c).astype(np.float64)
        T = self._arr(target, n * c, np.float32, "mse target").reshape(n, c).astype(np.float64)
        w = self._arr(weight, c, np.float32, "mse weight").astype(np.float64)
        d = O_ - T[None]
        self._arr(dout, E * n * c, np.float32, "mse dout")[:] = (2.0 * w * d / n).reshape(-1)
        if loss:
            self._arr(loss, E, np.float64, "mse loss")[:] = (w * d * d).sum(axis=(1, 2)) / n
        return 0

    def gtn_adamw(self, w, g, m, v, count, opts, step, stream):
        o = opts._obj
        st = self._arr(step, 1, np.float64, "adamw step")
        t = st[0] + 1.0
        frac = min((t - 1.0) / (o.total_steps - 1), 1.0) if o.total_steps > 1 else 0.0
        lr = o.lr_min + 0.5 * (o.lr_max - o.lr_min) * (1.0 + np.cos(np.pi * frac))
        W = self._arr(w, count, np.float32, "adamw w")
        G = self._arr(g, count, np.float32, "adamw g").astype(np.float64)
        Mv = self._arr(m, count, np.float32, "adamw m")
        Vv = self._arr(v, count, np.float32, "adamw v")
        mi = o.beta1 * Mv + (1 - o.beta1) * G
        vi = o.beta2 * Vv + (1 - o.beta2) * G * G
        Mv[:], Vv[:] = mi, vi
        W[:] = W * (1 - lr * o.weight_decay) - lr * (mi / (1 - o.beta1 ** t)) / (np.sqrt(vi / (1 - o.beta2 ** t)) + o.eps)
        st[0] += 1.0
        return 0

    def gtn_ensemble_stats(self, out, E, count, mean, sd, stream):
        X = self._arr(out, E * count, np.float32, "ens out").reshape(E, count).astype(np.float64)
        self._arr(mean, count, np.float32, "ens mean")[:] = X.mean(axis=0)
        if sd:
            self._arr(sd, count, np.float32, "ens sd")[:] = X.std(axis=0)
        return 0

    def gtn_gemm_grouped_noalloc(self, groups, n, dtype, prec, stream):
        return self.gtn_gemm_grouped(groups, n, dtype, prec, stream)

    def gtn_autotune_set(self, on):
        return 0

    def gtn_autotune_clear(self):
        return 0

    def gtn_autotune_report(self, out, n):
        return 0

    def gtn_gemm_noalloc_max_scratch(self):
        return 64 << 20

    # ---- streams and graph capture: calls issued during a capture are recorded with copies of their
    # ctypes arguments (the runtime copies kernel arguments when it records a launch) and replayed by
    # gtn_graph_launch; allocations run at once; a synchronous host copy inside a capture is an error
    def gtn_stream_create(self, p):
        p._obj.value = 0x5157
        return 0

    def gtn_stream_destroy(self, s):
        return 0

    def gtn_stream_synchronize(self, s):
        return 0

    def gtn_graph_begin(self, stream, pool):
        if getattr(self, "_rec", None) is not None:
            return self._fail("gtn_graph_begin: a capture is already open")
        self._rec = []
        self.pool_open = pool
        return 0

    def gtn_graph_end(self, stream, out):
        rec, self._rec = self._rec, None
        self.pool_open = 0
        if out is None:
            return 0
        self.graphs = getattr(self, "graphs", {})
        gid = len(self.graphs) + 1
        self.graphs[gid] = rec
        out._obj.value = gid
        return 0

    def gtn_graph_launch(self, g, stream):
        gid = g.value if isinstance(g, C.c_void_p) else g
        for name, args in self.graphs[gid]:
            rc = getattr(self, name)(*args)
            if rc:
                return rc
        return 0

    def gtn_graph_destroy(self, g):
        gid = g.value if isinstance(g, C.c_void_p) else g
        self.graphs.pop(gid, None)
        return 0

    def gtn_pool_release(self, pool):
        self.pool_released = getattr(self, "pool_released", set()) | {pool}
        for cls, blocks in getattr(self, "pool_free", {}).pop(pool, {}).items():
            for addr in blocks:
                del self.pool_of[addr]
                self.reg.free(addr)
        return 0

    def gtn_svd_workspace_bytes(self, blocks, n, opts):
        return 8

    def gtn_svd_blocks(self, blocks, n, opts, ws, wsb, stream):
        for i in range(n):
            b = blocks[i]
            self.calls.append(("svd", b.m, b.n, b.k))
            if b.k == 0:
                continue
            M = np.lib.stride_tricks.as_strided(self._arr(b.M, (b.m - 1) * b.lda + b.n, np.float64, "svd M"),
                                                (b.m, b.n), (b.lda * 8, 8)).copy()
            U_, s, Vh = np.linalg.svd(M, full_matrices=False)
            k = b.k
            if opts._obj.mode == 1:
                w, Z = np.linalg.eigh(M)
                order = np.argsort(-np.abs(w), kind="stable")
                s, U_ = w[order], Z[:, order]
                Vh = U_.T
            s0 = abs(s[0])
            keep = np.abs(s[:k] / (s0 + 1e-14)) > 1e-14
            U = U_[:, :k] * keep
            Vt = Vh[:k, :] * keep[:, None]
            for j in range(k):
                col = U[:, j]
                a = int(np.argmax(np.abs(col)))
                if col[a] < 0:
                    U[:, j] = -col
                    Vt[j, :] = -Vt[j, :]
            Ua = np.lib.stride_tricks.as_strided(self._arr(b.U, (b.m - 1) * b.ldu + k, np.float64, "svd U"),
                                                 (b.m, k), (b.ldu * 8, 8))
            Ua[...] = U
            self._arr(b.S, k, np.float64, "svd S")[:] = np.where(keep, s[:k], 0.0)
            if b.Vt:
                Va = np.lib.stride_tricks.as_strided(self._arr(b.Vt, (k - 1) * b.ldv + b.n, np.float64, "svd Vt"),
                                                     (k, b.n), (b.ldv * 8, 8))
                Va[...] = Vt
        return 0


_NOT_RECORDED = {"gtn_malloc", "gtn_free", "gtn_mem_info", "gtn_version", "gtn_last_error", "gtn_device_info", "gtn_set_dry_run",
                 "gtn_mlp_colsum_workspace", "gtn_autotune_set", "gtn_autotune_clear", "gtn_autotune_report",
                 "gtn_svd_workspace_bytes", "gtn_gemm_noalloc_max_scratch", "gtn_event_create", "gtn_event_destroy",
                 "gtn_stream_create", "gtn_stream_destroy", "gtn_graph_begin", "gtn_graph_end", "gtn_graph_launch",
                 "gtn_graph_destroy", "gtn_pool_release"}


def _copy_arg(a):
    if isinstance(a, (C.Structure, C.Array)):
        return type(a).from_buffer_copy(a)
    if type(a).__name__ == "CArgObject":
        return C.byref(_copy_arg(a._obj))
    if isinstance(a, C.c_void_p):
        return C.c_void_p(a.value)
    return a


class _Recorder:
    """The FakeLib seen through graph capture: while a capture is open, work calls are recorded."""

    def __init__(self, fake):
        self._fake = fake

    def __getattr__(self, name):
        f = getattr(self._fake, name)
        if name in _NOT_RECORDED or not name.startswith("gtn_"):
            return f
        fake = self._fake

        def call(*args):
            if getattr(fake, "_rec", None) is None:
                return f(*args)
            if name in ("gtn_memcpy",) and args[3] in (1, 2):
                return fake._fail("gtn_memcpy: a synchronous host copy inside a graph capture")
            if name in ("gtn_synchronize", "gtn_stream_synchronize", "gtn_event_elapsed_ms"):
                return fake._fail("%s inside a graph capture" % name)
            fake._rec.append((name, tuple(_copy_arg(a) for a in args)))
            return 0
        return call


@contextlib.contextmanager
def emulate(real_lib_path=None):
    """Run grassmanntn_amd on the CPU (tests only), either against the numpy re-statement of the
    kernels or, with real_lib_path, against a CPU build of the library's own HIP sources (whose
    "device" memory is host memory)."""
    from grassmanntn_amd import _lib as L
    reg = _Registry()
    if real_lib_path:
        fake = L._bind(C.CDLL(real_lib_path))
    else:
        fake = _Recorder(FakeLib(reg))
    saved = (L.lib, L.load_library, L._lib, L._stream)
    try:
        L.lib = lambda: fake
        L.load_library = lambda: fake
        L._lib = fake
        L._stream = None
        yield fake
    finally:
        L.lib, L.load_library, L._lib, L._stream = saved
