"""Coarse-graining drivers (gauge2d.py:1591-2145 of the reference): zcap, logZ, trg (Levin-Nave),
atrg2dy / atrg2dx, and the 3-D hotrg3dz / logZhotrg3dz.  These compose the device einsum / svd /
eig / hconjugate / sqrt."""
import numpy as np

from . import core as gtn
from .core import fail, FERMI


class _TrgCounters:
    """Process-wide counts of TRG steps run (`steps`) and of steps recomputed because an assumed
    rank count did not hold (`recomputed`; `misses` lists (step number, assumed counts, read-back
    counts) of the last few)."""
    steps = 0
    recomputed = 0
    misses = []


trg_counters = _TrgCounters()


def zcap(T):
    """Trace the two bosonic flavour legs with an all-ones capper (gauge2d.py:1591-1597,
    gauge2d_block.py:1590-1599)."""
    from .block import block
    n = T.shape[4]
    capper = gtn.dense(np.full((n, n), 1.0), statistics=(0, 0))
    if isinstance(T, block):
        capper = capper.toblock()
    return gtn.einsum("IJKLij,ij->IJKL", T, capper)


def logZ(T, boundary_conditions="periodic"):
    """ln of the supertrace einsum('IJIJ') (gauge2d.py:1599-1615).  The anti-periodic sign uses the
    popcount of the index in T's own encoder, as the reference does.  A negative real Z gives the
    complex logarithm (the reference's np.log returns nan there).  Block tensors get no anti-periodic
    signs, as in the block driver (gauge2d_block.py:1601-1615)."""
    from .block import block
    if isinstance(T, gtn.sparse):
        # the reference densifies a sparse tensor and takes it to the canonical encoder first
        T = gtn.dense(T)
        if T.encoder == "parity-preserving":
            T = T.switch_encoder()
    if boundary_conditions == "anti-periodic" and not isinstance(T, block):
        T = T.copy()
        # sign (-1)^popcount of the stored index along leg 1, as a signed gather
        from ._perm import Perm, strides
        P = Perm()
        st = strides(T.shape)
        ins = [P.add_in(T.shape[a], st[a]) for a in range(T.ndim)]
        for a in range(T.ndim):
            v = P.add_virt(T.shape[a])
            # the leg's stored index is its own-encoder index: popcount parity of the stored index
            P.add_comp(v, T.shape[a], [(ins[a], 1)], bit=0 if a == 1 else -1)
            P.add_out(T.shape[a], [(v, 1)])
        P.set_table([0, 1])
        T.data = P.run(T.data)
        T._even = None
    Z = gtn.einsum("IJIJ", T)
    Z = complex(Z) if isinstance(Z, complex) else float(Z)
    if isinstance(Z, float) and Z < 0:
        return np.log(complex(Z))
    return np.log(Z)


def trg(T, dcut=64, iternum=None, error_test=False, _speculate=True):
    """One Levin-Nave TRG step (gauge2d.py:1647-1755): returns (T', ||T'||[, err]).

    The step runs without waiting for the SVD's rank counts: it assumes every parity block keeps its
    int(dcut/2) values, so that the contractions queue behind the decomposition, and checks the
    counts (and the finiteness and parity checks of the SVD) when it reads the norm back; if a block
    kept fewer values the step is recomputed with the read-back counts."""
    if [T.shape[0], T.shape[1]] != [T.shape[2], T.shape[3]]:
        fail("Error[trg]: The shape must be of the form (m,n,m,n)!")
    if list(T.statistics) != [1, 1, -1, -1]:
        fail("Error[trg]: The statistics must be (1,1,-1,-1)!")
    if isinstance(T, gtn.dense):
        # the two permutations are folded into the decompositions' join gathers
        T1 = gtn.permuted_view(T, "ijkl->jkli")
        T2 = gtn.permuted_view(T, "ijkl->klij")
    else:
        T1 = gtn.einsum("ijkl->jkli", T)
        T2 = gtn.einsum("ijkl->klij", T)
    pending = None
    if isinstance(T1, gtn.dense):
        # both decompositions' parity blocks in one batched device call
        calls = [(T1, "ab|cd", dcut), (T2, "ab|cd", dcut)]
        if _speculate:
            ((U1, S1, V1), (U2, S2, V2)), pending = gtn.svd_many(calls, speculate=True)
        else:
            (U1, S1, V1), (U2, S2, V2) = gtn.svd_many(calls)
    else:
        U1, S1, V1 = T1.svd("ab|cd", dcut)
        U2, S2, V2 = T2.svd("ab|cd", dcut)
    r = gtn.sqrt(S1)
    U1 = gtn.einsum("abx,xc->abc", U1, r)
    V1 = gtn.einsum("ax,xbc->abc", r, V1)
    r = gtn.sqrt(S2)
    U2 = gtn.einsum("abx,xc->abc", U2, r)
    V2 = gtn.einsum("ax,xbc->abc", r, V2)
    VV = gtn.einsum("kwz,lxw->lxzk", V1, V2)
    UU = gtn.einsum("yxi,zyj->jzxi", U1, U2)
    T2n = gtn.einsum("lxzk,jzxi->ijkl", VV, UU)
    if pending is not None and not pending.verify():
        trg_counters.recomputed += 1
        trg_counters.misses = (trg_counters.misses + [(trg_counters.steps + 1, pending.assumed, pending.counts)])[-8:]
        return trg(T, dcut, iternum, error_test, _speculate=False)
    trg_counters.steps += 1
    err = None
    if error_test:
        Z1 = gtn.einsum("ijkl,klij", T, T)
        Z2 = gtn.einsum("ijij", T2n)
        err = abs(1 - Z2 / Z1)
    nrm = T2n.norm
    T2n = T2n / nrm
    return (T2n, nrm, err) if error_test else (T2n, nrm)


def atrg2dy(T1, T2, dcut=64, intermediate_dcut=None, iternum=None, error_test=False, alignment="y"):
    """Anisotropic TRG along y (gauge2d.py:1761-1869)."""
    T1o, T2o = T1, T2
    if intermediate_dcut is None:
        intermediate_dcut = dcut
    T1 = gtn.einsum("ijkl->li jk", T1)
    T2 = gtn.einsum("ijkl->li jk", T2)
    if isinstance(T1, gtn.dense):
        (U1, S1, V1), (U2, S2, V2) = gtn.svd_many([(T1, "li|jk", intermediate_dcut),
                                                    (T2, "li|jk", intermediate_dcut)])
    else:
        U1, S1, V1 = T1.svd("li|jk", intermediate_dcut)
        U2, S2, V2 = T2.svd("li|jk", intermediate_dcut)
    A = V1.copy()
    B = gtn.einsum("lia,ab->lib", U1, S1)
    Cc = gtn.einsum("ab,bjk->ajk", S2, V2)
    D = U2.copy()
    M = gtn.einsum("ajk,jib->aibk", Cc, B)
    U, S, V = M.svd("ai|bk", intermediate_dcut)
    r = gtn.sqrt(S)
    Y = gtn.einsum("abx,xc->abc", U, r)
    X = gtn.einsum("ax,xbc->abc", r, V)
    Q1 = gtn.einsum("iax,xbj->ijab", D, Y)
    Q2 = gtn.einsum("kya,ylb->abkl", X, A)
    Q = gtn.einsum("ijab,abkl->ijkl", Q1, Q2)
    U, S, V = Q.svd("ij|kl", dcut)
    r = gtn.sqrt(S)
    H = gtn.einsum("abx,xc->abc", U, r)
    G = gtn.einsum("ax,xbc->abc", r, V)
    H = gtn.einsum("lai->ila", H)
    G = gtn.einsum("kaj->ajk", G)
    T = gtn.einsum("ila,ajk->ijkl", H, G)
    err = None
    if error_test:
        Z1 = gtn.einsum("IJIK,iKiJ", T1o, T2o)
        Z2 = gtn.einsum("IJIJ", T)
        err = abs(1 - Z2 / Z1)
    nrm = T.norm
    T = T / nrm
    return (T, nrm, err) if error_test else (T, nrm)


def atrg2dx(T1, T2, dcut=64, intermediate_dcut=None, iternum=None, error_test=False):
    """atrg2dy with the x and y legs exchanged (gauge2d.py:1871-1885)."""
    def swap(T):
        return gtn.einsum("jikl->jilk", gtn.einsum("ijkl->jikl", T))
    res = atrg2dy(swap(T1), swap(T2), dcut, intermediate_dcut, iternum, error_test, alignment="x")
    return (swap(res[0]),) + tuple(res[1:])


def hotrg3dz(T1, T2, dcut=64, intermediate_dcut=None, iternum=None, error_test=False):
    """One 3-D HOTRG step along z (gauge2d.py:1891-2145) on site tensors (i1 i2 i3 i4 m n) with
    bosonic z legs m, n: two SVDs split each tensor into X (x legs), P, Q and Y (y legs); the
    isometries Ux, Uy are the eig of the squared four-leg blocks (of the two candidates, the one
    with fewer kept values, as the reference's S.shape[0] comparison picks); returns
    (T', ||T'||[, err]) with T' = (t1 t2 t3 t4 m n) / ||T'||.  The contractions and transpositions
    are the reference's sequence, so that every sign matches; the reference's svd_only / print_svd
    diagnostics are not provided."""
    T1o, T2o = T1, T2
    if intermediate_dcut is None:
        intermediate_dcut = dcut
    e = gtn.einsum

    def split(T):
        T = e("i1 i2 i3 i4 mn-> i1 i3 mn i2 i4", T)
        X, S, Y = T.svd("(i1 i3 m)(n i2 i4)", intermediate_dcut)
        r = gtn.sqrt(S)
        X = e("i1 i3 m a,ab->i1 i3 b m", X, r)
        Y = e("ab,b n i2 i4->n a i2 i4", r, Y)
        X, S, P = X.svd("(i1 i3)(a m)", intermediate_dcut)
        X = e("i1 i3 x, xy -> i1 i3 y", X, S)
        Q, S, Y = Y.svd("(na)(i2 i4)", intermediate_dcut)
        Y = e("xy, y i2 i4 -> x i2 i4", S, Y)
        return X, P, Q, Y

    X1, P1, Q1, Y1 = split(T1)
    X2, P2, Q2, Y2 = split(T2)
    XX = e("i1 i3 a, j1 j3 b -> i1 i3 ab j1 j3", X1, X2)
    XX = e("i1 i3 ab j1 j3 -> i1 i3 ab j3 j1 ", XX)
    XX = e("i1 i3 ab j3 j1 -> i1 i3 j3 ab j1 ", XX)
    XX = e("i1 i3 j3 ab j1 -> i3 j3 i1 ab j1 ", XX)
    XX = e("i3 j3 i1 ab j1 -> i3 j3 ab i1 j1 ", XX)
    c1 = XX.hconjugate("(i3 j3 ab)(i1 j1)")
    M1 = e(" I1 J1 i3 j3 ab, i3 j3 ab i1 j1 -> I1 J1 i1 j1", c1, XX)
    U1, S1, _ = M1.eig("(I J)(i j)", dcut)
    c3 = XX.hconjugate("(i3 j3)(ab i1 j1)")
    M3 = e(" I3 J3 ab i1 j1, ab i1 j1 i3 j3  -> I3 J3 i3 j3", XX, c3)
    U3, S3, _ = M3.eig("(I J)(i j)", dcut)
    Ux = U1 if S1.shape[0] < S3.shape[0] else U3
    cUx = Ux.hconjugate("ij|a")
    XX = e("i3 j3 ab i1 j1 -> j3 i3 ab i1 j1", XX)
    XX = e("j3 i3 ab i1 j1 -> j3 i3 ab j1 i1", XX)
    Xp = e("s i3 j3,j3 i3 kl j1 i1 -> s kl j1 i1", cUx, XX)
    Xp = e("s kl j1 i1, i1 j1 t -> s kl t", Xp, Ux)
    Xp = e("s kl t -> t s kl", Xp)
    Xp = e("t s kl , kam -> t s al m", Xp, P1)
    Xp = e("t s al m , lbm -> t s ab m", Xp, P2)
    YY = e("b j2 j4, a i2 i4 -> j2 j4 ba i2 i4 ", Y2, Y1)
    YY = e("j2 j4 ba i2 i4 -> j2 j4 i4 ba i2 ", YY)
    YY = e("j2 j4 i4 ba i2 -> j4 i4 j2 ba i2 ", YY)
    YY = e("j4 i4 j2 ba i2 -> j4 i4 ba i2 j2 ", YY)
    YY = e("j4 i4 ba i2 j2 -> i4 j4 ba i2 j2 ", YY)
    c2 = YY.hconjugate("(i4 j4 ba)(i2 j2)")
    M2 = e(" I2 J2 i4 j4 ba, i4 j4 ba i2 j2 -> I2 J2 i2 j2", c2, YY)
    U2, S2, _ = M2.eig("(I J)(i j)", dcut)
    c4 = YY.hconjugate("(i4 j4)(ba i2 j2)")
    M4 = e(" I4 J4 ba i2 j2, ba i2 j2 i4 j4  -> I4 J4 i4 j4", YY, c4)
    U4, S4, _ = M4.eig("(I J)(i j)", dcut)
    Uy = U2 if S2.shape[0] < S4.shape[0] else U4
    cUy = Uy.hconjugate("ij|a")
    YY = e("i4 j4 b a i2 j2 -> j4 i4 b a i2 j2", YY)
    YY = e("j4 i4 b a i2 j2 -> j4 i4 b a j2 i2", YY)
    Yp = e("s i4 j4,j4 i4 lk j2 i2 -> s lk j2 i2", cUy, YY)
    Yp = e("s lk j2 i2, i2 j2 t -> s lk t", Yp, Uy)
    Yp = e("s lk t -> lk t s", Yp)
    Yp = e("nak, lk t s -> n la t s", Q1, Yp)
    Yp = e("nbl, n la t s -> n ba t s", Q2, Yp)
    T = e(" t1 t3 kl m, n lk t2 t4 -> t1 t2 t3 t4 mn ", Xp, Yp)
    err = None
    if error_test:
        Z1 = e("IJIJmn,KLKLmn->mn", T1o, T2o)
        Z2 = e("IJIJmn->mn", T)
        err = (Z1 - Z2).norm / Z1.norm
    nrm = T.norm
    T = T / nrm
    return (T, nrm, err) if error_test else (T, nrm)


def _antiperiodic_J(T):
    """T with the sign (-1)^popcount(stored index) along leg 1, as a signed gather."""
    from ._perm import Perm, strides
    T = T.copy()
    P = Perm()
    st = strides(T.shape)
    ins = [P.add_in(T.shape[a], st[a]) for a in range(T.ndim)]
    for a in range(T.ndim):
        v = P.add_virt(T.shape[a])
        P.add_comp(v, T.shape[a], [(ins[a], 1)], bit=0 if a == 1 else -1)
        P.add_out(T.shape[a], [(v, 1)])
    P.set_table([0, 1])
    T.data = P.run(T.data)
    T._even = None
    return T


def logZhotrg3dz(T1, T2, boundary_conditions="periodic"):
    """ln of einsum('IJIJmn,KLKLmn') (gauge2d.py:1617-1641); anti-periodic: (-1)^popcount of the
    stored index along leg J of both tensors.  A negative real Z gives the complex logarithm."""
    def densify(T):
        if isinstance(T, gtn.sparse):
            T = gtn.dense(T)
            if T.encoder == "parity-preserving":
                T = T.switch_encoder()
        return T

    T1, T2 = densify(T1), densify(T2)
    if boundary_conditions == "anti-periodic":
        T1, T2 = _antiperiodic_J(T1), _antiperiodic_J(T2)
    Z = gtn.einsum("IJIJmn,KLKLmn", T1, T2)
    Z = complex(Z) if isinstance(Z, complex) else float(Z)
    if isinstance(Z, float) and Z < 0:
        return np.log(complex(Z))
    return np.log(Z)
