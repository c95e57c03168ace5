"""Coarse-graining drivers of the 2-D model (gauge2d.py:1591-1885 of the reference): zcap, logZ,
trg (Levin-Nave), atrg2dy / atrg2dx.  These compose the device einsum / svd / sqrt; the
graph-captured fast path of the TRG step lives in grassmanntn_amd.trg."""
import numpy as np

from . import core as gtn
from .core import fail, FERMI


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


def trg(T, dcut=64, iternum=None, error_test=False):
    """One Levin-Nave TRG step (gauge2d.py:1647-1755): returns (T', ||T'||[, err])."""
    if [T.shape[0], T.shape[1]] != [T.shape[2], T.shape[3]]:
        fail("Error[trg]: The shape must be of the form (m,n,m,n)!")
    if list(T.statistics) != [1, 1, -1, -1]:
        fail("Error[trg]: The statistics must be (1,1,-1,-1)!")
    T1 = gtn.einsum("ijkl->jkli", T)
    T2 = gtn.einsum("ijkl->klij", T)
    if isinstance(T1, gtn.dense):
        # both decompositions' parity blocks in one batched device call
        (U1, S1, V1), (U2, S2, V2) = gtn.svd_many([(T1, "ab|cd", dcut), (T2, "ab|cd", dcut)])
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
