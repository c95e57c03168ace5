"""Contraction patterns shared by the oracle and GPU parity tests: the TRG / ATRG / zcap einsums of
gauge2d.py plus general shapes (mixed statistics, bosonic legs, traces, three operands, outer
products)."""

EINSUMS = [
    ("ilk,jl->ijk", [(1, 1, -1), (-1, -1)], [(4, 8, 4), (8, 8)]),
    ("ij,jk->ik", [(1, 1), (-1, -1)], [(16, 8), (8, 16)]),
    ("ijk->kji", [(1, -1, 1)], [(4, 8, 2)]),
    ("ijkl->jkli", [(1, 1, -1, -1)], [(8, 4, 8, 4)]),
    ("ijkl->klij", [(1, 1, -1, -1)], [(8, 4, 8, 4)]),
    ("ijij", [(1, 1, -1, -1)], [(4, 8, 4, 8)]),
    ("ijkl,klij", [(1, 1, -1, -1), (1, 1, -1, -1)], [(4, 4, 4, 4), (4, 4, 4, 4)]),
    ("kwz,lxw->lxzk", [(-1, -1, 1), (-1, 1, 1)], [(8, 8, 8), (8, 8, 8)]),
    ("yxi,zyj->jzxi", [(1, -1, 1), (-1, -1, 1)], [(8, 8, 8), (8, 8, 8)]),
    ("lxzk,jzxi->ijkl", [(-1, 1, 1, -1), (1, -1, -1, 1)], [(8, 8, 8, 8), (8, 8, 8, 8)]),
    ("abx,xc->abc", [(1, -1, 1), (-1, 1)], [(4, 4, 8), (8, 8)]),
    ("ajk,jib->aibk", [(1, 1, -1), (-1, -1, 1)], [(4, 4, 4), (4, 4, 4)]),
    ("ab,bc,cd->ad", [(1, 1), (-1, 1), (-1, 1)], [(4, 4), (4, 4), (4, 4)]),
    ("iIaj,jIbk->akbi", [(1, 0, -1, 1), (-1, 0, 1, 1)], [(4, 3, 8, 4), (4, 3, 2, 8)]),
    ("IJKLij,ij->IJKL", [(1, 1, -1, -1, 0, 0), (0, 0)], [(4, 4, 4, 4, 2, 2), (2, 2)]),
    # outer products (no contracted index), as in hotrg3dz's XX / YY (gauge2d.py:1962, :2033)
    ("ia,jb->iajb", [(1, 1), (1, -1)], [(4, 8), (8, 4)]),
    ("i1 i3 a, j1 j3 b -> i1 i3 ab j1 j3", [(1, -1, 1), (1, -1, 1)], [(4, 4, 8), (4, 4, 2)]),
    ("ija,k->iakj", [(1, -1, 0), (1,)], [(4, 8, 3), (2,)]),
]
