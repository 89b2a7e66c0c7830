"""Generates lsm_element.h: exact element matrices of the lattice spring cube.

The reference sums bond and angle springs over per-bond / per-angle tables
(LSM2S/Cxxxx/XX_P.f:626-1262, force loops XX:1404-1500).  Every bond and angle
weight there is a sum over the solid voxel cubes that share it (XX:640-1018,
1023-1262), so the stored energy is a sum over solid cubes of one cube energy

    E_cube = sum_b  w_b alpha/2 (xhat_b . (u_p2 - u_p1))^2
           + sum_a  w_a beta/2  (f_j . (u_j - u_i) + f_k . (u_k - u_i))^2

over the 24 cube bonds and 72 cube angles of the cube_bon/cube_ang tables
with the cube_sbond/cube_abond weights (XX:1700-1714).  f_j, f_k are the
angular force directions anf(4:6), anf(7:9) (XX:200-306): dfi of XX:1450 is
-beta*w*(f_j.du_j + f_k.du_k) once the deformed arms r_t = c + du are
expanded (x_c cross c = 0).  Hence  F = -(alpha KA + beta KB) u  per cube.

The cube stress of calc_stress (XX:1780-1930) is linear in u as well:
sigma(d,c) = (alpha SA + beta SB) u_cube with the reference's sign rules.

All entries are rationals (|n|^2 in {1,3}); they are computed exactly with
fractions and written as correctly rounded doubles.  The header also carries
a generated node-centred force routine that gathers the 8 cubes around a
node and multiplies only the non-zero coefficients.

Run:  python3 gen_lsm_element.py > lsm_element.h
"""
from __future__ import annotations

import os
import sys
from fractions import Fraction as Fr

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from porescale_amd.geometry import cube_tables  # noqa: E402
from porescale_amd.io import BOND9, D3Q19  # noqa: E402

CORNER = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def vel(q):
    return tuple(int(D3Q19[d, q - 1]) for d in range(3))


def angle_dirs(t, qj, qk):
    """(gj, gk, n2, cj2, ck2): f_j = gj/(sqrt(n2) cj2), f_k = gk/(sqrt(n2) ck2)."""
    cj, ck = vel(qj), vel(qk)
    n = cross(cj, ck)
    if 9 <= t <= 16:  # reference orientation of the xz-plane ring
        n = tuple(-x for x in n)
    gj = cross(n, cj)
    gk = tuple(-x for x in cross(n, ck))
    return gj, gk, dot(n, n), dot(cj, cj), dot(ck, ck)


def build():
    bon, ang, sb, ab = cube_tables()
    KA = [[Fr(0)] * 24 for _ in range(24)]
    KB = [[Fr(0)] * 24 for _ in range(24)]
    SA = [[Fr(0)] * 24 for _ in range(9)]   # row d + 3c  (force comp d, bond dir c)
    SB = [[Fr(0)] * 24 for _ in range(9)]
    for (p1, p2, b), w in zip(bon, sb):
        c = tuple(int(BOND9[d, b - 1]) for d in range(3))
        c2 = dot(c, c)
        w = Fr(w).limit_denominator(8)
        # v = (-c at p1, +c at p2)/|c|; K += w v v^T  (v v^T = c c^T / c2)
        v = [Fr(0)] * 24
        for d in range(3):
            v[3 * (p1 - 1) + d] -= c[d]
            v[3 * (p2 - 1) + d] += c[d]
        for r in range(24):
            if v[r] == 0:
                continue
            for s in range(24):
                if v[s]:
                    KA[r][s] += w * v[r] * v[s] / c2
        # flin = w (xhat.(u2-u1)) xhat ; sigma(:,cc) += sgn(xhat_cc) flin
        for cc in range(3):
            sg = (c[cc] > 0) - (c[cc] < 0)
            if sg == 0:
                continue
            for d in range(3):
                for s in range(24):
                    if v[s]:
                        SA[d + 3 * cc][s] += sg * w * c[d] * v[s] / c2
    for (i, j, k, qj, qk, t), w in zip(ang, ab):
        w = Fr(w).limit_denominator(8)
        gj, gk, n2, cj2, ck2 = angle_dirs(t, qj, qk)
        # theta*sqrt(n2) = fj'.(uj-ui) + fk'.(uk-ui), fj' = gj/cj2, fk' = gk/ck2
        v = [Fr(0)] * 24
        for d in range(3):
            fj = Fr(gj[d], cj2)
            fk = Fr(gk[d], ck2)
            v[3 * (j - 1) + d] += fj
            v[3 * (k - 1) + d] += fk
            v[3 * (i - 1) + d] -= fj + fk
        for r in range(24):
            if v[r] == 0:
                continue
            for s in range(24):
                if v[s]:
                    KB[r][s] += w * v[r] * v[s] / n2
        # dfi = -w beta theta ; sigma(:,cc) -= sgn(cj_cc) dfi fj + sgn(ck_cc) dfi fk
        cj, ck = vel(qj), vel(qk)
        for cc in range(3):
            sj = (cj[cc] > 0) - (cj[cc] < 0)
            sk = (ck[cc] > 0) - (ck[cc] < 0)
            for d in range(3):
                coef = sj * Fr(gj[d], cj2) + sk * Fr(gk[d], ck2)
                if coef == 0:
                    continue
                for s in range(24):
                    if v[s]:
                        SB[d + 3 * cc][s] += w * coef * v[s] / n2
    return KA, KB, SA, SB


def dbl(x: Fr) -> str:
    return repr(float(x)) if x != 0 else "0.0"


def emit(KA, KB, SA, SB, out):
    w = out.write
    w("// Generated by gen_lsm_element.py -- exact element matrices of the lattice\n")
    w("// spring cube (see the generator's docstring).  Do not edit.\n")
    w("#pragma once\n#include <hip/hip_runtime.h>\n\nnamespace lsm_el {\n\n")
    for name, M, rows in (("KA", KA, 24), ("KB", KB, 24), ("SA", SA, 9), ("SB", SB, 9)):
        w(f"constexpr double {name}[{rows}][24] = {{\n")
        for r in range(rows):
            w("    {" + ", ".join(dbl(M[r][s]) for s in range(24)) + "},\n")
        w("};\n\n")
    # node-centred force: node n is corner p of the cube at voxel n - CORNER[p];
    # neighbour of corner q of that cube is at offset CORNER[q] - CORNER[p].
    nnz = 0
    w("// f[e] = -sum_p ka[p] (KA u_cube(p))[3p+e] + kb[p] (KB u_cube(p))[3p+e]\n")
    w("// u[s][e]: displacement of stencil node s = (dx+1) + 3(dy+1) + 9(dz+1).\n")
    w("// ka[p], kb[p]: spring scales of the voxel at node - CORNER[p].\n")
    w("template <typename T>\n__device__ __forceinline__ void node_force(const T (&u)[27][3], const T (&ka)[8],\n"
      "                                           const T (&kb)[8], T (&f)[3]) {\n")
    w("    T a0, a1, a2, b0, b1, b2;\n")
    for p in range(8):
        acc = {}
        for e in range(3):
            for mat, key in ((KA, "a"), (KB, "b")):
                terms = []
                for q in range(8):
                    off = tuple(CORNER[q][d] - CORNER[p][d] for d in range(3))
                    s = (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1)
                    for e2 in range(3):
                        v = mat[3 * p + e][3 * q + e2]
                        if v != 0:
                            terms.append((v, s, e2))
                nnz += len(terms)
                acc[(key, e)] = terms
        for key in ("a", "b"):
            for e in range(3):
                terms = acc[(key, e)]
                if not terms:
                    w(f"    {key}{e} = T(0);\n")
                    continue
                expr = f"T({dbl(terms[0][0])}) * u[{terms[0][1]}][{terms[0][2]}]"
                w(f"    {key}{e} = {expr};\n")
                for v, s, e2 in terms[1:]:
                    w(f"    {key}{e} += T({dbl(v)}) * u[{s}][{e2}];\n")
        for e in range(3):
            w(f"    f[{e}] -= ka[{p}] * a{e} + kb[{p}] * b{e};\n")
    w("}\n")
    w(f"// non-zero coefficients per node: {nnz}\n")
    w(f"constexpr int NODE_FORCE_NNZ = {nnz};\n\n")
    # diagonal 3x3 block of a node (for the block-Jacobi preconditioner)
    w("template <typename T>\n__device__ __forceinline__ void node_diag(const T (&ka)[8], const T (&kb)[8], T (&D)[3][3]) {\n")
    for e in range(3):
        for e2 in range(3):
            parts = []
            for p in range(8):
                va, vb = KA[3 * p + e][3 * p + e2], KB[3 * p + e][3 * p + e2]
                if va != 0:
                    parts.append(f"T({dbl(va)}) * ka[{p}]")
                if vb != 0:
                    parts.append(f"T({dbl(vb)}) * kb[{p}]")
            w(f"    D[{e}][{e2}] = " + (" + ".join(parts) if parts else "T(0)") + ";\n")
    w("}\n\n")
    # full 27-slot stencil of a node (assembled operator rows, for the
    # multigrid hierarchy): out[(s*9 + 3e + e2) * stride] = K(node, node+off(s))[e][e2]
    w("// K(node, node + offset(s))[e][e2] for the 27 offsets s = (dx+1) + 3(dy+1) + 9(dz+1),\n")
    w("// written to out[(9 s + 3 e + e2) * stride].\n")
    w("template <typename T>\n__device__ __forceinline__ void node_stencil(const T (&ka)[8], const T (&kb)[8], T *out,\n"
      "                                             int64_t stride) {\n")
    for sidx in range(27):
        off = (sidx % 3 - 1, (sidx // 3) % 3 - 1, sidx // 9 - 1)
        for e in range(3):
            for e2 in range(3):
                parts = []
                for pp in range(8):
                    q = [qq for qq in range(8) if all(CORNER[qq][d] - CORNER[pp][d] == off[d] for d in range(3))]
                    if not q:
                        continue
                    q = q[0]
                    va, vb = KA[3 * pp + e][3 * q + e2], KB[3 * pp + e][3 * q + e2]
                    if va != 0:
                        parts.append(f"T({dbl(va)}) * ka[{pp}]")
                    if vb != 0:
                        parts.append(f"T({dbl(vb)}) * kb[{pp}]")
                expr = " + ".join(parts) if parts else "T(0)"
                w(f"    out[{9 * sidx + 3 * e + e2} * stride] = {expr};\n")
    w("}\n\n} // namespace lsm_el\n")


if __name__ == "__main__":
    KA, KB, SA, SB = build()
    emit(KA, KB, SA, SB, sys.stdout)
