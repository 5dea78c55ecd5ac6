#!/usr/bin/env python3
"""Generates sf_lds_layout.h: the LDS strides of the 3D FP64 apply kernel.

The kernel keeps two per-cell tensor groups in LDS:
  A (x-pass output, 2 tensors): address e*CA + t*N*ZA + z*ZA + y*YA + i
  B (y-pass output, 3 tensors): address e*CB + t*N*ZB + z*ZB + k*Q  + i
(e the cell of the workgroup, doubles).  The stages touch them with four
lane patterns (lane = thread id, consecutive tasks on consecutive lanes):
  A by (e, z, y) with i fixed: x pass stores, x^T pass loads  (2Q + 2Q)
  A by (e, z, i) with y fixed: y pass loads,  y^T pass stores (2N + 2N)
  B by (e, z, i) with k fixed: y pass stores, y^T pass loads  (3Q + 3Q)
  B by (e, k, i) with z fixed: z pass loads and stores        (3N + 3N)
Bank model of 64-bit LDS accesses on gfx950 (measured with strided
ds_read_b64 loops): 64 banks of 4 bytes, 16 lanes served per pass, a pass
costs as many cycles as the most distinct addresses any bank receives.
The shared A / B region is assumed from N = 5 on (the kernel's stored-G
form at N = 5 splits them and keeps the dense layout).
For every (N, Q) the script picks (YA, ZA, CA, ZB, CB) minimising the
cycles of those accesses, keeping the cells per workgroup and the number
of workgroups per CU that the LDS allows (from N = 5 on, at most the 12
waves per CU that the kernels' register use admits), and never above
64 KiB.

    python3 gen_lds_layout.py > sf_lds_layout.h
"""
import numpy as np

NQ_PAIRS = [(n, q) for n in range(2, 11) for q in range(n - 1, n + 4)]
LDS_CU = 160 * 1024
GRANULE = 512


def cycles(addr, group=16, banks=64):
    """Cycles of one wave-wide 64-bit access pattern (addr in doubles)."""
    tot = 0
    for w0 in range(0, len(addr), 64):
        for g0 in range(w0, min(w0 + 64, len(addr)), group):
            a = np.unique(addr[g0:min(g0 + group, w0 + 64, len(addr))])
            b = np.concatenate([(2 * a) % banks, (2 * a + 1) % banks])
            tot += np.bincount(b, minlength=banks).max()
    return tot


def old_shape(N, Q):
    NN, NQ, QQ, NNN = N * N, N * Q, Q * Q, N * N * N
    QP = Q + 1 if Q % 2 == 0 else Q
    SA1, SB1 = NN * QP, N * QQ
    LC = N >= 5
    SC = max(3 * SB1, 2 * SA1 + NNN) if LC else 2 * SA1 + 3 * SB1
    MAXT = max(NN, NQ, QQ)
    EB = min(1 if MAXT > 100 else max(1, 256 // MAXT), max(1, 65536 // (SC * 8)))
    NT = (EB * MAXT + 63) // 64 * 64
    return dict(YA=QP, ZA=N * QP, CA=2 * SA1, ZB=QQ, CB=3 * SB1, EB=EB, LC=LC, NT=NT)


def lds_bytes(N, Q, L):
    EB, NNN = L['EB'], N * N * N
    RA, RX = EB * L['CA'], EB * NNN
    RC = max(EB * L['CB'], RA + RX) if L['LC'] else RA + max(EB * L['CB'], RX)
    return RC * 8 + EB * 24 * 8  # + the vertices of the fused-geometry form


def score_A(N, Q, EB, YA, ZA, CA):
    e, z, y = np.meshgrid(np.arange(EB), np.arange(N), np.arange(N), indexing='ij')
    zy = (e * CA + z * ZA + y * YA).ravel()
    e, z, i = np.meshgrid(np.arange(EB), np.arange(N), np.arange(Q), indexing='ij')
    zi = (e * CA + z * ZA + i).ravel()
    return 4 * Q * cycles(zy) + 4 * N * cycles(zi)


def score_B(N, Q, EB, ZB, CB):
    e, z, i = np.meshgrid(np.arange(EB), np.arange(N), np.arange(Q), indexing='ij')
    zi = (e * CB + z * ZB + i).ravel()
    e, k, i = np.meshgrid(np.arange(EB), np.arange(Q), np.arange(Q), indexing='ij')
    ki = (e * CB + k * Q + i).ravel()
    return 6 * Q * cycles(zi) + 6 * N * cycles(ki)


def rup(a, b):
    return (a + b - 1) // b * b


def pareto(cands):
    """The candidates no other one beats in both cycles and size (entry 1)."""
    out = []
    for c in sorted(cands, key=lambda c: (c[1], c[0])):
        if not out or c[0] < out[-1][0]:
            out.append(c)
    return out


def best(N, Q):
    base = old_shape(N, Q)
    EB = base['EB']
    b0 = lds_bytes(N, Q, base)
    wg0 = min(LDS_CU // rup(b0, GRANULE), 8)
    if N >= 5:
        # the FP64 kernels use more than 128 VGPRs from N = 5 on: at most 3
        # waves per SIMD (12 per CU) run, whatever the LDS allows
        wg0 = min(wg0, max(1, 12 // (base['NT'] // 64)))
    # A and B are scored independently; the joint LDS limit is checked after
    As = []
    for YA in range(Q, Q + 4):
        for ZA in range(N * YA, N * YA + 32):
            for pad in range(0, 32):
                CA = 2 * N * ZA + pad
                As.append((score_A(N, Q, EB, YA, ZA, CA), CA, YA, ZA))
    As = pareto(As)
    Bs = []
    for ZB in range(Q * Q, Q * Q + 32):
        for pad in range(0, 32):
            CB = 3 * N * ZB + pad
            Bs.append((score_B(N, Q, EB, ZB, CB), CB, ZB))
    Bs = pareto(Bs)
    s0 = score_A(N, Q, EB, base['YA'], base['ZA'], base['CA']) + score_B(N, Q, EB, base['ZB'], base['CB'])
    top = None
    for sa, CA, YA, ZA in As:
        for sb, CB, ZB in Bs:
            L = dict(YA=YA, ZA=ZA, CA=CA, ZB=ZB, CB=CB, EB=EB, LC=base['LC'])
            nb = lds_bytes(N, Q, L)
            if nb > 65536 or min(LDS_CU // rup(nb, GRANULE), 8) < wg0:
                continue
            key = (sa + sb, nb)
            if top is None or key < top[0]:
                top = (key, L)
    if top is None or top[0][0] >= s0:
        return base, s0, s0, b0, b0
    return top[1], s0, top[0][0], b0, top[0][1]


def main():
    rows = []
    for N, Q in NQ_PAIRS:
        L, s0, s1, b0, b1 = best(N, Q)
        rows.append((N, Q, L, s0, s1, b0, b1))
    print('// Generated by gen_lds_layout.py: LDS strides (doubles) of the 3D FP64')
    print('// apply, per (N, Q).  Columns: YA, ZA, CA (x-pass tensors: y, z and cell')
    print('// strides), ZB, CB (y-pass tensors: z and cell strides).  Modelled LDS')
    print('// cycles of the structured accesses per workgroup, before -> after, and')
    print('// the LDS bytes, before -> after, in the comment of each row.')
    print('#pragma once')
    print('#define SF_LDS_LAYOUT_F64(X) \\')
    for N, Q, L, s0, s1, b0, b1 in rows:
        print('  X(%d, %d, %d, %d, %d, %d, %d) /* cycles %d -> %d, bytes %d -> %d */ \\'
              % (N, Q, L['YA'], L['ZA'], L['CA'], L['ZB'], L['CB'], s0, s1, b0, b1))
    print('')


if __name__ == '__main__':
    main()
