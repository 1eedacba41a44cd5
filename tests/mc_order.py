"""Puts marching-cubes output in the GPU kernel's order.

The GPU emits one brick of 8 × 8 × 16 cubes per block: bricks in x, y, z
order; inside a brick, threads in x, y, z order, each thread a column of 4
cubes along z in z order; each cube's triangles in table order.  The
oracle emits in plain cube order (x fastest).  `to_brick_order` returns
the permutation that takes the oracle's triangles to the GPU's.
"""
import numpy as np

import oracle as O

BX, BY, BZ, ZPER = 8, 8, 16, 4


def _corner(i):
    dx = int(bool(i & 1) != bool(i & 2))
    dy = int(bool(i & 2))
    dz = 0 if i & 4 else 1
    return dx, dy, dz


def to_brick_order(field, res, iso):
    rx, ry, rz = (int(v) for v in res)
    f = np.asarray(field).reshape(rz, ry, rx)  # [z, y, x]
    case = np.zeros((rz - 1, ry - 1, rx - 1), np.int64)
    for i in range(8):
        dx, dy, dz = _corner(i)
        v = f[dz:dz + rz - 1, dy:dy + ry - 1, dx:dx + rx - 1]
        case |= (v > iso).astype(np.int64) << i
    counts = (O.mc_table().astype(np.uint64)[case.ravel()] >> np.uint64(60)).astype(np.int64)
    cube = np.repeat(np.arange(case.size), counts)  # cube id (x fastest) of every triangle
    x = cube % (rx - 1)
    y = (cube // (rx - 1)) % (ry - 1)
    z = cube // ((rx - 1) * (ry - 1))
    nbx, nby = (rx - 1 + BX - 1) // BX, (ry - 1 + BY - 1) // BY
    brick = (x // BX) + nbx * ((y // BY) + nby * (z // BZ))
    thread = (x % BX) + BX * ((y % BY) + BY * ((z % BZ) // ZPER))
    inner = thread * ZPER + z % ZPER
    return np.lexsort((np.arange(len(cube)), inner, brick))
