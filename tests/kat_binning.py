"""Hand-derived known-answer records for the binning step
(op/WProjPreProcessUVW.cu:111-286), shared by the oracle and GPU tests.

Geometry: nx = ny = 64, cell = 2^-13 rad  ->  du = dv = 1/(cell*64) = 128
(exact), half size 32; wplanes = 5 -> w_max = 0.25/cell = 2048,
dw = 2048/16 = 128 (exact); 1/lambda = 1.  All supports are 7 (S/2 = 3),
T = 5 for s = 4 and s = 5, plane base = 1225 k (25 * 49).
So pos = u/128 + 32, plane = round(sqrt(|w|/128)), corner = loc - 3,
offset = round((loc - pos) s) + s/2, conv_idx = 1225 k + 245 oy + 49 ox.
"""
import numpy as np

CELL = 2.0 ** -13
NX = 64
WPLANES = 5

# (u, v, w, flags[2]) -> expected (to_grid, plane, S, cx, cy, ox, oy, conj, conv_idx)
CASES_S4 = [
    # pixel centre: pos = (37, 29), offsets s/2 = 2
    ((640.0, -384.0, 0.0, (0, 0)), (1, 0, 7, 34, 26, 2, 2, 0, 588)),
    # w < 0: (u,v) negated -> pos (27, 30); plane round(sqrt(1)) = 1
    ((640.0, 256.0, -128.0, (0, 0)), (1, 1, 7, 24, 27, 2, 2, 1, 1813)),
    # exact .5 ties round away from zero: pos (35.5, 28.5) -> loc (36, 29),
    # offsets (0.5*4)+2 = 4
    ((448.0, -448.0, 0.0, (0, 0)), (1, 0, 7, 33, 26, 4, 4, 0, 1176)),
    # pos (37.25, 29.375): ox = -1+2 = 1, oy = round(-1.5)+2 = -2+2 = 0
    ((672.0, -336.0, 0.0, (0, 0)), (1, 0, 7, 34, 26, 1, 0, 0, 49)),
    # w = 128*1.5^2: sqrt = 1.5 exactly -> plane 2 (half away from zero)
    ((0.0, 0.0, 288.0, (0, 0)), (1, 2, 7, 29, 29, 2, 2, 0, 3038)),
    # w = 128*0.5^2 -> 0.5 -> plane 1 (banker's rounding would give 0)
    ((0.0, 0.0, 32.0, (0, 0)), (1, 1, 7, 29, 29, 2, 2, 0, 1813)),
    # plane boundaries k^2 dw and the clamp to W-1 = 4
    ((0.0, 0.0, 512.0, (0, 0)), (1, 2, 7, 29, 29, 2, 2, 0, 3038)),
    ((0.0, 0.0, 800.0, (0, 0)), (1, 3, 7, 29, 29, 2, 2, 0, 4263)),
    ((0.0, 0.0, 2048.0, (0, 0)), (1, 4, 7, 29, 29, 2, 2, 0, 5488)),
    ((0.0, 0.0, 12800.0, (0, 0)), (1, 4, 7, 29, 29, 2, 2, 0, 5488)),
    ((0.0, 0.0, -12800.0, (0, 0)), (1, 4, 7, 29, 29, 2, 2, 1, 5488)),
    # edge exclusion: corner must lie in [0, nx-1-S] = [0, 56]
    ((-3840.0, 0.0, 0.0, (0, 0)), (0, 0, 0, 0, 0, 0, 0, 0, 0)),      # loc 2 -> corner -1
    ((-3712.0, 0.0, 0.0, (0, 0)), (1, 0, 7, 0, 29, 2, 2, 0, 588)),   # loc 3 -> corner 0
    ((3456.0, 0.0, 0.0, (0, 0)), (1, 0, 7, 56, 29, 2, 2, 0, 588)),   # loc 59 -> corner 56
    ((3584.0, 0.0, 0.0, (0, 0)), (0, 0, 0, 0, 0, 0, 0, 0, 0)),       # loc 60 -> corner 57
    ((0.0, 3456.0, 0.0, (0, 0)), (1, 0, 7, 29, 56, 2, 2, 0, 588)),
    ((0.0, -3840.0, 0.0, (0, 0)), (0, 0, 0, 0, 0, 0, 0, 0, 0)),
    # flags: every correlation flagged -> not gridded; one clear -> gridded
    ((640.0, -384.0, 0.0, (1, 1)), (0, 0, 0, 0, 0, 0, 0, 0, 0)),
    ((640.0, -384.0, 0.0, (1, 0)), (1, 0, 7, 34, 26, 2, 2, 0, 588)),
]

CASES_S5 = [
    # odd sampling: tie 0.5*5 = 2.5 -> 3, +2 = 5 -> clamped to s-1 = 4
    ((448.0, -448.0, 0.0, (0, 0)), (1, 0, 7, 33, 26, 4, 4, 0, 1176)),
    # pos 37.4 -> loc 37: round(-0.4*5) = -2 -> 0; pos_v 29 -> offset 2
    ((691.2, -384.0, 0.0, (0, 0)), (1, 0, 7, 34, 26, 0, 2, 0, 490)),
]


def arrays(cases):
    uvw = np.array([c[0][:3] for c in cases], dtype=np.float64)
    flags = np.array([c[0][3] for c in cases], dtype=np.uint8)[:, None, :]
    exp = np.array([c[1] for c in cases], dtype=np.int64)
    return uvw, flags, exp
