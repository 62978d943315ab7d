"""Pins orc_kernels_build against the reference's own operator chain run with
numpy's FFT: casagridsf lm taper -> WProjPreFFTConvFunc screen written at
OriginToCorner positions -> forward 2-D FFT -> ZeroToCentre2DShift ->
SingleMultiply(1/divisor) -> ConvFunctionSupportFind -> ConvFunctionReOrganize
-> SumConvolutionFunctions.  The oracle instead evaluates a separable DFT
restricted to the kept window and finds supports from 1-D projections; this
test shows the two agree.  Index helpers below are the reference's integer
formulas (op/common.hcu:106-136), not the oracle's modular shortcut."""
import numpy as np
import pytest


def origin_to_corner(loc, dim):  # op/common.hcu:121-136
    if dim % 2:
        return loc + dim // 2 + 1 if loc < dim // 2 else loc - dim // 2
    return loc + dim // 2 if loc < dim // 2 else loc - dim // 2


def origin_to_centre(loc, dim):  # op/common.hcu:106-120
    if dim % 2:
        return loc + dim // 2 if loc <= dim // 2 else loc - dim // 2 - 1
    return loc + dim // 2 if loc < dim // 2 else loc - dim // 2


def reference_pipeline(orc, nx, cell, s, wplanes):
    St = nx // s
    if St % 2 == 0:
        St += 1
    L = St * s
    h = St // 2
    psi = np.array([orc.gridsf(abs(i - h) / h) for i in range(St)])
    taper = psi[:, None] * psi[None, :]               # casagridsf, n = St
    wmax = 0.25 / cell
    dw = wmax / (wplanes - 1) ** 2
    inc = cell * nx / St

    def screen_fft(w):
        buf = np.zeros((L, L), complex)
        for y in range(L):
            for x in range(L):
                cy, cx = y - L // 2, x - L // 2
                if abs(cy) > h or abs(cx) > h:
                    val = 0
                else:
                    rsq = (inc * cx) ** 2 + (inc * cy) ** 2
                    t = taper[cy + h, cx + h]
                    val = t * np.exp(1j * np.pi * 2 * w * (np.sqrt(1 - rsq) - 1)) if rsq < 1 else t
                buf[origin_to_corner(y, L), origin_to_corner(x, L)] = val
        F = np.fft.fft2(buf)
        out = np.zeros_like(F)
        for y in range(L):
            for x in range(L):
                out[origin_to_centre(y, L), origin_to_centre(x, L)] = F[y, x]
        return F, out

    F0, _ = screen_fft(0.0)
    divisor = F0[0, 0]
    assert abs(divisor.imag) < 1e-3
    T = s if s % 2 else s + 1
    planes, sups = [], []
    for k in range(wplanes):
        _, Y = screen_fft(k * k * dw)
        Y = Y * (1.0 / divisor.real)
        c = L // 2
        found = -1
        for loc in range(L - 1, c - 1, -1):
            if abs(Y[c, loc]) ** 2 > 1e-6 or abs(Y[loc, c]) ** 2 > 1e-6:
                found = loc
                break
        S = ((found - c + (s + 1) // 2) // s) * 2 + 1 if found >= 0 else 0
        S = min(S, St - 2)
        border = L // 2 - ((S // 2) * s + s // 2)
        tab = np.zeros((T, T, S, S), complex)
        n = S * T
        for ly in range(n):
            for lx in range(n):
                spy, spx = ly // T, lx // T
                sy, sx = ly % T, lx % T
                iy = (ly - spy if s % 2 == 0 else ly) + border
                ix = (lx - spx if s % 2 == 0 else lx) + border
                tab[sy, sx, spy, spx] = Y[iy, ix]
        planes.append(tab)
        sups.append(S)
    return divisor.real, dw, np.array(sups), planes


@pytest.mark.parametrize("s,nx", [(4, 64), (5, 64), (8, 128)])
def test_wproj_tables_match_fft_pipeline(orc, s, nx):
    cell, W = 0.005 * 64 / nx, 6
    div, dw, sups, planes = reference_pipeline(orc, nx, cell, s, W)
    p = orc.make_params(nx, nx, cell, cell, "wproj", W, s)
    K = orc.Kernels(p)
    assert K.w_increment == pytest.approx(dw, rel=1e-15)
    assert K.divisor == pytest.approx(div, rel=1e-12)
    np.testing.assert_array_equal(K.support, sups)
    assert sups[-1] > sups[0]  # the w-term must widen the kernel here
    for k in range(W):
        ref = planes[k]
        got = K.plane(k)
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
        np.testing.assert_allclose(K.sums[k], ref.real.sum(axis=(2, 3)).ravel(), rtol=1e-11, atol=1e-13)


def test_supports_only_entry_agrees(orc):
    p = orc.make_params(64, 64, 0.005, 0.005, "wproj", 6, 4)
    np.testing.assert_array_equal(orc.supports(p), orc.Kernels(p).support)


@pytest.mark.parametrize("S,s", [(7, 4), (7, 5), (9, 3)])
def test_taper_only_table(orc, S, s):
    """mode=normal: casagridsf (1-nu^2) psi(nu), nu = |i-c|/(3s), reorganised
    with ConvFunctionReOrganize (border 0)."""
    p = orc.make_params(64, 64, 1e-4, 1e-4, "normal", 1, s, S)
    K = orc.Kernels(p)
    T = s if s % 2 else s + 1
    dim = S * s if s % 2 else S * s + 1
    c = dim // 2

    def g(i):
        a = abs(i - c)
        if a >= 3 * s:
            return 0.0
        nu = a / (3 * s)
        return (1 - nu * nu) * orc.gridsf(nu)
    tab = K.plane(0)
    for oy in range(T):
        for ox in range(T):
            for sy in range(S):
                for sx in range(S):
                    assert tab[oy, ox, sy, sx] == pytest.approx(g(sy * s + oy) * g(sx * s + ox), abs=1e-15)
    # zero-offset sub-kernel of an odd-sampled table is centred: symmetric
    if s % 2:
        m = tab[s // 2, s // 2].real
        np.testing.assert_allclose(m, m[::-1, ::-1], atol=1e-15)
