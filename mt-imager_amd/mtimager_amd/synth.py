"""Seeded synthetic interferometer observations (MeasurementSet order).

Antennas uniform in a disk (r = R sqrt(U), theta = 2 pi U), local ENU ->
equatorial XYZ at latitude `lat`, earth-rotation uvw over an hour-angle range:

    u =  sin H X + cos H Y
    v = -sin d cos H X + sin d sin H Y + cos d Z
    w =  cos d cos H X - cos d sin H Y + sin d Z

Rows are time-major (every baseline, then the autocorrelations, per time
step) as in an MS; the imaging path sorts them baseline-major.
Visibilities follow the convention the mt-imager kernels imply (the uv kernel
is the forward FFT of taper * exp(+2 pi i w (n-1)) and the image is an inverse
FFT):  V = sum_s S_s exp(-2 pi i (u l + v m + w (n - 1))), u,v,w in
wavelengths.  Linear feeds: XX = I+Q, YY = I-Q, XY = U+iV, YX = U-iV.
"""
import numpy as np

C_LIGHT = 299792458.0
ARCSEC = np.pi / 180.0 / 3600.0
MJD0_S = 60000.0 * 86400.0  # 2023-02-25 00:00 UTC in MJD seconds


class Observation:
    """Arrays in MS order: uvw f64 [nrows][3] (metres), ant1/ant2 int32
    [nrows], freq f64 [nchan] (Hz), inv_lambda f64 [nchan], vis complex64
    [nrows][nchan][npol], flags u8 [nrows][nchan][npol], weight f32 [nrows][npol]."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def nrows(self):
        return self.uvw.shape[0]


def array_layout(nant, radius_m, seed):
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.random(nant))
    th = 2 * np.pi * rng.random(nant)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(nant)], axis=1)  # E, N, U


def uvw_tracks(enu, lat_deg, dec_deg, ha_hours, ntime, with_autocorr=True):
    lat, dec = np.deg2rad(lat_deg), np.deg2rad(dec_deg)
    E, N, U = enu[:, 0], enu[:, 1], enu[:, 2]
    X = -np.sin(lat) * N + np.cos(lat) * U
    Y = E
    Z = np.cos(lat) * N + np.sin(lat) * U
    xyz = np.stack([X, Y, Z], axis=1)
    nant = len(enu)
    a1, a2 = np.triu_indices(nant, k=1)
    if with_autocorr:
        a1 = np.concatenate([a1, np.arange(nant)])
        a2 = np.concatenate([a2, np.arange(nant)])
    B = xyz[a2] - xyz[a1]  # [nbl][3]
    H = np.deg2rad(15.0 * np.linspace(ha_hours[0], ha_hours[1], ntime))
    sH, cH = np.sin(H)[:, None], np.cos(H)[:, None]
    Bx, By, Bz = B[None, :, 0], B[None, :, 1], B[None, :, 2]
    u = sH * Bx + cH * By
    v = -np.sin(dec) * cH * Bx + np.sin(dec) * sH * By + np.cos(dec) * Bz
    w = np.cos(dec) * cH * Bx - np.cos(dec) * sH * By + np.sin(dec) * Bz
    uvw = np.stack([u, v, w], axis=-1).reshape(-1, 3)
    ant1 = np.tile(a1, ntime).astype(np.int32)
    ant2 = np.tile(a2, ntime).astype(np.int32)
    return uvw, ant1, ant2


def sky(nsrc, fov_rad, seed, pol_frac=0.0, flux=(0.01, 1.0)):
    rng = np.random.default_rng(seed + 1000)
    l = (rng.random(nsrc) - 0.5) * fov_rad
    m = (rng.random(nsrc) - 0.5) * fov_rad
    S = np.exp(rng.uniform(np.log(flux[0]), np.log(flux[1]), nsrc))
    chi = rng.uniform(0, np.pi, nsrc)
    Q = pol_frac * S * np.cos(2 * chi)
    Uq = pol_frac * S * np.sin(2 * chi)
    return dict(l=l, m=m, I=S, Q=Q, U=Uq, V=np.zeros(nsrc))


def predict_numpy(uvw, inv_lambda, src, npol):
    """Direct DFT prediction on the CPU: [nrows][nchan][npol] complex128."""
    n = np.sqrt(1.0 - src["l"] ** 2 - src["m"] ** 2)
    nrows, nchan = len(uvw), len(inv_lambda)
    out = np.zeros((nrows, nchan, npol), np.complex128)
    for c, il in enumerate(inv_lambda):
        ph = -2j * np.pi * il * (uvw[:, 0:1] * src["l"][None] + uvw[:, 1:2] * src["m"][None] +
                                 uvw[:, 2:3] * (n[None] - 1.0))
        e = np.exp(ph)  # [nrows][nsrc]
        I, Q, U, V = (e @ src[k] for k in ("I", "Q", "U", "V"))
        if npol == 1:
            out[:, c, 0] = I
        elif npol == 2:
            out[:, c, 0], out[:, c, 1] = I + Q, I - Q
        else:
            out[:, c, 0], out[:, c, 1] = I + Q, U + 1j * V
            out[:, c, 2], out[:, c, 3] = U - 1j * V, I - Q
    return out


def predict_torch(uvw, inv_lambda, src, npol, device="cuda", chunk=1 << 20):
    """Same as predict_numpy, evaluated on the GPU with torch (FP64 phases)."""
    import torch
    n = np.sqrt(1.0 - src["l"] ** 2 - src["m"] ** 2)
    lmn = torch.tensor(np.stack([src["l"], src["m"], n - 1.0]), dtype=torch.float64, device=device)
    fl = {k: torch.tensor(src[k], dtype=torch.float64, device=device) for k in ("I", "Q", "U", "V")}
    il = torch.tensor(np.asarray(inv_lambda), dtype=torch.float64, device=device)
    nrows, nchan = len(uvw), len(inv_lambda)
    out = torch.empty((nrows, nchan, npol), dtype=torch.complex64, device=device)
    U = torch.tensor(uvw, dtype=torch.float64, device=device)
    for r0 in range(0, nrows, chunk):
        ph0 = U[r0:r0 + chunk] @ lmn  # [rows][nsrc] metres
        for c in range(nchan):
            ph = (-2.0 * np.pi) * ph0 * il[c]
            e = torch.polar(torch.ones_like(ph), ph)
            I = e @ fl["I"].to(e.dtype)
            if npol == 1:
                out[r0:r0 + chunk, c, 0] = I.to(torch.complex64)
                continue
            Q = e @ fl["Q"].to(e.dtype)
            if npol == 2:
                out[r0:r0 + chunk, c, 0] = (I + Q).to(torch.complex64)
                out[r0:r0 + chunk, c, 1] = (I - Q).to(torch.complex64)
                continue
            Uu = e @ fl["U"].to(e.dtype)
            Vv = e @ fl["V"].to(e.dtype)
            out[r0:r0 + chunk, c, 0] = (I + Q).to(torch.complex64)
            out[r0:r0 + chunk, c, 1] = (Uu + 1j * Vv).to(torch.complex64)
            out[r0:r0 + chunk, c, 2] = (Uu - 1j * Vv).to(torch.complex64)
            out[r0:r0 + chunk, c, 3] = (I - Q).to(torch.complex64)
    return out


def make_observation(nant=27, radius_m=3000.0, lat_deg=-30.7, dec_deg=-30.0,
                     ha_hours=(-2.0, 2.0), ntime=300, freqs_hz=None, npol=4,
                     nsrc=20, fov_rad=None, pol_frac=0.0, noise=0.01, flag_frac=0.05,
                     seed=1, with_autocorr=True, predict=True):
    """Small-array defaults: 27 antennas in a 3 km disk, 300 steps over +-2 h at
    dec -30, 8 channels over 1.4-1.5 GHz, 20 point sources, sigma 0.01 Jy,
    5% flags, weights 1."""
    if freqs_hz is None:
        freqs_hz = np.linspace(1.4e9, 1.5e9, 8)
    freqs_hz = np.asarray(freqs_hz, dtype=np.float64)
    enu = array_layout(nant, radius_m, seed)
    uvw, a1, a2 = uvw_tracks(enu, lat_deg, dec_deg, ha_hours, ntime, with_autocorr)
    inv_lambda = freqs_hz / C_LIGHT
    nrows, nchan = len(uvw), len(freqs_hz)
    src = sky(nsrc, fov_rad if fov_rad is not None else 512 * 2 * ARCSEC, seed, pol_frac)
    rng = np.random.default_rng(seed + 7)
    if predict:
        vis = predict_numpy(uvw, inv_lambda, src, npol)
        if noise:
            vis = vis + noise * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
        vis = vis.astype(np.complex64)
    else:
        vis = None
    flags = (rng.random((nrows, nchan)) < flag_frac)[:, :, None].repeat(npol, axis=2).astype(np.uint8)
    weight = np.ones((nrows, npol), np.float32)
    # TIME column (MJD seconds, UTC): the hour-angle steps in solar seconds from
    # MJD 60000, one value per time step shared by its rows
    steps = np.linspace(ha_hours[0], ha_hours[1], ntime) * 3600.0 / 1.00273781191135448
    time = np.repeat(MJD0_S + steps - steps[0], nrows // ntime)
    return Observation(uvw=uvw, ant1=a1, ant2=a2, freq=freqs_hz, inv_lambda=inv_lambda,
                       vis=vis, flags=flags, weight=weight, nant=nant, src=src, npol=npol,
                       time=time, lat_deg=lat_deg, dec_deg=dec_deg)


def small_array_obs(seed=1, **kw):
    """27 antennas / 3 km / 300 steps / +-2 h / dec -30 / 8 ch 1.4-1.5 GHz /
    P=4 / 512^2 at 2 arcsec, 32 planes, s=4."""
    kw.setdefault("fov_rad", 512 * 2 * ARCSEC)
    return make_observation(seed=seed, **kw)


def large_array_geometry(seed=2):
    """64 antennas / 8 km / 1800 steps / +-4 h / dec -45 / 64 ch 0.9-1.7 GHz
    (uvw, antennas and frequencies only; visibilities are predicted on the GPU)."""
    return make_observation(nant=64, radius_m=8000.0, dec_deg=-45.0, ha_hours=(-4.0, 4.0),
                            ntime=1800, freqs_hz=np.linspace(0.9e9, 1.7e9, 64), npol=4,
                            fov_rad=4096 * 1.5 * ARCSEC, pol_frac=0.1, seed=seed, predict=False)


def uvw_baselines(enu, a1, a2, lat_deg, dec_deg, H):
    """uvw [ntime*nbl][3] (time-major rows) of the baselines (a1, a2) at hour
    angles H (radians)."""
    lat, dec = np.deg2rad(lat_deg), np.deg2rad(dec_deg)
    E, N, U = enu[:, 0], enu[:, 1], enu[:, 2]
    xyz = np.stack([-np.sin(lat) * N + np.cos(lat) * U, E, np.cos(lat) * N + np.sin(lat) * U], axis=1)
    B = xyz[a2] - xyz[a1]
    sH, cH = np.sin(H)[:, None], np.cos(H)[:, None]
    Bx, By, Bz = B[None, :, 0], B[None, :, 1], B[None, :, 2]
    u = sH * Bx + cH * By
    v = -np.sin(dec) * cH * Bx + np.sin(dec) * sH * By + np.cos(dec) * Bz
    w = np.cos(dec) * cH * Bx - np.cos(dec) * sH * By + np.sin(dec) * Bz
    return np.stack([u, v, w], axis=-1).reshape(-1, 3)


def snapshot_jobs(njobs=512, seed=4, min_vis=1e4, max_vis=5e6):
    """Independent snapshot imaging jobs (Stokes I, P=1): 32-2016 baselines of
    a 64-antenna 8 km array, 1-64 channels in 1.3-1.5 GHz, visibility counts
    log-uniform in [min_vis, max_vis] (time steps of 16 s fill the count),
    images of 256^2..2048^2 at a common 0.85 deg field, w-planes from
    {8, 16, 32, 64}, oversampling 4 or 8.  Returns host-side specs."""
    rng = np.random.default_rng(seed)
    enu = array_layout(64, 8000.0, seed)
    pa1, pa2 = np.triu_indices(64, k=1)
    jobs = []
    for j in range(njobs):
        nvis = float(np.exp(rng.uniform(np.log(min_vis), np.log(max_vis))))
        nbl = int(rng.integers(32, 2017))
        nch = int(rng.integers(1, 65))
        ntime = max(1, int(round(nvis / (nbl * nch))))
        nx = int(rng.choice([256, 512, 1024, 2048]))
        W = int(rng.choice([8, 16, 32, 64]))
        s = int(rng.choice([4, 8]))
        sel = rng.choice(len(pa1), nbl, replace=False)
        a1, a2 = pa1[sel].astype(np.int32), pa2[sel].astype(np.int32)
        h0 = rng.uniform(-4.0, 4.0 - ntime * 16.0 / 3600.0) if ntime * 16.0 < 8 * 3600 else -4.0
        H = np.deg2rad(15.0 * (h0 + np.arange(ntime) * 16.0 / 3600.0))
        uvw = uvw_baselines(enu, a1, a2, -30.7, -45.0, H)
        freqs = np.linspace(1.3e9, 1.5e9, nch) if nch > 1 else np.array([1.4e9])
        jobs.append(dict(nx=nx, cell=1.5 * 2048 / nx, wplanes=W, sampling=s, nchan=nch,
                         uvw=uvw, ant1=np.tile(a1, ntime), ant2=np.tile(a2, ntime),
                         inv_lambda=freqs / C_LIGHT, src=sky(20, 3072 * ARCSEC, seed + j),
                         seed=seed + j))
    return jobs
