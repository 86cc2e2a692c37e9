"""Pins of the oracle (CPU, no GPU): every check here is independent of the
oracle's own formulas -- known answers, closed forms, invariants, brute force
on tiny inputs, library routines -- so that a dropped term, a wrong sign or
index, or a transposed operand in the oracle fails one of them."""
import math

import numpy as np
import pytest

from oracle import ref


# --------------------------------------------------------------------------
# A3 Tukey window
# --------------------------------------------------------------------------
def test_tukey_known_answers():
    # borne1 = round(0.05f*1023/2) = round(25.575) = 26,
    # borne2 = round(1023*(1-0.025f)) = round(997.425) = 997
    t, b1, b2 = ref.tukey1d(1024, 0.05)
    assert (b1, b2) == (26, 997)
    assert np.all(t[27:998] == 1.0)
    # end values: 0.5*(1+cos(3.1415*(0-1))) and the mirrored end
    assert t[0] == pytest.approx(0.5 * (1 + math.cos(-3.1415)), rel=1e-12)
    assert t[1023] == pytest.approx(t[0], rel=1e-6)
    assert t[0] < 1e-8 and t[1023] < 1e-8
    # monotone rise / fall on the tapers
    assert np.all(np.diff(t[:27]) > 0) and np.all(np.diff(t[997:]) < 0)
    # alpha=0.1 on a 256 crop: borne1 = round(12.75) = 13, borne2 = round(242.25) = 242
    _, b1, b2 = ref.tukey1d(256, 0.1)
    assert (b1, b2) == (13, 242)


def test_tukey_float_division_quirk():
    """2*cpt/(alpha*(N-1)) is a float division: the taper differs from the
    double-precision formula by about float epsilon, never more."""
    t, b1, _ = ref.tukey1d(1024, 0.05)
    cpt = np.arange(b1 + 1)
    exact = 0.5 * (1 + np.cos(3.1415 * (2 * cpt / (0.05 * 1023) - 1)))
    assert np.max(np.abs(t[:b1 + 1] - exact)) < 1e-6
    assert np.max(np.abs(t[:b1 + 1] - exact)) > 0  # the quirk is really there


# --------------------------------------------------------------------------
# A1 config parsing
# --------------------------------------------------------------------------
CFG = (
    "# comment line\n"
    "NA 1.20\n"
    "N0 1.333\n"
    "LAMBDA 0.000000632\n"
    "NXMAX 125\n"
    "NXMAX 64\n"                 # last occurrence wins
    "#NB_HOLO 999\n"             # commented out
    "EMPTYKEY \n"
    "CRLF 2.5\r\n"
    "TEXT /tmp/some path/\n"
    "NOSPACE\n"
    "\n"
)


@pytest.fixture()
def cfg_path(tmp_path):
    p = tmp_path / "config_manip.txt"
    p.write_bytes(CFG.encode())
    return p


def test_config_known_answers(cfg_path):
    f32 = lambda s: float(np.float32(float(s)))
    assert ref.extract_val(cfg_path, "NA") == f32("1.20")
    assert ref.extract_val(cfg_path, "N0") == f32("1.333") != 1.333
    assert ref.extract_val(cfg_path, "LAMBDA") == f32("6.32e-7")
    assert ref.extract_val(cfg_path, "NXMAX") == 64.0
    assert ref.extract_val(cfg_path, "NB_HOLO") == 0.0
    assert ref.extract_val(cfg_path, "EMPTYKEY") == 0.0
    assert ref.extract_val(cfg_path, "EMPTYKEY", 1.0) == 1.0
    assert ref.extract_val(cfg_path, "COEF_NA_OBJ_LIMIT", 1.0) == 1.0
    assert ref.extract_val(cfg_path, "CRLF") == 2.5
    assert ref.extract_string(cfg_path, "TEXT") == "/tmp/some path/"
    assert ref.extract_val(cfg_path, "NOSPACE") == 0.0


def test_derive_geom_shipped_config():
    """Tomo_config/config_manip.txt + recon.txt values."""
    g = ref.derive_geom(1024, 125, 512, 247, 777, 1.40, 1.515, 0.000000632,
                        0.0000055, 0.15, 0.0020, 0.75)
    assert g.rayon == 135.0                        # round(125*1.515/1.4)=round(135.27)
    assert g.pix_holo == pytest.approx(5.5e-8, rel=1e-6)   # 5.5e-6 / (75/0.75)
    assert g.pix_uborn == pytest.approx(5.5e-8 * 1024 / 250, rel=1e-6)
    assert g.delta_f == pytest.approx(1 / (5.5e-8 * 1024), rel=1e-6)
    assert g.coin == (122, 652)


# --------------------------------------------------------------------------
# A4 centred FFT + crop
# --------------------------------------------------------------------------
def _tiny_geom(N=16, nx=2, cx=None, cy=None):
    cx = N // 2 - 3 if cx is None else cx
    cy = N // 2 + 3 if cy is None else cy
    pix = nx * 6.33e-7 / (1.2 * N)
    return ref.derive_geom(N, nx, 4 * nx, cx, cy, 1.2, 1.333, 6.33e-7, pix, 1, 1, 1)


def test_centred_fft_bruteforce():
    """out[ky][kx] = N^-2 sum h w exp(-2pi i((kx-N/2)(x-N/2)+(ky-N/2)(y-N/2))/N)."""
    rng = np.random.default_rng(0)
    g = _tiny_geom()
    N, C = g.n, g.C
    h = rng.uniform(0, 4095, size=(N, N))
    w = ref.tukey2d(N, g.alpha_holo)
    got = ref.holo_to_crop(h, g, "f64")
    x0, y0 = g.coin
    x = np.arange(N) - N / 2
    exp = np.empty((C, C), complex)
    for e in range(C):
        for d in range(C):
            kx, ky = x0 + d - N / 2, y0 + e - N / 2
            ph = np.exp(-2j * np.pi * (kx * x[None, :] + ky * x[:, None]) / N)
            exp[e, d] = (h * w * ph).sum() / N ** 2
    assert np.allclose(got, exp, rtol=0, atol=1e-9 * np.abs(exp).max())


def test_carrier_closed_form_and_parseval():
    """Unit window: a pure carrier gives 1/2 at N/2 + (kx, ky) times
    (-1)^(kx+ky) (phase origin at pixel N/2); Parseval holds for the full
    centred spectrum."""
    N, kx, ky = 32, -5, 7
    g = _tiny_geom(N=N, nx=4, cx=N // 2 + kx, cy=N // 2 + ky)
    y, x = np.mgrid[0:N, 0:N]
    h = np.cos(2 * np.pi * (kx * x + ky * y) / N)
    crop = ref.holo_to_crop(h, g, "f64", window=np.ones((N, N)))
    peak = crop[g.nxmax, g.nxmax]
    assert peak == pytest.approx(0.5 * (-1) ** (kx + ky), abs=1e-12)
    mask = np.ones_like(crop, bool)
    mask[g.nxmax, g.nxmax] = False
    assert np.abs(crop[mask]).max() < 1e-12
    rng = np.random.default_rng(1)
    h = rng.normal(size=(N, N))
    full = ref.fftshift2(np.fft.fft2(ref.fftshift2(h))) / N ** 2
    assert (np.abs(full) ** 2).sum() == pytest.approx((h ** 2).sum() / N ** 2, rel=1e-12)


# --------------------------------------------------------------------------
# A5 argmax
# --------------------------------------------------------------------------
def test_argmax_quirks():
    a = np.zeros((4, 4), complex)
    a[0, 0] = 2.0          # slot 0 compares |TF0|^2 = 4
    a[1, 1] = 3.0          # |TF| = 3 < 4 -> the quirk keeps slot 0
    assert ref.coord_spec(a) == 0
    a[0, 0] = 0.5          # |TF0|^2 = 0.25
    a[1, 1] = 0.0
    a[0, 3] = 0.4          # 0.4 > 0.25 although |TF0| = 0.5 > 0.4
    assert ref.coord_spec(a) == 3
    a[:] = 0
    a[0, 2] = 1j
    a[1, 3] = -1.0         # tie: the first index wins
    assert ref.coord_spec(a) == 2


def test_argmax_bruteforce_scan():
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = np.round(rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)), 1)
        flat = a.ravel()
        best, bi = abs(flat[0]) ** 2, 0
        for i in range(1, flat.size):
            if abs(flat[i]) > best:
                best, bi = abs(flat[i]), i
        assert ref.coord_spec(a) == bi


# --------------------------------------------------------------------------
# A6 demodulation
# --------------------------------------------------------------------------
def test_demod_plane_waves():
    C = 16
    s = 3.0 - 2.0j
    crop = np.zeros((C, C), complex)
    px, py = 11, 4
    crop[py, px] = s
    U = ref.demod(crop, px, py)
    assert np.allclose(U, s, atol=1e-12)
    # a second component dx bins away is a tilt exp(2 pi i dx (x - C/2)/C)
    dx = 3
    crop[py, (px + dx) % C] = 1.0
    U = ref.demod(crop, px, py)
    x = np.arange(C)
    expect = s + np.exp(2j * np.pi * dx * (x - C / 2) / C)[None, :]
    assert np.allclose(U, np.broadcast_to(expect, (C, C)), atol=1e-12)


# --------------------------------------------------------------------------
# A9-A11 least squares
# --------------------------------------------------------------------------
def _poly(X, Y, coef):
    return sum(c * X ** i * Y ** j for c, (i, j) in zip(coef, ref.poly_terms()))


def test_lsq_recovers_degree4_background():
    C = 128
    mask = np.full((C, C), 255, np.uint8)
    mask[40:90, 30:70] = 0
    plan = ref.LsqPlan.build(mask)
    assert len(plan.ys) == ((C + 4) // 5) ** 2 - int(((mask[::5, ::5]) <= 45).sum())
    rng = np.random.default_rng(3)
    coef = rng.normal(size=15) / (C ** np.array([i + j for i, j in ref.poly_terms()]))
    Y, X = np.mgrid[0:C, 0:C].astype(float)
    b = _poly(X, Y, coef)
    bg, _ = ref.lsq_background(b, plan)
    assert np.max(np.abs(bg - b)) < 1e-8 * np.max(np.abs(b))
    # a bump inside the masked region is ignored by the fit
    b2 = b.copy()
    b2[50:80, 40:60] += 5.0
    bg2, _ = ref.lsq_background(b2, plan)
    assert np.max(np.abs(bg2 - b)) < 1e-8 * np.max(np.abs(b))


def test_lsq_residual_orthogonal_and_matches_lstsq():
    C = 64
    plan = ref.LsqPlan.build(np.full((C, C), 255, np.uint8))
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(C, C))
    bg, coef = ref.lsq_background(vals, plan)
    r = vals[plan.ys, plan.xs] - bg[plan.ys, plan.xs]
    # normal equations: residual orthogonal to every (scaled) basis column
    A = plan.A / np.abs(plan.A).max(axis=0)
    assert np.max(np.abs(A.T @ r)) < 1e-8 * np.linalg.norm(r)
    c2 = np.linalg.lstsq(plan.A, vals[plan.ys, plan.xs], rcond=None)[0]
    assert np.allclose(plan.full @ c2, bg.ravel(), atol=1e-8)


# --------------------------------------------------------------------------
# A12 Born / Rytov field
# --------------------------------------------------------------------------
def test_born_field_closed_forms():
    C = 64
    g = _tiny_geom(N=256, nx=32)
    plan = ref.LsqPlan.build(np.full((C, C), 255, np.uint8))
    # flat field of any amplitude/phase -> U_B = a/(a + 1e-8) - 1 (the fits
    # are exact; 1e-8 is ampliCorr2's regulariser, Correction_aberration.cpp:182)
    U = np.full((C, C), 3.0 * np.exp(0.7j))
    ub = ref.form_field(U, plan, g, "f64")
    assert np.allclose(ub, 3.0 / (3.0 + 1e-8) - 1, rtol=0, atol=1e-12)
    # degree-2 phase ramp is removed entirely; a masked bump survives as
    # exp(i bump) - 1 in Born and i*bump in Rytov
    Y, X = np.mgrid[0:C, 0:C].astype(float)
    ramp = 0.01 * X - 0.02 * Y + 1e-4 * X * Y
    bump = np.zeros((C, C))
    bump[20:30, 24:40] = 0.3
    mask = np.full((C, C), 255, np.uint8)
    mask[18:32, 22:42] = 0
    plan = ref.LsqPlan.build(mask)
    U = 2.0 * np.exp(1j * (ramp + bump))
    ub = ref.form_field(U, plan, g, "f64")
    a = 2.0 / (2.0 + 1e-8)
    assert np.allclose(ub, a * np.exp(1j * bump) - 1, atol=1e-9)
    g.born = 0
    ur = ref.form_field(U, plan, g, "f64")
    assert np.allclose(ur, math.log(a) + 1j * bump, atol=1e-9)


# --------------------------------------------------------------------------
# A17 Ewald scatter
# --------------------------------------------------------------------------
def _ewald_geom(nx=4, V=16):
    N = 8 * nx
    return ref.derive_geom(N, nx, V, N // 2 - 2, N // 2 + 2, 1.2, 1.333, 6.33e-7,
                           nx * 6.33e-7 / (1.2 * N), 1, 1, 1)


def _brute_scatter(S, pos, g):
    """Triple loop in the reference's order (fonctions.cpp:906-932)."""
    A = S.shape[0]
    nx, V, R = g.nxmax, g.dim_final, g.rayon
    n0 = np.float32(g.n0)
    kv = np.float32(2 * 3.14159 / g.lam)
    k0 = np.float32(kv * n0)
    vol = np.zeros(V ** 3, complex)
    cnt = np.zeros(V ** 3, np.int64)
    for a in range(A):
        fxm0, fym0 = int(pos[a]) - nx, int(pos[a + A]) - nx
        fz2 = R * R - fxm0 * fxm0 - fym0 * fym0
        if ref.round_half_away(fz2) < 0:
            continue
        fzm0 = math.sqrt(fz2)
        for fdy in range(-nx, nx):
            for fdx in range(-nx, nx):
                if fdx * fdx + fdy * fdy < g.coef_na * nx * nx:
                    koz = float(ref.round_half_away(math.sqrt(R * R - fdx * fdx - fdy * fdy) - fzm0))
                    sdz = math.sqrt(R * R - fdx * fdx - fdy * fdy) * (1.0 / R)
                    vox = int((-fxm0 + fdx + V // 2) + (-fym0 + fdy + V // 2) * V
                              + ref.round_half_away((koz + V // 2) * V * V))
                    val = complex(0, float(k0) / 3.14159) * (-2 * 3.14159) * sdz \
                        * S[a, fdy + nx, fdx + nx]
                    vol[vox] += val
                    cnt[vox] += 1
    return vol, cnt


def test_ewald_bruteforce_tiny():
    g = _ewald_geom()
    nx, C = g.nxmax, g.C
    rng = np.random.default_rng(5)
    pos = np.array([nx, nx + 2, 0, 1, nx, nx - 3, 0, 7], dtype=np.int64)  # A=4
    S = rng.normal(size=(4, C, C)) + 1j * rng.normal(size=(4, C, C))
    vol, cnt = ref.ewald_scatter(S, pos, g, "f64")
    bv, bc = _brute_scatter(S, pos, g)
    assert np.array_equal(cnt, bc)
    assert np.allclose(vol, bv, rtol=1e-12, atol=1e-12 * np.abs(bv).max())


def test_ewald_counts_and_cap_closed_form():
    g = _ewald_geom(nx=6, V=24)
    nx, C, V, R = g.nxmax, g.C, g.dim_final, g.rayon
    P = sum(1 for y in range(-nx, nx) for x in range(-nx, nx) if x * x + y * y < nx * nx)
    # centred specular: the cap z = round(sqrt(R^2-r^2) - R)
    S = np.zeros((1, C, C), complex)
    fdx, fdy = 3, -2
    S[0, fdy + nx, fdx + nx] = 1.0 + 2.0j
    pos = np.array([nx, nx])
    vol, cnt = ref.ewald_scatter(S, pos, g, "f64")
    assert cnt.sum() == P
    z = round(math.sqrt(R * R - fdx * fdx - fdy * fdy) - R) + V // 2
    vox = (fdx + V // 2) + (fdy + V // 2) * V + z * V * V
    nz = np.flatnonzero(vol)
    assert list(nz) == [vox]
    # value = (i k0/PI)(-2 PI) sdz S = -2 i k0 sdz S with PI cancelling
    k0 = 2 * math.pi / g.lam * g.n0
    sdz = math.sqrt(R * R - fdx * fdx - fdy * fdy) / R
    assert vol[vox] == pytest.approx(-2j * k0 * sdz * (1 + 2j), rel=1e-6)
    # counts over several angles: sum of P per non-skipped angle
    pos = np.array([nx, 0, nx + 1, nx, 0, nx - 2])   # second angle at the corner
    skipped = (0 - nx) ** 2 + (0 - nx) ** 2 > R * R
    S = np.ones((3, C, C), complex)
    _, cnt = ref.ewald_scatter(S, pos, g, "f64")
    assert cnt.sum() == P * (3 - int(skipped))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ewald_sparse_equals_dense(prec):
    """The sparse form (for volumes too large to hold densely) holds exactly
    the dense form's nonzero voxels, bit for bit, with overlapping angles."""
    g = _ewald_geom(nx=6, V=24)
    nx, C = g.nxmax, g.C
    rng = np.random.default_rng(8)
    A = 9
    pos = np.concatenate([nx + rng.integers(-3, 4, A), nx + rng.integers(-3, 4, A)])
    S = rng.normal(size=(A, C, C)) + 1j * rng.normal(size=(A, C, C))
    vol, cnt = ref.ewald_scatter(S, pos, g, prec)
    vox, sums, c = ref.ewald_scatter_sparse(S, pos, g, prec)
    assert np.array_equal(vox, np.flatnonzero(cnt))
    assert np.array_equal(c, cnt[vox]) and int(c.max()) > 1
    assert np.array_equal(sums.view(np.uint8), vol[vox].view(np.uint8))


# --------------------------------------------------------------------------
# A18-A22 finalize
# --------------------------------------------------------------------------
def test_finalize_delta_is_plane_wave():
    g = _ewald_geom(nx=4, V=16)
    V = g.dim_final
    u0 = np.array([5, 9, 12])  # (x, y, z) voxel of the centred spectrum
    vol = np.zeros(V ** 3, complex)
    cnt = np.zeros(V ** 3, np.int32)
    i0 = u0[0] + u0[1] * V + u0[2] * V * V
    vol[i0] = 6.0 - 3.0j
    cnt[i0] = 3                       # normalised value 2 - 1j
    idx, otf = ref.finalize(vol, cnt, g, "f64")
    z, y, x = np.mgrid[0:V, 0:V, 0:V]
    k = u0 - V // 2
    ph = np.exp(2j * np.pi * (k[0] * (x - V // 2) + k[1] * (y - V // 2) + k[2] * (z - V // 2)) / V)
    kv = 2 * math.pi / g.lam
    cte = -1 / (2 * g.n0 * kv * kv)
    expect = cte * g.delta_f ** 3 * (2 - 1j) * ph
    assert np.allclose(idx, expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())
    assert otf.sum() == 1 and otf.ravel()[i0]


def test_finalize_parseval():
    g = _ewald_geom(nx=4, V=16)
    V = g.dim_final
    rng = np.random.default_rng(6)
    vol = rng.normal(size=V ** 3) + 1j * rng.normal(size=V ** 3)
    cnt = rng.integers(0, 4, size=V ** 3).astype(np.int32)
    idx, _ = ref.finalize(vol, cnt, g, "f64")
    t = vol / np.where(cnt == 0, 1, cnt)
    kv = 2 * math.pi / g.lam
    cte = -1 / (2 * g.n0 * kv * kv)
    lhs = (np.abs(idx) ** 2).sum()
    rhs = cte ** 2 * g.delta_f ** 6 * V ** 3 * (np.abs(t) ** 2).sum()
    assert lhs == pytest.approx(rhs, rel=1e-10)


# --------------------------------------------------------------------------
# phase unwrapping (DEROUL 1, VOLKOV 1)
# --------------------------------------------------------------------------
def test_kvect_and_mirror_known_answers():
    assert np.array_equal(ref.kvect_shift(8),
                          np.array([0, 1, 2, 3, -4, -3, -2, -1], dtype=np.float64) / 8)
    img = np.array([[1.0, 2.0], [3.0, 4.0]])
    # hand-written 4x4 even extension; the original is the lower-right quadrant
    want = np.array([[4, 3, 3, 4],
                     [2, 1, 1, 2],
                     [2, 1, 1, 2],
                     [4, 3, 3, 4]], dtype=np.float64)
    assert np.array_equal(ref.symmetrise_mirror(img), want)


def test_spectral_gradient_closed_form():
    D = 32
    y, x = np.mgrid[0:D, 0:D]
    f = np.cos(2 * np.pi * 3 * x / D) * np.cos(2 * np.pi * 2 * y / D) + 0.5 * np.sin(2 * np.pi * 5 * y / D)
    gx, gy = ref.gradient_fft(f)
    ex = -2 * np.pi * 3 / D * np.sin(2 * np.pi * 3 * x / D) * np.cos(2 * np.pi * 2 * y / D)
    ey = (-2 * np.pi * 2 / D * np.cos(2 * np.pi * 3 * x / D) * np.sin(2 * np.pi * 2 * y / D)
          + 0.5 * 2 * np.pi * 5 / D * np.cos(2 * np.pi * 5 * y / D))
    assert np.abs(gx - ex).max() < 1e-12
    assert np.abs(gy - ey).max() < 1e-12
    # a complex input differentiates term by term
    z = np.exp(1j * 2 * np.pi * 4 * x / D)
    zx, zy = ref.gradient_fft(z)
    assert np.abs(zx - 1j * 2 * np.pi * 4 / D * z).max() < 1e-12
    assert np.abs(zy).max() < 1e-12


def test_integration_inverts_the_gradient():
    D = 32
    y, x = np.mgrid[0:D, 0:D]
    f = (np.cos(2 * np.pi * 3 * x / D) * np.cos(2 * np.pi * 2 * y / D)
         + 0.5 * np.sin(2 * np.pi * 5 * y / D) + 0.25 * np.cos(2 * np.pi * (x + 2 * y) / D))
    gx, gy = ref.gradient_fft(f)
    r = ref.integ_grad(gx.real, gy.real)
    assert np.abs(r.real - (f - f.mean())).max() < 1e-12
    assert np.abs(r.imag).max() < 1e-12


@pytest.mark.parametrize("kind", ["bowl", "ramp", "bump"])
def test_volkov_recovers_smooth_phases(kind):
    C = 64
    y, x = np.mgrid[0:C, 0:C].astype(np.float64)
    true = {"bowl": 0.004 * ((x - 30) ** 2 + (y - 36) ** 2),
            "ramp": 0.3 * x + 0.2 * y,
            "bump": 12 * np.exp(-((x - 30) ** 2 + (y - 34) ** 2) / (2 * 9.0 ** 2))}[kind]
    assert true.max() - true.min() > 3 * np.pi  # several wraps
    wrapped = np.angle(np.exp(1j * true))
    for prec in ("f64", "f32"):
        u = ref.volkov_unwrap(wrapped, prec).astype(np.float64)
        d = u - true
        d -= d.mean()  # unwrapping is defined up to a constant
        assert np.abs(d).max() < 2e-3, (kind, prec, np.abs(d).max())
    # without wraps the 2 pi multiples vanish
    small = 0.4 * true / (true.max() - true.min())
    u = ref.volkov_unwrap(small)
    d = u - small
    assert np.abs(d - d.mean()).max() < 1e-3


def test_sampled_index_equals_dense_oracle():
    """The length tests' sampled oracle (the centred inverse transform as an
    explicit sum over the hit voxels) equals the dense oracle (numpy ifftn of
    the full volume) at sampled voxels, FP64 and FP32 pipelines."""
    import lengths as L
    from cases import make_case, rel_l2
    case = make_case("small")
    g = case.g
    V = g.dim_final
    idx = np.random.default_rng(0).integers(0, V ** 3, 300)
    for prec, tol in (("f64", 1e-12), ("f32", 1e-6)):
        f, p = ref.preprocess(case.acq.holos, g, case.acq.mask, prec)
        dense = ref.reconstruct(f, p, g, prec)[0].ravel()[idx]
        assert rel_l2(L.index_at(f, p, g, prec, idx), dense) < tol


def test_cpu_baseline_pipeline_equals_oracle():
    """The timed CPU baseline (oracle/cpu_pipeline.py: the FP64 arithmetic at
    the reference's stage boundaries, FFTs threaded) gives the FP64 oracle's
    speculars and index volume."""
    from oracle import cpu_pipeline
    from cases import make_case, rel_l2
    case = make_case("small")
    idx, pos, t, name = cpu_pipeline.run(case.acq.holos, case.g, case.acq.mask, 2)
    f64, p64 = ref.preprocess(case.acq.holos, case.g, case.acq.mask, "f64")
    want = ref.reconstruct(f64, p64, case.g, "f64")[0]
    assert np.array_equal(pos, p64)
    # FP64 rounding of another FFT library, through the normal-equation fit
    assert rel_l2(idx, want) < 1e-9
    assert set(t) == {"T_pre1", "T_pre2", "T_rec1", "T_rec2"} and name


# --------------------------------------------------------------------------
# Herraez unwrapping (DEROUL 1, VOLKOV 0)
# --------------------------------------------------------------------------
def test_glibc_rand_known_answers():
    """rand() with glibc's default seed: the first outputs every glibc
    program prints (1804289383, 846930886, ...); any start offset continues
    the same stream."""
    assert list(ref.glibc_rand(0, 5)) == [1804289383, 846930886, 1681692777, 1714636915,
                                          1957747793]
    assert list(ref.glibc_rand(3, 4)) == list(ref.glibc_rand(0, 7)[3:])


def test_herraez_border_reliability_wraps_like_int():
    """9999999 + rand() is a 32-bit int sum: past INT_MAX it wraps negative."""
    r = ref.herraez_border_reliability(np.array([0, 2137483647, 2137483649, 2147483647]))
    assert list(r) == [np.float32(9999999), np.float32(2147483646),
                       np.float32(-2147483648), np.float32(-2137483650)]


def _literal_mix(keys, L, R):
    i = j = 0
    out = []
    while i < len(L) and j < len(R):
        if keys[L[i]] <= keys[R[j]]:
            out.append(L[i]); i += 1
        else:
            out.append(R[j]); j += 1
    return out + list(L[i:]) + list(R[j:])


def test_herraez_mix_equals_two_pointer_merge():
    """The oracle's merge (stable merge of running maxima) against the
    literal two-pointer merge on unsorted runs with ties."""
    rng = np.random.default_rng(3)
    for _ in range(3000):
        h = int(rng.integers(1, 9))
        keys = rng.integers(0, 4, 2 * h).astype(np.float32)
        L, R = np.arange(h), np.arange(h, 2 * h)
        assert list(ref._herr_mix(keys, L, R)) == _literal_mix(keys, L, R)


def test_herraez_sort_known_answers():
    """Hand-traced runs of the reference's sort: an odd range leaves its last
    element in place and merges unsorted halves."""
    f = lambda k: list(ref.herraez_sort(np.array(k, np.float32)))
    assert f([3, 1, 2]) == [1, 0, 2]
    assert f([5, 4, 3, 2, 1]) == [3, 2, 1, 0, 4]
    assert f([6, 5, 4, 3, 2, 1]) == [4, 3, 5, 1, 0, 2]
    assert f([1, 2, 3, 4]) == [0, 1, 2, 3]
    assert f([2, 2, 1, 1]) == [2, 3, 0, 1]  # ties: left run first


def test_herraez_tree_is_the_rank_mst():
    """The grouping merges exactly the edges of the minimum spanning tree
    with the sort position as weight (scipy's MST on the same graph), and
    the increments satisfy inc(p2) = inc(p1) - w on every tree edge."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    rng = np.random.default_rng(8)
    C = 12
    phi = np.angle(np.exp(1j * rng.normal(scale=2.0, size=(C, C))))  # residues
    v = phi.astype(np.float32)
    rel = ref.herraez_reliability(phi, np.full((C, C), 1e7, np.float32))
    p1, p2, reliab, inc = ref.herraez_edges(v, rel)
    order = ref.herraez_sort(reliab)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    mst = minimum_spanning_tree(coo_matrix((rank + 1.0, (p1, p2)), shape=(C * C, C * C))).tocoo()
    tree = {(min(a, b), max(a, b)) for a, b in zip(mst.row, mst.col)}
    incr = ref.herraez_gather(order, p1, p2, inc, C * C)
    in_tree = [e for e in range(len(p1)) if (p1[e], p2[e]) in tree]
    assert len(in_tree) == C * C - 1
    for e in in_tree:
        assert incr[p2[e]] == incr[p1[e]] - inc[e]
    assert (incr == 0).any()  # the last surviving group head


def test_herraez_smooth_phase_recovery():
    """Smooth phases spanning several wraps (bowl + ramp, Gaussian bump) come
    back to float precision, up to one multiple of 2 pi."""
    C = 48
    y, x = np.mgrid[0:C, 0:C].astype(np.float64)
    for true in (0.004 * ((x - 17) ** 2 + (y - 30) ** 2) + 0.35 * x,
                 14.0 * np.exp(-((x - 24) ** 2 + (y - 20) ** 2) / 150.0)):
        for prec in (np.float64, np.float32):
            wr = np.angle(np.exp(1j * true)).astype(prec)
            u = ref.herraez_unwrap(wr, 3).astype(np.float64)
            k = np.round(np.mean(u - true) / (2 * np.pi))
            assert np.abs(u - true - 2 * np.pi * k).max() < 2e-5


def test_c_aber0_single_bin_closed_form():
    """C_ABER 0 (main.cpp:328-341): a crop whose only nonzero bin is the
    specular s demodulates to the constant s (the -posSpec shift moves it to
    bin 0, the unnormalised inverse FFT spreads it, the NXMAX shift keeps it
    constant); times conj(s)/|s| that is the constant |s|, real."""
    from cases import make_case
    case = make_case("small", c_aber=0)
    g = case.g
    C = g.C
    px, py = 19, 7
    s = 3.0 - 4.0j
    crop = np.zeros((C, C), np.complex128)
    crop[py, px] = s
    assert ref.coord_spec(crop) == py * C + px
    U = ref.demod(crop, px, py)
    assert np.allclose(U, s, rtol=0, atol=1e-12)
    plan = ref.LsqPlan.build(ref.load_mask(C, None))
    f = ref.form_field(U, plan, g, "f64", spec_value=crop[py, px])
    assert np.allclose(f, abs(s), rtol=0, atol=1e-12)


def test_arrayfire_crop_window():
    """The ArrayFire variant's crop (pretraitement_hors_axe_GPU
    src/GPU_fonctions.cpp:122): the plain crop times (1+i) T(x) T(y), T the
    pinned tukey1d factor: for a pure carrier the single peak becomes
    (1+i) T(bx) T(by) times the plain peak, every other bin stays 0."""
    import dataclasses
    from cases import make_case
    case = make_case("small")
    N, C = case.g.n, case.g.C
    ga = dataclasses.replace(case.g, alpha_holo=0.025, alpha_crop=0.025)
    gp = dataclasses.replace(case.g, alpha_holo=0.025)
    x0, y0 = ga.coin
    bx, by = 5, 9  # crop bin of the carrier
    kx, ky = x0 + bx - N // 2, y0 + by - N // 2
    y, x = np.mgrid[0:N, 0:N]
    h = 100 + np.cos(2 * np.pi * (kx * x + ky * y) / N)
    plain = ref.holo_to_crop(h, gp)
    af = ref.holo_to_crop(h, ga)
    T = ref.tukey1d(C, 0.025)[0]
    assert np.isclose(af[by, bx], (1 + 1j) * T[bx] * T[by] * plain[by, bx], rtol=1e-12)
    w = np.outer(T, T)
    assert np.allclose(af, (1 + 1j) * w * plain, rtol=0, atol=1e-12 * abs(plain).max())
