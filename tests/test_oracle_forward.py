"""Pins of the oracle's Born forward model and its adjoint (CPU, no GPU).

Each check is independent of the oracle's own formulas: a brute-force centred
DFT, the closed form of a point object, calcHolo's loop written literally in
the simulator's centred "human" coordinates (Point2D/Point3D::coordI flip y),
the algebraic value of calcHolo's constant product, the inner-product
definition of the adjoint, and the round trip through the (separately
pinned) reconstruction, which must give back the object on the hit voxels."""
import math

import numpy as np
import pytest

from oracle import ref


def _geom(nx=4, V=16):
    N = 8 * nx
    return ref.derive_geom(N, nx, V, N // 2 - 2, N // 2 + 2, 1.2, 1.333, 6.33e-7,
                           nx * 6.33e-7 / (1.2 * N), 1, 1, 1)


def _pos(nx, pts):
    xs = [p[0] for p in pts]
    ys = [p[1] for p in pts]
    return np.array(xs + ys, dtype=np.int64)


def test_contrast_spectrum_bruteforce_dft():
    """TF[k] = Tp^3 sum_x dn[x] exp(-2 pi i (k - V/2).(x - V/2) / V): the
    centred transform written as a plain sum on an 8^3 volume."""
    g = _geom(nx=2, V=8)
    V = g.dim_final
    rng = np.random.default_rng(3)
    dn = rng.normal(size=(V, V, V)) + 1j * rng.normal(size=(V, V, V))
    got = ref.contrast_spectrum(dn, g, "f64")
    c = np.arange(V) - V // 2
    E = np.exp(-2j * np.pi * np.outer(c, c) / V)  # E[k, x]
    want = np.einsum("az,by,cx,zyx->abc", E, E, E, dn) * ref.tomo_pixel(g) ** 3
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_calcholo_constant_is_minus_i_kv_over_sdz():
    """(-1/(2 PI)) (-i PI/k0) (-2 kv^2 n0) / sdz simplifies to -i kv / sdz
    (k0 = kv n0), with kv = 2*3.14159/lambda."""
    g = _geom(nx=4, V=16)
    nx, R = g.nxmax, g.rayon
    w = ref.born_weight(g)
    kv = 2 * 3.14159 / g.lam
    for fx, fy in [(0, 0), (3, -2), (-3, 1)]:
        sdz = math.sqrt(R * R - fx * fx - fy * fy) / R
        assert w[fy + nx, fx + nx] == pytest.approx(-1j * kv / sdz, rel=1e-13)
    assert w[0, 0] == 0  # corner: outside the pupil


def test_point_object_closed_form():
    """A point contrast at centred offset d: every cap sample is
    w(fd) Tp^3 exp(-2 pi i u.d / V), u = (fd - fm0, koz)."""
    g = _geom(nx=4, V=16)
    nx, V, R = g.nxmax, g.dim_final, g.rayon
    d = np.array([2, -3, 1])
    dn = np.zeros((V, V, V), complex)
    dn[V // 2 + d[2], V // 2 + d[1], V // 2 + d[0]] = 0.05
    pos = _pos(nx, [(nx, nx), (nx + 2, nx - 1), (1, nx + 3), (nx - 1, nx + 3)])
    S = ref.born_forward(dn, pos, g, "f64")
    A = 4
    kv = 2 * 3.14159 / g.lam
    tp3 = ref.tomo_pixel(g) ** 3
    for a in range(A):
        fxm0, fym0 = int(pos[a]) - nx, int(pos[a + A]) - nx
        if R * R - fxm0 ** 2 - fym0 ** 2 < 0:  # off the sphere: skipped
            assert not S[a].any()
            continue
        fzm0 = math.sqrt(R * R - fxm0 ** 2 - fym0 ** 2)
        for fy in range(-nx, nx):
            for fx in range(-nx, nx):
                got = S[a, fy + nx, fx + nx]
                if fx * fx + fy * fy >= nx * nx:
                    assert got == 0
                    continue
                h = math.sqrt(R * R - fx * fx - fy * fy)
                koz = math.floor(h - fzm0 + 0.5) if h - fzm0 >= 0 else -math.floor(fzm0 - h + 0.5)
                u = np.array([fx - fxm0, fy - fym0, koz])
                want = (-1j * kv / (h / R)) * tp3 * 0.05 * np.exp(-2j * np.pi * (u @ d) / V)
                assert got == pytest.approx(want, rel=1e-10, abs=1e-12 * abs(want))


def test_calcholo_literal_in_human_coordinates():
    """calcHolo (Simul_holo/src/fonctions.cpp:109-140) as written: centred
    coordinates with y up, fobj = round(fd - fi), indices through coordI
    (x + dim/2, -y + dim/2, z + dim/2); spec_H from tab_posSpec
    (main.cpp:203-204: x = spec.x + dim/2, y = -spec.y + dim/2)."""
    g = _geom(nx=6, V=24)
    nx, V, C, R = g.nxmax, g.dim_final, g.C, g.rayon
    assert R > nx
    rng = np.random.default_rng(11)
    dn = rng.normal(size=(V, V, V)) * 0.01
    pos = _pos(nx, [(nx, nx), (nx + 3, nx - 2), (2, nx + 1)])
    A = 3
    TF = ref.contrast_spectrum(dn, g, "f64").ravel()
    kv = 2 * 3.14159 / g.lam
    k0 = kv * g.n0
    cte = (-1 / (2 * 3.14159)) * (-1j * 3.14159 / k0) * (-2 * kv * kv * g.n0)
    S = ref.born_forward(dn, pos, g, "f64")
    for a in range(A):
        sx, sy = int(pos[a]) - C // 2, -(int(pos[a + A]) - C // 2)
        fiz = math.sqrt(R * R - sx * sx - sy * sy)
        hol = np.zeros((C, C), complex)
        for fdx in range(-nx, nx + 1):
            for fdy in range(-nx, nx + 1):
                if fdx * fdx + fdy * fdy < nx * nx:
                    fdz = math.sqrt(R * R - fdx * fdx - fdy * fdy)
                    ox = ref.round_half_away(fdx - sx)
                    oy = ref.round_half_away(fdy - sy)
                    oz = ref.round_half_away(fdz - fiz)
                    vox = int((oz + V // 2) * V * V + (-oy + V // 2) * V + (ox + V // 2))
                    hol[-fdy + C // 2, fdx + C // 2] = cte / (fdz / R) * TF[vox]
        assert np.allclose(S[a], hol, rtol=1e-12, atol=1e-12 * np.abs(hol).max())


def test_adjoint_inner_products():
    """<A x, y> = <x, A^H y> for random complex x, y (skipped angle included)."""
    g = _geom(nx=4, V=16)
    nx, V, C = g.nxmax, g.dim_final, g.C
    rng = np.random.default_rng(8)
    pos = _pos(nx, [(nx, nx), (nx + 2, nx - 3), (0, 0), (nx - 1, 2)])
    x = rng.normal(size=(V, V, V)) + 1j * rng.normal(size=(V, V, V))
    y = rng.normal(size=(4, C, C)) + 1j * rng.normal(size=(4, C, C))
    Ax = ref.born_forward(x, pos, g, "f64")
    AHy = ref.born_adjoint(y, pos, g)
    lhs = np.vdot(y, Ax)          # <Ax, y> as conj(y).Ax
    rhs = np.vdot(AHy, x)         # <x, A^H y> as conj(A^H y).x
    assert lhs == pytest.approx(rhs, rel=1e-12)
    assert abs(lhs) > 0


def test_round_trip_through_the_reconstruction():
    """ewald_scatter + finalize of the simulated spectra gives back the
    contrast projected on the hit voxels: each hit voxel's mean is its own
    TF value times (-2 k0 kv)(-1/(2 n0 kv^2)) = 1 up to the reference's
    mixed PI constants (3.14159 vs M_PI, float k0: ~1e-6)."""
    g = _geom(nx=6, V=24)
    nx, V = g.nxmax, g.dim_final
    rng = np.random.default_rng(9)
    zz, yy, xx = np.mgrid[0:V, 0:V, 0:V] - V // 2
    dn = 0.04 * np.exp(-((xx - 1) ** 2 + (yy + 2) ** 2 + zz ** 2) / 6.0) \
        + 0.002 * rng.normal(size=(V, V, V))
    pts = [(nx + int(round(0.8 * nx * math.cos(t))), nx + int(round(0.8 * nx * math.sin(t))))
           for t in np.linspace(0, 2 * math.pi, 12, endpoint=False)]
    pos = _pos(nx, pts)
    S = ref.born_forward(dn, pos, g, "f64")
    vol, cnt = ref.ewald_scatter(S, pos, g, "f64")
    idx, _ = ref.finalize(vol, cnt, g, "f64")
    TF = ref.contrast_spectrum(dn, g, "f64")
    masked = np.where(cnt.reshape(V, V, V) > 0, TF, 0)
    want = ref.fftshift3(np.fft.ifftn(ref.fftshift3(masked), norm="forward")) * g.delta_f ** 3
    err = np.linalg.norm(idx - want) / np.linalg.norm(want)
    assert err < 1e-5
    assert cnt.sum() > 0


def test_data_gradient_central_differences():
    """L(n) = 1/2 ||A n - y||^2 for a real n: its gradient Re(A^H (A n - y))
    (the adjoint-iterative form's data gradient) against central finite
    differences, h = 1e-6, along random directions (FP64)."""
    g = _geom(nx=6, V=24)
    nx, V, C = g.nxmax, g.dim_final, g.C
    rng = np.random.default_rng(14)
    pos = _pos(nx, [(nx + 1, nx - 2), (nx - 3, nx + 1), (nx, nx + 4), (nx + 4, nx)])
    n = 0.03 * rng.random(size=(V, V, V))
    y = 1e-9 * (rng.normal(size=(4, C, C)) + 1j * rng.normal(size=(4, C, C)))

    def loss(v):
        r = ref.born_forward(v, pos, g, "f64") - y
        return 0.5 * float(np.vdot(r, r).real)

    grad = ref.born_adjoint(ref.born_forward(n, pos, g, "f64") - y, pos, g).real
    h = 1e-6
    for k in range(3):
        d = rng.normal(size=(V, V, V))
        fd = (loss(n + h * d) - loss(n - h * d)) / (2 * h)
        an = float((grad * d).sum())
        assert fd == pytest.approx(an, rel=1e-6)


def test_field_synthesis_inverts_field_to_spectrum():
    """spectra_to_fields is the inverse of the reconstruction's per-angle
    spectrum with a flat window (shift, FFT x pixUBorn^2, shift): fields ->
    spectra -> fields returns the fields (C^2 pixUBorn^2 delta_f^2 = 1)."""
    g = _geom(nx=4, V=16)
    C = g.C
    rng = np.random.default_rng(15)
    U = rng.normal(size=(3, C, C)) + 1j * rng.normal(size=(3, C, C))
    pos = _pos(4, [(4, 4), (6, 2), (1, 5)])
    ones = np.ones((C, C))
    S = np.stack([ref.field_to_spectrum(U[a], int(pos[a]), int(pos[a + 3]), g, "f64", ones)
                  for a in range(3)])
    back = ref.spectra_to_fields(S, pos, g)
    assert np.allclose(back, U, rtol=1e-12, atol=1e-12)
    # a single spectral bin at the specular gives a flat field delta_f^2
    S = np.zeros((1, C, C), complex)
    S[0, 2, 6] = 1.0
    f = ref.spectra_to_fields(S, np.array([6, 2]), g)
    assert np.allclose(f, g.delta_f ** 2)
