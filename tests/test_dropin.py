"""The drop-in executables (tomo_pretraitement, tomo_reconstruction) driven
exactly like the reference: a HOME with .config/gui_tomo.conf pointing at
recon.txt / config_manip.txt, holograms i%03d.pgm and Image_mask.pgm, outputs
in CHEMIN_RESULT in the reference's file formats."""
import os
import subprocess

import numpy as np
import pytest

import mtd_transmission_amd as m
from cases import LAM, N0, NA, make_case, rel_l2
from oracle import ref
from tiffread import read_tiff_f32

PRE = os.path.join(m.BIN_DIR, "tomo_pretraitement")
REC = os.path.join(m.BIN_DIR, "tomo_reconstruction")


def write_pgm16(path, img):
    img = np.asarray(img, dtype=">u2")
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n65535\n" % (img.shape[1], img.shape[0]))
        f.write(img.tobytes())


def write_pgm8(path, img):
    img = np.asarray(img, dtype=np.uint8)
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(img.tobytes())


def setup_home(tmp_path, case, drop=(), deroul=0, born=1, volkov=1):
    home, acq, res, cfg = (tmp_path / d for d in ("home", "acquis", "result", "config"))
    for d in (home / ".config", acq, res, cfg):
        d.mkdir(parents=True)
    (home / ".config" / "gui_tomo.conf").write_text(
        f"CHEMIN_RESULT {res}\nCHEMIN_ACQUIS {acq}\nCHEMIN_CONFIG {cfg}\n"
        f"CHEMIN_CONFIG_PC_ACQUIS {cfg}\n")
    p = case.params
    (cfg / "config_manip.txt").write_text(
        f"NA {NA}\nN0 {N0}\nLAMBDA {LAM!r}\nF_TUBE 1\nF_OBJ 1\nTPCAM {p['pix']!r}\nRF 1\n"
        f"NXMAX {p['nxmax']}\nCIRCLE_CX {p['cx']}\nCIRCLE_CY {p['cy']}\n"
        f"NB_HOLO {len(case.acq.holos)}\n")
    (cfg / "recon.txt").write_text(
        f"CCD_ROIX {p['n']}\nDEROUL {deroul}\nVOLKOV {volkov}\nC_ABER 1\nBORN {born}\nNB_THREADS 4\n"
        f"PREMIER_ANGLE 0\nDIM_FINAL {p['v']}\nEXPORT_OTF 1\n")
    for k, h in enumerate(case.acq.holos):
        if k not in drop:
            write_pgm16(acq / ("i%03d.pgm" % k), h)
    if case.acq.mask is not None:
        write_pgm8(acq / "Image_mask.pgm", case.acq.mask)
    env = dict(os.environ, HOME=str(home))
    return env, res


def test_help_and_missing_inputs(tmp_path):
    for exe in (PRE, REC):
        r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "-polar" in r.stdout
    case = make_case("small")
    env, res = setup_home(tmp_path, case)
    r = subprocess.run([REC], env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "parametres.raw" in r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("deroul,born,volkov", [(0, 1, 1), (1, 0, 1), (1, 0, 0)])
def test_programs_round_trip(tmp_path, deroul, born, volkov):
    """(1, 0, 1): the shipped recon.txt (Volkov); (1, 0, 0): Herraez."""
    case = make_case("small", born=born, deroul=0 if not deroul else (1 if volkov else 2))
    drop = (3,)  # a missing hologram is skipped and the angles compacted
    print(f"drop-in round trip case=small deroul={deroul} born={born} volkov={volkov}",
          flush=True)
    env, res = setup_home(tmp_path, case, drop, deroul, born, volkov)
    r = subprocess.run([PRE], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    keep = [k for k in range(len(case.acq.holos)) if k not in drop]
    holos = case.acq.holos[keep]
    A, C, V = len(keep), case.g.C, case.g.dim_final
    dim = f"{C}x{C}x{A}"
    fields32, pos32 = ref.preprocess(holos, case.g, case.acq.mask, "f32")
    param = np.fromfile(res / "parametres.raw", "<f8")
    assert list(param[:4]) == [case.g.nxmax, A, case.g.rayon, case.g.n]
    assert param[4] == case.g.pix_holo
    assert np.array_equal(np.fromfile(res / "tab_posSpec.raw", "<f8"), pos32.astype(float))
    fre = np.fromfile(res / f"UBornfinal_Re{dim}.raw", "<f8")
    fim = np.fromfile(res / f"UBornfinal_Im{dim}.raw", "<f8")
    got = (fre + 1j * fim).reshape(A, C, C)
    assert rel_l2(got, fields32) < 1e-5
    centres = read_tiff_f32(res / "centres.tif")[0]
    assert centres[pos32[A - 1], pos32[2 * A - 1]] == A - 1

    r = subprocess.run([REC], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # the executable reads the f64 fields back; run the oracle on the same
    idx32, _, cnt32, otf32 = ref.reconstruct(got.astype(np.complex64), pos32, case.g, "f32")
    ire = read_tiff_f32(res / "indice.tif")
    iim = read_tiff_f32(res / "absorption.tif")
    assert ire.shape == (V, V, V)
    assert rel_l2(ire + 1j * iim, idx32) < 1e-5
    assert np.array_equal(read_tiff_f32(res / "OTF3D.tif").astype(bool), otf32)


H5DUMP = "/opt/conda/bin/h5dump"


def _h5_dataset(path, name, tmp_path):
    out = tmp_path / (name.strip("/").replace("/", "_") + ".bin")
    r = subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(out), str(path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, "<f4")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not in this image")
def test_programs_polar_hdf5(tmp_path):
    """-polar -etat /LC/polar0/: holograms from CHEMIN_RACINE/demosaic/<etat>,
    outputs and the reconstruction's configs in .../UBorn/, the index volume
    as indice_cplx_hdf5.hdf (/indice_complexe/Re, Im; Z, Y, X float32)."""
    case = make_case("small")
    print("drop-in polar case=small", flush=True)
    env, res = setup_home(tmp_path, case)
    home = tmp_path / "home"
    racine = tmp_path / "racine"
    etat = "/LC/polar0/"
    acq = racine / "demosaic" / "LC" / "polar0"
    ub = acq / "UBorn"
    ub.mkdir(parents=True)
    cfg = tmp_path / "config"
    (home / ".config" / "gui_tomo_polar.conf").write_text(
        f"CHEMIN_RESULT {res}\nCHEMIN_ACQUIS {tmp_path / 'acquis'}\nCHEMIN_CONFIG {cfg}\n"
        f"CHEMIN_RACINE {racine}\n")
    for k, h in enumerate(case.acq.holos):
        write_pgm16(acq / ("i%03d.pgm" % k), h)
    write_pgm8(racine / "Image_mask.pgm", case.acq.mask)
    for f in ("recon.txt", "config_manip.txt"):
        (ub / f).write_text((cfg / f).read_text())
    r = subprocess.run([PRE, "-polar", "-etat", etat], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    A, C, V = len(case.acq.holos), case.g.C, case.g.dim_final
    assert (ub / f"UBornfinal_Re{C}x{C}x{A}.raw").exists()
    r = subprocess.run([REC, "-polar", "-etat", etat], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    h5 = ub / "indice_cplx_hdf5.hdf"
    assert h5.exists() and not (ub / "indice.tif").exists()
    hdr = subprocess.run([H5DUMP, "-H", str(h5)], capture_output=True, text=True, timeout=60)
    assert f"( {V}, {V}, {V} )" in hdr.stdout and "H5T_IEEE_F32LE" in hdr.stdout
    re = _h5_dataset(h5, "/indice_complexe/Re", tmp_path).reshape(V, V, V)
    im = _h5_dataset(h5, "/indice_complexe/Im", tmp_path).reshape(V, V, V)
    fre = np.fromfile(ub / f"UBornfinal_Re{C}x{C}x{A}.raw", "<f8")
    fim = np.fromfile(ub / f"UBornfinal_Im{C}x{C}x{A}.raw", "<f8")
    fields = (fre + 1j * fim).reshape(A, C, C).astype(np.complex64)
    pos = np.fromfile(ub / "tab_posSpec.raw", "<f8").astype(np.int64)
    idx32, _, _, _ = ref.reconstruct(fields, pos, case.g, "f32")
    assert rel_l2(re + 1j * im, idx32) < 1e-5


SIM = os.path.join(m.BIN_DIR, "tomo_simul_holo")


@pytest.mark.gpu
def test_simulator_then_reconstruction(tmp_path):
    """tomo_simul_holo as Simul_holo (block object, 4-axis flower scan, fixed
    2048 hologram side) writes the fields, speculars and parameters that
    tomo_reconstruction then reads; fields against the oracle's forward
    model + field synthesis, index volume against the oracle's
    reconstruction of the written fields."""
    import math
    print("simulator -> reconstruction", flush=True)
    home, res, cfg = (tmp_path / d for d in ("home", "result", "config"))
    for d in (home / ".config", res, cfg):
        d.mkdir(parents=True)
    A, nx, N = 24, 16, 2048
    tpcam = 16.3 * LAM / (NA * N)      # NXMAX = (uint16)(NA N Tp / lambda) = 16
    (home / ".config" / "gui_tomo_polar.conf").write_text(
        f"CHEMIN_RESULT {res}\nCHEMIN_CONFIG {cfg}\n")
    (cfg / "config_manip.txt").write_text(
        f"NB_HOLO {A}\nN0 {N0}\nNA_OBJ {NA}\nNA_COND {NA}\nLAMBDA {LAM!r}\nF_TUBE 1\nF_OBJ 1\n"
        f"RF 1\nTPCAM {tpcam!r}\n")
    env = dict(os.environ, HOME=str(home))
    r = subprocess.run([SIM], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    C, V = 2 * nx, 4 * nx
    f32 = lambda v: float(np.float32(v))
    rayon = float(ref.round_half_away(nx / math.sin(math.asin(f32(NA) / f32(N0)))))
    param = np.fromfile(res / "parametres.raw", "<f8")
    assert list(param[:4]) == [nx, A, rayon, N]
    # the flower pattern, from OTF::bFleur's formula
    th = 2 * math.pi / A * np.arange(A)
    sx = ref.round_half_away((nx - 1) * np.cos(4 * th) * np.cos(th))
    sy = ref.round_half_away((nx - 1) * np.cos(4 * th) * np.sin(th))
    pos = np.concatenate([sx + C // 2, -sy + C // 2]).astype(np.int64)
    assert np.array_equal(np.fromfile(res / "tab_posSpec.raw", "<f8"), pos.astype(float))
    # the block of genere_barre on the reconstruction grid
    g = ref.derive_geom(N, nx, V, N // 2, N // 2, NA, N0, LAM, tpcam, 1, 1, 1)
    g.rayon = rayon
    tp = ref.tomo_pixel(g)
    lo = [int(ref.round_half_away(v / tp)) for v in (-2.5e-6, -2.5e-6, -2.5e-6)]
    hi = [int(ref.round_half_away(v / tp)) for v in (2.5e-6, 2.5e-6, 2.5e-6)]
    obj = np.zeros((V, V, V))
    obj[lo[2] + V // 2:hi[2] + V // 2, lo[1] + V // 2:hi[1] + V // 2,
        lo[0] + V // 2:hi[0] + V // 2] = f32(1.395 - f32(N0))
    want = ref.spectra_to_fields(ref.born_forward(obj, pos, g, "f64"), pos, g)
    dim = f"{C}x{C}x{A}"
    got = (np.fromfile(res / f"UBornfinal_Re{dim}.raw", "<f8")
           + 1j * np.fromfile(res / f"UBornfinal_Im{dim}.raw", "<f8")).reshape(A, C, C)
    e = rel_l2(got, want)
    print(f"  simulated fields rel-L2 vs f64 {e:.3e}", flush=True)
    assert e < 1e-5
    assert (res.parent / (res.name + "centres_simul.tif")).exists()
    # tomo_reconstruction on the simulator's output
    cfg2 = tmp_path / "config2"
    cfg2.mkdir()
    (home / ".config" / "gui_tomo.conf").write_text(
        f"CHEMIN_RESULT {res}\nCHEMIN_ACQUIS {tmp_path}\nCHEMIN_CONFIG {cfg2}\n")
    (cfg2 / "config_manip.txt").write_text(
        f"NA {NA}\nN0 {N0}\nLAMBDA {LAM!r}\nF_TUBE 1\nF_OBJ 1\nTPCAM {tpcam!r}\nRF 1\n"
        f"NXMAX {nx}\nCIRCLE_CX {N // 2}\nCIRCLE_CY {N // 2}\nNB_HOLO {A}\n")
    (cfg2 / "recon.txt").write_text(
        f"CCD_ROIX {N}\nDEROUL 0\nVOLKOV 1\nC_ABER 1\nBORN 1\nNB_THREADS 4\nPREMIER_ANGLE 0\n"
        f"DIM_FINAL {V}\nEXPORT_OTF 0\n")
    r = subprocess.run([REC], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    gr = ref.derive_geom(N, nx, V, N // 2, N // 2, NA, N0, LAM, tpcam, 1, 1, 1)
    idx32, _, _, _ = ref.reconstruct(got.astype(np.complex64), pos, gr, "f32")
    ire = read_tiff_f32(res / "indice.tif")
    iim = read_tiff_f32(res / "absorption.tif")
    assert rel_l2(ire + 1j * iim, idx32) < 1e-5
    # and the reconstruction sees the block
    c = float((ire * obj).sum() / (np.linalg.norm(ire) * np.linalg.norm(obj)))
    print(f"  reconstructed index vs block: correlation {c:.3f}", flush=True)
    assert c > 0.5


# The two configurations the reference ships, key by key (values only):
# Tomo_config/{config_manip,recon}.txt and
# Tomo_prog/Simul_holo/config/Simul_holo/{config_manip,recon}.txt.
SHIPPED = {
    "Tomo_config": dict(
        manip=dict(NA=1.40, N0=1.515, LAMBDA=6.32e-7, F_TUBE=0.15, F_OBJ=0.0020,
                   TPCAM=5.5e-6, RF=0.75, NXMAX=125, CIRCLE_CX=247, CIRCLE_CY=777),
        recon=dict(CCD_ROIX=1024, DEROUL=1, VOLKOV=1, C_ABER=1, BORN=0, NB_THREADS=4,
                   PREMIER_ANGLE=0, DIM_FINAL=512, EXPORT_OTF=1)),
    "Simul_holo": dict(
        manip=dict(NA=1.40, N0=1.515, LAMBDA=6.33e-7, F_TUBE=0.15, F_OBJ=0.0018,
                   TPCAM=5.5e-6, RF=0.7, NXMAX=61, CIRCLE_CX=773, CIRCLE_CY=210),
        recon=dict(CCD_ROIX=1024, DEROUL=1, VOLKOV=1, C_ABER=1, BORN=0, NB_THREADS=3,
                   PREMIER_ANGLE=1, DIM_FINAL=416, EXPORT_OTF=1)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name,exe", [("Tomo_config", "tomo_pretraitement"),
                                      ("Simul_holo", "tomo_pretraitement_GPU")])
def test_shipped_configurations(tmp_path, name, exe):
    """The reference's shipped configurations unchanged (crop 250 with a
    512^3 volume; crop 122, a Bluestein length, with DIM_FINAL 416; both
    DEROUL 1 / VOLKOV 1 / BORN 0 / C_ABER 1) through the drop-in programs on
    seeded synthetic holograms: speculars bit-exact; the index volume (at 400
    sampled voxels of the oracle's exact sum) within rel-L2 1e-5 of the FP32
    oracle; the Rytov fields within 1e-5 of the FP32 oracle wherever the
    phase is defined. Under the ArrayFire variant's name the oracle runs
    with that variant's windows (alpha 0.025, the (1+i) crop window), and its
    two extra outputs are checked: phasePI_atan2.raw against atan2 of the
    oracle's demodulated fields (wherever the phase is defined), and
    UBornAmpFinal.raw against exp(Re) of the last field (Rytov). The phase, the Rytov field's imaginary part, is
    ill-conditioned where the demodulated field nearly vanishes (the phase
    singularities of a strongly scattering object): an error du moves it by
    du/|u|. On these inputs over 99% of the squared difference from the FP64
    oracle, for the FP32 oracle as for the GPU, sits on the 0.7% of pixels
    whose |u| is below 1% of the median, which the field comparison leaves
    out (the index volume keeps every pixel). The simulator's configuration
    runs through the GPU variant's installed name, whose outputs equal the
    CPU program's."""
    import lengths as L
    from mtd_transmission_amd import synth
    cfgv = SHIPPED[name]
    mp, rc = cfgv["manip"], cfgv["recon"]
    n, nx, V = rc["CCD_ROIX"], mp["NXMAX"], rc["DIM_FINAL"]
    C = 2 * nx
    A = 5
    nb_holo = A + rc["PREMIER_ANGLE"]  # NbAngle = NB_HOLO - PREMIER_ANGLE
    print(f"shipped configuration {name}: N={n} crop={C} V={V} A={A} via {exe}", flush=True)
    go = ref.derive_geom(n, nx, V, mp["CIRCLE_CX"], mp["CIRCLE_CY"], mp["NA"], mp["N0"],
                         mp["LAMBDA"], mp["TPCAM"], mp["F_TUBE"], mp["F_OBJ"], mp["RF"],
                         born=rc["BORN"], c_aber=rc["C_ABER"], deroul=rc["DEROUL"])
    # holograms whose sideband fills the crop: the generator's pixel is
    # NXMAX lambda / (NA N) (the configuration's optics only scale the result)
    pix = nx * mp["LAMBDA"] / (mp["NA"] * n)
    ph = synth.scale_phantom(synth.make_phantom(61, 3), 0.45 * n * pix / 33.8e-6)
    kx, ky = synth.scan_fermat(A, nx, 0.9)
    h, mask = synth.simulate_torch(n, nx, mp["CIRCLE_CX"], mp["CIRCLE_CY"], pix, mp["LAMBDA"],
                                   mp["N0"], kx, ky, ph, seed=62)
    holos = h.cpu().numpy()
    home, acq, res, cfg = (tmp_path / d for d in ("home", "acquis", "result", "config"))
    for d in (home / ".config", acq, res, cfg):
        d.mkdir(parents=True)
    (home / ".config" / "gui_tomo.conf").write_text(
        f"CHEMIN_RESULT {res}\nCHEMIN_ACQUIS {acq}\nCHEMIN_CONFIG {cfg}\n"
        f"CHEMIN_CONFIG_PC_ACQUIS {cfg}\n")
    (cfg / "config_manip.txt").write_text(
        "".join(f"{k} {v!r}\n" for k, v in mp.items()) + f"NB_HOLO {nb_holo}\n")
    (cfg / "recon.txt").write_text("".join(f"{k} {v!r}\n" for k, v in rc.items()))
    for k in range(A):
        write_pgm16(acq / ("i%03d.pgm" % k), holos[k])
    write_pgm8(acq / "Image_mask.pgm", mask)
    env = dict(os.environ, HOME=str(home))
    r = subprocess.run([os.path.join(m.BIN_DIR, exe)], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    variant = exe == "tomo_pretraitement_GPU"
    if variant:  # pretraitement_hors_axe_GPU/main.cpp:55-60
        import dataclasses
        go = dataclasses.replace(go, alpha_holo=0.025, alpha_crop=0.025)
    f32, p32, c32 = ref.preprocess(holos, go, mask, "f32", return_crops=True)
    assert np.array_equal(np.fromfile(res / "tab_posSpec.raw", "<f8"), p32.astype(float))
    dim = f"{C}x{C}x{A}"
    got = (np.fromfile(res / f"UBornfinal_Re{dim}.raw", "<f8")
           + 1j * np.fromfile(res / f"UBornfinal_Im{dim}.raw", "<f8")).reshape(A, C, C)
    # |u| of the FP64 oracle's demodulated fields: where the phase is defined
    _, p64, crops64 = ref.preprocess(holos, go, mask, "f64", return_crops=True)
    au = np.abs(np.stack([ref.demod(crops64[a], int(p64[a]), int(p64[a + A]), "f64")
                          for a in range(A)]))
    defined = au >= 0.01 * np.median(au)
    ef = rel_l2(got[defined], f32[defined])
    print(f"  phase defined at {defined.mean():.4f} of the pixels", flush=True)
    if variant:
        u32 = np.stack([ref.demod(c32[a], int(p32[a]), int(p32[a + A]), "f32")
                        for a in range(A)])
        ph = np.fromfile(res / "phasePI_atan2.raw", "<f4").reshape(A, C, C)
        dph = np.angle(np.exp(1j * (ph - np.angle(u32))))  # difference mod 2 pi
        eph = float(np.abs(dph[defined]).max())
        amp = np.fromfile(res / "UBornAmpFinal.raw", "<f4").reshape(C, C)
        ea = rel_l2(amp, np.exp(f32[-1].real))
        print(f"  phasePI_atan2 max |diff| {eph:.2e} rad; UBornAmpFinal rel-L2 {ea:.2e}",
              flush=True)
        assert eph < 1e-4 and ea < 1e-5
    r = subprocess.run([REC], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ire = read_tiff_f32(res / "indice.tif")
    iim = read_tiff_f32(res / "absorption.tif")
    assert ire.shape == (V, V, V)
    idx = np.random.default_rng(5).integers(0, V ** 3, size=400)
    want = L.index_at(got.astype(np.complex64), p32, go, "f32", idx)
    ei = rel_l2(ire.ravel()[idx] + 1j * iim.ravel()[idx], want)
    print(f"  fields rel-L2 vs f32 {ef:.3e}; index at 400 voxels {ei:.3e}", flush=True)
    assert ef < 1e-5 and ei < 1e-5
    if variant:  # the CPU name with -arrayfire gives the same fields
        res2 = tmp_path / "result_cpu_name"
        res2.mkdir()
        (home / ".config" / "gui_tomo.conf").write_text(
            f"CHEMIN_RESULT {res2}\nCHEMIN_ACQUIS {acq}\nCHEMIN_CONFIG {cfg}\n"
            f"CHEMIN_CONFIG_PC_ACQUIS {cfg}\n")
        r = subprocess.run([PRE, "-arrayfire"], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        for f in (f"UBornfinal_Re{dim}.raw", f"UBornfinal_Im{dim}.raw", "tab_posSpec.raw",
                  "parametres.raw"):
            assert (res / f).read_bytes() == (res2 / f).read_bytes(), f
