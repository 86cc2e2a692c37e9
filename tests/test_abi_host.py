"""CPU tests of the C ABI's host side: the library loads without a GPU and
exports every entry point of include/mtd.h; the C++ config parser, geometry
derivation and Tukey table match the same known answers as the oracle."""
import os
import re

import numpy as np
import pytest

import mtd_transmission_amd as m
from oracle import ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mtd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mtd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = m.lib()  # loads torch's HIP runtime first, as every user of the binding does
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(syms) <= set(m.SIGNATURES), set(syms) - set(m.SIGNATURES)


def test_version_and_null_context():
    assert m.lib().mtd_version().startswith(b"mtd-mi355x")
    assert m.lib().mtd_last_error(None) == b"null context"
    assert m.lib().mtd_ctx_sync(None) != 0


CFG = (
    "# comment line\n"
    "NA 1.20\n"
    "N0 1.333\n"
    "LAMBDA 0.000000632\n"
    "NXMAX 125\n"
    "NXMAX 64\n"
    "#NB_HOLO 999\n"
    "EMPTYKEY \n"
    "CRLF 2.5\r\n"
    "TEXT /tmp/some path/\n"
    "NOSPACE\n"
    "\n"
)


def test_cpp_config_parser_known_answers(tmp_path):
    p = tmp_path / "c.txt"
    p.write_bytes(CFG.encode())
    f32 = lambda s: float(np.float32(float(s)))
    cases = [("NA", None, f32("1.20"), True), ("N0", None, f32("1.333"), True),
             ("LAMBDA", None, f32("6.32e-7"), True), ("NXMAX", None, 64.0, True),
             ("NB_HOLO", None, 0.0, False), ("EMPTYKEY", None, 0.0, False),
             ("EMPTYKEY", 1.0, 1.0, False), ("COEF_NA_OBJ_LIMIT", 1.0, 1.0, False),
             ("CRLF", None, 2.5, True), ("NOSPACE", None, 0.0, True)]
    for key, dflt, want, found in cases:
        v, f = m.extract_val(p, key, dflt)
        assert v == want, key
        assert f == found, key
        assert ref.extract_val(p, key, dflt) == want, key
    assert m.extract_string(p, "TEXT") == "/tmp/some path/"
    assert m.extract_string(p, "MISSING") == ""
    with pytest.raises(m.MTDError):
        m.extract_val(tmp_path / "absent.txt", "NA")


def test_cpp_geom_derive_shipped_config():
    g = m.geom_derive(1024, 125, 512, 247, 777, 1.40, 1.515, 0.000000632, 0.0000055,
                      0.15, 0.0020, 0.75)
    assert g.rayon == 135.0
    assert g.pix_holo == pytest.approx(5.5e-8, rel=1e-6)
    assert g.n0 == float(np.float32(1.515))
    o = ref.derive_geom(1024, 125, 512, 247, 777, 1.40, 1.515, 0.000000632, 0.0000055,
                        0.15, 0.0020, 0.75)
    for k in ("n", "nxmax", "dim_final", "circle_cx", "circle_cy", "n0", "na",
              "pix_holo", "rayon", "pix_uborn", "delta_f", "coef_na"):
        assert getattr(g, k) == getattr(o, k), k
    assert g.lambda_ == o.lam


def test_cpp_tukey_matches_known_answers():
    t = np.array(m.tukey1d(1024, 0.05))
    to, b1, b2 = ref.tukey1d(1024, 0.05)
    assert np.array_equal(t, to)
    assert np.all(t[b1 + 1:b2 + 1] == 1.0) and t[b1] < 1.0 and t[b2 + 1] < 1.0
    assert (b1, b2) == (26, 997)
    t = np.array(m.tukey1d(256, 0.1))
    assert np.array_equal(t, ref.tukey1d(256, 0.1)[0])


def test_context_requires_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(m.MTDError, match="no GPU"):
        m.Context(0)


def test_synthetic_cases_are_seeded_and_well_posed():
    from cases import make_case
    a = make_case("small")
    from mtd_transmission_amd import synth
    b = synth.simulate(128, 16, a.params["cx"], a.params["cy"], a.g.pix_holo, a.g.lam,
                       a.g.n0, a.acq.kx, a.acq.ky, a.acq.phantom, seed=8)
    assert np.array_equal(a.acq.holos, b.holos)
    assert int((a.acq.mask[::5, ::5] > 45).sum()) >= 15
    fields, pos = ref.preprocess(a.acq.holos, a.g, a.acq.mask, "f64")
    A = len(a.acq.kx)
    assert np.array_equal(pos[:A] - a.g.nxmax, a.acq.kx)
    assert np.array_equal(pos[A:] - a.g.nxmax, a.acq.ky)


def test_tiff_writer_classic_and_bigtiff(tmp_path):
    """The float32 TIFF writer: small files are classic TIFF (magic 42) and
    read back exactly, for 1-, 2- and 3-row pages; BigTIFF (magic 43, 64-bit
    offsets, values of up to 8 bytes stored in the entries) reads back
    exactly too; a classic file that would pass 4 GiB (a 1024^3 float
    volume) is refused before anything is written, where the default format
    switches to BigTIFF."""
    from tiffread import read_tiff_f32
    rng = np.random.default_rng(3)
    for h in (1, 2, 3):
        a = rng.normal(size=(4, h, 5)).astype(np.float32)
        for fmt, kind in ((0, "tiff"), (1, "tiff"), (2, "bigtiff")):
            p = tmp_path / f"v{h}_{fmt}.tif"
            assert m.write_tiff_f32(p, a, 1e-7, 4, fmt) == kind
            back, k = read_tiff_f32(p, with_format=True)
            assert k == kind and np.array_equal(back, a)
    big = np.zeros((1, 1, 1), np.float32)  # never read: refused on size
    p = tmp_path / "huge.tif"
    rc = m.lib().mtd_write_tiff_f32(str(p).encode(), big.ctypes.data_as(m.ctypes.c_void_p),
                                    1024, 1024, 1024, 1e-7, 4, 1)
    assert rc == -2 and not p.exists()
    # 1023 pages: the pixels alone fit 32-bit offsets, with each page's strip
    # tables (8 KiB) the file does not
    rc = m.lib().mtd_write_tiff_f32(str(p).encode(), big.ctypes.data_as(m.ctypes.c_void_p),
                                    1024, 1024, 1023, 1e-7, 4, 1)
    assert rc == -2 and not p.exists()
