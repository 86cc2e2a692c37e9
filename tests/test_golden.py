"""The committed golden vectors (tests/golden/tiny_plain.npz, written by
tests/golden/make_golden.py from the FP64 oracle): the oracle still
reproduces them (CPU), and the HIP path meets them (GPU: speculars
bit-exact, fields and index volume within rel-L2 1e-5)."""
import os

import numpy as np
import pytest

from cases import rel_l2
from oracle import ref

LAM, N0, NA = 6.33e-7, 1.333, 1.2
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_plain.npz")


def _load():
    d = np.load(GOLDEN)  # allow_pickle stays False
    n, nx, V, cx, cy, c_aber = (int(v) for v in d["geom"])
    return d, (n, nx, V, cx, cy, float(d["pix"]), c_aber)


def test_oracle_reproduces_golden():
    d, (n, nx, V, cx, cy, pix, c_aber) = _load()
    g = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1, c_aber=c_aber)
    fields, pos = ref.preprocess(d["holos"], g, None, "f64")
    assert np.array_equal(pos, d["posspec"])
    assert rel_l2(fields, d["fields"]) < 1e-12
    index, _, cnt, _ = ref.reconstruct(fields, pos, g, "f64")
    assert int(cnt.sum()) == int(d["counts_total"])
    assert rel_l2(index, d["index"]) < 1e-6  # stored as complex64


@pytest.mark.gpu
def test_gpu_meets_golden(gpu_ctx):
    import torch
    import mtd_transmission_amd as m
    d, (n, nx, V, cx, cy, pix, c_aber) = _load()
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix, 1.0, 1.0, 1.0, 1.0, 1, c_aber, 0)
    holos = torch.from_numpy(d["holos"]).cuda()
    A, C = holos.shape[0], 2 * nx
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    fields = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    gpu_ctx.preprocess(gl, holos, None, pos, fields)
    re = torch.empty((V, V, V), device="cuda")
    im = torch.empty_like(re)
    pos2 = torch.empty_like(pos)
    gpu_ctx.reconstruct_plain(gl, holos, None, re, im, pos2)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), d["posspec"])
    assert np.array_equal(pos2.cpu().numpy(), d["posspec"])
    ef = rel_l2(fields.cpu().numpy(), d["fields"])
    ei = rel_l2(re.cpu().numpy() + 1j * im.cpu().numpy(), d["index"])
    print(f"golden: fields rel-L2 {ef:.3e}, index rel-L2 {ei:.3e}", flush=True)
    assert ef < 1e-5 and ei < 1e-5
