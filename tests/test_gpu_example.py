"""GPU: the reference's example (examples/cnt_8-0, copied to tests/golden/cnt_8-0) run end
to end on the engine: iteration 0 of the one-dimensional driver over its 2000 energies
(liteqt_amd/example_rgf.py), writing the driver's output files."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cnt_8-0")


def _cols(path):
    rows = [ln for ln in open(path).read().split("\n") if ln.strip()]
    from liteqt_amd.writers import parse_fortran_e
    return np.array([[parse_fortran_e(ln[k:k + 18]) for k in range(0, len(ln), 18)] for ln in rows])


def test_cnt_example_iteration_zero(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import example_rgf
    r = example_rgf.run(EX, str(tmp_path))
    nen, length, ns = 2000, 18, 4
    assert len(r["en"]) == nen and np.isclose(r["en"][0], -15.0 + 20.0 / nen)
    for name in ("gw_ldos0000.dat", "gw_ndos0000.dat", "gw_pdos0000.dat"):
        text = open(tmp_path / name).read()
        assert text.count("\n") == nen * (length * ns + 1)
        c = _cols(tmp_path / name)
        assert c.shape == (nen * length * ns, 4)
        # -2 Im G^r, Im G^<, -Im G^> are spectral weights: >= 0 up to rounding (eta_dev = 0
        # in the reference's parameters, so inside a gap they are rounding noise)
        assert c[:, 3].min() >= -1e-9 * c[:, 3].max(), (name, c[:, 3].min())
        assert np.array_equal(np.unique(c[:, 0]).size, length * ns)
    tr = _cols(tmp_path / "gw_trR_0000.dat")
    tre = _cols(tmp_path / "gw_trL_0000.dat")
    assert tr.shape == (nen, 2) and tre.shape == (nen, 2)
    assert np.allclose(tr[:, 1], 2.0 * r["tr"], rtol=1e-3, atol=1e-30)
    assert np.allclose(tre[:, 1], 2.0 * r["tre"], rtol=1e-3, atol=1e-30)
    assert np.all(np.isfinite(r["tr"])) and r["tr"].max() > 1.0  # the tube conducts in its bands
    j = open(tmp_path / "gw_Jdens_0000.dat").read()
    assert j.count("\n") == nen * (length * ns + 1)
    lines = open(tmp_path / "gw_Id_iteration.dat").read().splitlines()
    assert len(lines) == 1 and len(lines[0]) == 4 + 32
