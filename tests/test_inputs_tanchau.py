"""The real-case loader on the reference's TanChau inputs (Inputs/TanChau):
the transforms of Global_Variables.py:27-133 give h.shape == (467, 327) and
N, M == 464, 324, as the reference's recorded notebook run printed
(Main_Notebook.ipynb cells 1 and 6); the 335 x 465 viscosity map is taken as
the reference's device reads it; and the case initialises and steps on the
CPU checker.  Skipped where the reference checkout is absent."""
import os

import numpy as np
import pytest

import oracle
from hydist_amd import inputs
from hydist_amd.case import BND_SERIES, params_from_dict, parse_case_header

REF = "/root/reference"
TC = os.path.join(REF, "Inputs", "TanChau")
pytestmark = pytest.mark.skipif(not os.path.isdir(TC), reason="reference inputs not present")

FILES = dict(topography="dosauTanChau.txt", visco_map="bandohsnhotroiA.txt", roughness_map="bandohsnham.txt",
             friction_map="bandoFw.txt", boundary_type="boundary_type.txt", up="bientren_Q_TC.txt",
             down="bienduoi.txt", left="bientrai.txt", right="bienphai_Z_TC.txt", CC_u="FS_bientren.txt",
             CC_d="FS_bienduoi.txt", CC_l="FS_bientrai.txt", CC_r="FS_bienphai.txt")


@pytest.fixture(scope="module")
def tanchau(tmp_path_factory):
    d = tmp_path_factory.mktemp("tanchau")
    dirs = d / "DirsOfInputs.txt"
    dirs.write_text("".join(f"{k} {os.path.join(TC, v)}\n" for k, v in FILES.items()))
    header = open(os.path.join(REF, "constant_TanChau.cuh")).read()
    with pytest.warns(UserWarning, match="visco_map has shape 465x335"):
        case = inputs.load_case(str(dirs), header, name="TanChau")
    return case, str(dirs), header


def test_shapes_match_the_recorded_run(tanchau):
    case, _, _ = tanchau
    assert case.arrays["h"].shape == (467, 327)
    assert (case.N, case.M) == (464, 324)
    for k in ("hsnham", "VISCOINDX", "Fw"):
        assert case.arrays[k].shape == (467, 327), k


def test_transform_and_edge_padding(tanchau):
    case, _, _ = tanchau
    raw = np.loadtxt(os.path.join(TC, FILES["topography"]))
    t = np.flip(raw.T, 1)
    h = case.arrays["h"]
    assert np.array_equal(h[1:-2, 1:-2], t)
    assert np.array_equal(h[0], h[1]) and np.array_equal(h[-1], h[-3]) and np.array_equal(h[:, -1], h[:, -3])


def test_viscosity_map_read_as_the_device_reads_it(tanchau):
    case, dirs, header = tanchau
    raw = np.loadtxt(os.path.join(TC, FILES["visco_map"]))
    padded = np.pad(np.flip(raw.T, 1), ((1, 2), (1, 2)), "edge")
    assert padded.shape == (468, 338)
    want = padded.reshape(-1)[:467 * 327].reshape(467, 327)
    assert np.array_equal(case.arrays["VISCOINDX"], want)
    with pytest.raises(ValueError, match="visco_map has shape"):
        inputs.load_case(dirs, header, mismatched_maps="error")


def test_boundaries_and_series(tanchau):
    case, _, _ = tanchau
    bt = case.arrays["boundary_type"]
    assert list(bt[:, 0]) == [2, 0, 0, 1]             # up Q, right Z (boundary_type.txt)
    assert list(case.arrays["bienQ"]) == [1, 0, 0, 0]
    # len(bc_up) / one Q segment (Global_Variables.py:54); the notebook's
    # printed 8763 (Main_Notebook.ipynb cell 0) is of Inputs/TanChau2, whose
    # list DirsOfInputs_1.txt names but which the reference does not ship
    assert case.total_time == 4246
    assert case.arrays["bc_up"][:3].tolist() == [-6910.0, -4910.0, -3520.0]


def test_case_header_and_checker_steps(tanchau):
    """The TanChau header (segment_limit 10, dX = dY = 10, dT = 2) and a short
    run on the CPU checker: the wet area of the map and finite fields."""
    case, _, header = tanchau
    p = params_from_dict(parse_case_header(header))
    assert (p.segment_limit, p.dX, p.dY, p.dT) == (10, 10.0, 10.0, 2.0)
    case.boundary_mode = BND_SERIES
    o = oracle.Oracle(case)
    o.init()
    wet = o.wet_count()
    frac = wet / (case.N * case.M)
    assert 0.2 < frac < 0.6, frac
    o.run(0.0, 20, case.opts())
    for k in ("u", "v", "z"):
        assert np.isfinite(o.get(k)).all(), k
    assert np.abs(o.get("z")).max() < 5.0
