"""constant.cuh-format parser: C typing, IEEE evaluation as written,
single-precision cases rejected."""
import json
import os

import pytest

from hydist_amd.case import CaseHeaderError, make_case_header, params_from_dict, parse_case_header

HERE = os.path.dirname(os.path.abspath(__file__))


def test_c_typing_rules():
    d = parse_case_header("#define DOUBLE double\n__constant__ DOUBLE a = 2 / 3;\n"
                          "__constant__ DOUBLE b = 2.0 / 3;\n__constant__ int c = 7 / 2;\n"
                          "__constant__ DOUBLE e = -7 / 2;\n__constant__ DOUBLE f = 1.0 / 3600;")
    assert d["a"] == 0.0 and d["b"] == 2.0 / 3 and d["c"] == 3 and d["e"] == -3.0
    assert d["f"] == 1.0 / 3600


def test_derived_literals_kept_as_written():
    # a derived literal that does not equal the recomputed one keeps its own value
    d = parse_case_header("#define DOUBLE double\n__constant__ DOUBLE dT = 0.125;\n"
                          "__constant__ DOUBLE dX = 1.25;\n__constant__ DOUBLE dTchia2dX = 0.12500 / (2 * 1.25);\n"
                          "__constant__ DOUBLE HaiChiadT = 2.0 / 0.125;")
    assert d["dTchia2dX"] == 0.125 / 2.5 and d["HaiChiadT"] == 16.0


def test_float_case_rejected():
    with pytest.raises(CaseHeaderError):
        parse_case_header("#define DOUBLE float\n__constant__ DOUBLE dX = 1.0;")


def test_reference_headers_golden():
    """The parser against the reference's case headers as g++ evaluates them
    (constexpr of each expression as written; tests/golden/make_case_headers.py
    wrote the fixture)."""
    with open(os.path.join(HERE, "golden", "case_headers.json")) as f:
        golden = json.load(f)
    ref = "/root/reference"
    if not os.path.isdir(ref):
        pytest.skip("reference headers not present")
    for name, vals in golden.items():
        with open(os.path.join(ref, name)) as f:
            d = parse_case_header(f.read())
        for k, v in vals.items():
            assert d[k] == v, (name, k)


def test_generated_header_roundtrip():
    txt = make_case_header(20.0, 20.0, 5.0, 20, H_TINH=0.05)
    p = params_from_dict(parse_case_header(txt))
    assert p.dX == 20.0 and p.dT == 5.0 and p.dTchia2dX == 5.0 / (2 * 20.0)
    assert p.HaiChiadT == 2.0 / 5.0 and p.H_TINH == 0.05 and p.segment_limit == 20
