"""Input files of the reference's model drivers (InputGroup,
src/util/input.cc:98-331; RealTensorRundmrg's constructor and Sweep,
real_tensor_rundmrg.cc:3-113).  The files under tests/golden/ are written in
the reference's format; the expected values below are what they say."""
import os

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import sweep_table
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SPINLESS = os.path.join(GOLDEN, "input_spinless_small")
SPINFUL = os.path.join(GOLDEN, "input_spinful_small")


def test_run_config_and_sweep_fill_in():
    c = dm.input_run_config(SPINLESS)
    assert (c["first_block"], c["first_dim"]) == (20, 4)
    assert not c["disk_cache"] and not c["cache_resume"]
    # 3 rows for 6 half sweeps: the rest repeats the last row with maxm taken
    # from its maxb column (real_tensor_rundmrg.cc:94)
    rows = [(60, 400, 10, 1e-8, 2.0), (60, 400, 10, 1e-8, 2.5), (80, 300, 8, 1e-9, -1.0)]
    assert c["table"] == rows + [(300, 300, 8, 1e-9, -1.0)] * 3
    o = sweep_table(rows, 3)
    assert [r[:4] for r in o] == [r[:4] for r in c["table"]]
    assert np.allclose([r[4] for r in o], [0.1 ** r[4] for r in c["table"]])


def test_zero_row_ends_the_table():
    c = dm.input_run_config(SPINFUL)
    assert c["disk_cache"]
    # rows after the 0 row are ignored; fill-in from the last row read
    assert c["table"] == ([(100, 400, 10, 1e-8, 2.0)] * 2 + [(100, 400, 10, 1e-8, 3.0)] * 2
                          + [(400, 400, 10, 1e-8, 3.0)] * 2)


def test_model_from_input_equals_direct_model():
    # keys inside another group are not seen; period_y = 1 reads as true
    m = dm.InputModel(SPINLESS, "real_spinless_bose_square")
    ref = dm.SpinlessModel(3, 2, 6, 3, hop_x=-1.0, hop_y=-0.75, U=2.0, mu=0.1, period_y=True)
    o = SpinlessBoseSquare(3, 2, 6, 3, jx=-1.0, jy=-0.75, U=2.0, mu=0.1, period_y=True)
    W, _ = o.mpo()
    assert (m.L, m.d, m.nq, m.complex) == (6, 3, 1, False)
    for s in range(m.L):
        assert np.array_equal(m.mpo_site(s), ref.mpo_site(s))
        assert np.array_equal(m.mpo_site(s), W[s])
    mz = dm.InputModel(SPINLESS, "complex_spinless_bose_square")
    assert mz.complex and np.array_equal(mz.mpo_site(2).real, m.mpo_site(2))


def test_spinful_model_from_input():
    m = dm.InputModel(SPINFUL, "complex_spinful_bose_square")
    o = SpinfulBoseSquare(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
    W, qt = o.mpo()
    assert (m.L, m.d, m.nq, m.complex) == (8, 3, 2, True)
    for s in range(m.L):
        assert np.array_equal(m.mpo_site(s), W[s])
    for b in range(m.L + 1):
        assert [tuple(q) for q in m.space_table(b)] == list(o.space_tables()[b])


def test_missing_key_and_bad_driver(tmp_path):
    p = tmp_path / "input_missing"
    txt = open(SPINLESS).read().replace("    U = 2\n", "")
    p.write_text(txt)
    with pytest.raises(dm.DmrgError, match="there is no item: U"):
        dm.InputModel(str(p), "real_spinless_bose_square")
    with pytest.raises(dm.DmrgError, match="unknown model driver"):
        dm.InputModel(SPINLESS, "no_such_model")
    with pytest.raises(dm.DmrgError, match="can't open input file"):
        dm.InputModel(str(tmp_path / "nope"), "real_spinless_bose_square")


def test_default_physics_dim_and_missing_table(tmp_path):
    p = tmp_path / "input_default"
    txt = open(SPINLESS).read().replace("    physics_dim = 3\n", "")
    p.write_text(txt)
    assert dm.InputModel(str(p), "real_spinless_bose_square").d == 2   # getInt("physics_dim", 2)
    q = tmp_path / "input_notable"
    q.write_text(txt.replace("    sweeps\n", "    tables\n"))
    with pytest.raises(dm.DmrgError, match="Couldn't find table sweeps"):
        dm.input_run_config(str(q))


def test_echo_lines(capfd, tmp_path):
    p = tmp_path / "input_default"
    p.write_text(open(SPINLESS).read().replace("    physics_dim = 3\n", ""))
    dm.InputModel(str(p), "real_spinless_bose_square", echo=True)
    dm.input_run_config(str(p), echo=True)
    out = capfd.readouterr().out.splitlines()
    assert "Got input.num_site_x = 3" in out
    assert "Def input.physics_dim = 2" in out
    assert "Got input.Jy = -0.75" in out
    assert "Got input.period_y = 1" in out
    assert "No  input.resume_name = None" in out
    assert "Got input.num_sweep = 3" in out


RUN = """    disk_cache = false
    cache_name = c
    cache_record = false
    record_name = r
    cache_resume = false
    resume_name = q
    record_process = false
    process_name = p
    first_block = 20
    first_dim = 4.9
    num_sweep = 2
"""


def test_grammar_of_keys(tmp_path):
    # keys are line heads of the group body; the first head equal to the key
    # decides; nested groups and other top-level groups are not looked into
    p = tmp_path / "input_grammar"
    p.write_text("other { a = wrong\n}\n"
                 "input { a = hello\n"
                 "    b =   \n"
                 "    c hello\n"
                 "    c = later\n"
                 "    inner\n    {\n        d = 1\n    }\n"
                 "    e =\n    next\n"
                 "    f=g h\n"
                 "}\n"
                 "z = 9\n")
    s = lambda k: dm.input_string(str(p), k)
    assert s("a") == "hello"         # on the line of the group's bracket
    assert s("b") == ""              # empty value: found, empty
    assert s("c") is None            # no '=' after the first "c": absent
    assert s("d") is None            # inside a nested group
    assert s("e") == ""              # nothing after '=' on its line
    assert s("f") == "g"             # one whitespace-delimited word
    assert s("z") is None            # past the group's closing bracket


def test_file_without_final_newline_and_missing_keys(tmp_path):
    p = tmp_path / "input_nonl"
    p.write_text("input\n{\n    x = 1\n    y = two")
    assert dm.input_string(str(p), "y") == "two"
    assert dm.input_string(str(p), "nope") is None


def test_int_values_and_partial_table_rows(tmp_path):
    # ints go through the stream extractor ("4.9" reads 4); a row the
    # extractor fails in keeps zeros for the fields it did not reach, the
    # next row ends the table, and the fill-in repeats the last row
    p = tmp_path / "input_partial"
    p.write_text("input\n{\n" + RUN + "    sweeps\n    {\n        maxm maxb niter canpre noise\n"
                 "        60 400 10 1E-8 2\n        70 300\n    }\n}\n")
    c = dm.input_run_config(str(p))
    assert c["first_dim"] == 4
    assert c["table"] == [(60, 400, 10, 1e-8, 2.0), (70, 300, 0, 0.0, 0.0), (300, 300, 0, 0.0, 0.0),
                          (300, 300, 0, 0.0, 0.0)]


def test_nested_group_search_runs_past_the_parent(tmp_path):
    # a nested group is sought line by line from the parent's body start,
    # and the scan does not stop at the parent's closing bracket: a table
    # after the group is found
    p = tmp_path / "input_after"
    p.write_text("input\n{\n" + RUN + "}\nsweeps\n{\n    maxm maxb niter canpre noise\n"
                 "    50 100 5 1E-7 1\n}\n")
    c = dm.input_run_config(str(p))
    assert c["table"] == [(50, 100, 5, 1e-7, 1.0)] + [(100, 100, 5, 1e-7, 1.0)] * 3
    q = tmp_path / "input_badgroup"
    q.write_text("input = 3\n")
    with pytest.raises(dm.DmrgError, match="bracket does not follow group name"):
        dm.input_string(str(q), "x")
    r = tmp_path / "input_empty_table"
    r.write_text("input\n{\n" + RUN + "    sweeps\n    {\n        maxm\n        0 0 0 0 0\n    }\n}\n")
    with pytest.raises(dm.DmrgError, match="empty sweep table"):
        dm.input_run_config(str(r))
