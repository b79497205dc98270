"""Reference driver input files run end to end on the device (InputGroup ->
model -> Run) against the oracle's sweeps on the same model and table."""
import os

import numpy as np
import pytest

import dmrgcc_amd as dm
from oracle.dmrg import DMRG
from oracle.model import SpinfulBoseSquare, SpinlessBoseSquare

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,driver", [("input_spinless_small", "real_spinless_bose_square"),
                                         ("input_spinless_small", "complex_spinless_bose_square"),
                                         ("input_spinful_small", "complex_spinful_bose_square")])
def test_run_input_matches_oracle(name, driver):
    path = os.path.join(GOLDEN, name)
    r = dm.run_input(path, driver)
    c = dm.input_run_config(path)
    if driver == "complex_spinful_bose_square":
        om = SpinfulBoseSquare(2, 2, 2, 2, 3, flux=0.25, U_uu=4.0, U_dd=4.0, U_ud=2.0)
    else:
        om = SpinlessBoseSquare(3, 2, 6, 3, jx=-1.0, jy=-0.75, U=2.0, mu=0.1, period_y=True,
                                complex=driver.startswith("complex"))
    nhalf = len(c["table"])
    o = DMRG(om, c["table"], nhalf // 2, first_block=c["first_block"], first_dim=c["first_dim"])
    e_ref = np.array(o.run())
    assert r["rand_draws"] == o.rng.count
    assert np.abs(r["energies"] - e_ref).max() < 1e-10 * np.abs(e_ref).max()
