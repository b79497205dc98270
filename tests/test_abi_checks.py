"""Argument checks that refuse bad calls before any device work: the Python
vector checks of HeffPlan / HeffBatch (the C ABI takes no vector lengths) and
the batch's refusal of a repeated plan.  No GPU is touched."""
import ctypes

import numpy as np
import pytest
import torch

import dmrgcc_amd as dm


def _fake_plan(words, cplx=False):
    # a plan object with no native handle: every check below must fire
    # before the library is called
    p = dm.HeffPlan.__new__(dm.HeffPlan)
    p._h = None
    p.complex = cplx
    p.words = words
    return p


def test_vector_checks_refuse_before_the_call():
    p = _fake_plan(16)
    good = torch.zeros(16, dtype=torch.float64)  # CPU: refused as not CUDA
    with pytest.raises(dm.DmrgError, match="CUDA"):
        p.apply(good, good)
    with pytest.raises(dm.DmrgError, match="dtype"):
        p.apply(torch.zeros(16, dtype=torch.float32), good)
    with pytest.raises(dm.DmrgError, match="torch tensor"):
        p.apply(np.zeros(16), good)
    with pytest.raises(dm.DmrgError, match="dtype"):
        p.lanczos(torch.zeros(8, dtype=torch.complex128), 4)  # complex vector, real plan
    with pytest.raises(dm.DmrgError, match="dtype"):
        dm.HeffBatch([p], [torch.zeros(16, dtype=torch.int64)], [good])
    with pytest.raises(dm.DmrgError, match="one x and one y"):
        dm.HeffBatch([p, p], [good], [good])


def test_vector_length_and_layout_checks():
    # the length and contiguity rules on their own (device-independent part)
    class FakeCuda(torch.Tensor):
        @property
        def is_cuda(self):
            return True

    short = torch.zeros(15, dtype=torch.float64).as_subclass(FakeCuda)
    with pytest.raises(dm.DmrgError, match="needs 16"):
        dm._vec_ptr(short, 16, False, "x")
    strided = torch.zeros(32, dtype=torch.float64)[::2].as_subclass(FakeCuda)
    with pytest.raises(dm.DmrgError, match="contiguous"):
        dm._vec_ptr(strided, 16, False, "x")
    # a complex plan takes complex128 (psi_dim entries) or 2 psi_dim float64
    z = torch.zeros(8, dtype=torch.complex128).as_subclass(FakeCuda)
    assert dm._vec_ptr(z, 16, True, "x").value == z.data_ptr()
    with pytest.raises(dm.DmrgError, match="needs 16"):
        dm._vec_ptr(torch.zeros(7, dtype=torch.complex128).as_subclass(FakeCuda), 16, True, "x")


def test_batch_refuses_a_repeated_plan():
    # The duplicate check runs before any handle is dereferenced, so opaque
    # non-null handle values are enough to exercise it on the CPU.
    lib = dm.lib()
    vp = ctypes.c_void_p
    fake = [vp(0x1000), vp(0x2000), vp(0x1000)]
    xs = (vp * 3)(vp(0x10), vp(0x20), vp(0x30))
    ys = (vp * 3)(vp(0x40), vp(0x50), vp(0x60))
    out = vp()
    rc = lib.dmrgcc_heff_batch_create(3, (vp * 3)(*fake), xs, ys, ctypes.byref(out))
    assert rc == 3
    assert "more than once" in lib.dmrgcc_last_error().decode()
    assert not out.value


def test_dropin_harness_builds_and_exports():
    # the C++ harness of the drop-in class headers is built with the library
    # and loads next to it (its calls need a GPU: tests/test_gpu_dropin.py)
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    dm.lib()
    h = ctypes.CDLL(os.path.join(here, "cxx", "libdropin_test.so"))
    for f in ("dropin_site_create", "dropin_site_env", "dropin_site_heff", "dropin_site_lanczos",
              "dropin_site_env_update", "dropin_site_canonical", "dropin_site_destroy", "dropin_run"):
        assert hasattr(h, f), f
