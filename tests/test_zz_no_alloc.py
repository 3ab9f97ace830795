"""ppm_batch_forward_noalloc (the library's scratch array instead of a caller workspace) against
ppm_batch_forward on the same batch, through the C ABI: FEs, probe positions, df and the status
(error word, per-molecule window flags) bit for bit, on two calls in a row with the second
batch smaller than the first, for both schedules; window flags also when probes leave their
windows.  Arguments are rejected before any device work (nmol above the maximum, reset 0)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spammm_amd import _lib                    # noqa: E402
from spammm_amd import batch as B              # noqa: E402
from spammm_amd import workloads as W          # noqa: E402


def make_batch(idx, **kw):
    return B.BatchForward([W.small_batch_molecule(i) for i in idx], **kw)


def run(bf, cfg, noalloc):
    bf.FEs.zero_()
    bf.PPs.zero_()
    bf.df.zero_()
    bf.launch(cfg, noalloc=noalloc)
    torch.cuda.synchronize()
    err, flags = bf.status()
    return bf.FEs.clone(), bf.PPs.clone(), bf.df.clone(), err, flags


@pytest.mark.parametrize("cfg", [B.LaunchConfig(8, 8, 2, "fused"), B.LaunchConfig(4, 8, 3, "two_phase")])
def test_noalloc_equals_workspace_entry_point_twice_shrinking(cfg):
    for idx in ((0, 1, 2, 3, 4, 5, 6), (7, 8, 9)):     # the second call: a smaller batch
        bf = make_batch(idx)
        print(f"[noalloc] {cfg.key()} batch of {len(idx)}", flush=True)
        ref = run(bf, cfg, noalloc=False)
        got = run(bf, cfg, noalloc=True)
        for a, b in zip(ref[:3], got[:3]):
            assert torch.equal(a, b)
        assert ref[3] == got[3] == 0
        np.testing.assert_array_equal(ref[4], got[4])
        assert not got[4].any()


def test_noalloc_reports_window_flags():
    bf = make_batch((0, 1, 2, 3), z_margin=-0.5)      # windows thinner than the probe range
    cfg = B.LaunchConfig()
    ref = run(bf, cfg, noalloc=False)
    got = run(bf, cfg, noalloc=True)
    assert ref[4].any()
    np.testing.assert_array_equal(ref[4], got[4])
    for a, b in zip(ref[:3], got[:3]):
        assert torch.equal(a, b)


def test_noalloc_rejects_bad_arguments_before_device_work():
    L = _lib.load()
    bf = make_batch((0, 1))
    cfg = B.LaunchConfig()
    bf._set_items(cfg)
    P = bf._params(cfg, B.MODES["fused"])
    args = (bf.descs_dev.data_ptr(), ctypes.addressof(P), bf.atoms.data_ptr(), bf.cpar.data_ptr(),
            bf.grids.data_ptr(), bf.FEs.data_ptr(), None, None, None, None, bf.status_buf.data_ptr(), 4)
    assert L.ppm_batch_forward_noalloc(*args, 0, None) == _lib.PPM_ERR_INVALID_ARG
    assert b"reset=0" in L.ppm_last_error()
    P.nmol = _lib.PPM_BATCH_NOALLOC_MAX_MOL + 1
    assert L.ppm_batch_forward_noalloc(*args, 1, None) == _lib.PPM_ERR_INVALID_ARG
    assert b"allocation-free maximum" in L.ppm_last_error()
