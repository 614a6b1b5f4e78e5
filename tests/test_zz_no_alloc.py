"""The allocation-free DFA entry point against the allocating one, through the C ABI.

``gfsm_dfa_run_device_static`` runs on a library array of GFSM_STATIC_SCRATCH_BYTES per GPU
instead of allocating scratch. Its verdicts must equal ``gfsm_dfa_run_device``'s (and the
oracle's), bit for bit, on two calls in a row where the second batch is smaller (it finds
the first call's scratch contents), for every technique; a run needing more scratch than
the array holds is refused before anything is enqueued.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gpufsm import workloads as W
from gpufsm.backends.hip import _lib
from gpufsm.backends.hip.handles import DeviceDFA
from oracle import native

pytestmark = pytest.mark.gpu


def _rare_dfa(seed):
    """256 states, one accepting state entered from 24 of the 65,536 (state, byte) pairs:
    first matches land deep inside long inputs, many inputs never match."""
    dfa = W.dense_dfa(256, 0.0, seed=seed)
    trans = np.asarray(dfa.trans).reshape(256, 256)
    trans[trans == 3] = 4
    rng = np.random.default_rng(seed)
    trans.reshape(-1)[rng.choice(256 * 256, size=24, replace=False)] = 3
    dfa.trans[:] = trans.reshape(-1)
    dfa.accept[3] = True
    return dfa


def _ragged(seed, count, max_len):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, size=count)
    data = rng.integers(0, 256, size=max(int(lens.sum()), 1), dtype=np.uint8)
    off = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return data, off


class _Dev:
    """One batch's device buffers (torch tensors as plumbing)."""

    def __init__(self, torch, data, off):
        self.count = int(off.size) - 1
        self.total = int(off[-1])
        self.data = torch.from_numpy(data).cuda()
        self.off = torch.from_numpy(off).cuda()
        self.flags = torch.full((max(self.count, 1),), -7, dtype=torch.int32, device="cuda")
        self.lens = torch.full((max(self.count, 1),), -7, dtype=torch.int32, device="cuda")

    def out(self):
        return self.flags[: self.count].cpu().numpy(), self.lens[: self.count].cpu().numpy()


def _static(lib, dev, tech, b, stream):
    return lib.gfsm_dfa_run_device_static(
        dev.handle, tech, None, ctypes.c_void_p(b.data.data_ptr()),
        ctypes.c_void_p(b.off.data_ptr()), b.count, b.total, ctypes.c_void_p(b.flags.data_ptr()),
        ctypes.c_void_p(b.lens.data_ptr()), ctypes.c_void_p(stream))


@pytest.mark.parametrize("technique", ["split", "auto", "gather", "lds"])
def test_static_scratch_matches_the_allocating_entry_point(technique):
    import torch

    lib = _lib.load()
    tech = _lib.DFA_TECHNIQUES[technique]
    dfa = _rare_dfa(41)
    stream = torch.cuda.current_stream().cuda_stream
    # first a large batch (2048 x 16 KiB plus ragged short inputs), then a smaller one
    big_d, big_o = W.uniform_batch(2048, 16384, seed=42)
    rag_d, rag_o = _ragged(43, 900, 700)
    d1 = np.concatenate([big_d, rag_d])
    o1 = np.concatenate([big_o, big_o[-1] + rag_o[1:]])
    d2, o2 = _ragged(44, 300, 5000)
    with DeviceDFA(dfa) as dev:
        for k, (d, o) in enumerate(((d1, o1), (d2, o2))):
            ref, st = _Dev(torch, d, o), _Dev(torch, d, o)
            dev.run_device(ref.data.data_ptr(), ref.off.data_ptr(), ref.count, ref.total,
                           ref.flags.data_ptr(), ref.lens.data_ptr(), stream, tech)
            rc = _static(lib, dev, tech, st, stream)
            assert rc == _lib.OK, lib.gfsm_last_error().decode()
            torch.cuda.synchronize()
            rf, rn = ref.out()
            sf, sn = st.out()
            np.testing.assert_array_equal(sf, rf, err_msg=f"call {k}: flags")
            np.testing.assert_array_equal(sn, rn, err_msg=f"call {k}: lens")
            wf, wn, _ = native.dfa_batch(dfa, d, o)
            np.testing.assert_array_equal(sf, wf, err_msg=f"call {k}: oracle flags")
            np.testing.assert_array_equal(sn, wn, err_msg=f"call {k}: oracle lens")
            assert wf.any() and not wf.all()


def test_static_scratch_refuses_runs_beyond_its_size():
    """16 GiB of input would need far more than 32 MiB of split scratch: refused on the
    host, nothing enqueued (the buffers below are never read)."""
    import torch

    lib = _lib.load()
    d, o = _ragged(45, 4, 100)
    b = _Dev(torch, d, o)
    b.total = 16 << 30
    with DeviceDFA(_rare_dfa(46)) as dev:
        rc = _static(lib, dev, _lib.DFA_SPLIT, b, torch.cuda.current_stream().cuda_stream)
        assert rc == _lib.EUNSUPPORTED
        assert b"static scratch" in lib.gfsm_last_error()
    torch.cuda.synchronize()
    assert (b.flags.cpu().numpy() == -7).all()  # untouched
