"""CPU checks of the parity harness itself: every synthetic slot of the GPU parity cases fits
the slot capacity the library enforces (getDataBytes plus four frames), and every FFT window
of every station lies inside its slot buffer, so no case is rejected or silently flagged."""
import numpy as np
import pytest

import fxoracle as fo
import fxcases as fc
from difx_amd import synth

CASES = {
    "cfg0": lambda: fc.case_cfg0(B=128), "cfg0_lsb": lambda: fc.case_cfg0(B=64, lsb=True),
    "cfg1": lambda: fc.case_cfg1(B=32), "cfg2": lambda: fc.case_cfg2(B=16),
    "cfg2_8": lambda: fc.case_cfg2(B=8, nstations=8), "cfg3": lambda: fc.case_cfg3(B=32),
    "long": lambda: fc.case_long_ffts(B=8), "fourstep": lambda: fc.case_fourstep_ffts(B=8),
    "cplx_fourstep": lambda: fc.case_complex_fourstep(B=4), "offsets": lambda: fc.case_offsets(B=32),
    "mark5b": lambda: fc.case_mark5b(B=32), "real8": lambda: fc.case_real8(B=32),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_slots_fit_and_cover(name):
    cfg = CASES[name]()
    offsets = [0, 1, 0]
    data, ctrl = fc.slot(cfg, offsets)
    B = cfg.blockspersend()
    for d, ds in enumerate(cfg.datastreams):
        cap = synth.data_bytes(cfg, d) + 4 * ds.framebytes
        assert len(data[d]) <= cap, (name, d, len(data[d]), cap)
        st = fo.process_station(cfg, d, offsets, data[d], ctrl[d], B)
        assert np.all(st.weight > 0), (name, d, np.flatnonzero(st.weight == 0)[:8])


def test_damaged_fourstep_slot_takes_both_decode_paths():
    """The damaged slot of the four-step GPU case holds FFTs with every sample valid (weight 1,
    the gathered decode) and FFTs with some invalid samples (0 < weight < 1, the per-sample
    decode), beside flagged ones (weight 0)."""
    cfg = fc.case_fourstep_ffts(B=4)
    offsets = [0, 1, 0]
    data, ctrl = fc.damaged_slot(cfg, offsets)
    B = cfg.blockspersend()
    w = np.concatenate([fo.process_station(cfg, d, offsets, data[d], ctrl[d], B).weight
                        for d in range(len(cfg.datastreams))])
    assert (w == 1).sum() > 0 and ((w > 0) & (w < 1)).sum() > 0 and (w == 0).sum() > 0
