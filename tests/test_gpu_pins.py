"""The hand-derived scenarios of test_oracle_pins.py on the device: the map
kernel's vote table (first-min replacement when full, the count <= 1 guard),
KEEP ties, the vote_fast extra pass and the FW/RC merge give the records
worked out by hand, strand records included."""

import numpy as np
import pytest

import mialign
from mialign._lib import default_params

import test_oracle_pins as pins

pytestmark = pytest.mark.gpu


def _rec(d, i=0):
    return tuple(int(d[f][i]) for f in ("cid", "pos", "flag", "nm", "mapq", "sl", "sr"))


@pytest.mark.parametrize("name", list(pins.SCENARIOS))
def test_gpu_hand_derived_scenario(name):
    print(f"test_gpu_hand_derived_scenario[{name}]", flush=True)
    contigs, cases, _ = pins.SCENARIOS[name]()
    seq = np.frombuffer(b"".join(bytes(c) for c in contigs), np.uint8).copy()
    off = np.zeros(len(contigs) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in contigs])
    gix = mialign.GpuIndex.build(seq, off, 21)
    try:
        for read, kw, want, want_fw, want_rc in cases:
            b = np.frombuffer(read, np.uint8).reshape(1, -1)
            mg, fw, rc = mialign.map_batch(gix, b, np.array([len(read)], np.uint32), len(read),
                                           params=default_params(k=21, **kw), strands=True)
            assert _rec(mg) == want, (name, _rec(fw), _rec(rc))
            if want_fw is not None:
                assert _rec(fw) == want_fw
            if want_rc is not None:
                assert _rec(rc) == want_rc
    finally:
        gix.close()
