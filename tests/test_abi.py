"""The C ABI library without a GPU: it loads, exports every function include/*.h declares, and its
host-side checks (description validation, footprint arithmetic) behave; no compute is called."""
import ctypes
import os
import re

import pytest

from braingrid_amd import bghip as B
from braingrid_amd import model as MD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bg(?:hip|host)_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["bghip.h", "bghost.h"])
def test_library_exports_every_declared_function(header):
    lib = ctypes.CDLL(B.LIB_PATH)
    names = _declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_python_binding_covers_the_header():
    from braingrid_amd import host
    assert set(_declared("bghip.h")) == set(B.EXPORTED)
    assert set(_declared("bghost.h")) == set(host.EXPORTED)


def _desc(**kw):
    d = MD.net_desc(N=kw.get("N", 100), M=kw.get("M", 10), neuron_model=0, synapse_model=1, max_spikes=200)
    for k, v in kw.items():
        if k not in ("N", "M"):
            setattr(d, k, v)
    return d


def test_footprint_scales_with_slots_and_rejects_bad_descriptions():
    f1 = B.footprint(_desc(N=1000, M=100))
    f2 = B.footprint(_desc(N=1000, M=200))
    assert f1 > 0 and f2 > f1
    # per extra slot: meta, phase, psr, W, ring word, DS r/u/U/D/F/lastSpike, incoming and outgoing maps
    per_slot = (f2 - f1) / (1000 * 100)
    assert 50 <= per_slot <= 80
    assert B.footprint(_desc(ring_bits=48)) == -1
    assert B.footprint(_desc(N=0)) == -1
    assert B.footprint(_desc(N=300000)) == -1


def test_synapse_type_table_matches_create_synapse():
    t = MD.synapse_types(1e-4)
    assert list(t.total_delay) == [9, 9, 9, 16]
    import math
    for k, tau in enumerate((6e-3, 6e-3, 3e-3, 3e-3)):
        assert abs(t.decay[k] - math.exp(-1e-4 / tau)) < 1e-7


def test_context_creation_fails_cleanly_without_gpu():
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(B.BgHipError):
        B.Context(0)
