"""Host-side configuration of the host model on the CPU: the stimulus file reader
(braingrid_amd/host.py read_stimulus) against the oracle's own parser, and the parameter block."""
import os

import pytest

from braingrid_amd import bghip as B
from braingrid_amd import host
from oracle import oracle_py as O
from tests import helpers as H


@pytest.mark.parametrize("name", list(H.INPUT_STIM))
def test_read_stimulus_matches_oracle_parser(name):
    path = os.path.join(H.CONFIG_DIR, H.INPUT_STIM[name])
    assert host.read_stimulus(path, H.CONFIG_DIR) == O.parse_input_xml(path, H.CONFIG_DIR)


def test_input_parameters():
    c = host.read_config(os.path.join(H.CONFIG_DIR, "input-lif-regular-async.xml"), H.CONFIG_DIR)
    c["input"] = host.read_stimulus(os.path.join(H.CONFIG_DIR, H.INPUT_STIM["input-lif-regular-async"]))
    p, keep = host.make_params(c)
    assert p.input_kind == B.INPUT_REGULAR and p.input_sync == host.INPUT_SYNC["no"]
    assert abs(p.input_duration - 0.005) < 1e-9
    assert list(p.input_values) == pytest.approx(c["input"]["values"], rel=1e-7, abs=0)
    c["input"] = dict(kind="poisson", fr_mean=10.0, weight=25.0, mask=[1, 5, 7])
    p, keep = host.make_params(c)
    assert p.input_kind == B.INPUT_POISSON and p.n_input_mask == 3 and [p.input_mask[k] for k in range(3)] == [1, 5, 7]
    c["input"] = dict(kind="poisson", fr_mean=10.0, weight=25.0, mask=None)
    assert host.make_params(c)[0].n_input_mask == -1
    c["input"] = dict(kind="regular", duration=0.01, interval=0.02, sync="sometimes", values=[0.0] * 100)
    with pytest.raises(ValueError, match="sync"):
        host.make_params(c)


def test_static_connection_parameters():
    c = host.read_config(os.path.join(H.CONFIG_DIR, "static-lif-ds.xml"), H.CONFIG_DIR)
    p, _ = host.make_params(c)
    assert c["conn"] == "ConnStatic" and p.conn_kind == host.CONN_KINDS["ConnStatic"]
    assert p.static_nconns == 14 and abs(p.static_thresh - 3.0) < 1e-7
    assert tuple(p.static_exc_w) == pytest.approx((0.0, 2e-7)) and tuple(p.static_inh_w) == pytest.approx((-1e-7, 0.0))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and a JSON message, before
    touching any device."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
