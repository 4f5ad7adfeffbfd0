"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker. It also parses BrainGrid XML simulation configs (the subset the hot path needs:
SimInfoParams, AllLIFNeurons/AllIZHNeurons ranges, the synapse class, ConnGrowth parameters,
FixedLayout/DynamicLayout) and reads the state dumps written by oracle/_ref/bgref_dump.
"""
from __future__ import annotations

import ctypes
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

LIF, IZH = 0, 1
SPIKING, DS, STDP, DYNSTDP = 0, 1, 2, 3
LAYOUT_FIXED, LAYOUT_DYNAMIC = 0, 1
CONN_NONE, CONN_GROWTH, CONN_STATIC = 0, 1, 2

SYN_CLASSES = {"AllSpikingSynapses": SPIKING, "AllDSSynapses": DS, "AllSTDPSynapses": STDP,
               "AllDynamicSTDPSynapses": DYNSTDP}
NEURON_CLASSES = {"AllLIFNeurons": LIF, "AllIZHNeurons": IZH}

_F2 = ctypes.c_float * 2


class Params(ctypes.Structure):
    """Mirror of bgo_params (oracle/bg_oracle.h)."""
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("neuron_kind", ctypes.c_int32), ("syn_kind", ctypes.c_int32),
        ("layout_kind", ctypes.c_int32), ("conn_kind", ctypes.c_int32),
        ("max_synapses_per_neuron", ctypes.c_int32), ("max_firing_rate", ctypes.c_int32),
        ("num_epochs", ctypes.c_int32), ("ring_bits", ctypes.c_int32),
        ("dt", ctypes.c_float), ("epoch_duration", ctypes.c_float),
        ("seed", ctypes.c_uint32),
        ("noise_mode", ctypes.c_int32), ("math_mode", ctypes.c_int32), ("sum_mode", ctypes.c_int32),
        ("inh_list", ctypes.POINTER(ctypes.c_int32)), ("n_inh", ctypes.c_int32),
        ("active_list", ctypes.POINTER(ctypes.c_int32)), ("n_active", ctypes.c_int32),
        ("frac_exc", ctypes.c_float), ("frac_starter", ctypes.c_float),
        ("Iinject", _F2), ("Inoise", _F2), ("Vthresh", _F2), ("Vresting", _F2), ("Vreset", _F2),
        ("Vinit", _F2), ("starter_Vthresh", _F2), ("starter_Vreset", _F2),
        ("excA", _F2), ("excB", _F2), ("excC", _F2), ("excD", _F2),
        ("inhA", _F2), ("inhB", _F2), ("inhC", _F2), ("inhD", _F2),
        ("epsilon", ctypes.c_float), ("beta", ctypes.c_float), ("rho", ctypes.c_float),
        ("target_rate", ctypes.c_float), ("min_radius", ctypes.c_float), ("start_radius", ctypes.c_float),
        ("stdp_override", ctypes.c_int32),
        ("stdp_Apos", ctypes.c_float), ("stdp_Aneg", ctypes.c_float), ("stdp_taupos", ctypes.c_float),
        ("stdp_tauneg", ctypes.c_float), ("stdp_tauspre", ctypes.c_float), ("stdp_tauspost", ctypes.c_float),
        ("stdp_gap", ctypes.c_float), ("stdp_Wex", ctypes.c_float), ("stdp_mupos", ctypes.c_float),
        ("stdp_muneg", ctypes.c_float), ("stdp_use_froemke_dan", ctypes.c_int32),
        ("delay_override", ctypes.c_float * 4),
        ("input_kind", ctypes.c_int32), ("input_fr_mean", ctypes.c_float), ("input_weight", ctypes.c_float),
        ("input_mask", ctypes.POINTER(ctypes.c_int32)), ("n_input_mask", ctypes.c_int32),
        ("input_duration", ctypes.c_float), ("input_interval", ctypes.c_float), ("input_sync", ctypes.c_int32),
        ("input_values", ctypes.c_float * 100),
        ("static_nconns", ctypes.c_int32), ("static_thresh", ctypes.c_float),
        ("static_exc_w", _F2), ("static_inh_w", _F2),
    ]


INPUT_NONE, INPUT_POISSON, INPUT_REGULAR = 0, 1, 2
SYNC = {"yes": 0, "no": 1, "wave": 2}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -f oracle/Makefile`)")
        L = ctypes.CDLL(LIB_PATH)
        L.bgo_create.restype = ctypes.c_void_p
        L.bgo_create.argtypes = [ctypes.POINTER(Params), ctypes.c_char_p, ctypes.c_int]
        L.bgo_destroy.argtypes = [ctypes.c_void_p]
        for f in ("bgo_update_connections", "bgo_update_history", "bgo_run_epoch", "bgo_rebuild_index_map"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        L.bgo_advance.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.bgo_advance.restype = ctypes.c_int
        L.bgo_update_weights.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.bgo_update_weights.restype = ctypes.c_int
        L.bgo_add_synapse.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float]
        L.bgo_add_synapse.restype = ctypes.c_int64
        L.bgo_set_radii.argtypes = [ctypes.c_void_p, ctypes.c_float]
        L.bgo_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char),
                                ctypes.POINTER(ctypes.c_int64)]
        L.bgo_array.restype = ctypes.c_void_p
        L.bgo_step.argtypes = [ctypes.c_void_p]
        L.bgo_step.restype = ctypes.c_int64
        L.bgo_total_synapses.argtypes = [ctypes.c_void_p]
        L.bgo_total_synapses.restype = ctypes.c_int64
        L.bgo_last_error.argtypes = [ctypes.c_void_p]
        L.bgo_last_error.restype = ctypes.c_char_p
        L.bgo_poisson_isi.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float,
                                      ctypes.c_float]
        L.bgo_poisson_isi.restype = ctypes.c_int32
        _lib = L
    return _lib


_DT = {b"f": np.float32, b"d": np.float64, b"i": np.int32, b"I": np.uint32, b"q": np.uint64, b"b": np.uint8}


def default_config() -> dict:
    """The parameter set of configfiles/test-small.xml (LIF + DS + ConnGrowth)."""
    return dict(
        width=10, height=10, neuron="AllLIFNeurons", synapse="AllDSSynapses", conn="ConnGrowth",
        layout="FixedLayout", inh_list=[], active_list=[], frac_exc=0.9, frac_starter=0.1,
        max_synapses_per_neuron=200, max_firing_rate=200, num_epochs=1, epoch_duration=1.0,
        dt=1e-4, seed=1,
        Iinject=(13.5e-9, 13.5e-9), Inoise=(1.0e-9, 1.5e-9), Vthresh=(15.0e-3, 15.0e-3),
        Vresting=(0.0, 0.0), Vreset=(13.5e-3, 13.5e-3), Vinit=(13.0e-3, 13.0e-3),
        starter_Vthresh=(13.565e-3, 13.655e-3), starter_Vreset=(13.0e-3, 13.0e-3),
        excA=(0.02, 0.02), excB=(0.2, 0.2), excC=(-65, -50), excD=(2, 8),
        inhA=(0.02, 0.1), inhB=(0.2, 0.25), inhC=(-65, -65), inhD=(2, 2),
        epsilon=0.6, beta=0.1, rho=0.0001, target_rate=1.9, min_radius=0.1, start_radius=0.4,
        ring_bits=32, noise_mode=0, math_mode=0, sum_mode=0, stdp=None, delay_override=None, input=None,
    )


def _text(el):
    return (el.text or "").strip()


def _read_nlist(path: str, tag: str) -> list[int]:
    root = ET.parse(path).getroot()
    el = root if root.tag == tag else root.find(tag)
    return [int(v) for v in _text(el).split()] if el is not None else []


def parse_config_xml(path: str, base_dir: str) -> dict:
    """Parse a BrainGrid simulation XML into the dict used by make_params().

    Follows the visitors in Core/SimulationInfo.cpp:53-179, Neurons/AllIFNeurons.cpp:129-339,
    Neurons/AllIZHNeurons.cpp:90-200, Connections/ConnGrowth.cpp:125-206,
    Layouts/FixedLayout.cpp:136-205 and Layouts/DynamicLayout.cpp:59-74. NList paths are
    resolved against base_dir, as the reference resolves them against its working directory."""
    cfg = default_config()
    root = ET.parse(path).getroot()
    sim = root.find("SimInfoParams")
    cfg["width"] = int(_text(sim.find("PoolSize/x")))
    cfg["height"] = int(_text(sim.find("PoolSize/y")))
    cfg["epoch_duration"] = float(_text(sim.find("SimParams/Tsim")))
    cfg["num_epochs"] = int(float(_text(sim.find("SimParams/numSims"))))
    cfg["max_firing_rate"] = int(_text(sim.find("SimConfig/maxFiringRate")))
    cfg["max_synapses_per_neuron"] = int(_text(sim.find("SimConfig/maxSynapsesPerNeuron")))
    cfg["seed"] = int(_text(sim.find("Seed/value")))
    model = root.find("ModelParams")
    neu = model.find("NeuronsParams")
    cfg["neuron"] = neu.get("class")
    for key, tag in (("Iinject", "Iinject"), ("Inoise", "Inoise"), ("Vthresh", "Vthresh"),
                     ("Vresting", "Vresting"), ("Vreset", "Vreset"), ("Vinit", "Vinit"),
                     ("starter_Vthresh", "starter_vthresh"), ("starter_Vreset", "starter_vreset")):
        el = neu.find(tag)
        if el is not None:
            cfg[key] = (float(_text(el.find("min"))), float(_text(el.find("max"))))
    for tag, key in (("Aconst", "A"), ("Bconst", "B"), ("Cconst", "C"), ("Dconst", "D")):
        el = neu.find(tag)
        if el is not None:
            cfg["exc" + key] = (float(_text(el.find("minExc"))), float(_text(el.find("maxExc"))))
            cfg["inh" + key] = (float(_text(el.find("minInh"))), float(_text(el.find("maxInh"))))
    cfg["synapse"] = model.find("SynapsesParams").get("class")
    conn = model.find("ConnectionsParams")
    cfg["conn"] = conn.get("class")
    g = conn.find("GrowthParams")
    if g is not None:
        for key, tag in (("epsilon", "epsilon"), ("beta", "beta"), ("rho", "rho"), ("target_rate", "targetRate"),
                         ("min_radius", "minRadius"), ("start_radius", "startRadius")):
            cfg[key] = float(_text(g.find(tag)))
    sp = conn.find("StaticConnectionsParams")   # Connections/ConnStatic.cpp:120-210
    if sp is not None:
        cfg["static_nconns"] = int(_text(sp.find("nConnsPerNeuron")))
        cfg["static_thresh"] = float(_text(sp.find("threshConnsRadius")))
        sw = conn.find("StaticConnectionsWeight")
        cfg["static_exc_w"] = (float(_text(sw.find("minExc"))), float(_text(sw.find("maxExc"))))
        cfg["static_inh_w"] = (float(_text(sw.find("minInh"))), float(_text(sw.find("maxInh"))))
    lay = model.find("LayoutParams")
    cfg["layout"] = lay.get("class")
    files = lay.find("LayoutFiles")
    if cfg["layout"] == "FixedLayout":
        a = files.find("activeNListFileName")
        i = files.find("inhNListFileName")
        cfg["active_list"] = _read_nlist(os.path.join(base_dir, _text(a)), "A") if a is not None and _text(a) else []
        cfg["inh_list"] = _read_nlist(os.path.join(base_dir, _text(i)), "I") if i is not None and _text(i) else []
    else:
        cfg["frac_exc"] = float(_text(files.find("frac_EXC")))
        cfg["frac_starter"] = float(_text(files.find("starter_neurons")))
    return cfg


def parse_input_xml(path: str, base_dir: str) -> dict:
    """Parse a stimulus input file (Inputs/FSInput.cpp:40-97): SInputPoisson
    (Inputs/SInputPoisson.cpp:24-108) or SInputRegular (Inputs/SInputRegular.cpp:26-110)."""
    root = ET.parse(path).getroot()
    parms = root if root.tag == "InputParams" else root.find("InputParams")
    name = parms.find("IMethod").get("name")
    ip = parms.find("IntParams")
    if name == "SInputPoisson":
        inp = dict(kind="poisson", fr_mean=float(ip.get("fr_mean")), weight=float(ip.get("weight")), mask=None)
        masks = parms.find("Masks")
        if masks is not None:
            for node in masks:
                if node.tag == "M":
                    inp["mask"] = [int(float(v)) for v in _text(node).split()]
                elif node.tag == "LayoutFiles" and node.get("maskNListFileName"):
                    inp["mask"] = [int(float(v)) for v in
                                   _text(ET.parse(os.path.join(base_dir, node.get("maskNListFileName"))).getroot()).split()]
        return inp
    if name == "SInputRegular":
        vals = [float(v) for v in _text(parms.find("Values/I")).split()]
        if len(vals) != 100:
            raise ValueError("SInputRegular needs a 10x10 <I> table")
        return dict(kind="regular", duration=float(ip.get("duration")), interval=float(ip.get("interval")),
                    sync=ip.get("sync"), values=vals)
    raise ValueError(f"unknown input method {name}")


def make_params(cfg: dict) -> tuple[Params, list]:
    """Build a Params struct; returns (params, keepalive) where keepalive holds the list buffers."""
    p = Params()
    p.width, p.height = cfg["width"], cfg["height"]
    p.neuron_kind = NEURON_CLASSES[cfg["neuron"]]
    p.syn_kind = SYN_CLASSES[cfg["synapse"]]
    p.layout_kind = LAYOUT_FIXED if cfg["layout"] == "FixedLayout" else LAYOUT_DYNAMIC
    p.conn_kind = {"ConnGrowth": CONN_GROWTH, "ConnStatic": CONN_STATIC}.get(cfg["conn"], CONN_NONE)
    if p.conn_kind == CONN_STATIC:
        p.static_nconns, p.static_thresh = cfg["static_nconns"], cfg["static_thresh"]
        p.static_exc_w[0], p.static_exc_w[1] = cfg["static_exc_w"]
        p.static_inh_w[0], p.static_inh_w[1] = cfg["static_inh_w"]
    p.max_synapses_per_neuron = cfg["max_synapses_per_neuron"]
    p.max_firing_rate = cfg["max_firing_rate"]
    p.num_epochs = cfg["num_epochs"]
    p.ring_bits = cfg.get("ring_bits", 32)
    p.dt = cfg["dt"]
    p.epoch_duration = cfg["epoch_duration"]
    p.seed = cfg["seed"] & 0xFFFFFFFF
    p.noise_mode, p.math_mode, p.sum_mode = cfg["noise_mode"], cfg["math_mode"], cfg["sum_mode"]
    inh = (ctypes.c_int32 * max(1, len(cfg["inh_list"])))(*cfg["inh_list"])
    act = (ctypes.c_int32 * max(1, len(cfg["active_list"])))(*cfg["active_list"])
    p.inh_list, p.n_inh = ctypes.cast(inh, ctypes.POINTER(ctypes.c_int32)), len(cfg["inh_list"])
    p.active_list, p.n_active = ctypes.cast(act, ctypes.POINTER(ctypes.c_int32)), len(cfg["active_list"])
    p.frac_exc, p.frac_starter = cfg["frac_exc"], cfg["frac_starter"]
    for k in ("Iinject", "Inoise", "Vthresh", "Vresting", "Vreset", "Vinit", "starter_Vthresh", "starter_Vreset",
              "excA", "excB", "excC", "excD", "inhA", "inhB", "inhC", "inhD"):
        getattr(p, k)[0], getattr(p, k)[1] = cfg[k]
    for k in ("epsilon", "beta", "rho", "target_rate", "min_radius", "start_radius"):
        setattr(p, k, cfg[k])
    st = cfg.get("stdp")
    if st:
        p.stdp_override = 1
        p.stdp_Apos, p.stdp_Aneg = st["Apos"], st["Aneg"]
        p.stdp_taupos, p.stdp_tauneg = st["taupos"], st["tauneg"]
        p.stdp_tauspre, p.stdp_tauspost = st.get("tauspre", 34e-3), st.get("tauspost", 75e-3)
        p.stdp_gap, p.stdp_Wex = st.get("gap", 2e-3), st.get("Wex", 1.0)
        p.stdp_mupos, p.stdp_muneg = st.get("mupos", 0.0), st.get("muneg", 0.0)
        p.stdp_use_froemke_dan = 1 if st.get("froemke_dan", False) else 0
    if cfg.get("delay_override"):
        for k in range(4):
            p.delay_override[k] = cfg["delay_override"][k]
    keep = [inh, act]
    inp = cfg.get("input")
    if inp and inp["kind"] == "poisson":
        p.input_kind, p.input_fr_mean, p.input_weight = INPUT_POISSON, inp["fr_mean"], inp["weight"]
        if inp.get("mask") is None:
            p.n_input_mask = -1
        else:
            m = (ctypes.c_int32 * max(1, len(inp["mask"])))(*inp["mask"])
            keep.append(m)
            p.input_mask, p.n_input_mask = ctypes.cast(m, ctypes.POINTER(ctypes.c_int32)), len(inp["mask"])
    elif inp and inp["kind"] == "regular":
        p.input_kind, p.input_duration, p.input_interval = INPUT_REGULAR, inp["duration"], inp["interval"]
        p.input_sync = SYNC.get(inp["sync"], -1)
        for k, v in enumerate(inp["values"]):
            p.input_values[k] = v
    return p, keep


class OracleNet:
    """One oracle network (Model + SingleThreadedSpikingModel state)."""

    def __init__(self, cfg: dict):
        self.cfg = dict(cfg)
        self._p, self._keep = make_params(self.cfg)
        err = ctypes.create_string_buffer(256)
        h = lib().bgo_create(ctypes.byref(self._p), err, 256)
        if not h:
            raise RuntimeError("bgo_create: " + err.value.decode())
        self.h = h

    def close(self):
        if self.h:
            lib().bgo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"{what}: {lib().bgo_last_error(self.h).decode()}")

    def advance(self, n_steps: int):
        self._check(lib().bgo_advance(self.h, n_steps), "advance")

    def update_connections(self):
        self._check(lib().bgo_update_connections(self.h), "update_connections")

    def update_weights(self, W: np.ndarray):
        """ConnGrowth::updateSynapsesWeights with W[src, dest] in place of the overlap areas."""
        W = np.ascontiguousarray(W, dtype=np.float32)
        n = self.cfg["width"] * self.cfg["height"]
        assert W.shape == (n, n)
        self._check(lib().bgo_update_weights(self.h, W.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "update_weights")

    def update_history(self):
        self._check(lib().bgo_update_history(self.h), "update_history")

    def run_epoch(self):
        self._check(lib().bgo_run_epoch(self.h), "run_epoch")

    def add_synapse(self, src: int, dest: int, w: float) -> int:
        i = lib().bgo_add_synapse(self.h, src, dest, w)
        self._check(i, "add_synapse")
        return i

    def rebuild_index_map(self):
        self._check(lib().bgo_rebuild_index_map(self.h), "rebuild_index_map")

    def set_radii(self, r: float):
        lib().bgo_set_radii(self.h, r)

    @property
    def step(self) -> int:
        return lib().bgo_step(self.h)

    @property
    def total_synapses(self) -> int:
        return lib().bgo_total_synapses(self.h)

    def array(self, name: str) -> np.ndarray:
        """A copy of the named state array."""
        return self.view(name).copy()

    def view(self, name: str) -> np.ndarray:
        """The named state array in place (writable; valid until the array is reallocated)."""
        dt = ctypes.c_char()
        cnt = ctypes.c_int64()
        ptr = lib().bgo_array(self.h, name.encode(), ctypes.byref(dt), ctypes.byref(cnt))
        if not ptr:
            raise KeyError(name)
        t = _DT[dt.value]
        n = cnt.value
        if n == 0:
            return np.zeros(0, t)
        buf = (ctypes.c_char * (n * np.dtype(t).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=t)


def read_dump(path: str) -> list[tuple[str, int, dict]]:
    """Parse a bgref_dump file into [(marker, epoch, {name: array})]."""
    codes = {b"f": np.float32, b"i": np.int32, b"I": np.uint32, b"q": np.uint64, b"b": np.uint8}
    out = []
    with open(path, "rb") as f:
        assert f.read(8) == b"BGDUMP1\n"
        data = f.read()
    pos = 0
    pending: dict = {}
    header: dict = {}
    cur = None
    while pos < len(data):
        (nl,) = struct.unpack_from("<I", data, pos); pos += 4
        name = data[pos:pos + nl].decode(); pos += nl
        code = data[pos:pos + 1]; pos += 1
        (cnt,) = struct.unpack_from("<Q", data, pos); pos += 8
        t = codes[code]
        nbytes = cnt * np.dtype(t).itemsize
        arr = np.frombuffer(data, dtype=t, count=cnt, offset=pos).copy() if cnt else np.zeros(0, t)
        pos += nbytes
        if name.startswith("@"):
            if cur is not None:
                out.append(cur)
            cur = (name[1:], int(arr[0]), {})
            if not out and pending:
                header = pending
            pending = {}
            continue
        if cur is None:
            pending[name] = arr
        else:
            cur[2][name] = arr
    if cur is not None:
        out.append(cur)
    if header:
        out.insert(0, ("header", 0, header))
    return out


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "bgref_dump"))
