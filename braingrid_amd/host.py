"""ctypes binding of the host model (include/bghost.h) and a BrainGrid XML config reader.

`Model` drives HipSpikingModel: Model::setupSim on the host, every step and growth update on the
GPU. `read_config` accepts the reference's simulation XML files (the SimInfoParams / ModelParams
schema parsed by Core/SimulationInfo.cpp:53-179 and the FClassOfCategory visitors).
"""
from __future__ import annotations

import ctypes
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import bghip as B

_F2 = ctypes.c_float * 2
NEURON_CLASSES = {"AllLIFNeurons": B.NEURON_LIF, "AllIZHNeurons": B.NEURON_IZH}
SYN_CLASSES = {"AllSpikingSynapses": B.SYN_SPIKING, "AllDSSynapses": B.SYN_DS, "AllSTDPSynapses": B.SYN_STDP,
               "AllDynamicSTDPSynapses": B.SYN_DYNSTDP}


class Params(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("neuron_model", ctypes.c_int32),
                ("synapse_model", ctypes.c_int32), ("layout_kind", ctypes.c_int32), ("conn_kind", ctypes.c_int32),
                ("max_synapses_per_neuron", ctypes.c_int32), ("max_firing_rate", ctypes.c_int32),
                ("num_epochs", ctypes.c_int32), ("ring_bits", ctypes.c_int32), ("dt", ctypes.c_float),
                ("epoch_duration", ctypes.c_float), ("seed", ctypes.c_uint32),
                ("inh_list", ctypes.POINTER(ctypes.c_int32)), ("n_inh", ctypes.c_int32),
                ("active_list", ctypes.POINTER(ctypes.c_int32)), ("n_active", ctypes.c_int32),
                ("frac_exc", ctypes.c_float), ("frac_starter", ctypes.c_float)] + \
               [(k, _F2) for k in ("Iinject", "Inoise", "Vthresh", "Vresting", "Vreset", "Vinit", "starter_Vthresh",
                                   "starter_Vreset", "excA", "excB", "excC", "excD", "inhA", "inhB", "inhC", "inhD")] + \
               [(k, ctypes.c_float) for k in ("epsilon", "beta", "rho", "target_rate", "min_radius", "start_radius")] + \
               [("block_steps", ctypes.c_int32), ("static_nconns", ctypes.c_int32), ("static_thresh", ctypes.c_float),
                ("static_exc_w", _F2), ("static_inh_w", _F2),
                ("input_kind", ctypes.c_int32), ("input_fr_mean", ctypes.c_float), ("input_weight", ctypes.c_float),
                ("input_mask", ctypes.POINTER(ctypes.c_int32)), ("n_input_mask", ctypes.c_int32),
                ("input_duration", ctypes.c_float), ("input_interval", ctypes.c_float), ("input_sync", ctypes.c_int32),
                ("input_values", ctypes.c_float * 100)]

INPUT_SYNC = {"yes": 0, "no": 1, "wave": 2}

CONN_KINDS = {"ConnGrowth": 1, "ConnStatic": 2}


def base_config() -> dict:
    """Parameters of configfiles/test-small.xml: LIF + DS + ConnGrowth (the benchmark family)."""
    return dict(width=10, height=10, neuron="AllLIFNeurons", synapse="AllDSSynapses", conn="ConnGrowth",
                layout="DynamicLayout", inh_list=[], active_list=[], frac_exc=0.9, frac_starter=0.1,
                max_synapses_per_neuron=200, max_firing_rate=200, num_epochs=10, epoch_duration=1.0, dt=1e-4, seed=1,
                Iinject=(13.5e-9, 13.5e-9), Inoise=(1.0e-9, 1.5e-9), Vthresh=(15.0e-3, 15.0e-3),
                Vresting=(0.0, 0.0), Vreset=(13.5e-3, 13.5e-3), Vinit=(13.0e-3, 13.0e-3),
                starter_Vthresh=(13.565e-3, 13.655e-3), starter_Vreset=(13.0e-3, 13.0e-3),
                excA=(0.02, 0.02), excB=(0.2, 0.2), excC=(-65, -50), excD=(2, 8),
                inhA=(0.02, 0.1), inhB=(0.2, 0.25), inhC=(-65, -65), inhD=(2, 2),
                epsilon=0.6, beta=0.1, rho=1e-4, target_rate=1.9, min_radius=0.1, start_radius=0.4,
                ring_bits=32, block_steps=0)


def _t(el):
    return (el.text or "").strip()


def read_config(path: str, base_dir: str | None = None) -> dict:
    base_dir = base_dir or os.path.dirname(os.path.abspath(path))
    c = base_config()
    root = ET.parse(path).getroot()
    s = root.find("SimInfoParams")
    c.update(width=int(_t(s.find("PoolSize/x"))), height=int(_t(s.find("PoolSize/y"))),
             epoch_duration=float(_t(s.find("SimParams/Tsim"))), num_epochs=int(float(_t(s.find("SimParams/numSims")))),
             max_firing_rate=int(_t(s.find("SimConfig/maxFiringRate"))),
             max_synapses_per_neuron=int(_t(s.find("SimConfig/maxSynapsesPerNeuron"))),
             seed=int(_t(s.find("Seed/value"))))
    m = root.find("ModelParams")
    neu = m.find("NeuronsParams")
    c["neuron"] = neu.get("class")
    for key, tag in (("Iinject", "Iinject"), ("Inoise", "Inoise"), ("Vthresh", "Vthresh"), ("Vresting", "Vresting"),
                     ("Vreset", "Vreset"), ("Vinit", "Vinit"), ("starter_Vthresh", "starter_vthresh"),
                     ("starter_Vreset", "starter_vreset")):
        el = neu.find(tag)
        if el is not None:
            c[key] = (float(_t(el.find("min"))), float(_t(el.find("max"))))
    for tag, k in (("Aconst", "A"), ("Bconst", "B"), ("Cconst", "C"), ("Dconst", "D")):
        el = neu.find(tag)
        if el is not None:
            c["exc" + k] = (float(_t(el.find("minExc"))), float(_t(el.find("maxExc"))))
            c["inh" + k] = (float(_t(el.find("minInh"))), float(_t(el.find("maxInh"))))
    c["synapse"] = m.find("SynapsesParams").get("class")
    conn = m.find("ConnectionsParams")
    c["conn"] = conn.get("class")
    g = conn.find("GrowthParams")
    if g is not None:
        for key, tag in (("epsilon", "epsilon"), ("beta", "beta"), ("rho", "rho"), ("target_rate", "targetRate"),
                         ("min_radius", "minRadius"), ("start_radius", "startRadius")):
            c[key] = float(_t(g.find(tag)))
    sp = conn.find("StaticConnectionsParams")   # ConnStatic::readParameters (Connections/ConnStatic.cpp:120-210)
    if sp is not None:
        c["static_nconns"] = int(_t(sp.find("nConnsPerNeuron")))
        c["static_thresh"] = float(_t(sp.find("threshConnsRadius")))
        sw = conn.find("StaticConnectionsWeight")
        c["static_exc_w"] = (float(_t(sw.find("minExc"))), float(_t(sw.find("maxExc"))))
        c["static_inh_w"] = (float(_t(sw.find("minInh"))), float(_t(sw.find("maxInh"))))
    lay = m.find("LayoutParams")
    c["layout"] = lay.get("class")
    files = lay.find("LayoutFiles")
    if c["layout"] == "FixedLayout":
        def nlist(tag, root_tag):
            el = files.find(tag)
            if el is None or not _t(el):
                return []
            r = ET.parse(os.path.join(base_dir, _t(el))).getroot()
            node = r if r.tag == root_tag else r.find(root_tag)
            return [int(v) for v in _t(node).split()]
        c["active_list"] = nlist("activeNListFileName", "A")
        c["inh_list"] = nlist("inhNListFileName", "I")
    else:
        c["frac_exc"] = float(_t(files.find("frac_EXC")))
        c["frac_starter"] = float(_t(files.find("starter_neurons")))
    return c


_lib = None


def load():
    global _lib
    if _lib is None:
        L = B.load()
        vp = ctypes.c_void_p
        for name, res, args in (
                ("bghost_model_create", ctypes.c_int, [vp, ctypes.POINTER(Params), ctypes.POINTER(vp)]),
                ("bghost_model_destroy", None, [vp]), ("bghost_model_net", vp, [vp]),
                ("bghost_model_advance", ctypes.c_int, [vp, ctypes.c_uint64]),
                ("bghost_model_advance_epoch", ctypes.c_int, [vp]),
                ("bghost_model_update_connections", ctypes.c_int, [vp, ctypes.POINTER(B.GrowthStats)]),
                ("bghost_model_update_history", ctypes.c_int, [vp]), ("bghost_model_run_epoch", ctypes.c_int, [vp]),
                ("bghost_model_save_xml", ctypes.c_int, [vp, ctypes.c_char_p]),
                ("bghost_model_set_radii", ctypes.c_int, [vp, ctypes.c_float]),
                ("bghost_model_array", vp, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
                ("bghost_model_epoch", ctypes.c_int32, [vp]),
                ("bghost_model_serialize", ctypes.c_int, [vp, ctypes.c_char_p]),
                ("bghost_model_deserialize", ctypes.c_int, [vp, ctypes.c_char_p])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


EXPORTED = ("bghost_model_create", "bghost_model_destroy", "bghost_model_net", "bghost_model_advance",
            "bghost_model_advance_epoch", "bghost_model_update_connections", "bghost_model_update_history",
            "bghost_model_run_epoch", "bghost_model_save_xml", "bghost_model_set_radii", "bghost_model_array",
            "bghost_model_epoch", "bghost_model_serialize", "bghost_model_deserialize")

_ARRAY_DTYPES = {"neuron_type_map": np.int32, "starter_map": np.uint8}


def read_stimulus(path: str, base_dir: str | None = None) -> dict:
    """A stimulus input file (the -s option; Inputs/FSInput.cpp:40-97): SInputPoisson
    (Inputs/SInputPoisson.cpp:24-108: IntParams fr_mean and weight, optional <Masks> list or mask
    file) or SInputRegular (Inputs/SInputRegular.cpp:26-110: IntParams duration, interval, sync and
    the 10x10 <Values><I> table). The result goes into a config as cfg["input"]."""
    base_dir = base_dir or os.path.dirname(os.path.abspath(path))
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
                    inp["mask"] = [int(float(v)) for v in _t(node).split()]
                elif node.tag == "LayoutFiles" and node.get("maskNListFileName"):
                    f = os.path.join(base_dir, node.get("maskNListFileName"))
                    inp["mask"] = [int(float(v)) for v in _t(ET.parse(f).getroot()).split()]
        return inp
    if name == "SInputRegular":
        vals = [float(v) for v in _t(parms.find("Values/I")).split()]
        if len(vals) != 100:
            raise ValueError("SInputRegular needs a 10x10 <I> table")
        return dict(kind="regular", duration=float(ip.get("duration")), interval=float(ip.get("interval")),
                    sync=ip.get("sync"), values=vals)
    raise ValueError(f"unknown input method {name}")


def make_params(c: dict):
    p = Params()
    p.width, p.height = c["width"], c["height"]
    p.neuron_model, p.synapse_model = NEURON_CLASSES[c["neuron"]], SYN_CLASSES[c["synapse"]]
    p.layout_kind = 0 if c["layout"] == "FixedLayout" else 1
    p.conn_kind = CONN_KINDS.get(c["conn"], 0)
    if p.conn_kind == CONN_KINDS["ConnStatic"]:
        p.static_nconns, p.static_thresh = c["static_nconns"], c["static_thresh"]
        p.static_exc_w[0], p.static_exc_w[1] = c["static_exc_w"]
        p.static_inh_w[0], p.static_inh_w[1] = c["static_inh_w"]
    p.max_synapses_per_neuron, p.max_firing_rate = c["max_synapses_per_neuron"], c["max_firing_rate"]
    p.num_epochs, p.ring_bits, p.dt, p.epoch_duration = c["num_epochs"], c["ring_bits"], c["dt"], c["epoch_duration"]
    p.seed = c["seed"] & 0xFFFFFFFF
    inh = (ctypes.c_int32 * max(1, len(c["inh_list"])))(*c["inh_list"])
    act = (ctypes.c_int32 * max(1, len(c["active_list"])))(*c["active_list"])
    p.inh_list, p.n_inh = ctypes.cast(inh, ctypes.POINTER(ctypes.c_int32)), len(c["inh_list"])
    p.active_list, p.n_active = ctypes.cast(act, ctypes.POINTER(ctypes.c_int32)), len(c["active_list"])
    p.frac_exc, p.frac_starter = c["frac_exc"], c["frac_starter"]
    for k in ("Iinject", "Inoise", "Vthresh", "Vresting", "Vreset", "Vinit", "starter_Vthresh", "starter_Vreset",
              "excA", "excB", "excC", "excD", "inhA", "inhB", "inhC", "inhD"):
        getattr(p, k)[0], getattr(p, k)[1] = c[k]
    for k in ("epsilon", "beta", "rho", "target_rate", "min_radius", "start_radius"):
        setattr(p, k, c[k])
    p.block_steps = c.get("block_steps", 0)
    keep = [inh, act]
    inp = c.get("input")
    if inp and inp["kind"] == "poisson":
        p.input_kind, p.input_fr_mean, p.input_weight = B.INPUT_POISSON, inp["fr_mean"], inp["weight"]
        if inp.get("mask") is None:
            p.n_input_mask = -1
        else:
            m = (ctypes.c_int32 * max(1, len(inp["mask"])))(*inp["mask"])
            keep.append(m)
            p.input_mask, p.n_input_mask = ctypes.cast(m, ctypes.POINTER(ctypes.c_int32)), len(inp["mask"])
    elif inp and inp["kind"] == "regular":
        p.input_kind, p.input_duration, p.input_interval = B.INPUT_REGULAR, inp["duration"], inp["interval"]
        if inp["sync"] not in INPUT_SYNC:
            raise ValueError(f"SInputRegular sync {inp['sync']!r} is not one of yes / no / wave")
        p.input_sync = INPUT_SYNC[inp["sync"]]
        for k, v in enumerate(inp["values"]):
            p.input_values[k] = v
    elif inp:
        raise ValueError(f"unknown input kind {inp['kind']!r}")
    return p, tuple(keep)


class Model:
    """HipSpikingModel through bghost.h: setupSim on construction."""

    def __init__(self, ctx: B.Context, cfg: dict):
        L = load()
        self.ctx, self.cfg = ctx, dict(cfg)
        self._p, self._keep = make_params(self.cfg)
        h = ctypes.c_void_p()
        ctx.check(L.bghost_model_create(ctx.h, ctypes.byref(self._p), ctypes.byref(h)), "bghost_model_create")
        self.h = h
        self.N = cfg["width"] * cfg["height"]
        self.net = B.Net.wrap(ctx, L.bghost_model_net(h))

    def close(self):
        if self.h:
            load().bghost_model_destroy(self.h)
            self.h = None
            self.net.h = None

    def advance(self, n):
        self.ctx.check(load().bghost_model_advance(self.h, n), "advance")

    def advance_epoch(self):
        self.ctx.check(load().bghost_model_advance_epoch(self.h), "advance_epoch")

    def update_connections(self, stats: bool = True):
        """The growth update; with stats=False it does not wait for the device."""
        if not stats:
            self.ctx.check(load().bghost_model_update_connections(self.h, None), "update_connections")
            return None
        st = B.GrowthStats()
        self.ctx.check(load().bghost_model_update_connections(self.h, ctypes.byref(st)), "update_connections")
        return st

    def update_history(self):
        self.ctx.check(load().bghost_model_update_history(self.h), "update_history")

    def run_epoch(self):
        self.ctx.check(load().bghost_model_run_epoch(self.h), "run_epoch")

    def save_xml(self, path):
        self.ctx.check(load().bghost_model_save_xml(self.h, path.encode()), "save_xml")

    def serialize(self, path: str):
        """IModel::serialize: the model's network and recorder as a checkpoint file."""
        self.ctx.check(load().bghost_model_serialize(self.h, os.fsencode(path)), "serialize")

    def deserialize(self, path: str):
        """IModel::deserialize: continue from a checkpoint of a model with this configuration."""
        self.ctx.check(load().bghost_model_deserialize(self.h, os.fsencode(path)), "deserialize")
        self.net.h = ctypes.c_void_p(load().bghost_model_net(self.h))

    def set_radii(self, r):
        self.ctx.check(load().bghost_model_set_radii(self.h, r), "set_radii")

    @property
    def epoch(self):
        return load().bghost_model_epoch(self.h)

    def array(self, name: str) -> np.ndarray:
        cnt = ctypes.c_int64()
        ptr = load().bghost_model_array(self.h, name.encode(), ctypes.byref(cnt))
        if not ptr:
            raise KeyError(name)
        dt = _ARRAY_DTYPES.get(name, np.float32)
        n = cnt.value
        if n == 0:
            return np.zeros(0, dt)
        buf = (ctypes.c_char * (n * np.dtype(dt).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dt).copy()
