"""ctypes binding of libbghip.so (include/bghip.h) with numpy conveniences.

This is plumbing: every computation happens in the HIP kernels behind the C ABI. Importing this
module never falls back to anything else: if the shared library is missing, `load()` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BGHIP_LIB") or os.path.join(HERE, "lib", "libbghip.so")

NEURON_LIF, NEURON_IZH = 0, 1
SYN_SPIKING, SYN_DS, SYN_STDP, SYN_DYNSTDP = 0, 1, 2, 3
ERR_RING_COLLISION, ERR_SPIKE_OVERFLOW, ERR_SLOT_OVERFLOW = 1, 2, 4

_F4 = ctypes.c_float * 4
_I4 = ctypes.c_int32 * 4
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint32)
_qp = ctypes.POINTER(ctypes.c_uint64)
_bp = ctypes.POINTER(ctypes.c_uint8)


class SynapseTypes(ctypes.Structure):
    _fields_ = [("tau", _F4), ("decay", _F4), ("total_delay", _I4), ("U", _F4), ("D", _F4), ("F", _F4),
                ("u0", ctypes.c_float), ("r0", ctypes.c_float),
                ("Apos", ctypes.c_float), ("Aneg", ctypes.c_float), ("STDPgap", ctypes.c_float),
                ("tauspost", ctypes.c_float), ("tauspre", ctypes.c_float), ("taupos", ctypes.c_float),
                ("tauneg", ctypes.c_float), ("Wex", ctypes.c_float), ("mupos", ctypes.c_float),
                ("muneg", ctypes.c_float), ("useFroemkeDanSTDP", ctypes.c_int32)]


class NetDesc(ctypes.Structure):
    _fields_ = [("num_neurons", ctypes.c_int32), ("max_synapses_per_neuron", ctypes.c_int32),
                ("neuron_model", ctypes.c_int32), ("synapse_model", ctypes.c_int32),
                ("ring_bits", ctypes.c_int32), ("max_spikes", ctypes.c_int32), ("dt", ctypes.c_float),
                ("noise_seed", ctypes.c_uint64), ("block_steps", ctypes.c_int32), ("types", SynapseTypes)]


class Neurons(ctypes.Structure):
    _fields_ = [(n, _fp) for n in ("Vm", "Vthresh", "Vreset", "I0", "Inoise", "C1", "C2", "Trefract",
                                   "summation_map")] + \
               [("nStepsInRefr", _ip)] + \
               [(n, _fp) for n in ("Aconst", "Bconst", "Cconst", "Dconst", "u", "C3")] + \
               [("spikeCount", _ip), ("spikeCountOffset", _ip), ("spike_history", _qp)]


class Synapses(ctypes.Structure):
    _fields_ = [("in_use", _bp), ("sourceNeuronIndex", _ip), ("destNeuronIndex", _ip), ("type", _ip),
                ("W", _fp), ("psr", _fp), ("decay", _fp), ("total_delay", _ip), ("delayIdx", _ip),
                ("delayQueue", _up), ("r", _fp), ("u", _fp), ("U", _fp), ("D", _fp), ("F", _fp),
                ("lastSpike", _qp), ("delayQueuePost", _up), ("delayIdxPost", _ip)] + \
               [(n, _fp) for n in ("Apos", "Aneg", "STDPgap", "tauspost", "tauspre", "taupos", "tauneg", "Wex",
                                   "mupos", "muneg")] + \
               [("useFroemkeDanSTDP", _bp), ("synapse_counts", _up)]


class IndexMap(ctypes.Structure):
    _fields_ = [(n, _up) for n in ("incomingSynapseIndexMap", "incomingSynapseBegin", "incomingSynapseCount",
                                   "outgoingSynapseIndexMap", "outgoingSynapseBegin", "outgoingSynapseCount")]


class Growth(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("epsilon", "beta", "rho", "targetRate", "maxRate", "minRadius",
                                              "startRadius", "epochDuration")]


class Input(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("fr_mean", ctypes.c_float), ("weight", ctypes.c_float),
                ("mask", _bp), ("neuron_type", _ip), ("duration", ctypes.c_float), ("interval", ctypes.c_float),
                ("values", _fp), ("shift", _ip)]


class InputState(ctypes.Structure):
    _fields_ = [("nISIs", _ip), ("psr", _fp), ("r", _fp), ("u", _fp), ("lastSpike", _qp), ("delayQueue", _up),
                ("delayIdx", _ip), ("steps_in_cycle", ctypes.c_int32)]


INPUT_NONE, INPUT_POISSON, INPUT_REGULAR = 0, 1, 2
INPUT_STATE_FIELDS = {"nISIs": np.int32, "psr": np.float32, "r": np.float32, "u": np.float32,
                      "lastSpike": np.uint64, "delayQueue": np.uint32, "delayIdx": np.int32}


class BatchShape(ctypes.Structure):
    """bghip_batch_shape: one batch group's kind and launch shape."""
    _fields_ = [("mode", ctypes.c_int32), ("neuron_model", ctypes.c_int32), ("synapse_model", ctypes.c_int32),
                ("ring_words", ctypes.c_int32), ("networks", ctypes.c_uint32), ("work_items", ctypes.c_uint32),
                ("block", ctypes.c_int32), ("dests_per_wave", ctypes.c_int32), ("block_stream", ctypes.c_int32),
                ("syn_blocks_per_shard", ctypes.c_int32), ("sum_blocks_per_shard", ctypes.c_int32),
                ("tuned", ctypes.c_int32), ("step_us", ctypes.c_float)]


class GrowthStats(ctypes.Structure):
    _fields_ = [("total_synapses", ctypes.c_int64), ("created", ctypes.c_int64), ("erased", ctypes.c_int64)]


NEURON_FIELDS_F = ("Vm", "Vthresh", "Vreset", "I0", "Inoise", "C1", "C2", "Trefract", "summation_map")
IZH_FIELDS = ("Aconst", "Bconst", "Cconst", "Dconst", "u", "C3")
SYN_DTYPES = {"in_use": np.uint8, "sourceNeuronIndex": np.int32, "destNeuronIndex": np.int32, "type": np.int32,
              "W": np.float32, "psr": np.float32, "decay": np.float32, "total_delay": np.int32,
              "delayIdx": np.int32, "delayQueue": np.uint32, "r": np.float32, "u": np.float32,
              "U": np.float32, "D": np.float32, "F": np.float32, "lastSpike": np.uint64,
              "delayQueuePost": np.uint32, "delayIdxPost": np.int32, "Apos": np.float32, "Aneg": np.float32,
              "STDPgap": np.float32, "tauspost": np.float32, "tauspre": np.float32, "taupos": np.float32,
              "tauneg": np.float32, "Wex": np.float32, "mupos": np.float32, "muneg": np.float32,
              "useFroemkeDanSTDP": np.uint8}
DS_FIELDS = ("r", "u", "U", "D", "F", "lastSpike")
STDP_FIELDS = ("delayQueuePost", "delayIdxPost", "Apos", "Aneg", "STDPgap", "tauspost", "tauspre", "taupos",
               "tauneg", "Wex", "mupos", "muneg", "useFroemkeDanSTDP")

_lib = None


def load():
    """Load libbghip.so; raise if it has not been built (there is no other implementation)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "bghip_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "bghip_ctx_destroy": (None, [vp]),
        "bghip_last_error": (ctypes.c_char_p, [vp]),
        "bghip_ctx_mem_info": (ctypes.c_int, [vp, _qp, _qp]),
        "bghip_net_create": (ctypes.c_int, [vp, ctypes.POINTER(NetDesc), ctypes.POINTER(vp)]),
        "bghip_net_destroy": (None, [vp]),
        "bghip_net_footprint": (ctypes.c_int64, [ctypes.POINTER(NetDesc)]),
        "bghip_upload_neurons": (ctypes.c_int, [vp, ctypes.POINTER(Neurons)]),
        "bghip_upload_synapses": (ctypes.c_int, [vp, ctypes.POINTER(Synapses), ctypes.c_uint64]),
        "bghip_upload_layout": (ctypes.c_int, [vp, _fp, _fp, _ip]),
        "bghip_upload_growth": (ctypes.c_int, [vp, ctypes.POINTER(Growth), _fp, _fp]),
        "bghip_advance": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "bghip_update_connections": (ctypes.c_int, [vp, ctypes.POINTER(GrowthStats)]),
        "bghip_update_weights": (ctypes.c_int, [vp, _fp, ctypes.POINTER(GrowthStats)]),
        "bghip_update_weights_static": (ctypes.c_int, [vp, _fp, ctypes.POINTER(GrowthStats)]),
        "bghip_net_implicit": (ctypes.c_int, [vp]),
        "bghip_download_spikes": (ctypes.c_int, [vp, _ip, _ip, _qp]),
        "bghip_clear_spike_counts": (ctypes.c_int, [vp]),
        "bghip_update_history": (ctypes.c_int, [vp]),
        "bghip_sync": (ctypes.c_int, [vp]),
        "bghip_wait": (ctypes.c_int, [vp]),
        "bghip_net_error_bits": (ctypes.c_uint32, [vp]),
        "bghip_net_step": (ctypes.c_uint64, [vp]),
        "bghip_download_neurons": (ctypes.c_int, [vp, ctypes.POINTER(Neurons)]),
        "bghip_download_synapses": (ctypes.c_int, [vp, ctypes.POINTER(Synapses)]),
        "bghip_total_synapses": (ctypes.c_int64, [vp]),
        "bghip_download_index_map": (ctypes.c_int, [vp, ctypes.POINTER(IndexMap)]),
        "bghip_download_growth": (ctypes.c_int, [vp, _fp, _fp]),
        "bghip_kernel_timing": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_int64)]),
        "bghip_upload_input": (ctypes.c_int, [vp, ctypes.POINTER(Input), ctypes.POINTER(InputState)]),
        "bghip_recorder_enable": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                 _fp, _fp]),
        "bghip_recorder_download": (ctypes.c_int, [vp, _fp, _fp, _fp, _fp]),
        "bghip_net_save": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "bghip_add_synapses": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int32), _fp]),
        "bghip_add_synapses_ex": (ctypes.c_int, [vp, ctypes.c_int64, _ip, _ip, _fp, _ip, _fp]),
        "bghip_net_get_desc": (ctypes.c_int, [vp, ctypes.POINTER(NetDesc)]),
        "bghip_recorder_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                               ctypes.POINTER(ctypes.c_int32)]),
        "bghip_net_load": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "bghip_download_input": (ctypes.c_int, [vp, ctypes.POINTER(InputState)]),
        "bghip_batch_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
        "bghip_batch_destroy": (None, [vp]),
        "bghip_batch_set": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.c_int]),
        "bghip_batch_advance": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "bghip_batch_sync": (ctypes.c_int, [vp]),
        "bghip_batch_autotune": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "bghip_batch_group_count": (ctypes.c_int, [vp]),
        "bghip_batch_get_shape": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(BatchShape)]),
        "bghip_batch_set_shape": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(BatchShape)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ("bghip_ctx_create", "bghip_ctx_destroy", "bghip_last_error", "bghip_ctx_mem_info", "bghip_net_create",
            "bghip_net_destroy", "bghip_net_footprint", "bghip_upload_neurons", "bghip_upload_synapses",
            "bghip_upload_layout", "bghip_upload_growth", "bghip_advance", "bghip_update_connections",
            "bghip_update_history", "bghip_sync", "bghip_wait", "bghip_net_error_bits", "bghip_net_step",
            "bghip_download_neurons", "bghip_download_synapses", "bghip_total_synapses",
            "bghip_download_index_map", "bghip_download_growth", "bghip_kernel_timing", "bghip_batch_create",
            "bghip_batch_destroy", "bghip_batch_set", "bghip_batch_advance", "bghip_batch_sync",
            "bghip_upload_input", "bghip_download_input", "bghip_recorder_enable", "bghip_recorder_download",
            "bghip_net_save", "bghip_net_load", "bghip_net_get_desc", "bghip_recorder_info",
            "bghip_add_synapses", "bghip_add_synapses_ex", "bghip_batch_autotune", "bghip_batch_group_count", "bghip_batch_get_shape",
            "bghip_batch_set_shape", "bghip_update_weights", "bghip_download_spikes", "bghip_clear_spike_counts",
            "bghip_update_weights_static", "bghip_net_implicit")


def _ptr(a: np.ndarray | None, ct):
    if a is None:
        return ctypes.cast(None, ct)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ct)


_CT = {np.float32: _fp, np.int32: _ip, np.uint32: _up, np.uint64: _qp, np.uint8: _bp}


class BgHipError(RuntimeError):
    pass


class Context:
    def __init__(self, device: int = 0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.bghip_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise BgHipError(f"bghip_ctx_create({device}) = {rc}")
        self.h = h

    def error(self) -> str:
        return load().bghip_last_error(self.h).decode()

    def mem_info(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(load().bghip_ctx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)), "mem_info")
        return f.value, t.value

    def check(self, rc: int, what: str):
        if rc != 0:
            raise BgHipError(f"{what} -> {rc}: {self.error()}")

    def close(self):
        if self.h:
            load().bghip_ctx_destroy(self.h)
            self.h = None


class Net:
    """One device network: the GPU-side replacement of GPUSpikingModel's device structs."""

    def __init__(self, ctx: Context, desc: NetDesc):
        self.ctx = ctx
        self.desc = desc
        h = ctypes.c_void_p()
        ctx.check(load().bghip_net_create(ctx.h, ctypes.byref(desc), ctypes.byref(h)), "net_create")
        self.h = h
        self.N = desc.num_neurons
        self.M = desc.max_synapses_per_neuron
        self.W32 = desc.ring_bits // 32

    def close(self):
        if self.h:
            load().bghip_net_destroy(self.h)
            self.h = None

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def add_synapses(self, src, dest, W, total_delay=None, Wex=None):
        """bghip_add_synapses(_ex): addSynapse for each src[i] -> dest[i] in order, then W[iSyn] = W[i]
        and, when given, the synapse's total delay (steps) and STDP Wex."""
        s = np.ascontiguousarray(src, np.int32)
        d = np.ascontiguousarray(dest, np.int32)
        w = np.ascontiguousarray(W, np.float32)
        td = None if total_delay is None else np.ascontiguousarray(total_delay, np.int32)
        wx = None if Wex is None else np.ascontiguousarray(Wex, np.float32)
        if not (s.shape == d.shape == w.shape and s.ndim == 1) or any(a is not None and a.shape != s.shape
                                                                      for a in (td, wx)):
            raise ValueError("src, dest, W (and total_delay, Wex) must be 1-D arrays of one length")
        _i32 = ctypes.POINTER(ctypes.c_int32)
        if td is None and wx is None:
            self._c(load().bghip_add_synapses(self.h, s.size, _ptr(s, _i32), _ptr(d, _i32), _ptr(w, _fp)),
                    "add_synapses")
        else:
            self._c(load().bghip_add_synapses_ex(self.h, s.size, _ptr(s, _i32), _ptr(d, _i32), _ptr(w, _fp),
                                                 _ptr(td, _i32), _ptr(wx, _fp)), "add_synapses")

    # ---- checkpoints
    def save(self, path: str):
        """bghip_net_save: the whole network state as a versioned binary file."""
        self._c(load().bghip_net_save(self.h, os.fsencode(path)), "net_save")

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Net":
        """bghip_net_load: a new network holding a checkpoint's state."""
        h = ctypes.c_void_p()
        ctx.check(load().bghip_net_load(ctx.h, os.fsencode(path), ctypes.byref(h)), "net_load")
        return cls.wrap(ctx, h.value)

    @classmethod
    def wrap(cls, ctx: Context, handle) -> "Net":
        """A Net object over an existing handle; close() destroys the handle."""
        net = cls.__new__(cls)
        net.ctx, net.h = ctx, ctypes.c_void_p(handle)
        d = NetDesc()
        ctx.check(load().bghip_net_get_desc(net.h, ctypes.byref(d)), "net_get_desc")
        net.desc, net.N, net.M, net.W32 = d, d.num_neurons, d.max_synapses_per_neuron, d.ring_bits // 32
        return net

    def recorder_info(self):
        """(n_burst, n_spikes_hist, num_epochs, with_growth, epoch) of the on-device recorder, or
        None when it is not enabled."""
        nb, nsh = ctypes.c_int64(), ctypes.c_int64()
        E, g, e = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._c(load().bghip_recorder_info(self.h, ctypes.byref(nb), ctypes.byref(nsh), ctypes.byref(E),
                                           ctypes.byref(g), ctypes.byref(e)), "recorder_info")
        return None if e.value < 0 else (nb.value, nsh.value, E.value, bool(g.value), e.value)

    # ---- uploads
    def upload_neurons(self, arrs: dict):
        nb = Neurons()
        keep = []
        for f in NEURON_FIELDS_F + IZH_FIELDS:
            if f in arrs and arrs[f] is not None:
                a = np.ascontiguousarray(arrs[f], dtype=np.float32)
                keep.append(a)
                setattr(nb, f, _ptr(a, _fp))
        for f in ("nStepsInRefr", "spikeCount", "spikeCountOffset"):
            if f in arrs:
                a = np.ascontiguousarray(arrs[f], dtype=np.int32)
                keep.append(a)
                setattr(nb, f, _ptr(a, _ip))
        if "spike_history" in arrs:
            a = np.ascontiguousarray(arrs["spike_history"], dtype=np.uint64)
            keep.append(a)
            nb.spike_history = _ptr(a, _qp)
        self._c(load().bghip_upload_neurons(self.h, ctypes.byref(nb)), "upload_neurons")

    def upload_synapses(self, arrs: dict, step: int):
        sb = Synapses()
        keep = []
        for f, dt in SYN_DTYPES.items():
            if f in arrs and arrs[f] is not None:
                a = np.ascontiguousarray(arrs[f], dtype=dt)
                keep.append(a)
                setattr(sb, f, _ptr(a, _CT[dt]))
        a = np.ascontiguousarray(arrs["synapse_counts"], dtype=np.uint32)
        keep.append(a)
        sb.synapse_counts = _ptr(a, _up)
        self._c(load().bghip_upload_synapses(self.h, ctypes.byref(sb), step), "upload_synapses")

    def upload_layout(self, xloc, yloc, ntype):
        x = np.ascontiguousarray(xloc, np.float32)
        y = np.ascontiguousarray(yloc, np.float32)
        t = np.ascontiguousarray(ntype, np.int32)
        self._c(load().bghip_upload_layout(self.h, _ptr(x, _fp), _ptr(y, _fp), _ptr(t, _ip)), "upload_layout")

    def upload_growth(self, g: Growth, radii, rates=None):
        r = np.ascontiguousarray(radii, np.float32)
        q = np.ascontiguousarray(rates, np.float32) if rates is not None else None
        self._c(load().bghip_upload_growth(self.h, ctypes.byref(g), _ptr(r, _fp), _ptr(q, _fp)), "upload_growth")

    # ---- simulation
    def advance(self, n_steps: int):
        self._c(load().bghip_advance(self.h, n_steps), "advance")

    def update_connections(self) -> GrowthStats:
        st = GrowthStats()
        self._c(load().bghip_update_connections(self.h, ctypes.byref(st)), "update_connections")
        return st

    def update_weights(self, W: np.ndarray) -> GrowthStats:
        """ConnGrowth::updateSynapsesWeights with W[src, dest] in place of the overlap areas."""
        W = np.ascontiguousarray(W, dtype=np.float32)
        st = GrowthStats()
        self._c(load().bghip_update_weights(self.h, _ptr(W, _fp), ctypes.byref(st)), "update_weights")
        return st

    def update_weights_static(self, W: np.ndarray) -> GrowthStats:
        """update_weights without device allocation (bghip_update_weights_static: the matrix goes
        through the library's static per-GPU scratch, so N <= 2048)."""
        W = np.ascontiguousarray(W, dtype=np.float32)
        st = GrowthStats()
        self._c(load().bghip_update_weights_static(self.h, _ptr(W, _fp), ctypes.byref(st)), "update_weights_static")
        return st

    def update_history(self):
        self._c(load().bghip_update_history(self.h), "update_history")

    def sync(self):
        self._c(load().bghip_sync(self.h), "sync")

    def wait(self):
        """Waits for the network's queued work; device errors are left for sync / error_bits."""
        self._c(load().bghip_wait(self.h), "wait")

    def error_bits(self) -> int:
        return load().bghip_net_error_bits(self.h)

    @property
    def step(self) -> int:
        return load().bghip_net_step(self.h)

    @property
    def implicit(self) -> bool:
        """True when the network runs the implicit-ring kernels (bghip_net_implicit)."""
        return bool(load().bghip_net_implicit(self.h))

    @property
    def total_synapses(self) -> int:
        return load().bghip_total_synapses(self.h)

    # ---- downloads
    def download_neurons(self, max_spikes: int) -> dict:
        N = self.N
        izh = self.desc.neuron_model == NEURON_IZH
        out = {f: np.zeros(N, np.float32) for f in NEURON_FIELDS_F}
        if izh:
            out.update({f: np.zeros(N, np.float32) for f in IZH_FIELDS})
        for f in ("nStepsInRefr", "spikeCount", "spikeCountOffset"):
            out[f] = np.zeros(N, np.int32)
        out["spike_history"] = np.zeros(N * max_spikes, np.uint64)
        nb = Neurons()
        for f, a in out.items():
            setattr(nb, f, _ptr(a, _CT[a.dtype.type]))
        self._c(load().bghip_download_neurons(self.h, ctypes.byref(nb)), "download_neurons")
        return out

    def download_spikes(self) -> dict:
        """spikeCount, spikeCountOffset and spike_history (bghip_download_spikes)."""
        N, ms = self.N, self.desc.max_spikes
        out = {"spikeCount": np.zeros(N, np.int32), "spikeCountOffset": np.zeros(N, np.int32),
               "spike_history": np.zeros(N * ms, np.uint64)}
        self._c(load().bghip_download_spikes(self.h, _ptr(out["spikeCount"], _ip), _ptr(out["spikeCountOffset"], _ip),
                                             _ptr(out["spike_history"], _qp)), "download_spikes")
        return out

    def clear_spike_counts(self):
        self._c(load().bghip_clear_spike_counts(self.h), "clear_spike_counts")

    def download_synapses(self) -> dict:
        S = self.N * self.M
        sk = self.desc.synapse_model
        fields = ["in_use", "sourceNeuronIndex", "destNeuronIndex", "type", "W", "psr", "decay", "total_delay",
                  "delayIdx", "delayQueue"]
        if sk in (SYN_DS, SYN_DYNSTDP):
            fields += list(DS_FIELDS)
        if sk in (SYN_STDP, SYN_DYNSTDP):
            fields += list(STDP_FIELDS)
        out = {}
        sb = Synapses()
        for f in fields:
            n = S * self.W32 if f == "delayQueue" else S
            out[f] = np.zeros(n, SYN_DTYPES[f])
            setattr(sb, f, _ptr(out[f], _CT[SYN_DTYPES[f]]))
        out["synapse_counts"] = np.zeros(self.N, np.uint32)
        sb.synapse_counts = _ptr(out["synapse_counts"], _up)
        self._c(load().bghip_download_synapses(self.h, ctypes.byref(sb)), "download_synapses")
        return out

    def download_index_map(self) -> dict:
        T = self.total_synapses
        N = self.N
        out = {"incomingSynapseIndexMap": np.zeros(T, np.uint32), "incomingSynapseBegin": np.zeros(N, np.uint32),
               "incomingSynapseCount": np.zeros(N, np.uint32), "outgoingSynapseIndexMap": np.zeros(T, np.uint32),
               "outgoingSynapseBegin": np.zeros(N, np.uint32), "outgoingSynapseCount": np.zeros(N, np.uint32)}
        im = IndexMap()
        for f, a in out.items():
            setattr(im, f, _ptr(a, _up))
        self._c(load().bghip_download_index_map(self.h, ctypes.byref(im)), "download_index_map")
        return out

    def download_growth(self):
        r = np.zeros(self.N, np.float32)
        q = np.zeros(self.N, np.float32)
        self._c(load().bghip_download_growth(self.h, _ptr(r, _fp), _ptr(q, _fp)), "download_growth")
        return r, q

    def recorder_enable(self, n_burst: int, n_spikes_hist: int, num_epochs: int, with_growth: bool = False,
                        rates0=None, radii0=None):
        """bghip_recorder_enable: spike histograms (and, with_growth, rates/radii rows) kept on the device."""
        r0 = None if rates0 is None else np.ascontiguousarray(rates0, np.float32)
        q0 = None if radii0 is None else np.ascontiguousarray(radii0, np.float32)
        self._c(load().bghip_recorder_enable(self.h, n_burst, n_spikes_hist, num_epochs, 1 if with_growth else 0,
                                             _ptr(r0, _fp), _ptr(q0, _fp)), "recorder_enable")

    def recorder_download(self) -> dict:
        info = self.recorder_info()
        if info is None:
            raise BgHipError("recorder_download: the recorder is not enabled")
        nb, nsh, E, growth, _ = info
        out = {"burstinessHist": np.zeros(nb, np.float32), "spikesHistory": np.zeros(nsh, np.float32)}
        if growth:
            out["ratesHistory"] = np.zeros((E + 1) * self.N, np.float32)
            out["radiiHistory"] = np.zeros((E + 1) * self.N, np.float32)
        self._c(load().bghip_recorder_download(self.h, _ptr(out["burstinessHist"], _fp), _ptr(out["spikesHistory"], _fp),
                                               _ptr(out.get("ratesHistory"), _fp), _ptr(out.get("radiiHistory"), _fp)),
                "recorder_download")
        return out

    def upload_input(self, kind: int, neuron_type=None, fr_mean: float = 0.0, weight: float = 0.0, mask=None,
                     duration: float = 0.0, interval: float = 0.0, values=None, shift=None, state: dict | None = None):
        """bghip_upload_input: SInputPoisson (fr_mean Hz, weight x 1e-8, optional 0/1 mask) or
        SInputRegular (duration/interval s, per-neuron values and shifts)."""
        inp = Input()
        keep = []

        def arr(a, dt, ct):
            if a is None:
                return ctypes.cast(None, ct)
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return _ptr(a, ct)
        inp.kind, inp.fr_mean, inp.weight = kind, fr_mean, weight
        inp.mask = arr(mask, np.uint8, _bp)
        inp.neuron_type = arr(neuron_type, np.int32, _ip)
        inp.duration, inp.interval = duration, interval
        inp.values = arr(values, np.float32, _fp)
        inp.shift = arr(shift, np.int32, _ip)
        st = None
        if state is not None:
            st = InputState()
            for f, dt in INPUT_STATE_FIELDS.items():
                if state.get(f) is not None:
                    setattr(st, f, arr(state[f], dt, _CT[dt]))
            st.steps_in_cycle = int(state.get("steps_in_cycle", 0))
        self._c(load().bghip_upload_input(self.h, ctypes.byref(inp), ctypes.byref(st) if st is not None else None),
                "upload_input")

    def download_input(self) -> dict:
        out = {f: np.zeros(self.N, dt) for f, dt in INPUT_STATE_FIELDS.items()}
        st = InputState()
        for f, a in out.items():
            setattr(st, f, _ptr(a, _CT[a.dtype.type]))
        self._c(load().bghip_download_input(self.h, ctypes.byref(st)), "download_input")
        out["steps_in_cycle"] = st.steps_in_cycle
        return out

    def timing(self, enable: bool):
        avg = ctypes.c_double()
        cnt = ctypes.c_int64()
        self._c(load().bghip_kernel_timing(self.h, 1 if enable else 0, ctypes.byref(avg), ctypes.byref(cnt)),
                "kernel_timing")
        return avg.value, cnt.value


class Batch:
    """Many networks advanced by shared launches (bghip_batch_*)."""

    def __init__(self, ctx: Context):
        h = ctypes.c_void_p()
        ctx.check(load().bghip_batch_create(ctx.h, ctypes.byref(h)), "batch_create")
        self.ctx, self.h, self.members = ctx, h, []

    def set(self, nets):
        self.members = list(nets)
        arr = (ctypes.c_void_p * max(1, len(nets)))(*[n.h.value if isinstance(n.h, ctypes.c_void_p) else n.h
                                                      for n in nets])
        self.ctx.check(load().bghip_batch_set(self.h, arr, len(nets)), "batch_set")

    def advance(self, n_steps: int):
        self.ctx.check(load().bghip_batch_advance(self.h, n_steps), "batch_advance")

    def sync(self):
        self.ctx.check(load().bghip_batch_sync(self.h), "batch_sync")

    def autotune(self, trial_steps: int):
        """Turns the launch-shape autotuner on (trial_steps timed steps per candidate) or off (0)."""
        self.ctx.check(load().bghip_batch_autotune(self.h, trial_steps), "batch_autotune")

    def shapes(self) -> list[dict]:
        """Launch shape of every group, as dicts of the bghip_batch_shape fields."""
        n = load().bghip_batch_group_count(self.h)
        if n < 0:
            self.ctx.check(n, "batch_group_count")
        out = []
        for i in range(n):
            s = BatchShape()
            self.ctx.check(load().bghip_batch_get_shape(self.h, i, ctypes.byref(s)), "batch_get_shape")
            out.append({f: getattr(s, f) for f, _ in BatchShape._fields_})
        return out

    def set_shape(self, group: int, **fields):
        """Sets group `group`'s launch shape; fields not given keep their current values."""
        s = BatchShape()
        self.ctx.check(load().bghip_batch_get_shape(self.h, group, ctypes.byref(s)), "batch_get_shape")
        for k, v in fields.items():
            setattr(s, k, v)
        self.ctx.check(load().bghip_batch_set_shape(self.h, group, ctypes.byref(s)), "batch_set_shape")

    def close(self):
        if self.h:
            load().bghip_batch_destroy(self.h)
            self.h = None


def footprint(desc: NetDesc) -> int:
    return load().bghip_net_footprint(ctypes.byref(desc))
