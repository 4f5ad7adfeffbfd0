"""Host-side helpers shared by the Python drivers (tests, bench, serving loop).

`synapse_types` fills the per-type creation table of the C ABI exactly as the reference's
createSynapse chain assigns per-type constants (Synapses/AllSpikingSynapses.cpp:191-233,
Synapses/AllDSSynapses.cpp:160-218, Synapses/AllSTDPSynapses.cpp:252-273). The decay is the
host C library's expf(-dt/tau) (AllSpikingSynapses.cpp:344-354), called through ctypes so the bits
match the reference's host computation.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

from . import bghip as B

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]

F32 = np.float32

TAU = (6e-3, 6e-3, 3e-3, 3e-3)          # II, IE, EI, EE
DELAY = (0.8e-3, 0.8e-3, 0.8e-3, 1.5e-3)
DS_U = (0.32, 0.25, 0.05, 0.5)
DS_D = (0.144, 0.7, 0.125, 1.1)
DS_F = (0.06, 0.02, 1.2, 0.05)
STDP_DEFAULTS = dict(Apos=0.5, Aneg=-0.5, STDPgap=2e-3, tauspost=75e-3, tauspre=34e-3, taupos=15e-3,
                     tauneg=35e-3, Wex=1.0, mupos=0.0, muneg=0.0, useFroemkeDanSTDP=1)


def expf(x) -> np.float32:
    return F32(_libm.expf(float(F32(x))))


def total_delay(delay_s: float, dt: float) -> int:
    """static_cast<int>(delay / deltaT) + 1 in float (AllSpikingSynapses.cpp:227)."""
    return int(F32(F32(delay_s) / F32(dt))) + 1


def synapse_types(dt: float, stdp: dict | None = None, delay_override=None) -> B.SynapseTypes:
    t = B.SynapseTypes()
    dtf = F32(dt)
    for k in range(4):
        tau = F32(TAU[k])
        t.tau[k] = tau
        t.decay[k] = expf(F32(-dtf / tau))
        d = delay_override[k] if delay_override and delay_override[k] > 0 else DELAY[k]
        t.total_delay[k] = total_delay(d, dt)
        t.U[k], t.D[k], t.F[k] = F32(DS_U[k]), F32(DS_D[k]), F32(DS_F[k])
    t.u0, t.r0 = F32(0.4), F32(1.0)
    s = dict(STDP_DEFAULTS)
    if stdp:
        s.update(Apos=stdp["Apos"], Aneg=stdp["Aneg"], taupos=stdp["taupos"], tauneg=stdp["tauneg"],
                 tauspre=stdp.get("tauspre", 34e-3), tauspost=stdp.get("tauspost", 75e-3),
                 STDPgap=stdp.get("gap", 2e-3), Wex=stdp.get("Wex", 1.0), mupos=stdp.get("mupos", 0.0),
                 muneg=stdp.get("muneg", 0.0), useFroemkeDanSTDP=1 if stdp.get("froemke_dan", False) else 0)
    for k, v in s.items():
        setattr(t, k, F32(v) if k != "useFroemkeDanSTDP" else int(v))
    return t


def net_desc(N: int, M: int, neuron_model: int, synapse_model: int, max_spikes: int, dt: float = 1e-4,
             seed: int = 1, ring_bits: int = 32, block_steps: int = 0, stdp=None, delay_override=None) -> B.NetDesc:
    d = B.NetDesc()
    d.num_neurons, d.max_synapses_per_neuron = N, M
    d.neuron_model, d.synapse_model = neuron_model, synapse_model
    d.ring_bits, d.max_spikes, d.dt = ring_bits, max_spikes, F32(dt)
    d.noise_seed, d.block_steps = seed, block_steps
    d.types = synapse_types(dt, stdp, delay_override)
    return d


def growth_params(epsilon, beta, rho, target_rate, min_radius, start_radius, epoch_duration) -> B.Growth:
    g = B.Growth()
    g.epsilon, g.beta, g.rho = F32(epsilon), F32(beta), F32(rho)
    g.targetRate, g.minRadius, g.startRadius = F32(target_rate), F32(min_radius), F32(start_radius)
    g.maxRate = F32(F32(target_rate) / F32(epsilon))   # Connections/ConnGrowth.cpp:201-203
    g.epochDuration = F32(epoch_duration)
    return g
