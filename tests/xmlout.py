"""Writes recorder output in the byte format of XmlGrowthRecorder::saveSimData
(Recorders/XmlGrowthRecorder.cpp:104-154) and, without the radii/rates histories, of
XmlRecorder::saveSimData (Recorders/XmlRecorder.cpp:111-158), with Matrix/CompleteMatrix.cpp:225-249
and Matrix/VectorMatrix.cpp:173-190: ostream's default float formatting is C's "%g"."""
import numpy as np


def _g(v) -> str:
    return "%g" % float(v)


def matrix(name, rows):
    s = f'   <Matrix name="{name}" type="complete" rows="{len(rows)}" columns="{len(rows[0])}" multiplier="1.0">\n   '
    for r in rows:
        s += "".join(_g(v) + " " for v in r) + "\n"
    return s + "\n</Matrix>\n"


def vector(name, v):
    s = f'   <Matrix name="{name}" type="complete" rows="1" columns="{len(v)}" multiplier="1.0">\n   '
    return s + "".join(_g(x) + " " for x in v) + "\n</Matrix>\n"


def growth_xml(radii_hist, rates_hist, burst, spikes_hist, xloc, yloc, ntype, starter_map, vthresh, tsim, end_time):
    out = '<?xml version="1.0" standalone="no"?>\n<!-- State output file for the DCT growth modeling-->\n<SimState>\n'
    if radii_hist is not None:
        out += matrix("radiiHistory", radii_hist) + matrix("ratesHistory", rates_hist)
    out += vector("burstinessHist", burst) + vector("spikesHistory", spikes_hist)
    out += vector("xloc", xloc) + vector("yloc", yloc) + vector("neuronTypes", ntype)
    st = np.nonzero(starter_map)[0]
    if len(st):
        out += vector("starterNeurons", st)
    out += vector("neuronThresh", vthresh)
    out += f'   <Matrix name="Tsim" type="complete" rows="1" columns="1" multiplier="1.0">\n   {_g(tsim)}\n</Matrix>\n'
    out += (f'   <Matrix name="simulationEndTime" type="complete" rows="1" columns="1" multiplier="1.0">\n'
            f'   {_g(end_time)}\n</Matrix>\n')
    return out + "</SimState>\n"


def oracle_growth_xml(net, cfg) -> str:
    N = cfg["width"] * cfg["height"]
    E = cfg["num_epochs"]
    f32 = np.float32
    end = f32(f32(net.step) * f32(cfg["dt"]))
    growth = cfg["conn"] == "ConnGrowth"
    return growth_xml(net.array("radiiHistory").reshape(E + 1, N) if growth else None,
                      net.array("ratesHistory").reshape(E + 1, N) if growth else None,
                      net.array("burstinessHist"), net.array("spikesHistory"), net.array("xloc"),
                      net.array("yloc"), net.array("neuron_type_map"), net.array("starter_map"),
                      net.array("Vthresh"), f32(cfg["epoch_duration"]), end)
