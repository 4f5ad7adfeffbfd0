"""Writes the small simulation configs used to pin the oracle against the compiled reference.

Each config is a BrainGrid XML parameter file (the format of configfiles/test-small.xml) chosen so
that synapses exist for most of the run: ConnGrowth with a start radius above 0.5 connects
grid neighbours at the first growth update, so epochs 2.. exercise the delay ring, the synapse
advance, the summation and further growth. Run from the repository root:
    python tests/golden/make_configs.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "configs")

LIF = """    <NeuronsParams class="AllLIFNeurons" name="NeuronsParams">
      <Iinject name="Iinject"><min name="min">13.5e-09</min><max name="max">13.5e-09</max></Iinject>
      <Inoise name="Inoise"><min name="min">1.0e-09</min><max name="max">1.5e-09</max></Inoise>
      <Vthresh name="Vthresh"><min name="min">15.0e-03</min><max name="max">15.0e-03</max></Vthresh>
      <Vresting name="Vresting"><min name="min">0.0</min><max name="max">0.0</max></Vresting>
      <Vreset name="Vreset"><min name="min">13.5e-03</min><max name="max">13.5e-03</max></Vreset>
      <Vinit name="Vinit"><min name="min">13.0e-03</min><max name="max">13.0e-03</max></Vinit>
      <starter_vthresh name="starter_vthresh"><min name="min">13.565e-3</min><max name="max">13.655e-3</max></starter_vthresh>
      <starter_vreset name="starter_vreset"><min name="min">13.0e-3</min><max name="max">13.0e-3</max></starter_vreset>
    </NeuronsParams>
"""
IZH = """    <NeuronsParams class="AllIZHNeurons" name="NeuronsParams">
      <Iinject name="Iinject"><min name="min">13.5e-09</min><max name="max">13.5e-09</max></Iinject>
      <Inoise name="Inoise"><min name="min">0.5e-06</min><max name="max">0.7329e-06</max></Inoise>
      <Vthresh name="Vthresh"><min name="min">30.0e-03</min><max name="max">30.0e-03</max></Vthresh>
      <Vresting name="Vresting"><min name="min">0.0</min><max name="max">0.0</max></Vresting>
      <Vreset name="Vreset"><min name="min">-0.065</min><max name="max">-0.065</max></Vreset>
      <Vinit name="Vinit"><min name="min">-0.065</min><max name="max">-0.065</max></Vinit>
      <starter_vthresh name="starter_vthresh"><min name="min">13.565e-3</min><max name="max">13.655e-3</max></starter_vthresh>
      <starter_vreset name="starter_vreset"><min name="min">13.0e-3</min><max name="max">13.0e-3</max></starter_vreset>
      <Aconst name="Aconst"><minExc name="minExc">0.02</minExc><maxExc name="maxExc">0.02</maxExc><minInh name="minInh">0.02</minInh><maxInh name="maxInh">0.1</maxInh></Aconst>
      <Bconst name="Bconst"><minExc name="minExc">0.2</minExc><maxExc name="maxExc">0.2</maxExc><minInh name="minInh">0.2</minInh><maxInh name="maxInh">0.25</maxInh></Bconst>
      <Cconst name="Cconst"><minExc name="minExc">-65</minExc><maxExc name="maxExc">-50</maxExc><minInh name="minInh">-65</minInh><maxInh name="maxInh">-65</maxInh></Cconst>
      <Dconst name="Dconst"><minExc name="minExc">2</minExc><maxExc name="maxExc">8</maxExc><minInh name="minInh">2</minInh><maxInh name="maxInh">2</maxInh></Dconst>
    </NeuronsParams>
"""


def fixed_layout(active, inh):
    return f"""    <LayoutParams class="FixedLayout" name="LayoutParams">
      <LayoutFiles name="LayoutFiles">
        <activeNListFileName name="activeNListFileName" type="InputFile">{active}</activeNListFileName>
        <inhNListFileName name="inhNListFileName" type="InputFile">{inh}</inhNListFileName>
        <probedNListFileName name="prbNListFileName" type="InputFile"/>
      </LayoutFiles>
    </LayoutParams>
"""


def dynamic_layout(fexc, fst):
    return f"""    <LayoutParams class="DynamicLayout" name="LayoutParams">
      <LayoutFiles name="LayoutFiles">
        <frac_EXC name="frac_EXC">{fexc}</frac_EXC>
        <starter_neurons name="starter_neurons">{fst}</starter_neurons>
      </LayoutFiles>
    </LayoutParams>
"""


def growth_conn(start_radius):
    return f"""    <ConnectionsParams class="ConnGrowth" name="ConnectionsParams">
      <GrowthParams name="GrowthParams">
        <epsilon name="epsilon">0.60</epsilon>
        <beta name="beta">0.10</beta>
        <rho name="rho">0.0001</rho>
        <targetRate name="targetRate">1.9</targetRate>
        <minRadius name="minRadius">0.1</minRadius>
        <startRadius name="startRadius">{start_radius}</startRadius>
      </GrowthParams>
    </ConnectionsParams>
"""


def static_conn(nconns, thresh, exc=(0, 0.5e-7), inh=(-0.5e-7, 0)):
    """ConnStatic (the block of configfiles/static_izh_1000.xml)."""
    return f"""    <ConnectionsParams class="ConnStatic" name="ConnectionsParams">
      <StaticConnectionsParams name="StaticConnectionsParams">
        <nConnsPerNeuron name="nConnsPerNeuron">{nconns}</nConnsPerNeuron>
        <threshConnsRadius name="threshConnsRadius">{thresh}</threshConnsRadius>
        <pRewiring name="pRewiring">0</pRewiring>
      </StaticConnectionsParams>
      <StaticConnectionsWeight name="StaticConnectionsWeight">
        <minExc name="minExc">{exc[0]}</minExc>
        <maxExc name="maxExc">{exc[1]}</maxExc>
        <minInh name="minInh">{inh[0]}</minInh>
        <maxInh name="maxInh">{inh[1]}</maxInh>
      </StaticConnectionsWeight>
    </ConnectionsParams>
"""


def config(name, x, y, tsim, nsims, maxsyn, neurons, synapse, layout, start_radius=None, seed=1, max_rate=200,
           conn=None):
    conn = conn if conn is not None else growth_conn(start_radius)
    return f"""<?xml version="1.0" encoding="UTF-8" standalone="no"?>
<BGSimParams>
  <SimInfoParams name="SimInfoParams">
    <PoolSize name="PoolSize"><x name="x">{x}</x><y name="y">{y}</y><z name="z">1</z></PoolSize>
    <SimParams name="SimParams"><Tsim name="Tsim">{tsim}</Tsim><numSims name="numSims">{nsims}</numSims></SimParams>
    <SimConfig name="SimConfig"><maxFiringRate name="maxFiringRate">{max_rate}</maxFiringRate><maxSynapsesPerNeuron name="maxSynapsesPerNeuron">{maxsyn}</maxSynapsesPerNeuron></SimConfig>
    <Seed name="Seed"><value name="value">{seed}</value></Seed>
    <OutputParams name="OutputParams"><stateOutputFileName name="stateOutputFileName">{name}-out.xml</stateOutputFileName></OutputParams>
  </SimInfoParams>
  <ModelParams>
{neurons}    <SynapsesParams class="{synapse}" name="SynapsesParams">
    </SynapsesParams>
{conn}{layout}  </ModelParams>
</BGSimParams>
"""


NLISTS = {
    "NList/A5x5.xml": ("A", [3, 12, 21]),
    "NList/I5x5.xml": ("I", [7, 18]),
    "NList/A6x6.xml": ("A", [2, 14, 22, 33]),
    "NList/I6x6.xml": ("I", [5, 19, 30]),
    "NList/A10x10.xml": ("A", [4, 15, 26, 37, 48, 51, 62, 73, 84, 95]),
    "NList/I10x10.xml": ("I", [9, 18, 27, 33, 45, 56, 60, 71, 82, 90]),
}

CONFIGS = {
    # 4-neighbour DS synapses from the first growth update on
    "conn-lif-ds": dict(x=5, y=5, tsim=0.5, nsims=4, maxsyn=40, neurons=LIF, synapse="AllDSSynapses",
                        layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.6),
    # up to ~28 DS synapses per neuron, complete and partial overlaps
    "conn-lif-ds-dense": dict(x=6, y=6, tsim=0.4, nsims=3, maxsyn=40, neurons=LIF, synapse="AllDSSynapses",
                              layout=fixed_layout("NList/A6x6.xml", "NList/I6x6.xml"), start_radius=1.6),
    "conn-lif-spiking": dict(x=5, y=5, tsim=0.5, nsims=3, maxsyn=40, neurons=LIF, synapse="AllSpikingSynapses",
                             layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.8),
    "conn-lif-stdp": dict(x=5, y=5, tsim=0.5, nsims=4, maxsyn=40, neurons=LIF, synapse="AllSTDPSynapses",
                          layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.9),
    "conn-lif-dynstdp": dict(x=5, y=5, tsim=0.5, nsims=4, maxsyn=40, neurons=LIF, synapse="AllDynamicSTDPSynapses",
                             layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.9),
    "conn-izh-ds": dict(x=5, y=5, tsim=0.5, nsims=3, maxsyn=40, neurons=IZH, synapse="AllDSSynapses",
                        layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.9, max_rate=1000),
    "conn-izh-stdp": dict(x=5, y=5, tsim=0.5, nsims=3, maxsyn=40, neurons=IZH, synapse="AllSTDPSynapses",
                          layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.9, max_rate=1000),
    # DynamicLayout (fractions 0.9 / 0.1): the reference harness swaps the layout object in, see
    # oracle/ref_dump.cpp; the XML itself carries an empty FixedLayout
    "conn-lif-ds-dynamic": dict(x=8, y=8, tsim=0.5, nsims=3, maxsyn=40, neurons=LIF, synapse="AllDSSynapses",
                                layout=fixed_layout("", ""), start_radius=1.2, seed=7),
}

# ConnStatic nearest-k (Connections/ConnStatic.cpp:33-88) on the 10x10 grid: interior neurons have
# 4 neighbours at distance 1, 4 at sqrt 2, 4 at 2 and 8 at sqrt 5, so k = 10 and k = 14 cut through
# a group of equal distances and the standard library's tie order decides which neighbours connect
# and which weight draw each one gets; the all-to-all case is configfiles/static_izh_1000.xml's
# (every other neuron within the radius, k = N - 1) on 8x8.
STATIC_CONFIGS = {
    "static-izh-spiking": dict(x=10, y=10, tsim=0.5, nsims=2, maxsyn=40, neurons=IZH, synapse="AllSpikingSynapses",
                               layout=fixed_layout("NList/A10x10.xml", "NList/I10x10.xml"), max_rate=1000,
                               conn=static_conn(10, 2.5)),
    "static-lif-ds": dict(x=10, y=10, tsim=0.5, nsims=2, maxsyn=40, neurons=LIF, synapse="AllDSSynapses",
                          layout=fixed_layout("NList/A10x10.xml", "NList/I10x10.xml"), seed=3,
                          conn=static_conn(14, 3.0, exc=(0, 2.0e-7), inh=(-1.0e-7, 0))),
    "static-izh-all": dict(x=8, y=8, tsim=0.5, nsims=2, maxsyn=64, neurons=IZH, synapse="AllSpikingSynapses",
                           layout=fixed_layout("", ""), max_rate=1000, seed=777, conn=static_conn(63, 50)),
}
CONFIGS.update(STATIC_CONFIGS)


# Stimulus inputs (Inputs/SInputPoisson.cpp, Inputs/SInputRegular.cpp). The reference's CPU Poisson
# input indexes its one-slot-per-neuron synapse object with maxSynapsesPerNeuron * neuron, which
# stays inside the object only for maxSynapsesPerNeuron = 1, so the Poisson configs use M = 1 and
# a start radius that forms no synapses.
def poisson(fr, weight, mask=None):
    m = f"""   <Masks><M>{" ".join(map(str, mask))}</M></Masks>\n""" if mask else ""
    return f"""<?xml version="1.0" standalone="no"?>
<InputParams>
   <IMethod name="SInputPoisson"/>
   <IntParams fr_mean="{fr}" weight="{weight}"/>
{m}</InputParams>
"""


def regular(duration, interval, sync, on):
    vals = ["1e-8" if k in on else "0.0" for k in range(100)]
    rows = "\n".join("        " + " ".join(vals[r * 10:(r + 1) * 10]) for r in range(10))
    return f"""<?xml version="1.0" standalone="no"?>
<InputParams>
   <IMethod name="SInputRegular"/>
   <IntParams duration="{duration}" interval="{interval}" sync="{sync}"/>
   <Values>
    <I>
{rows}
    </I>
   </Values>
</InputParams>
"""


STIMULI = {
    "stim/poisson-50hz-w25-mask.xml": poisson(50, 25, [3, 7, 11, 12, 13, 40, 41, 57, 58, 59, 60, 61, 62, 63, 64, 65,
                                                       77, 80, 88, 99]),
    "stim/poisson-10hz-w25.xml": poisson(10, 25),
    "stim/regular-5ms-20ms-async.xml": regular(0.005, 0.020, "no", {0, 5, 23, 27, 44, 45, 54, 55, 72, 77, 99}),
    "stim/regular-10ms-30ms-wave.xml": regular(0.010, 0.030, "wave", set(range(0, 100, 3))),
}

INPUT_CONFIGS = {
    "input-lif-poisson": (dict(x=10, y=10, tsim=0.5, nsims=2, maxsyn=1, neurons=LIF, synapse="AllDSSynapses",
                               layout=fixed_layout("NList/A10x10.xml", "NList/I10x10.xml"), start_radius=0.4),
                          "stim/poisson-50hz-w25-mask.xml"),
    "input-izh-poisson": (dict(x=5, y=5, tsim=0.5, nsims=2, maxsyn=1, neurons=IZH, synapse="AllDSSynapses",
                               layout=fixed_layout("NList/A5x5.xml", "NList/I5x5.xml"), start_radius=0.4,
                               max_rate=1000), "stim/poisson-10hz-w25.xml"),
    "input-lif-regular-async": (dict(x=10, y=10, tsim=0.4, nsims=3, maxsyn=40, neurons=LIF, synapse="AllDSSynapses",
                                     layout=fixed_layout("NList/A10x10.xml", "NList/I10x10.xml"), start_radius=0.6),
                                "stim/regular-5ms-20ms-async.xml"),
    "input-lif-regular-wave": (dict(x=10, y=10, tsim=0.5, nsims=2, maxsyn=40, neurons=LIF,
                                    synapse="AllDynamicSTDPSynapses",
                                    layout=fixed_layout("NList/A10x10.xml", "NList/I10x10.xml"), start_radius=0.9,
                                    max_rate=1000), "stim/regular-10ms-30ms-wave.xml"),
}


def main():
    os.makedirs(os.path.join(OUT, "NList"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "stim"), exist_ok=True)
    for rel, text in STIMULI.items():
        with open(os.path.join(OUT, rel), "w") as f:
            f.write(text)
    for name, (kw, _stim) in INPUT_CONFIGS.items():
        with open(os.path.join(OUT, name + ".xml"), "w") as f:
            f.write(config(name, **kw))
    for rel, (tag, vals) in NLISTS.items():
        with open(os.path.join(OUT, rel), "w") as f:
            f.write(f'<?xml version="1.0" encoding="UTF-8"?>\n<{tag}>{" ".join(map(str, vals))}</{tag}>\n')
    for name, kw in CONFIGS.items():
        with open(os.path.join(OUT, name + ".xml"), "w") as f:
            f.write(config(name, **kw))
    print("wrote", len(CONFIGS) + len(INPUT_CONFIGS), "configs to", OUT)


if __name__ == "__main__":
    main()
