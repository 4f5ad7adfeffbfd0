/*
 * HipModelAdapter: IModel on the MI355X engine. Life cycle and array mapping as described in
 * HipModelAdapter.h; each method cites the GPUSpikingModel method it replaces.
 */
#include "HipModelAdapter.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>
#include <unistd.h>

#include "AllDSSynapses.h"
#include "AllDynamicSTDPSynapses.h"
#include "AllIZHNeurons.h"
#include "AllLIFNeurons.h"
#include "AllSTDPSynapses.h"
#include "AllSpikingNeurons.h"
#include "AllSpikingSynapses.h"
#include "ConnGrowth.h"
#include "Layout.h"

HipModelAdapter::HipModelAdapter(Connections *conns, IAllNeurons *neurons, IAllSynapses *synapses, Layout *layout,
                                 int device)
    : Model(conns, neurons, synapses, layout), m_device(device), m_ctx(NULL), m_net(NULL), m_pending(0),
      m_step_base(0), m_max_spikes(0), m_total_synapses(0)
{
}

HipModelAdapter::~HipModelAdapter()
{
    if (m_net) bghip_net_destroy(m_net);
    if (m_ctx) bghip_ctx_destroy(m_ctx);
}

/* Errors end the run, as HANDLE_ERROR does in the reference (Utils/Book.h:21-30). */
void HipModelAdapter::check(int rc, const char *what) const
{
    if (rc == BGHIP_OK) return;
    fprintf(stderr, "! ERROR: %s failed (%d): %s\n", what, rc, m_ctx ? bghip_last_error(m_ctx) : "no context");
    exit(1);
}

/* Per-type creation constants of createSynapse (Synapses/AllSpikingSynapses.cpp:191-233,
 * Synapses/AllDSSynapses.cpp:180-218, Synapses/AllSTDPSynapses.cpp:252-273), so that synapses the
 * device creates during growth carry the values the host classes would give them. decay is the
 * host's exp(-dt/tau) (Synapses/AllSpikingSynapses.cpp:344-354). */
void HipModelAdapter::fillTypes(bghip_synapse_types &t, BGFLOAT deltaT) const
{
    memset(&t, 0, sizeof t);
    const BGFLOAT tau[4] = {6e-3, 6e-3, 3e-3, 3e-3};          /* II, IE, EI, EE */
    const BGFLOAT delay[4] = {0.8e-3, 0.8e-3, 0.8e-3, 1.5e-3};
    const BGFLOAT U[4] = {0.32, 0.25, 0.05, 0.5};
    const BGFLOAT D[4] = {0.144, 0.7, 0.125, 1.1};
    const BGFLOAT F[4] = {0.06, 0.02, 1.2, 0.05};
    for (int k = 0; k < 4; k++) {
        t.tau[k] = tau[k];
        t.decay[k] = expf(-deltaT / tau[k]);
        t.total_delay[k] = (int)(delay[k] / deltaT) + 1;
        t.U[k] = U[k];
        t.D[k] = D[k];
        t.F[k] = F[k];
    }
    t.u0 = DEFAULT_U;                       /* Utils/Global.h:128, as Synapses/AllDSSynapses.cpp:164 */
    t.r0 = 1.0;
    t.Apos = 0.5; t.Aneg = -0.5; t.STDPgap = 2e-3; t.tauspost = 75e-3; t.tauspre = 34e-3;
    t.taupos = 15e-3; t.tauneg = 35e-3; t.Wex = 1.0; t.mupos = 0; t.muneg = 0;
    t.useFroemkeDanSTDP = 1;
}

/* GPUSpikingModel::setupSim (Core/GPUSpikingModel.cu:106-152): the reference's setup, then one
 * upload of everything it produced. */
void HipModelAdapter::setupSim(SimulationInfo *sim_info)
{
    Model::setupSim(sim_info);

    const int N = sim_info->totalNeurons;
    m_max_spikes = (int)(sim_info->epochDuration * sim_info->maxFiringRate);
    m_pending = 0;
    m_step_base = 0;

    bghip_net_desc d;
    memset(&d, 0, sizeof d);
    d.num_neurons = N;
    d.max_synapses_per_neuron = sim_info->maxSynapsesPerNeuron;
    if (dynamic_cast<AllIZHNeurons *>(m_neurons)) d.neuron_model = BGHIP_NEURON_IZH;
    else if (dynamic_cast<AllLIFNeurons *>(m_neurons)) d.neuron_model = BGHIP_NEURON_LIF;
    else check(BGHIP_EUNSUPPORTED, "neuron class");
    if (dynamic_cast<AllDynamicSTDPSynapses *>(m_synapses)) d.synapse_model = BGHIP_SYN_DYNSTDP;
    else if (dynamic_cast<AllSTDPSynapses *>(m_synapses)) d.synapse_model = BGHIP_SYN_STDP;
    else if (dynamic_cast<AllDSSynapses *>(m_synapses)) d.synapse_model = BGHIP_SYN_DS;
    else if (dynamic_cast<AllSpikingSynapses *>(m_synapses)) d.synapse_model = BGHIP_SYN_SPIKING;
    else check(BGHIP_EUNSUPPORTED, "synapse class");
    d.ring_bits = 32;                       /* LENGTH_OF_DELAYQUEUE: one uint32 per slot */
    d.max_spikes = m_max_spikes;
    d.dt = sim_info->deltaT;
    d.noise_seed = (uint64_t)sim_info->seed;
    d.block_steps = 0;
    fillTypes(d.types, sim_info->deltaT);

    check(bghip_ctx_create(m_device, &m_ctx), "bghip_ctx_create");
    check(bghip_net_create(m_ctx, &d, &m_net), "bghip_net_create");
    uploadNeurons(sim_info);

    std::vector<int32_t> types(N);
    for (int i = 0; i < N; i++) types[i] = (int32_t)m_layout->neuron_type_map[i];
    check(bghip_upload_layout(m_net, &(*m_layout->xloc)[0], &(*m_layout->yloc)[0], &types[0]), "bghip_upload_layout");

    if (ConnGrowth *g = dynamic_cast<ConnGrowth *>(m_conns)) {
        bghip_growth gp;
        gp.epsilon = g->m_growth.epsilon;
        gp.beta = g->m_growth.beta;
        gp.rho = g->m_growth.rho;
        gp.targetRate = g->m_growth.targetRate;
        gp.maxRate = g->m_growth.maxRate;
        gp.minRadius = g->m_growth.minRadius;
        gp.startRadius = g->m_growth.startRadius;
        gp.epochDuration = sim_info->epochDuration;
        check(bghip_upload_growth(m_net, &gp, &(*g->radii)[0], &(*g->rates)[0]), "bghip_upload_growth");
    }
    uploadSynapses(sim_info);
    m_total_synapses = bghip_total_synapses(m_net);
}

/* AllIFNeurons / AllIZHNeurons / AllSpikingNeurons arrays (copyNeuronHostToDevice,
 * Neurons/AllIFNeurons_d.cu:93-128 and Neurons/AllIZHNeurons_d.cu). */
void HipModelAdapter::uploadNeurons(const SimulationInfo *sim_info)
{
    const int N = sim_info->totalNeurons;
    AllIFNeurons &n = dynamic_cast<AllIFNeurons &>(*m_neurons);
    std::vector<uint64_t> hist((size_t)N * m_max_spikes);
    for (int i = 0; i < N; i++)
        memcpy(&hist[(size_t)i * m_max_spikes], n.spike_history[i], sizeof(uint64_t) * m_max_spikes);
    bghip_neurons h;
    memset(&h, 0, sizeof h);
    h.Vm = n.Vm; h.Vthresh = n.Vthresh; h.Vreset = n.Vreset; h.I0 = n.I0; h.Inoise = n.Inoise;
    h.C1 = n.C1; h.C2 = n.C2; h.Trefract = n.Trefract; h.summation_map = n.summation_map;
    h.nStepsInRefr = n.nStepsInRefr;
    h.spikeCount = n.spikeCount; h.spikeCountOffset = n.spikeCountOffset;
    h.spike_history = hist.empty() ? NULL : &hist[0];
    if (AllIZHNeurons *z = dynamic_cast<AllIZHNeurons *>(m_neurons)) {
        h.Aconst = z->Aconst; h.Bconst = z->Bconst; h.Cconst = z->Cconst; h.Dconst = z->Dconst;
        h.u = z->u; h.C3 = z->C3;
    }
    check(bghip_upload_neurons(m_net, &h), "bghip_upload_neurons");
}

/* AllSpikingSynapses / AllDSSynapses / AllSTDPSynapses slot arrays as the connection class left
 * them (copySynapseHostToDevice, Synapses/AllSpikingSynapses_d.cu:74-107): empty for ConnGrowth,
 * ConnStatic's synapses otherwise. bool and enum arrays are widened to the ABI's byte and int. */
void HipModelAdapter::uploadSynapses(const SimulationInfo *sim_info)
{
    const size_t NM = (size_t)sim_info->totalNeurons * sim_info->maxSynapsesPerNeuron;
    AllSpikingSynapses &s = dynamic_cast<AllSpikingSynapses &>(*m_synapses);
    std::vector<uint8_t> in_use(NM), fd;
    std::vector<int32_t> type(NM);
    for (size_t i = 0; i < NM; i++) {
        in_use[i] = s.in_use[i] ? 1 : 0;
        type[i] = (int32_t)s.type[i];
    }
    bghip_synapses h;
    memset(&h, 0, sizeof h);
    h.in_use = &in_use[0];
    h.sourceNeuronIndex = s.sourceNeuronIndex; h.destNeuronIndex = s.destNeuronIndex; h.type = &type[0];
    h.W = s.W; h.psr = s.psr; h.decay = s.decay; h.total_delay = s.total_delay; h.delayIdx = s.delayIdx;
    h.delayQueue = s.delayQueue; h.synapse_counts = s.synapse_counts;
    if (AllDSSynapses *ds = dynamic_cast<AllDSSynapses *>(m_synapses)) {
        h.r = ds->r; h.u = ds->u; h.U = ds->U; h.D = ds->D; h.F = ds->F;
        h.lastSpike = (uint64_t *)ds->lastSpike;
    }
    if (AllSTDPSynapses *st = dynamic_cast<AllSTDPSynapses *>(m_synapses)) {
        fd.resize(NM);
        for (size_t i = 0; i < NM; i++) fd[i] = st->useFroemkeDanSTDP[i] ? 1 : 0;
        h.delayQueuePost = st->delayQueuePost; h.delayIdxPost = st->delayIdxPost;
        h.Apos = st->Apos; h.Aneg = st->Aneg; h.STDPgap = st->STDPgap; h.tauspost = st->tauspost;
        h.tauspre = st->tauspre; h.taupos = st->taupos; h.tauneg = st->tauneg; h.Wex = st->Wex;
        h.mupos = st->mupos; h.muneg = st->muneg; h.useFroemkeDanSTDP = &fd[0];
    }
    if (AllDynamicSTDPSynapses *dy = dynamic_cast<AllDynamicSTDPSynapses *>(m_synapses)) {
        h.r = dy->r; h.u = dy->u; h.U = dy->U; h.D = dy->D; h.F = dy->F;
        h.lastSpike = (uint64_t *)dy->lastSpike;
    }
    check(bghip_upload_synapses(m_net, &h, g_simulationStep), "bghip_upload_synapses");
}

/* GPUSpikingModel::advance (Core/GPUSpikingModel.cu:188-225): one step, queued. The Simulator
 * increments g_simulationStep after each call; the engine keeps its own step counter in step with
 * it because every queued step is handed over before anything reads state. */
void HipModelAdapter::advance(const SimulationInfo *sim_info)
{
    if (sim_info->pInput != NULL && m_pending == 0) {
        fprintf(stderr, "! ERROR: a per-step host stimulus (ISInput::inputStimulus) cannot reach the device; "
                        "configure the input with bghip_upload_input instead\n");
        exit(1);
    }
    m_pending++;
}

void HipModelAdapter::flushSteps()
{
    if (m_pending == 0) return;
    check(bghip_advance(m_net, m_pending), "bghip_advance");
    m_pending = 0;
}

/* copyNeuronDeviceSpikeCountsToHost / copyNeuronDeviceSpikeHistoryToHost
 * (Neurons/AllSpikingNeurons_d.cu:17-50): counts, offsets and the history rows the recorder bins. */
void HipModelAdapter::downloadSpikes(const SimulationInfo *sim_info)
{
    const int N = sim_info->totalNeurons;
    AllSpikingNeurons &n = dynamic_cast<AllSpikingNeurons &>(*m_neurons);
    std::vector<uint64_t> hist((size_t)N * m_max_spikes);
    check(bghip_download_spikes(m_net, n.spikeCount, n.spikeCountOffset, hist.empty() ? NULL : &hist[0]),
          "bghip_download_spikes");
    for (size_t k = 0; k < hist.size(); k++)
        if (hist[k] != ~(uint64_t)0 && hist[k] >= m_step_base) hist[k] -= m_step_base;
    for (int i = 0; i < N; i++)
        memcpy(n.spike_history[i], &hist[(size_t)i * m_max_spikes], sizeof(uint64_t) * m_max_spikes);
}

/* GPUSpikingModel::updateConnections (Core/GPUSpikingModel.cu:246-259): growth runs on the device
 * (rates, radii, frontiers, overlap, weights, create/erase, index maps); the host objects receive
 * what the recorder reads. */
void HipModelAdapter::updateConnections(const SimulationInfo *sim_info)
{
    flushSteps();
    if (ConnGrowth *g = dynamic_cast<ConnGrowth *>(m_conns)) {
        bghip_growth_stats st;
        check(bghip_update_connections(m_net, &st), "bghip_update_connections");
        m_total_synapses = st.total_synapses;
        check(bghip_download_growth(m_net, &(*g->radii)[0], &(*g->rates)[0]), "bghip_download_growth");
    }
    downloadSpikes(sim_info);
}

/* GPUSpikingModel::updateHistory (Core/GPUSpikingModel.cu:266-272): the reference recorder compiles
 * its histories from the host arrays (binning, spike-count clear, minRadius clamp,
 * Recorders/XmlGrowthRecorder.cpp:72-97); the device applies the same clamp and clear. */
void HipModelAdapter::updateHistory(const SimulationInfo *sim_info)
{
    flushSteps();
    Model::updateHistory(sim_info);
    check(bghip_update_history(m_net), "bghip_update_history");
    check(bghip_sync(m_net), "bghip_sync");
}

void HipModelAdapter::saveData(SimulationInfo *sim_info)
{
    flushSteps();
    check(bghip_sync(m_net), "bghip_sync");
    Model::saveData(sim_info);
}

/* GPUSpikingModel::cleanupSim (Core/GPUSpikingModel.cu:161-176). */
void HipModelAdapter::cleanupSim(SimulationInfo *sim_info)
{
    if (m_net) {
        flushSteps();
        check(bghip_sync(m_net), "bghip_sync");
        bghip_net_destroy(m_net);
        m_net = NULL;
    }
    if (m_ctx) {
        bghip_ctx_destroy(m_ctx);
        m_ctx = NULL;
    }
    Model::cleanupSim(sim_info);
}

/* Model::serialize (Core/Model.cpp:85-98) as written by BGDriver -w: the neuron count and a NUL,
 * which Simulator::deserialize reads back first (Core/Simulator.cpp:206-214), then the engine's
 * checkpoint (the whole device network: state, rings, STDP state, input state). */
void HipModelAdapter::serialize(ostream &output, const SimulationInfo *sim_info)
{
    flushSteps();
    char path[] = "/tmp/bghip_ckpt_XXXXXX";
    int fd = mkstemp(path);
    if (fd < 0) check(BGHIP_EINVAL, "mkstemp");
    close(fd);
    check(bghip_net_save(m_net, path), "bghip_net_save");
    output << sim_info->totalNeurons << ends;
    std::ifstream in(path, std::ios::binary);
    output << in.rdbuf();
    in.close();
    remove(path);
    output.flush();
}

/* Model::deserialize (Core/Model.cpp:62-76), called by BGDriver -r after setupSim with the stream
 * past the neuron count: a checkpoint written by serialize() for the same configuration replaces
 * the network setupSim created, and the host radii and spike bookkeeping follow it. The engine
 * continues from the saved step (its noise and input streams are keyed on the step, so the run
 * continues exactly as an uninterrupted one would), while the reference restarts
 * g_simulationStep at 0 and its recorders bin spike times from there (Recorders/XmlRecorder.cpp:
 * 71-104): the host sees spike steps relative to the resumed run's start. */
void HipModelAdapter::deserialize(istream &input, const SimulationInfo *sim_info)
{
    char path[] = "/tmp/bghip_ckpt_XXXXXX";
    int fd = mkstemp(path);
    if (fd < 0) check(BGHIP_EINVAL, "mkstemp");
    close(fd);
    {
        std::ofstream out(path, std::ios::binary);
        out << input.rdbuf();
    }
    bghip_net *loaded = NULL;
    int rc = bghip_net_load(m_ctx, path, &loaded);
    remove(path);
    check(rc, "bghip_net_load");
    bghip_net_desc a, b;
    bghip_net_get_desc(m_net, &a);
    bghip_net_get_desc(loaded, &b);
    if (a.num_neurons != b.num_neurons || a.max_synapses_per_neuron != b.max_synapses_per_neuron ||
        a.neuron_model != b.neuron_model || a.synapse_model != b.synapse_model || a.max_spikes != b.max_spikes ||
        a.dt != b.dt || a.ring_bits != b.ring_bits) {
        bghip_net_destroy(loaded);
        check(BGHIP_EINVAL, "checkpoint of another configuration");
    }
    bghip_net_destroy(m_net);
    m_net = loaded;
    m_pending = 0;
    m_step_base = bghip_net_step(m_net) - g_simulationStep;
    m_total_synapses = bghip_total_synapses(m_net);
    if (ConnGrowth *g = dynamic_cast<ConnGrowth *>(m_conns))
        check(bghip_download_growth(m_net, &(*g->radii)[0], &(*g->rates)[0]), "bghip_download_growth");
    downloadSpikes(sim_info);
}
