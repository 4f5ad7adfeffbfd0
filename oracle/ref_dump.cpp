// State-dump harness around the compiled reference simulator (test infrastructure only).
//
// It links the reference's own objects (built from the sources under /root/reference by
// oracle/Makefile.ref) and runs exactly the reference's single-threaded flow:
//   LoadAllParameters (Core/BGDriver.cpp:219-258) -> Simulator::setup (Core/Simulator.cpp:33-44)
//   -> per epoch: IModel::advance once per step (Core/Simulator.cpp:155-186)
//   -> IModel::updateConnections -> IModel::updateHistory (Core/Simulator.cpp:129-131).
// After setup, after every epoch's last step and after every growth update it writes the full
// model state, bit for bit, to a tagged binary file that tests/ reads with numpy. Nothing here
// changes the reference's arithmetic; the `private/protected=public` compile flags only give this
// translation unit read access to the members it dumps.
//
// usage: bgref_dump <config.xml> <dump.bin> [--max-epochs K] [--no-history]
//                   [--dynamic-layout FRAC_EXC FRAC_STARTER] [--stim STIMULUS.xml]
//                   [--time-pregrown RADIUS STEPS]
// --dynamic-layout replaces the configured layout object by a DynamicLayout with the given
// fractions before setup: the reference's XML visitor for DynamicLayout rejects its own
// parameters (Layouts/DynamicLayout.cpp:29-77 returns false after parsing them), so this is the
// only way to run Layouts/DynamicLayout.cpp:98-165.
// --time-pregrown sets every neurite radius to RADIUS after setup, runs the reference's own growth
// update once (ConnGrowth::updateConns/updateFrontiers/updateOverlap/updateSynapsesWeights and the
// synapse index map), then times STEPS calls of IModel::advance (the reference's CPU step) and
// prints {"steps", "seconds", "synapses"} as one JSON line instead of writing a dump: the CPU
// baseline of the pre-grown benchmark network, measured on the reference's own code.
#include <ctime>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "Global.h"
#include "SimulationInfo.h"
#include "Simulator.h"
#include "Model.h"
#include "IRecorder.h"
#include "FSInput.h"
#include "AllSpikingNeurons.h"
#include "AllIFNeurons.h"
#include "AllIZHNeurons.h"
#include "AllSynapses.h"
#include "AllSpikingSynapses.h"
#include "AllDSSynapses.h"
#include "AllSTDPSynapses.h"
#include "AllDynamicSTDPSynapses.h"
#include "ConnGrowth.h"
#include "Layout.h"
#include "DynamicLayout.h"
#include "SynapseIndexMap.h"
#include "SInputPoisson.h"
#include "SInputRegular.h"

bool LoadAllParameters(SimulationInfo *simInfo);

static FILE *g_out = NULL;

static void rec(const char *name, char code, const void *data, unsigned long long count, size_t esize)
{
    unsigned int nl = (unsigned int)strlen(name);
    fwrite(&nl, 4, 1, g_out);
    fwrite(name, 1, nl, g_out);
    fwrite(&code, 1, 1, g_out);
    fwrite(&count, 8, 1, g_out);
    if (count) fwrite(data, esize, count, g_out);
}
static void recf(const char *n, const float *p, size_t c) { rec(n, 'f', p, c, 4); }
static void reci(const char *n, const int *p, size_t c) { rec(n, 'i', p, c, 4); }
static void recu(const char *n, const unsigned int *p, size_t c) { rec(n, 'I', p, c, 4); }
static void recq(const char *n, const unsigned long long *p, size_t c) { rec(n, 'q', p, c, 8); }
static void recb(const char *n, const bool *p, size_t c)
{
    std::vector<unsigned char> v(c);
    for (size_t i = 0; i < c; i++) v[i] = p[i] ? 1 : 0;
    rec(n, 'b', c ? &v[0] : NULL, c, 1);
}
static void marker(const char *tag, int epoch)
{
    int e = epoch;
    rec(tag, 'i', &e, 1, 4);
}

// Stimulus input state: SInputPoisson's interval counters and its one-slot-per-neuron DS input
// synapses (Inputs/SInputPoisson.h), SInputRegular's cycle position (Inputs/SInputRegular.h).
static void dump_input(SimulationInfo *si)
{
    int N = si->totalNeurons;
    SInputPoisson *sp = dynamic_cast<SInputPoisson *>(si->pInput);
    if (sp && sp->m_synapses) {
        AllDSSynapses *ds = dynamic_cast<AllDSSynapses *>(sp->m_synapses);
        reci("in_nISIs", sp->nISIs, N);
        recf("in_psr", ds->psr, N);
        recf("in_r", ds->r, N);
        recf("in_u", ds->u, N);
        recf("in_W", ds->W, N);
        recq("in_lastSpike", (const unsigned long long *)ds->lastSpike, N);
        recu("in_delayQueue", ds->delayQueue, N);
        reci("in_delayIdx", ds->delayIdx, N);
    }
    SInputRegular *sr = dynamic_cast<SInputRegular *>(si->pInput);
    if (sr && sr->values) {
        recf("in_values", sr->values, N);
        reci("in_shift", sr->nShiftValues, N);
        reci("in_steps_in_cycle", &sr->nStepsInCycle, 1);
    }
}

static void dump_state(SimulationInfo *si, Model *model, bool with_history)
{
    int N = si->totalNeurons;
    int max_spikes = (int)(si->epochDuration * si->maxFiringRate);
    unsigned long long step = g_simulationStep;
    recq("g_simulationStep", &step, 1);
    dump_input(si);

    AllIFNeurons *ifn = dynamic_cast<AllIFNeurons *>(model->m_neurons);
    if (ifn) {
        recf("Vm", ifn->Vm, N);
        recf("Vthresh", ifn->Vthresh, N);
        recf("Vreset", ifn->Vreset, N);
        recf("summation_map", ifn->summation_map, N);
        reci("nStepsInRefr", ifn->nStepsInRefr, N);
        reci("spikeCount", ifn->spikeCount, N);
        reci("spikeCountOffset", ifn->spikeCountOffset, N);
        if (with_history) {
            std::vector<unsigned long long> h((size_t)N * max_spikes);
            for (int i = 0; i < N; i++)
                for (int j = 0; j < max_spikes; j++) h[(size_t)i * max_spikes + j] = ifn->spike_history[i][j];
            recq("spike_history", &h[0], h.size());
        }
    }
    AllIZHNeurons *izh = dynamic_cast<AllIZHNeurons *>(model->m_neurons);
    if (izh) recf("u", izh->u, N);

    AllSynapses *syn = dynamic_cast<AllSynapses *>(model->m_synapses);
    size_t S = (size_t)N * si->maxSynapsesPerNeuron;
    if (syn && S) {
        recb("in_use", syn->in_use, S);
        reci("sourceNeuronIndex", syn->sourceNeuronIndex, S);
        reci("destNeuronIndex", syn->destNeuronIndex, S);
        recf("W", syn->W, S);
        recf("psr", syn->psr, S);
        std::vector<int> t(S);
        for (size_t i = 0; i < S; i++) t[i] = syn->in_use[i] ? (int)syn->type[i] : -1;
        reci("type", &t[0], S);
        recu("synapse_counts", syn->synapse_counts, N);
    }
    AllSpikingSynapses *ss = dynamic_cast<AllSpikingSynapses *>(model->m_synapses);
    if (ss && S) {
        recf("decay", ss->decay, S);
        reci("total_delay", ss->total_delay, S);
        recu("delayQueue", ss->delayQueue, S);
        reci("delayIdx", ss->delayIdx, S);
    }
    AllDSSynapses *ds = dynamic_cast<AllDSSynapses *>(model->m_synapses);
    if (ds && S) {
        recf("r", ds->r, S);
        recf("u_syn", ds->u, S);
        recq("lastSpike", (const unsigned long long *)ds->lastSpike, S);
    }
    AllSTDPSynapses *st = dynamic_cast<AllSTDPSynapses *>(model->m_synapses);
    if (st && S) {
        recu("delayQueuePost", st->delayQueuePost, S);
        reci("delayIdxPost", st->delayIdxPost, S);
    }
    AllDynamicSTDPSynapses *dst = dynamic_cast<AllDynamicSTDPSynapses *>(model->m_synapses);
    if (dst && S) {
        recf("r", dst->r, S);
        recf("u_syn", dst->u, S);
        recq("lastSpike", (const unsigned long long *)dst->lastSpike, S);
    }

    ConnGrowth *cg = dynamic_cast<ConnGrowth *>(model->m_conns);
    if (cg) {
        std::vector<float> r(N), q(N);
        for (int i = 0; i < N; i++) { r[i] = (*cg->radii)[i]; q[i] = (*cg->rates)[i]; }
        recf("radii", &r[0], N);
        recf("rates", &q[0], N);
    }
    SynapseIndexMap *m = model->m_synapseIndexMap;
    if (m && syn && syn->total_synapse_counts) {
        size_t T = syn->total_synapse_counts;
        recu("incomingSynapseIndexMap", m->incomingSynapseIndexMap, T);
        recu("incomingSynapseBegin", m->incomingSynapseBegin, N);
        recu("incomingSynapseCount", m->incomingSynapseCount, N);
        recu("outgoingSynapseIndexMap", m->outgoingSynapseIndexMap, T);
        recu("outgoingSynapseBegin", m->outgoingSynapseBegin, N);
        recu("outgoingSynapseCount", m->outgoingSynapseCount, N);
    }
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <config.xml> <dump.bin> [--max-epochs K] [--no-history]\n", argv[0]);
        return 2;
    }
    int max_epochs = -1;
    bool with_history = true;
    bool dyn = false;
    float fexc = 0, fst = 0;
    const char *stim = NULL;
    float time_radius = 0;
    int time_steps = 0;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "--max-epochs") && i + 1 < argc) max_epochs = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--no-history")) with_history = false;
        else if (!strcmp(argv[i], "--dynamic-layout") && i + 2 < argc) {
            dyn = true;
            fexc = (float)atof(argv[++i]);
            fst = (float)atof(argv[++i]);
        } else if (!strcmp(argv[i], "--stim") && i + 1 < argc) {
            stim = argv[++i];
        } else if (!strcmp(argv[i], "--time-pregrown") && i + 2 < argc) {
            time_radius = (float)atof(argv[++i]);
            time_steps = atoi(argv[++i]);
        }
    }
    SimulationInfo *si = new SimulationInfo();
    si->stateInputFileName = argv[1];
    std::string xml_out = std::string(argv[2]) + ".xml";
    si->stateOutputFileName = xml_out;
    if (!LoadAllParameters(si)) return 1;
    si->simRecorder = si->model->getConnections()->createRecorder(si);
    if (stim) si->stimulusInputFileName = stim;   // the -s option (Core/BGDriver.cpp:320)
    si->pInput = FSInput::get()->CreateInstance(si);

    g_out = fopen(argv[2], "wb");
    if (!g_out) return 1;
    fwrite("BGDUMP1\n", 1, 8, g_out);

    Model *model = dynamic_cast<Model *>(si->model);
    if (dyn) {
        DynamicLayout *dl = new DynamicLayout();
        dl->m_frac_excitatory_neurons = fexc;
        dl->m_frac_starter_neurons = fst;
        model->m_layout = dl;
    }
    Simulator sim;
    sim.setup(si);

    int N = si->totalNeurons;
    if (time_steps > 0) {
        ConnGrowth *g = dynamic_cast<ConnGrowth *>(model->m_conns);
        if (!g) return 1;
        for (int i = 0; i < N; i++) (*g->radii)[i] = time_radius;
        si->currentStep = 1;
        si->model->updateConnections(si);
        AllSynapses *syn = dynamic_cast<AllSynapses *>(model->m_synapses);
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        for (int k = 0; k < time_steps; k++) {
            si->model->advance(si);
            g_simulationStep++;
        }
        clock_gettime(CLOCK_MONOTONIC, &b);
        printf("{\"steps\": %d, \"seconds\": %.6f, \"synapses\": %llu}\n", time_steps,
               (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec),
               (unsigned long long)syn->total_synapse_counts);
        fclose(g_out);
        return 0;
    }
    Layout *L = model->m_layout;
    {
        std::vector<float> x(N), y(N);
        std::vector<int> t(N);
        for (int i = 0; i < N; i++) { x[i] = (*L->xloc)[i]; y[i] = (*L->yloc)[i]; t[i] = L->neuron_type_map[i]; }
        recf("xloc", &x[0], N);
        recf("yloc", &y[0], N);
        reci("neuron_type_map", &t[0], N);
        recb("starter_map", L->starter_map, N);
        AllIFNeurons *ifn = dynamic_cast<AllIFNeurons *>(model->m_neurons);
        recf("Iinject", ifn->Iinject, N);
        recf("Inoise", ifn->Inoise, N);
        recf("Vinit", ifn->Vinit, N);
        recf("Trefract", ifn->Trefract, N);
        recf("I0", ifn->I0, N);
        recf("C1", ifn->C1, N);
        recf("C2", ifn->C2, N);
        AllIZHNeurons *izh = dynamic_cast<AllIZHNeurons *>(model->m_neurons);
        if (izh) {
            recf("Aconst", izh->Aconst, N);
            recf("Bconst", izh->Bconst, N);
            recf("Cconst", izh->Cconst, N);
            recf("Dconst", izh->Dconst, N);
            recf("C3", izh->C3, N);
        }
    }
    marker("@setup", 0);
    dump_state(si, model, with_history);

    int epochs = si->maxSteps;
    if (max_epochs >= 0 && max_epochs < epochs) epochs = max_epochs;
    for (int ep = 1; ep <= epochs; ep++) {
        si->currentStep = ep;
        uint64_t endStep = g_simulationStep + static_cast<uint64_t>(si->epochDuration / si->deltaT);
        while (g_simulationStep < endStep) {
            if (si->pInput != NULL) si->pInput->inputStimulus(si);
            si->model->advance(si);
            g_simulationStep++;
        }
        marker("@advanced", ep);
        dump_state(si, model, with_history);
        si->model->updateConnections(si);
        si->model->updateHistory(si);
        marker("@grown", ep);
        dump_state(si, model, with_history);
    }
    fclose(g_out);
    if (si->pInput != NULL) { si->pInput->term(si); delete si->pInput; }
    sim.saveData(si);
    if (si->simRecorder != NULL) si->simRecorder->term();
    return 0;
}
