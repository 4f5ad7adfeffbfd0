/*
 * HipModelAdapter -- BrainGrid's IModel backed by the MI355X engine (libbghip.so).
 *
 * This is the class a BrainGrid build adds next to GPUSpikingModel (Core/GPUSpikingModel.h) and
 * selects in createAllModelClassInstances (Core/BGDriver.cpp:204-208). It keeps every reference
 * object as it is -- SimulationInfo, the neuron, synapse, connection and layout classes, the
 * recorder, the Simulator loop -- and moves their host arrays across the C ABI of include/bghip.h:
 *
 *   setupSim          Model::setupSim (the reference's own draws, layout and connections), then one
 *                     upload of the neuron, synapse, layout and growth arrays
 *                     (GPUSpikingModel::setupSim, Core/GPUSpikingModel.cu:106-152);
 *   advance           called once per step by Simulator::advanceUntilGrowth
 *                     (Core/Simulator.cpp:155-186); steps are queued and handed to bghip_advance
 *                     as one call before anything reads state, so the fused kernels run whole
 *                     delay blocks per launch;
 *   updateConnections growth on the device (bghip_update_connections), then the spike bookkeeping
 *                     and the growth radii/rates downloaded into the host objects the recorder
 *                     reads (GPUSpikingModel::updateConnections, Core/GPUSpikingModel.cu:246-259);
 *   updateHistory     the reference recorder's compileHistories on the host arrays, then the same
 *                     minRadius clamp and spike-count clear on the device
 *                     (GPUSpikingModel::updateHistory, Core/GPUSpikingModel.cu:266-272);
 *   serialize /       the engine's own checkpoint (bghip_net_save / bghip_net_load) streamed
 *   deserialize       through the reference's -w / -r streams (Core/BGDriver.cpp:88-116); a
 *                     resumed run continues the saved run exactly, its recorder counting time
 *                     from the resume as the reference's does.
 *
 * Neurons: AllLIFNeurons, AllIZHNeurons. Synapses: AllSpikingSynapses, AllDSSynapses,
 * AllSTDPSynapses, AllDynamicSTDPSynapses. Connections: ConnGrowth (growth on the device) and any
 * class that only places synapses at setup (ConnStatic). Stimulus input is configured on the
 * device with bghip_upload_input; a Simulator that still calls ISInput::inputStimulus per step is
 * refused at the first step, because that input would never reach the device.
 *
 * The file builds with the reference's own flags (C++98) against its headers.
 */
#ifndef HIP_MODEL_ADAPTER_H
#define HIP_MODEL_ADAPTER_H

#include "Model.h"
#include "bghip.h"

class HipModelAdapter : public Model
{
    public:
        HipModelAdapter(Connections *conns, IAllNeurons *neurons, IAllSynapses *synapses, Layout *layout,
                        int device = 0);
        virtual ~HipModelAdapter();

        virtual void setupSim(SimulationInfo *sim_info);
        virtual void advance(const SimulationInfo *sim_info);
        virtual void updateConnections(const SimulationInfo *sim_info);
        virtual void updateHistory(const SimulationInfo *sim_info);
        virtual void cleanupSim(SimulationInfo *sim_info);
        virtual void saveData(SimulationInfo *sim_info);
        virtual void serialize(ostream &output, const SimulationInfo *sim_info);
        virtual void deserialize(istream &input, const SimulationInfo *sim_info);

        /* The engine's network (NULL before setupSim and after cleanupSim). */
        bghip_net *net() { return m_net; }
        /* Synapses after the last growth update (bghip_growth_stats). */
        int64_t totalSynapses() const { return m_total_synapses; }

    private:
        void check(int rc, const char *what) const;
        void flushSteps();
        void uploadNeurons(const SimulationInfo *sim_info);
        void uploadSynapses(const SimulationInfo *sim_info);
        void downloadSpikes(const SimulationInfo *sim_info);
        void fillTypes(bghip_synapse_types &t, BGFLOAT deltaT) const;

        int m_device;
        bghip_ctx *m_ctx;
        bghip_net *m_net;
        uint64_t m_pending;       /* steps queued since the last bghip_advance */
        uint64_t m_step_base;     /* engine step at g_simulationStep 0 (nonzero after -r) */
        int m_max_spikes;
        int64_t m_total_synapses;
};

#endif
