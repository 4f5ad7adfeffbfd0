/*
 * BrainGrid's driver with the MI355X engine behind IModel: the `growth` program of a BrainGrid
 * build whose createAllModelClassInstances takes the USE_HIP branch (Core/BGDriver.cpp:204-208)
 * and constructs a HipModelAdapter over the classes the XML names.
 *
 * Everything else is the reference's own code, linked from its objects: the command line
 * (parseCommandLine, Core/BGDriver.cpp:289-330), SimulationInfo::readParameters, the class factory
 * (FClassOfCategory), the recorder the connection class creates (XmlGrowthRecorder for ConnGrowth),
 * the Simulator loop (Core/Simulator.cpp:33-186) and the -r / -w memory images. The steps of
 * main() follow Core/BGDriver.cpp:49-146.
 *
 *   growth_hip -t config.xml -o state-out.xml [-r mem-in] [-w mem-out]
 *
 * The device is BGHIP_DEVICE (default 0).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>

#include "FClassOfCategory.h"
#include "FSInput.h"
#include "Global.h"
#include "IRecorder.h"
#include "Simulator.h"
#include "HipModelAdapter.h"

using namespace std;

/* Core/BGDriver.cpp (linked from the reference's driver object, whose own main is renamed). */
bool parseCommandLine(int argc, char *argv[], SimulationInfo *simInfo);

/* LoadAllParameters + createAllModelClassInstances (Core/BGDriver.cpp:160-258) with the model
 * branch that builds the engine-backed IModel. */
static bool loadWithEngine(SimulationInfo *simInfo, int device)
{
    TiXmlDocument simDoc(simInfo->stateInputFileName.c_str());
    if (!simDoc.LoadFile()) {
        cerr << "Failed loading simulation parameter file " << simInfo->stateInputFileName << ":\n\t"
             << simDoc.ErrorDesc() << endl;
        return false;
    }
    if (!simInfo->readParameters(&simDoc)) return false;

    TiXmlElement *parms = simDoc.FirstChildElement()->FirstChildElement("ModelParams");
    if (parms == NULL) {
        cerr << "Could not find <ModelParams> in simulation parameter file" << endl;
        return false;
    }
    IAllNeurons *neurons = NULL;
    IAllSynapses *synapses = NULL;
    Connections *conns = NULL;
    Layout *layout = NULL;
    const TiXmlNode *pNode = NULL;
    while ((pNode = parms->IterateChildren(pNode)) != NULL) {
        if (strcmp(pNode->Value(), "NeuronsParams") == 0) neurons = FClassOfCategory::get()->createNeurons(pNode);
        else if (strcmp(pNode->Value(), "SynapsesParams") == 0) synapses = FClassOfCategory::get()->createSynapses(pNode);
        else if (strcmp(pNode->Value(), "ConnectionsParams") == 0) conns = FClassOfCategory::get()->createConnections(pNode);
        else if (strcmp(pNode->Value(), "LayoutParams") == 0) layout = FClassOfCategory::get()->createLayout(pNode);
    }
    if (neurons == NULL || synapses == NULL || conns == NULL || layout == NULL) {
        cerr << "!ERROR: failed to create classes" << endl;
        return false;
    }
    simInfo->model = new HipModelAdapter(conns, neurons, synapses, layout, device);

    if (!FClassOfCategory::get()->readParameters(&simDoc)) return false;
    if (simInfo->stateOutputFileName.empty()) {
        cerr << "! ERROR: no stateOutputFileName is specified." << endl;
        return false;
    }
    return true;
}

int main(int argc, char *argv[])
{
    const char *dev = getenv("BGHIP_DEVICE");
    const int device = dev ? atoi(dev) : 0;
    SimulationInfo *simInfo = new SimulationInfo();
    if (!parseCommandLine(argc, argv, simInfo)) {
        cerr << "! ERROR: failed during command line parse" << endl;
        return -1;
    }
    if (!loadWithEngine(simInfo, device)) {
        cerr << "! ERROR: failed while parsing simulation parameters." << endl;
        return -1;
    }
    simInfo->simRecorder = simInfo->model->getConnections()->createRecorder(simInfo);
    if (simInfo->simRecorder == NULL) {
        cerr << "! ERROR: invalid state output file name extension." << endl;
        return -1;
    }
    simInfo->pInput = FSInput::get()->CreateInstance(simInfo);

    time_t start_time, end_time;
    time(&start_time);
    Simulator *simulator = new Simulator();
    simulator->setup(simInfo);
    if (!simInfo->memInputFileName.empty()) {
        ifstream memory_in(simInfo->memInputFileName.c_str(), ofstream::binary | ofstream::in);
        simulator->deserialize(memory_in, simInfo);
    }
    simulator->simulate(simInfo);
    if (simInfo->pInput != NULL) {
        simInfo->pInput->term(simInfo);
        delete simInfo->pInput;
        simInfo->pInput = NULL;
    }
    simulator->saveData(simInfo);
    if (!simInfo->memOutputFileName.empty()) {
        ofstream memory_out(simInfo->memOutputFileName.c_str(), ofstream::binary | ofstream::trunc);
        simulator->serialize(memory_out, simInfo);
    }
    simulator->finish(simInfo);
    if (simInfo->simRecorder != NULL) simInfo->simRecorder->term();

    time(&end_time);
    const double elapsed = difftime(end_time, start_time);
    cout << "time simulated: " << simInfo->epochDuration * simInfo->maxSteps << endl;
    cout << "time elapsed: " << elapsed << endl;

    delete simInfo->model;
    simInfo->model = NULL;
    if (simInfo->simRecorder != NULL) {
        delete simInfo->simRecorder;
        simInfo->simRecorder = NULL;
    }
    delete simInfo;
    delete simulator;
    return 0;
}
