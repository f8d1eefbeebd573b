// match: command-line driver with the flags and output of the reference's
// bin/match.cpp (flags :199-284, run loop :380-470, output :473-511), over
// the drop-in classes of this library.
//
//   match [-m "types"] [-t] [-r] [-d device] [-a repetitions] [-b barrier 0..16]
//         [--gnuplot file] [--scaledata] [--threads "counts"] [--intermediate] graph.{graph,mtx}
//
// Types: 0 CPU random, 1 CPU static minimum degree, 2 CPU Karp--Sipser, 3 CPU
// dynamic minimum degree, 4 CPU weighted, 5 CPU weighted (edge-based), 6/7 GPU
// random (maximal), 8/9 GPU weighted (maximal), 10/11 CPU round matchers on
// host threads (the reference's TBB types).  stdout: a header line
// "<file> <vertices> <edges> <barrier>", then per type "<type> <threads>
// <size avg dev> <weight avg dev> <total ms avg dev> <match ms avg dev>".
// --intermediate prints, for GPU types, "round<TAB>weight<TAB>size" after
// every round (the reference's MATCH_INTERMEDIATE_COUNT build).
// *.graph.bz2 / *.mtx.bz2 inputs are decompressed while they are read.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mm_match.h"
#include "graph.h"
#include "matchcpu.h"
#include "matchgpu.h"
#include "matchtbb.h"

using namespace std;
using namespace mtc;

namespace
{

string readGraph(Graph &graph, const string &fileName)
{
    string shortName = fileName.substr(fileName.find_last_of("/\\") + 1);
    const bool metis = fileName.find(".graph") != string::npos, market = fileName.find(".mtx") != string::npos;
    if (!metis && !market) {
        cerr << "Unknown filename extension for '" << fileName << "'!" << endl;
        throw exception();
    }
    string error;
    unique_ptr<istream> in = openGraphInput(fileName, error);
    if (!in) {
        cerr << error << endl;
        throw exception();
    }
    if (metis) graph.readMETIS(*in);
    else graph.readMatrixMarket(*in);
    if (in->bad()) {
        cerr << "Error while reading '" << fileName << "'!" << endl;
        throw exception();
    }
    // The reported name drops the compression suffix (bin/match.cpp:96, 109).
    if (shortName.find(".bz2") != string::npos) shortName = shortName.substr(0, shortName.find(".bz2"));
    return shortName;
}

GraphMatching *getMatcher(const Graph &graph, int type, int threadsPerBlock, unsigned int barrier)
{
    switch (type) {
    case 0: return new GraphMatchingCPURandom(graph);
    case 1: return new GraphMatchingCPUStatMinDeg(graph);
    case 2: return new GraphMatchingCPUKarpSipser(graph);
    case 3: return new GraphMatchingCPUMinDeg(graph);
    case 4: return new GraphMatchingCPUWeighted(graph);
    case 5: return new GraphMatchingCPUWeightedEdge(graph);
    case 6: return new GraphMatchingGPURandom(graph, threadsPerBlock, barrier);
    case 7: return new GraphMatchingGPURandomMaximal(graph, threadsPerBlock, barrier);
    case 8: return new GraphMatchingGPUWeighted(graph, threadsPerBlock, barrier);
    case 9: return new GraphMatchingGPUWeightedMaximal(graph, threadsPerBlock, barrier);
    case 10: return new GraphMatchingTBBRandom(graph, barrier);
    case 11: return new GraphMatchingTBBWeighted(graph, barrier);
    default: cerr << "Unknown matching type!" << endl; throw exception();
    }
}

void getStats(double &avg, double &dev, const vector<double> &data)
{
    double s = 0.0;
    for (double x : data) s += x;
    avg = s / (double)data.size();
    s = 0.0;
    for (double x : data) s += (x - avg) * (x - avg);
    dev = sqrt(s / (double)data.size());
}

// Interval timer: HIP events on the null stream when a device is present
// (as the reference times every matcher), the host clock otherwise.
struct Stamp {
    bool dev = false;
    hipEvent_t ev = nullptr;
    chrono::steady_clock::time_point host;
    // Every HIP call sits in a RandStreamGuard: the runtime may draw from
    // rand(), which would move the seeds the matchers take from it.
    explicit Stamp(bool d) : dev(d)
    {
        RandStreamGuard guard;
        if (dev && hipEventCreate(&ev) != hipSuccess) ev = nullptr;
    }
    ~Stamp()
    {
        RandStreamGuard guard;
        if (ev) (void)hipEventDestroy(ev);
    }
    void mark()
    {
        host = chrono::steady_clock::now();
        if (ev) {
            RandStreamGuard guard;
            (void)hipEventRecord(ev, 0);
            (void)hipEventSynchronize(ev);
        }
    }
    static float ms(const Stamp &a, const Stamp &b)
    {
        float t = 0.0f;
        RandStreamGuard guard;
        if (a.ev && b.ev && hipEventElapsedTime(&t, a.ev, b.ev) == hipSuccess) return t;
        return chrono::duration<float, milli>(b.host - a.host).count();
    }
};

bool takesValue(const string &a)
{
    return a == "-d" || a == "--device" || a == "-a" || a == "--repetitions" || a == "-b" || a == "--barrier" ||
           a == "--gnuplot" || a == "--threads" || a == "-m" || a == "--match";
}

// Per-round weight and size of one GPU type: the rounds are deterministic in
// the seed table, so round r's state is that of an r-round run.  Returns the
// rounds executed (the seeds the caller then consumes) or -1.
int intermediate(const Graph &graph, int type, unsigned int barrier, int device)
{
    const bool weighted = type >= 8;
    const bool maximal = type == 7 || type == 9;
    const bool has_w = graph.neighbourWeights.size() == graph.neighbours.size() && !graph.neighbours.empty();
    vector<unsigned int> seeds = peekRand(MM_NR_MAX_MATCH_ROUNDS);
    RandStreamGuard guard;  // the library's HIP calls below must not move the stream
    mm_graph *g = nullptr;
    if (mm_graph_create(device, graph.nrVertices, (int64_t)graph.neighbours.size(),
                        reinterpret_cast<const int32_t *>(graph.neighbourRanges.data()), graph.neighbours.data(),
                        has_w ? graph.neighbourWeights.data() : nullptr, &g) != MM_OK) {
        cerr << mm_last_error() << endl;
        return -1;
    }
    mm_opts o;
    memset(&o, 0, sizeof(o));
    o.kind = weighted ? MM_WEIGHTED : MM_RANDOM;
    o.maximal = maximal;
    o.barrier = barrier;
    o.seeds = seeds.data();
    o.n_seeds = (int32_t)seeds.size();
    mm_stats st;
    if (mm_match(g, &o, nullptr, &st) != MM_OK) {
        cerr << mm_last_error() << endl;
        mm_graph_destroy(g);
        return -1;
    }
    const int R = maximal ? st.rounds : MM_NR_MATCH_ROUNDS;
    vector<int> match((size_t)graph.nrVertices);
    cout << "0\t0\t0" << endl;
    for (int r = 1; r <= R; ++r) {
        o.maximal = 0;
        o.max_rounds = r;
        if (mm_match(g, &o, match.data(), nullptr) != MM_OK) {
            cerr << mm_last_error() << endl;
            mm_graph_destroy(g);
            return -1;
        }
        double weight = 0.0;
        long size = 0;
        GraphMatching::getWeight(weight, size, match, graph);
        cout << r << "\t" << weight << "\t" << size << endl;
    }
    mm_graph_destroy(g);
    return R;
}

}  // namespace

int main(int argc, char **argv)
{
    string fileName, gnuplotFileName;
    set<int> matchTypes{0};
    set<int> cpuThreads{(int)max(1u, thread::hardware_concurrency())};
    int repetitions = 1, device = 0;
    const int threadsPerBlock = 256;
    unsigned int barrier = MM_DEFAULT_BARRIER;  // p = 0.534059...
    bool scaleData = false, performTest = false, randomise = false, interm = false;

    try {
        for (int k = 1; k < argc; ++k) {
            string a = argv[k], v;
            const size_t eq = a.find('=');
            if (a.rfind("--", 0) == 0 && eq != string::npos) {
                v = a.substr(eq + 1);
                a = a.substr(0, eq);
            } else if (takesValue(a)) {
                if (k + 1 >= argc) throw exception();
                v = argv[++k];
            }
            if (a == "-h" || a == "--help") {
                cerr << "match [-m \"types\"] [-t] [-r] [-d device] [-a repetitions] [-b barrier(0..16)] "
                        "[--gnuplot file] [--scaledata] [--threads \"counts\"] [--intermediate] input-file" << endl;
                return 1;
            } else if (a == "-t" || a == "--test") performTest = true;
            else if (a == "-r" || a == "--random") randomise = true;
            else if (a == "--scaledata") scaleData = true;
            else if (a == "--intermediate") interm = true;
            else if (a == "-d" || a == "--device") device = max(0, stoi(v));
            else if (a == "-a" || a == "--repetitions") repetitions = max(1, stoi(v));
            else if (a == "--gnuplot") gnuplotFileName = v;
            else if (a == "-b" || a == "--barrier") {
                const int b = max(0, stoi(v));
                barrier = b >= 16 ? 0xffffffffu : (unsigned int)((long)b * 0x10000000L);
            } else if (a == "--threads" || a == "-m" || a == "--match") {
                istringstream in(v);
                set<int> vals;
                int t;
                while (in >> t) vals.insert(t);
                if (a == "--threads") cpuThreads = vals;
                else matchTypes = vals;
            } else if (!a.empty() && a[0] == '-') throw exception();
            else fileName = a;
        }
        if (fileName.empty()) {
            cerr << "You have to specify an input file!" << endl;
            throw exception();
        }
    } catch (exception &) {
        cerr << "Invalid command line arguments!" << endl;
        return -1;
    }

    cerr << fileName << "..." << endl;
    Graph graph;
    string shortName;
    try {
        shortName = readGraph(graph, fileName);
    } catch (exception &) {
        cerr << "An exception occured when reading " << fileName << " from disk!" << endl;
        return -1;
    }

    // The runtime comes up before srand(12345) (bin/match.cpp:43-79, 382);
    // without a device the CPU types still run, timed by the host clock.
    int nDev = 0;
    const bool haveDevice = hipGetDeviceCount(&nDev) == hipSuccess && nDev > 0;
    bool needDevice = false;
    for (int t : matchTypes) needDevice = needDevice || (t >= 6 && t <= 9);
    if (haveDevice) {
        if (device >= nDev || hipSetDevice(device) != hipSuccess || hipFree(nullptr) != hipSuccess) {
            cerr << "Invalid GPU device index (" << device << ")!" << endl;
            return -1;
        }
    } else if (needDevice) {
        cerr << "Unable to initialise the GPU!" << endl;
        return -1;
    }
    cerr << "Using up to " << thread::hardware_concurrency() << " CPU threads." << endl;

    ofstream gnuplotFile;
    if (!gnuplotFileName.empty()) {
        gnuplotFile.open(gnuplotFileName.c_str(), ios_base::out | ios_base::app);
        if (!gnuplotFile.good()) {
            cerr << "Unable to open '" << gnuplotFileName << "' for output!" << endl;
            return -1;
        }
        if (!scaleData) {
            int minDeg = INT_MAX, maxDeg = INT_MIN;
            for (const int2 &r : graph.neighbourRanges) {
                minDeg = min(minDeg, r.y - r.x);
                maxDeg = max(maxDeg, r.y - r.x);
            }
            gnuplotFile << shortName << "\t" << graph.nrVertices << "\t" << graph.nrEdges << "\t" << minDeg << "\t"
                        << maxDeg;
        }
    }
    if (!interm) cout << shortName << " " << graph.nrVertices << " " << graph.nrEdges << " " << barrier << endl;
    else {
        cerr << "Warning: Running in INTERMEDIATE mode!" << endl;
        repetitions = 1;
    }

    for (int type : matchTypes) {
        for (int threads : cpuThreads) {
            const bool hostRounds = type >= 10;
            if (hostRounds) setHostThreads(threads);
            vector<double> sizes((size_t)repetitions), weights((size_t)repetitions), totals((size_t)repetitions),
                times((size_t)repetitions);
            Graph graph2 = graph;
            srand(12345);
            for (int k = 0; k < repetitions; ++k) {
                if (randomise) graph2.random_shuffle();
                if (interm && type >= 6 && type <= 9) {
                    const int R = intermediate(graph2, type, barrier, device);
                    if (R < 0) return -1;
                    for (int r = 0; r < R; ++r) (void)rand();  // one seed per executed round
                    continue;
                }
                Stamp s0(haveDevice), s1(haveDevice), s2(haveDevice), s3(haveDevice);
                s0.mark();
                vector<int> match;
                try {
                    GraphMatching *matcher = getMatcher(graph2, type, threadsPerBlock, barrier);
                    match = matcher->initialMatching();
                    s1.host = chrono::steady_clock::now();
                    matcher->performMatching(match, s1.ev, s2.ev);
                    s2.host = chrono::steady_clock::now();
                    delete matcher;
                } catch (exception &) {
                    cerr << "An exception occured when matching!" << endl;
                    return -1;
                }
                s3.mark();
                if (performTest && !GraphMatching::testMatching(match, graph2)) {
                    cerr << "Invalid matching!" << endl;
                    return -1;
                }
                double weight = 0.0;
                long size = 0;
                GraphMatching::getWeight(weight, size, match, graph2);
                sizes[(size_t)k] = (double)size;
                weights[(size_t)k] = weight;
                totals[(size_t)k] = Stamp::ms(s0, s3);
                times[(size_t)k] = Stamp::ms(s1, s2);
            }
            if (!interm) {
                double avg = 0.0, dev = 0.0;
                cout << type << " ";
                if (!scaleData) gnuplotFile << "\t";
                const int shown = hostRounds ? threads : threadsPerBlock;
                cout << shown << " ";
                gnuplotFile << shown << "\t";
                for (const vector<double> *d : {&sizes, &weights, &totals, &times}) {
                    getStats(avg, dev, *d);
                    cout << avg << " " << dev << " ";
                    gnuplotFile << avg << "\t" << dev << "\t";
                }
                cout << endl;
                gnuplotFile << (double)barrier / (double)UINT_MAX;
                if (scaleData) gnuplotFile << endl;
                gnuplotFile.flush();
            }
            if (!hostRounds) break;
        }
    }
    setHostThreads(0);
    gnuplotFile << endl;
    gnuplotFile.close();
    return 0;
}
