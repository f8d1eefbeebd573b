// Matcher base class of the drop-in interface: the reference's
// mtc::GraphMatching (src/matchcpu.h:24-48) with the same static helpers and
// state encoding.  performMatching takes hipEvent_t where the reference takes
// cudaEvent_t; the caller records nothing itself, the matcher records t1
// after set-up and t2 after the last round (src/matchgpu.cu:500-527).
#ifndef MM_DROPIN_MATCHCPU_H
#define MM_DROPIN_MATCHCPU_H

#include <hip/hip_runtime.h>

#include <vector>

#include "graph.h"

#define NR_MATCH_ROUNDS 20
#define NR_MAX_MATCH_ROUNDS 256

namespace mtc
{

class GraphMatching
{
    public:
        GraphMatching(const Graph &);
        virtual ~GraphMatching();

        static void getWeight(double &, long &, const std::vector<int> &, const Graph &);
        static bool testMatching(const std::vector<int> &, const Graph &);

        static inline int matchVal(const int &i, const int &j) {return 4 + (i < j ? i : j);};
        static inline bool isMatched(const int &m) {return m >= 4;};

        std::vector<int> initialMatching() const;
        virtual void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const = 0;

    protected:
        const Graph &graph;
};

// Sequential CPU matchers (src/matchcpu.h:50-108): type 0 random greedy,
// 1 static minimum degree, 2 Karp--Sipser, 3 dynamic (two-sided) minimum
// degree, 4 weighted vertex-order greedy, 5 edge-sorted greedy.
class GraphMatchingCPURandom : public GraphMatching
{
    public:
        GraphMatchingCPURandom(const Graph &);
        ~GraphMatchingCPURandom();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingCPUMinDeg : public GraphMatching
{
    public:
        GraphMatchingCPUMinDeg(const Graph &);
        ~GraphMatchingCPUMinDeg();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingCPUStatMinDeg : public GraphMatching
{
    public:
        GraphMatchingCPUStatMinDeg(const Graph &);
        ~GraphMatchingCPUStatMinDeg();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingCPUKarpSipser : public GraphMatching
{
    public:
        GraphMatchingCPUKarpSipser(const Graph &);
        ~GraphMatchingCPUKarpSipser();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingCPUWeighted : public GraphMatching
{
    public:
        GraphMatchingCPUWeighted(const Graph &);
        ~GraphMatchingCPUWeighted();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingCPUWeightedEdge : public GraphMatching
{
    public:
        GraphMatchingCPUWeightedEdge(const Graph &);
        ~GraphMatchingCPUWeightedEdge();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

// Records ev on the null stream and waits for it when a HIP device is
// present; the CPU matchers time themselves with these events like the
// reference (src/matchcpu.cu:266-267), and skip them without a device.
void recordHostEvent(hipEvent_t &ev);

}  // namespace mtc

#endif
