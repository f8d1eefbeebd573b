// GPU matchers of the drop-in interface: the reference's GraphMatchingGPU
// family (src/matchgpu.h:29-84, factory codes 6-9 in bin/match.cpp:145-148)
// with the same constructor arguments, backed by the gfx950 kernels through
// the C ABI of include/mm_match.h.
//
// Seeds: like the reference, each executed round consumes one rand() from the
// process-wide glibc stream.  The maximal variants read the next 255 values
// without consuming them (a copy of glibc's random() state), run on the
// device, then consume exactly the number of rounds that ran.
#ifndef MM_DROPIN_MATCHGPU_H
#define MM_DROPIN_MATCHGPU_H

#include <vector>

#include "../../include/mm_match.h"
#include "graph.h"
#include "matchcpu.h"

namespace mtc
{

class GraphMatchingGPU : public GraphMatching
{
    public:
        GraphMatchingGPU(const Graph &, const int &, const unsigned int &);
        virtual ~GraphMatchingGPU();

        virtual void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const = 0;

        // Rounds executed by the last performMatching call.
        int lastRounds() const {return rounds;};

    protected:
        void run(std::vector<int> &, hipEvent_t &, hipEvent_t &, int kind, bool maximal) const;

        const int threadsPerBlock;
        const unsigned int selectBarrier;
        mm_graph *dgraph;
        mutable int rounds;
};

class GraphMatchingGPURandom : public GraphMatchingGPU
{
    public:
        GraphMatchingGPURandom(const Graph &, const int &, const unsigned int &);
        ~GraphMatchingGPURandom();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingGPURandomMaximal : public GraphMatchingGPU
{
    public:
        GraphMatchingGPURandomMaximal(const Graph &, const int &, const unsigned int &);
        ~GraphMatchingGPURandomMaximal();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingGPUWeighted : public GraphMatchingGPU
{
    public:
        GraphMatchingGPUWeighted(const Graph &, const int &, const unsigned int &);
        ~GraphMatchingGPUWeighted();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingGPUWeightedMaximal : public GraphMatchingGPU
{
    public:
        GraphMatchingGPUWeightedMaximal(const Graph &, const int &, const unsigned int &);
        ~GraphMatchingGPUWeightedMaximal();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

// Reads the next `count` values of glibc's rand() stream without consuming them.
std::vector<unsigned int> peekRand(int count);

// Keeps the process-wide rand() stream out of reach of the HIP runtime, which
// may draw from it: the stream's state is copied on construction and put back
// on destruction.  Every HIP call a driver makes between srand() and the last
// rand() it needs belongs inside one (the matcher classes use it internally).
class RandStreamGuard
{
    public:
        RandStreamGuard();
        ~RandStreamGuard();
        RandStreamGuard(const RandStreamGuard &) = delete;
        RandStreamGuard &operator=(const RandStreamGuard &) = delete;

    private:
        char saved[256];
        size_t bytes;
};

}  // namespace mtc

#endif
