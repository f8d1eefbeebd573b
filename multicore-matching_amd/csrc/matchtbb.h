// Parallel CPU round matchers of the drop-in interface: the reference's
// GraphMatchingTBB classes (src/matchtbb.h:27-60, types 10 and 11), here on
// the library's host threads.  They run the same hash-seeded rounds as the
// GPU matchers' maximal variants (types 7 and 9) -- one rand() per executed
// round, at most 255 rounds -- and return the same match array.
#ifndef MM_DROPIN_MATCHTBB_H
#define MM_DROPIN_MATCHTBB_H

#include <hip/hip_runtime.h>

#include <vector>

#include "graph.h"
#include "matchcpu.h"

namespace mtc
{

class GraphMatchingTBB : public GraphMatching
{
    public:
        GraphMatchingTBB(const Graph &, const unsigned int &);
        ~GraphMatchingTBB();

        virtual void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const = 0;

    protected:
        void rounds(std::vector<int> &, bool weighted, hipEvent_t &, hipEvent_t &) const;
        const unsigned int selectBarrier;
};

class GraphMatchingTBBRandom : public GraphMatchingTBB
{
    public:
        GraphMatchingTBBRandom(const Graph &, const unsigned int &);
        ~GraphMatchingTBBRandom();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

class GraphMatchingTBBWeighted : public GraphMatchingTBB
{
    public:
        GraphMatchingTBBWeighted(const Graph &, const unsigned int &);
        ~GraphMatchingTBBWeighted();

        void performMatching(std::vector<int> &, hipEvent_t &, hipEvent_t &) const;
};

// Host threads of the CPU round matchers (the reference's
// task_scheduler_init(n)); 0 restores the default (all hardware threads).
void setHostThreads(int);

}  // namespace mtc

#endif
