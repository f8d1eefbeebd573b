// Graph CSR type of the drop-in interface.  Same members and layout as the
// reference's mtc::Graph (src/graph.h:30-60): int2 neighbour ranges
// (x = begin, y = end), int neighbour ids, float weights, the undirected edge
// list.  int2 comes from HIP's vector types.
#ifndef MM_DROPIN_GRAPH_H
#define MM_DROPIN_GRAPH_H

#include <hip/hip_runtime.h>

#include <iostream>
#include <memory>
#include <string>
#include <vector>

namespace mtc
{

class Edge
{
    public:
        Edge();
        Edge(const int &, const int &, const float & = 1.0);
        ~Edge();

        int x, y;
        float w;
};

class Graph
{
    public:
        Graph();
        ~Graph();

        void clear();
        std::istream &readMETIS(std::istream &);
        std::istream &readMatrixMarket(std::istream &);
        std::vector<int> random_shuffle();
        bool empty() const;

    public:
        int nrVertices, nrVertexWeights, nrEdges;
        std::vector<int> vertexWeights;
        std::vector<int2> neighbourRanges;
        std::vector<int> neighbours;
        std::vector<float> neighbourWeights;
        std::vector<Edge> edges;
};

// Opens a graph file for readMETIS / readMatrixMarket.  A name containing
// ".bz2" is decompressed while it is read (bin/match.cpp:85-110 reads such
// files through a bzip2 filter); libbz2 is loaded at run time.  Returns null
// and sets error when the file cannot be opened.
std::unique_ptr<std::istream> openGraphInput(const std::string &fileName, std::string &error);

}  // namespace mtc

#endif
