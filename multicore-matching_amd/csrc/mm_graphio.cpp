// Graph ingest for the drop-in mtc::Graph: METIS and Matrix Market text
// readers with the reference's conventions, so real inputs produce the same
// CSR (and therefore the same matching) as the reference would:
//   METIS (src/graph.cu:351-511): 1-based ids, optional edge/vertex weights
//     from the fmt flags, self-loops and repeated neighbours dropped (first
//     occurrence kept), rows kept in file order; edges = (i, j) with i > j.
//   Matrix Market (src/graph.cu:86-349): symmetric / skew / Hermitian square
//     matrices only; entries strictly below the diagonal become edges with
//     weight |z| = sqrt(re^2 + im^2); each edge is appended to both rows in
//     edge order.
// Compressed inputs (bin/match.cpp:81-110 reads *.graph.bz2 and *.mtx.bz2
// through a bzip2 filter) are decompressed while they are read, by libbz2's
// public streaming API, loaded when the first such file is opened.
#include <dlfcn.h>

#include <cctype>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <cmath>
#include <exception>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "graph.h"

using namespace std;

namespace mtc
{

namespace
{

[[noreturn]] void fail(const char *msg)
{
    cerr << msg << endl;
    throw exception();
}

string strip_leading(string line)
{
    size_t k = 0;
    while (k < line.size() && isspace((unsigned char)line[k])) ++k;
    return line.substr(k);
}

// bzip2 streaming decompression (libbz2's documented low-level interface:
// a bz_stream, BZ2_bzDecompressInit / BZ2_bzDecompress / BZ2_bzDecompressEnd).
struct BzStream {
    char *next_in;
    unsigned int avail_in, total_in_lo32, total_in_hi32;
    char *next_out;
    unsigned int avail_out, total_out_lo32, total_out_hi32;
    void *state;
    void *(*bzalloc)(void *, int, int);
    void (*bzfree)(void *, void *);
    void *opaque;
};
constexpr int BZ_OK_ = 0, BZ_STREAM_END_ = 4;

struct Bz2Api {
    int (*init)(BzStream *, int, int) = nullptr;
    int (*run)(BzStream *) = nullptr;
    int (*end)(BzStream *) = nullptr;
};

const Bz2Api *bz2_api()
{
    static Bz2Api api;
    static std::once_flag once;
    std::call_once(once, [] {
        void *h = dlopen("libbz2.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("libbz2.so.1.0", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        api.init = reinterpret_cast<int (*)(BzStream *, int, int)>(dlsym(h, "BZ2_bzDecompressInit"));
        api.run = reinterpret_cast<int (*)(BzStream *)>(dlsym(h, "BZ2_bzDecompress"));
        api.end = reinterpret_cast<int (*)(BzStream *)>(dlsym(h, "BZ2_bzDecompressEnd"));
    });
    return api.init && api.run && api.end ? &api : nullptr;
}

// A read-only streambuf over a .bz2 file: compressed bytes in 1 MiB chunks,
// decompressed into a 4 MiB window on each underflow.  Concatenated bzip2
// streams (as written by parallel compressors) are read one after another.
class Bz2Buf : public std::streambuf
{
    public:
        Bz2Buf(const string &path, const Bz2Api *api) : api_(api), file_(path.c_str(), ios_base::binary), in_(1 << 20), out_(4 << 20)
        {
            memset(&bz_, 0, sizeof(bz_));
            ok_ = file_.good() && api_->init(&bz_, 0, 0) == BZ_OK_;
            open_ = ok_;
        }
        ~Bz2Buf() override
        {
            if (open_) api_->end(&bz_);
        }
        bool ok() const { return ok_; }

    protected:
        int_type underflow() override
        {
            if (gptr() < egptr()) return traits_type::to_int_type(*gptr());
            while (ok_) {
                if (bz_.avail_in == 0 && !eof_) {
                    file_.read(in_.data(), (streamsize)in_.size());
                    const streamsize got = file_.gcount();
                    if (got <= 0) eof_ = true;
                    bz_.next_in = in_.data();
                    bz_.avail_in = (unsigned int)max<streamsize>(got, 0);
                }
                if (!open_) {
                    // Between streams: more input starts another one.
                    if (bz_.avail_in == 0) return traits_type::eof();
                    if (api_->init(&bz_, 0, 0) != BZ_OK_) { ok_ = false; break; }
                    open_ = true;
                }
                bz_.next_out = out_.data();
                bz_.avail_out = (unsigned int)out_.size();
                const int rc = api_->run(&bz_);
                const size_t produced = out_.size() - bz_.avail_out;
                if (rc == BZ_STREAM_END_) {
                    api_->end(&bz_);
                    open_ = false;
                } else if (rc != BZ_OK_ || (produced == 0 && bz_.avail_in == 0 && eof_)) {
                    // Corrupt or truncated input: the istream turns this into badbit.
                    ok_ = false;
                    throw ios_base::failure("corrupt bzip2 input");
                }
                if (produced > 0) {
                    setg(out_.data(), out_.data(), out_.data() + produced);
                    return traits_type::to_int_type(*gptr());
                }
                if (!open_ && bz_.avail_in == 0 && eof_) return traits_type::eof();
            }
            return traits_type::eof();
        }

    private:
        const Bz2Api *api_;
        ifstream file_;
        vector<char> in_, out_;
        BzStream bz_;
        bool ok_ = false, open_ = false, eof_ = false;
};

class Bz2Istream : public std::istream
{
    public:
        Bz2Istream(const string &path, const Bz2Api *api) : std::istream(nullptr), buf_(path, api) { rdbuf(&buf_); }
        bool ok() const { return buf_.ok(); }

    private:
        Bz2Buf buf_;
};

}  // namespace

std::unique_ptr<std::istream> openGraphInput(const std::string &fileName, std::string &error)
{
    if (fileName.find(".bz2") == string::npos) {
        std::unique_ptr<std::istream> in(new ifstream(fileName.c_str()));
        if (!in->good()) {
            error = "Unable to open '" + fileName + "'!";
            return nullptr;
        }
        return in;
    }
    const Bz2Api *api = bz2_api();
    if (!api) {
        error = "Unable to load libbz2 to read '" + fileName + "'!";
        return nullptr;
    }
    std::unique_ptr<Bz2Istream> in(new Bz2Istream(fileName, api));
    if (!in->ok()) {
        error = "Unable to open '" + fileName + "' as a bzip2 file!";
        return nullptr;
    }
    return std::unique_ptr<std::istream>(in.release());
}

istream &Graph::readMETIS(istream &in)
{
    clear();
    int flags = 0;
    bool have_header = false;
    while (!have_header && in.good()) {
        string line;
        getline(in, line);
        line = strip_leading(line);
        if (line.length() < 2 || line[0] == '%') continue;
        istringstream sline(line);
        if (!(sline >> nrVertices >> nrEdges)) fail("Unable to read METIS header!");
        nrVertexWeights = 0;
        sline >> flags >> nrVertexWeights;
        have_header = true;
    }
    const bool weightedEdges = (flags % 10) != 0, weightedVertices = ((flags / 10) % 10) != 0;
    if (nrVertexWeights != 0 && !weightedVertices) fail("Invalid vertex weight specification!");
    if (weightedVertices && nrVertexWeights == 0) nrVertexWeights = 1;
    if (nrVertices < 0) fail("Unable to read METIS header!");

    neighbourRanges.assign(nrVertices, make_int2(0, 0));
    vertexWeights.assign(nrVertexWeights > 0 ? (size_t)nrVertexWeights * nrVertices : 0, 0);
    bool warnedSelf = false, warnedDuplicates = false;
    for (int i = 0; i < nrVertices; ++i) {
        if (!in.good()) fail("Read error!");
        string line;
        getline(in, line);
        istringstream sline(line);
        for (int k = 0; weightedVertices && k < nrVertexWeights; ++k) sline >> vertexWeights[(size_t)nrVertexWeights * i + k];
        set<int> seen;
        neighbourRanges[i].x = (int)neighbours.size();
        int id;
        while (sline >> id) {
            if (id < 1 || id > nrVertices) fail("Invalid neighbour index!");
            float w = 1.0f;
            if (weightedEdges) sline >> w;
            if (id - 1 == i) {
                if (!warnedSelf) {
                    warnedSelf = true;
                    cerr << "Warning, skipping v--v edges!" << endl;
                }
                continue;
            }
            if (!seen.insert(id).second) {
                if (!warnedDuplicates) {
                    warnedDuplicates = true;
                    cerr << "Warning, duplicate neighbours!" << endl;
                }
                continue;
            }
            neighbours.push_back(id - 1);
            neighbourWeights.push_back(w);
        }
        neighbourRanges[i].y = (int)neighbours.size();
    }
    for (int i = 0; i < nrVertices; ++i)
        for (int j = neighbourRanges[i].x; j < neighbourRanges[i].y; ++j)
            if (i > neighbours[j]) edges.push_back(Edge(i, neighbours[j], neighbourWeights[j]));
    if (nrEdges != (int)edges.size()) {
        cerr << "Warning: Invalid edge count!" << endl;
        nrEdges = (int)edges.size();
    }
    return in;
}

istream &Graph::readMatrixMarket(istream &in)
{
    enum { NoStorage, Coordinate, Array } storage = NoStorage;
    enum { NoData, Real, Complex, Integer, Pattern } data = NoData;
    enum { NoLayout, General, Symmetric, Skew, Hermitian } layout = NoLayout;
    bool haveDims = false;
    int nrRows = 0, nrCols = 0, nrEntries = 0, arrRow = 1, arrCol = 1;
    clear();
    while (in.good()) {
        string line;
        getline(in, line);
        line = strip_leading(line);
        if (line.length() < 2) continue;
        istringstream sline(line);
        if (line[0] == '%' && line[1] == '%') {
            string word;
            while (sline >> word) {
                if (word == "%%MatrixMarket" || word == "matrix") continue;
                if (word == "coordinate") storage = Coordinate;
                else if (word == "array") storage = Array;
                else if (word == "float" || word == "real" || word == "double") data = Real;
                else if (word == "complex") data = Complex;
                else if (word == "int" || word == "integer") data = Integer;
                else if (word == "pattern") data = Pattern;
                else if (word == "general") layout = General;
                else if (word == "symmetric") layout = Symmetric;
                else if (word == "skew" || word == "skew-symmetric") layout = Skew;
                else if (word == "hermitean" || word == "Hermitean" || word == "hermitian") layout = Hermitian;
                else cerr << "Unknown Matrix Market header tag '" << word << "'!" << endl;
            }
            continue;
        }
        if (line[0] == '%') continue;
        if (storage == NoStorage || data == NoData || layout == NoLayout) fail("Incomplete Matrix Market header!");
        if (layout == General) fail("Matrix with unsymmetric nonzero pattern!");
        if (!haveDims) {
            sline >> nrRows >> nrCols;
            if (nrRows != nrCols) fail("Non-square matrix with symmetry flags!");
            if (storage == Coordinate) sline >> nrEntries;
            else nrEntries = nrRows * nrCols;
            if (nrEntries == 0 || nrRows == 0) fail("Invalid Matrix Market dimensions!");
            nrVertices = nrRows;
            neighbourRanges.assign(nrVertices, make_int2(0, 0));
            edges.reserve(nrEntries);
            haveDims = true;
            continue;
        }
        int row, col;
        double re = 1.0, im = 0.0;
        if (storage == Coordinate) {
            sline >> row >> col;
            if (row < 1 || row > nrRows || col < 1 || col > nrCols) {
                cerr << "Invalid row and/or column index " << row << ", " << col << "!" << endl;
                throw exception();
            }
            --row;
            --col;
            if (data != Pattern) {
                sline >> re;
                if (data == Complex) sline >> im;
            }
        } else {
            row = arrRow - 1;
            col = arrCol - 1;
            sline >> re;
            if (data == Complex) sline >> im;
            if (++arrRow >= nrRows) {
                arrRow = 1;
                arrCol++;
            }
        }
        if (row > col) edges.push_back(Edge(row, col, (float)sqrt(re * re + im * im)));
    }
    // Both directions of every edge, appended in edge order.
    nrEdges = (int)edges.size();
    vector<int> deg(nrVertices, 0);
    for (const Edge &e : edges) {
        ++deg[e.x];
        ++deg[e.y];
    }
    int off = 0;
    for (int i = 0; i < nrVertices; ++i) {
        neighbourRanges[i] = make_int2(off, off);
        off += deg[i];
    }
    neighbours.assign(2 * edges.size(), 0);
    neighbourWeights.assign(2 * edges.size(), 1.0f);
    for (const Edge &e : edges) {
        neighbours[neighbourRanges[e.x].y] = e.y;
        neighbourWeights[neighbourRanges[e.x].y++] = e.w;
        neighbours[neighbourRanges[e.y].y] = e.x;
        neighbourWeights[neighbourRanges[e.y].y++] = e.w;
    }
    return in;
}

}  // namespace mtc
