// Graph ingest for the drop-in mtc::Graph: METIS and Matrix Market text
// readers, written from the two formats' published descriptions (the METIS
// manual's graph file format; NIST's "The Matrix Market Exchange Formats").
// Row order decides ties in the matching rounds, so the CSR each reader builds
// follows the conventions the reference's readers use for the same files
// (src/graph.cu:86-349 Matrix Market, 351-470 METIS):
//   METIS: 1-based ids; rows in file order; a row keeps the first occurrence
//     of each neighbour and drops self-loops; the edge list holds (i, j, w)
//     for i > j in row order.
//   Matrix Market: square matrices with a symmetric, skew-symmetric or
//     Hermitian pattern; each entry strictly below the diagonal is an edge of
//     weight |z|; every edge is appended to both of its rows in entry order.
// Compressed inputs (bin/match.cpp:81-110 reads *.graph.bz2 and *.mtx.bz2
// through a bzip2 filter) are decompressed while they are read, by libbz2's
// public streaming API, loaded when the first such file is opened.
#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "graph.h"

using namespace std;

namespace mtc
{

namespace
{

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


namespace
{

[[noreturn]] void reject(const string &what)
{
    cerr << what << endl;
    throw exception();
}

// Splits an istream into lines; counts them for messages.
class Lines
{
    public:
        explicit Lines(istream &in) : in_(in) {}
        // Next line (without its newline); false at the end of the input.
        bool next(string &line)
        {
            if (!getline(in_, line)) return false;
            ++number_;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            return true;
        }
        long number() const { return number_; }

    private:
        istream &in_;
        long number_ = 0;
};

// Reads whitespace-separated numbers and words off one line.
class Tokens
{
    public:
        explicit Tokens(const string &line) : p_(line.c_str()), end_(line.c_str() + line.size()) {}
        bool done()
        {
            skip();
            return p_ == end_;
        }
        bool integer(long long &v)
        {
            skip();
            if (p_ == end_) return false;
            char *stop = nullptr;
            errno = 0;
            const long long x = strtoll(p_, &stop, 10);
            if (stop == p_ || errno == ERANGE || !ends_token(stop)) return false;
            p_ = stop;
            v = x;
            return true;
        }
        bool real(double &v)
        {
            skip();
            if (p_ == end_) return false;
            char *stop = nullptr;
            const double x = strtod(p_, &stop);
            if (stop == p_ || !ends_token(stop)) return false;
            p_ = stop;
            v = x;
            return true;
        }
        bool word(string &w)
        {
            skip();
            const char *b = p_;
            while (p_ != end_ && !isspace((unsigned char)*p_)) ++p_;
            w.assign(b, p_);
            return !w.empty();
        }

    private:
        void skip()
        {
            while (p_ != end_ && isspace((unsigned char)*p_)) ++p_;
        }
        bool ends_token(const char *q) const { return q == end_ || isspace((unsigned char)*q); }
        const char *p_, *end_;
};

bool blank(const string &line)
{
    for (char c : line)
        if (!isspace((unsigned char)c)) return false;
    return true;
}

// First character that is not white space ('\0' for a blank line).
char lead(const string &line)
{
    for (char c : line)
        if (!isspace((unsigned char)c)) return c;
    return '\0';
}

string lower(string s)
{
    for (char &c : s) c = (char)tolower((unsigned char)c);
    return s;
}

// Fills g's rows from its edge list: edge k = (x, y, w) appends y to row x
// and x to row y, in edge order.
void rows_from_edges(Graph &g)
{
    const int n = g.nrVertices;
    vector<int64_t> start((size_t)n + 1, 0);
    for (const Edge &e : g.edges) {
        ++start[(size_t)e.x + 1];
        ++start[(size_t)e.y + 1];
    }
    for (int i = 0; i < n; ++i) start[(size_t)i + 1] += start[i];
    if (start[n] > INT32_MAX) reject("Graph too large: more than 2^31 - 1 row entries.");
    g.neighbours.assign((size_t)start[n], 0);
    g.neighbourWeights.assign((size_t)start[n], 1.0f);
    vector<int64_t> fill(start.begin(), start.end() - 1);
    auto put = [&](int row, int other, float w) {
        const int64_t k = fill[row]++;
        g.neighbours[k] = other;
        g.neighbourWeights[k] = w;
    };
    for (const Edge &e : g.edges) {
        put(e.x, e.y, e.w);
        put(e.y, e.x, e.w);
    }
    g.neighbourRanges.resize(n);
    for (int i = 0; i < n; ++i) g.neighbourRanges[i] = make_int2((int)start[i], (int)start[(size_t)i + 1]);
    g.nrEdges = (int)g.edges.size();
}

}  // namespace

// METIS graph file: a header "n m [fmt [ncon]]" after optional '%' comment
// lines, then one line per vertex: [size] [ncon vertex weights] and its
// neighbours (1-based), each followed by a weight when fmt asks for edge
// weights.  fmt's decimal digits, from the right: edge weights, vertex
// weights, vertex sizes.  Rows may omit a neighbour's mirror entry; the
// matchers check the CSR they are given.
istream &Graph::readMETIS(istream &in)
{
    clear();
    Lines lines(in);
    string line;
    long long n = -1, m = 0, fmt = 0, ncon = 0;
    while (lines.next(line)) {
        if (blank(line) || lead(line) == '%') continue;
        Tokens t(line);
        if (!t.integer(n) || !t.integer(m) || n < 0 || n > INT32_MAX)
            reject("METIS: the header line must start with the vertex and edge counts.");
        if (!t.done() && !t.integer(fmt)) reject("METIS: unreadable fmt field in the header.");
        if (!t.done() && !t.integer(ncon)) reject("METIS: unreadable ncon field in the header.");
        break;
    }
    if (n < 0) reject("METIS: no header line.");
    const bool edge_w = fmt % 10 != 0, vertex_w = (fmt / 10) % 10 != 0, vertex_size = (fmt / 100) % 10 != 0;
    if (ncon != 0 && !vertex_w) reject("METIS: ncon is given but fmt has no vertex weights.");
    if (vertex_w && ncon == 0) ncon = 1;
    if (ncon < 0 || ncon > 1024) reject("METIS: ncon out of range.");

    nrVertices = (int)n;
    nrVertexWeights = (int)ncon;
    vertexWeights.assign((size_t)ncon * (size_t)n, 0);
    neighbourRanges.assign((size_t)n, make_int2(0, 0));
    neighbours.reserve((size_t)max<long long>(0, 2 * m));
    neighbourWeights.reserve((size_t)max<long long>(0, 2 * m));
    // last_row[u] = the row that last listed u, for the duplicate test
    vector<int> last_row((size_t)n, -1);
    long self_loops = 0, repeats = 0;
    for (int i = 0; i < nrVertices; ++i) {
        bool got = lines.next(line);
        while (got && lead(line) == '%') got = lines.next(line);
        if (!got) {
            // A file may end without the line of an isolated last vertex.
            if (i + 1 == nrVertices) {
                neighbourRanges[i] = make_int2((int)neighbours.size(), (int)neighbours.size());
                break;
            }
            reject("METIS: the file ends after " + to_string(i) + " of " + to_string(n) + " vertex lines.");
        }
        Tokens t(line);
        long long x;
        if (vertex_size && !t.integer(x)) reject("METIS: missing vertex size on line " + to_string(lines.number()) + ".");
        for (long long k = 0; k < ncon; ++k) {
            if (!t.integer(x)) reject("METIS: missing vertex weight on line " + to_string(lines.number()) + ".");
            vertexWeights[(size_t)ncon * i + k] = (int)x;
        }
        const int row_begin = (int)neighbours.size();
        long long id;
        while (t.integer(id)) {
            if (id < 1 || id > n)
                reject("METIS: neighbour " + to_string(id) + " out of range on line " + to_string(lines.number()) + ".");
            double w = 1.0;
            if (edge_w && !t.real(w)) reject("METIS: missing edge weight on line " + to_string(lines.number()) + ".");
            const int u = (int)(id - 1);
            if (u == i) {
                ++self_loops;
                continue;
            }
            if (last_row[u] == i) {
                ++repeats;
                continue;
            }
            last_row[u] = i;
            neighbours.push_back(u);
            neighbourWeights.push_back((float)w);
        }
        if (!t.done()) reject("METIS: unreadable entry on line " + to_string(lines.number()) + ".");
        if (neighbours.size() > (size_t)INT32_MAX) reject("METIS: more than 2^31 - 1 row entries.");
        neighbourRanges[i] = make_int2(row_begin, (int)neighbours.size());
    }
    if (self_loops) cerr << "Warning: dropped " << self_loops << " self-loop (v--v) entries." << endl;
    if (repeats) cerr << "Warning: dropped " << repeats << " duplicate neighbour entries." << endl;
    for (int i = 0; i < nrVertices; ++i)
        for (int k = neighbourRanges[i].x; k < neighbourRanges[i].y; ++k)
            if (neighbours[k] < i) edges.push_back(Edge(i, neighbours[k], neighbourWeights[k]));
    nrEdges = (int)edges.size();
    if ((long long)nrEdges != m)
        cerr << "Warning: the header gives " << m << " edges, the rows hold " << nrEdges << "." << endl;
    return in;
}

// Matrix Market file: the banner "%%MatrixMarket matrix <format> <field>
// <symmetry>" (case-insensitive), '%' comment lines, a size line
// ("rows cols entries" for coordinate, "rows cols" for array format) and the
// entries ("i j [re [im]]", or one value per line for array format).  Only
// square matrices with a symmetric, skew-symmetric or Hermitian pattern are
// graphs here.
//
// Array format: entry k is placed at row k mod (rows - 1), column
// k div (rows - 1) (0-based), which is where the reference's reader puts it
// (src/graph.cu:277-281 wraps the row counter at rows - 1), so both read the
// same edges from such a file.
istream &Graph::readMatrixMarket(istream &in)
{
    clear();
    Lines lines(in);
    string line;
    enum class Format { coordinate, array } format = Format::coordinate;
    enum class Field { real, complex, integer, pattern } field = Field::real;
    bool banner = false, sized = false;
    long long rows = 0, cols = 0, declared = 0, seen = 0;
    while (lines.next(line)) {
        if (blank(line)) continue;
        if (lead(line) == '%') {
            if (banner || line.compare(0, 2, "%%") != 0) continue;  // comment
            Tokens t(line);
            string tag, object, fmt_word, field_word, sym_word;
            t.word(tag);
            if (lower(tag) != "%%matrixmarket") continue;  // some other %% comment
            if (!t.word(object) || !t.word(fmt_word) || !t.word(field_word) || !t.word(sym_word))
                reject("Matrix Market: the banner needs object, format, field and symmetry.");
            object = lower(object);
            fmt_word = lower(fmt_word);
            field_word = lower(field_word);
            sym_word = lower(sym_word);
            if (object != "matrix") reject("Matrix Market: object '" + object + "' is not a matrix.");
            if (fmt_word == "coordinate") format = Format::coordinate;
            else if (fmt_word == "array") format = Format::array;
            else reject("Matrix Market: unknown format '" + fmt_word + "'.");
            if (field_word == "real") field = Field::real;
            else if (field_word == "complex") field = Field::complex;
            else if (field_word == "integer") field = Field::integer;
            else if (field_word == "pattern") field = Field::pattern;
            else reject("Matrix Market: unknown field '" + field_word + "'.");
            if (sym_word == "general")
                reject("Matrix Market: a general (unsymmetric) matrix is not an undirected graph.");
            if (sym_word != "symmetric" && sym_word != "skew-symmetric" && sym_word != "hermitian")
                reject("Matrix Market: unknown symmetry '" + sym_word + "'.");
            if (field == Field::pattern && format == Format::array)
                reject("Matrix Market: pattern matrices must be in coordinate format.");
            banner = true;
            continue;
        }
        if (!banner) reject("Matrix Market: data before the %%MatrixMarket banner.");
        Tokens t(line);
        if (!sized) {
            if (!t.integer(rows) || !t.integer(cols)) reject("Matrix Market: unreadable size line.");
            if (rows != cols) reject("Matrix Market: a symmetric matrix must be square.");
            if (format == Format::coordinate) {
                if (!t.integer(declared)) reject("Matrix Market: the size line needs an entry count.");
            } else {
                declared = rows * cols;
            }
            if (rows <= 0 || rows > INT32_MAX || declared <= 0) reject("Matrix Market: empty or oversized matrix.");
            nrVertices = (int)rows;
            edges.reserve((size_t)min<long long>(declared, 1LL << 28));
            sized = true;
            continue;
        }
        long long r, c;
        double re = 1.0, im = 0.0;
        if (format == Format::coordinate) {
            if (!t.integer(r) || !t.integer(c)) reject("Matrix Market: unreadable entry on line " + to_string(lines.number()) + ".");
            if (r < 1 || r > rows || c < 1 || c > cols)
                reject("Matrix Market: entry (" + to_string(r) + ", " + to_string(c) + ") outside the matrix.");
            --r;
            --c;
        } else {
            const long long period = rows > 1 ? rows - 1 : 1;
            r = seen % period;
            c = seen / period;
        }
        if (field != Field::pattern) {
            if (!t.real(re)) reject("Matrix Market: missing value on line " + to_string(lines.number()) + ".");
            if (field == Field::complex && !t.real(im))
                reject("Matrix Market: missing imaginary part on line " + to_string(lines.number()) + ".");
        }
        ++seen;
        if (r > c) edges.push_back(Edge((int)r, (int)c, (float)sqrt(re * re + im * im)));
    }
    if (!banner) reject("Matrix Market: no %%MatrixMarket banner.");
    if (!sized) reject("Matrix Market: no size line.");
    if (format == Format::coordinate && seen != declared)
        cerr << "Warning: the size line gives " << declared << " entries, the file holds " << seen << "." << endl;
    rows_from_edges(*this);
    return in;
}

}  // namespace mtc
