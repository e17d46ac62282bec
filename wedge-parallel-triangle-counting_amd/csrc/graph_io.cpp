// graph_io.cpp -- graph ingestion: the reference's text reader (tc.cu:561-685)
// as a parallel parser over a memory-mapped file, and a binary CSR file that
// lets later runs skip reading and preprocessing.
//
// Text protocol of read_graph: header lines are skipped with fgets into a
// 256-byte buffer until one starts with a digit (tc.cu:571-574); Matrix
// Market reads one more line (the size line, tc.cu:577-579); the first pair
// is scanned from that buffer (tc.cu:590-606) and the rest of the file with
// fscanf("%d %d\n") until a conversion fails (tc.cu:608-620).  After the
// first pair the format is therefore a stream of %d tokens separated by any
// whitespace, paired in order, ending at the first token that does not
// convert.  That stream is cut at whitespace into one piece per thread; each
// piece is tokenised on its own and the pieces are joined up to the first
// one that stopped early, which gives the sequential scan's tokens exactly.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "host_util.h"
#include "wtc.h"

namespace wtc {
namespace {

// Read-only mapping of a whole file (empty files map to nothing).
struct MappedFile {
    const char *p = nullptr;
    size_t size = 0;
    int fd = -1;
    int open(const char *path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return WTC_ERR_IO;
        struct stat st;
        if (fstat(fd, &st) != 0) return WTC_ERR_IO;
        size = (size_t)st.st_size;
        if (size == 0) return WTC_OK;
        void *m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) return WTC_ERR_IO;
        p = static_cast<const char *>(m);
        return WTC_OK;
    }
    ~MappedFile() {
        if (p) munmap(const_cast<char *>(p), size);
        if (fd >= 0) ::close(fd);
    }
};

// fgets(buf, 256, f) over [p, end)
bool get_line(const char *&p, const char *end, char *buf, size_t cap) {
    if (p >= end) return false;
    size_t k = 0;
    while (k + 1 < cap && p < end) {
        const char c = *p++;
        buf[k++] = c;
        if (c == '\n') break;
    }
    buf[k] = 0;
    return true;
}

inline bool space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

// One %d conversion into a UINT_t (value through long, then narrowed).
inline bool scan_d(const char *&p, const char *end, uint32_t &out) {
    while (p < end && space(*p)) p++;
    bool neg = false;
    if (p < end && (*p == '+' || *p == '-')) neg = (*p++ == '-');
    if (p >= end || (unsigned)(*p - '0') > 9u) return false;
    uint64_t v = 0;
    bool big = false;
    for (; p < end && (unsigned)(*p - '0') <= 9u; p++) {
        if (!big) {
            v = v * 10u + (uint64_t)(*p - '0');
            big = v > (uint64_t)INT64_MAX;
        }
    }
    const int64_t sv = big ? (neg ? INT64_MIN : INT64_MAX) : (neg ? -(int64_t)v : (int64_t)v);
    out = (uint32_t)(uint64_t)sv;
    return true;
}

// Tokens of [p, end) up to the first one that does not convert, in order,
// into out[lead...] (out[0..lead) left for the caller).  Pieces per thread of
// at least kPieceBytes.
constexpr uint64_t kPieceBytes = 1u << 20;

int tokenize(const char *p, const char *end, size_t lead, std::vector<uint32_t> &out) {
    const uint64_t len = (uint64_t)(end - p);
    const int T = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)host_threads(), len / kPieceBytes));
    std::vector<const char *> cut(T + 1);
    cut[0] = p;
    cut[T] = end;
    for (int t = 1; t < T; t++) {
        const char *c = std::max(p + len * (uint64_t)t / (uint64_t)T, cut[t - 1]);
        while (c < end && !space(*c)) c++;  // tokens never contain whitespace
        cut[t] = c;
    }
    struct Piece {
        std::vector<uint32_t> tok;
        bool stopped = false;
        bool nomem = false;
    };
    std::vector<Piece> pc(T);
    auto body = [&](int t) {
        Piece &x = pc[t];
        const char *q = cut[t], *qe = cut[t + 1];
        try {
            x.tok.reserve((size_t)(qe - q) / 8 + 16);
            for (;;) {
                while (q < qe && space(*q)) q++;
                if (q >= qe) break;
                uint32_t v;
                if (!scan_d(q, qe, v)) {
                    x.stopped = true;
                    break;
                }
                x.tok.push_back(v);
            }
        } catch (const std::bad_alloc &) {
            x.nomem = true;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(body, t);
    body(0);
    for (auto &x : th) x.join();
    uint64_t total = 0;
    int last = T - 1;
    for (int t = 0; t < T; t++) {
        if (pc[t].nomem) return WTC_ERR_NOMEM;
        total += pc[t].tok.size();
        if (pc[t].stopped) {
            last = t;
            break;
        }
    }
    try {
        out.resize(lead + total);
    } catch (const std::bad_alloc &) {
        return WTC_ERR_NOMEM;
    }
    std::vector<uint64_t> off(last + 2, 0);
    for (int t = 0; t <= last; t++) off[t + 1] = off[t] + pc[t].tok.size();
    par_chunks((uint64_t)last + 1, 1, [&](uint64_t t0, uint64_t t1, int) {
        for (uint64_t t = t0; t < t1; t++) {
            if (!pc[t].tok.empty())
                std::memcpy(out.data() + lead + off[t], pc[t].tok.data(), pc[t].tok.size() * 4);
            std::vector<uint32_t>().swap(pc[t].tok);
        }
    });
    return WTC_OK;
}

// ------------------------------------------------------------- CSR files
// Layout (little-endian): 64-byte header, rowPtr as u64[n+1], colInd as
// u32[m].  The checksum covers rowPtr and colInd.
constexpr char kMagic[8] = {'W', 'T', 'C', 'C', 'S', 'R', '0', '1'};
struct CsrHeader {
    char magic[8];
    uint64_t n, m;
    uint64_t tag;       // caller's key for what the CSR was built from
    uint64_t checksum;  // payload_checksum of the two arrays
    uint64_t reserved[3];
};
static_assert(sizeof(CsrHeader) == 64, "CSR file header is 64 bytes");

// Position-dependent sum of mixed 8-byte words: sum_i mix64(w_i ^ (i * K)),
// where w_i are the bytes 8i..8i+7 of the payload (zero-padded).  A sum, so
// pieces can be hashed in parallel.
uint64_t words_hash(const unsigned char *b, uint64_t bytes, uint64_t word0) {
    uint64_t h = 0;
    const uint64_t full = bytes / 8;
    for (uint64_t i = 0; i < full; i++) {
        uint64_t w;
        std::memcpy(&w, b + 8 * i, 8);
        h += mix64(w ^ ((word0 + i) * 0x9e3779b97f4a7c15ULL));
    }
    if (bytes % 8) {
        uint64_t w = 0;
        std::memcpy(&w, b + 8 * full, bytes % 8);
        h += mix64(w ^ ((word0 + full) * 0x9e3779b97f4a7c15ULL));
    }
    return h;
}

uint64_t array_hash(const void *p, uint64_t bytes, uint64_t salt) {
    const uint64_t kPiece = 1u << 22;  // bytes per task (a multiple of 8)
    const uint64_t pieces = (bytes + kPiece - 1) / kPiece;
    std::vector<uint64_t> part(pieces, 0);
    const unsigned char *b = static_cast<const unsigned char *>(p);
    par_chunks(pieces, 1, [&](uint64_t k0, uint64_t k1, int) {
        for (uint64_t k = k0; k < k1; k++) {
            const uint64_t s = k * kPiece, e = std::min(bytes, s + kPiece);
            part[k] = words_hash(b + s, e - s, s / 8);
        }
    });
    uint64_t h = mix64(salt ^ bytes);
    for (uint64_t x : part) h += x;
    return h;
}

uint64_t payload_checksum(uint64_t n, uint64_t m, const uint64_t *rowPtr, const uint32_t *colInd) {
    return array_hash(rowPtr, (n + 1) * 8, 1) ^ mix64(array_hash(colInd, m * 4, 2));
}

}  // namespace
}  // namespace wtc

using namespace wtc;

extern "C" {

int wtc_read_graph(const char *filename, int matrix_market, int zero_indexed, int preprocess_style,
                   wtc_big_graph **out) {
    if (!filename || !out) return WTC_ERR_ARG;
    if (preprocess_style < WTC_PREPROCESS_CPU || preprocess_style > WTC_PREPROCESS_GPU_CONSTRAINED)
        return WTC_ERR_ARG;
    *out = nullptr;
    MappedFile f;
    int rc = f.open(filename);
    if (rc) return rc;
    const char *p = f.p, *end = f.p + f.size;
    char line[256];
    do {
        if (!get_line(p, end, line, sizeof line)) return WTC_ERR_FORMAT;
    } while (line[0] < '0' || line[0] > '9');
    if (matrix_market && !get_line(p, end, line, sizeof line)) return WTC_ERR_FORMAT;

    // the first pair comes from the line buffer; the stream only if it converted
    std::vector<uint32_t> tok;
    const char *lp = line, *le = line + std::strlen(line);
    uint32_t v0, w0;
    if (scan_d(lp, le, v0) && scan_d(lp, le, w0)) {
        if ((rc = tokenize(p, end, 2, tok))) return rc;
        tok[0] = v0;
        tok[1] = w0;
    }
    const uint64_t np = tok.size() / 2;  // an unpaired last token is dropped, like fscanf's failed %d
    const uint32_t shift = zero_indexed ? 0u : 1u;
    // id 0 in a 1-indexed file is an error wherever it occurs among the pairs
    // read (tc.cu:599-602,612-615); the largest id after the shift sets n
    const int T = host_threads();
    std::vector<uint32_t> mx(T, 0);
    std::atomic<int> zero{0};
    par_chunks(np, 1u << 18, [&](uint64_t i0, uint64_t i1, int t) {
        uint32_t m = mx[t];
        for (uint64_t i = i0; i < i1; i++) {
            const uint32_t v = tok[2 * i], w = tok[2 * i + 1];
            if (!zero_indexed && (v == 0 || w == 0)) zero = 1;
            if (v != w) m = std::max(m, std::max(v - shift, w - shift));
        }
        mx[t] = m;
    });
    if (zero) return WTC_ERR_ZERO_ID;
    const uint32_t max_vertex = *std::max_element(mx.begin(), mx.end());
    if ((uint64_t)max_vertex + 1 > 0xffffffffULL) return WTC_ERR_FORMAT;
    if (preprocess_style == WTC_PREPROCESS_CPU)
        return build_symmetric_tokens((uint64_t)max_vertex + 1, tok.data(), np, shift, out);
    // -p 1/2: the sort and deduplication run on the GPU (tc.cu:628-644), on
    // shifted, loop-free pairs
    uint64_t kept = 0;
    for (uint64_t i = 0; i < np; i++) {
        const uint32_t v = tok[2 * i] - shift, w = tok[2 * i + 1] - shift;
        if (v != w) {
            tok[2 * kept] = v;
            tok[2 * kept + 1] = w;
            kept++;
        }
    }
    return wtc_gpu_symmetric_default(tok.data(), kept, max_vertex + 1, out);
}

int wtc_write_csr(const char *path, const wtc_big_graph *g, uint64_t tag) {
    if (!path || !g || !g->rowPtr || (g->numEdges && !g->colInd)) return WTC_ERR_ARG;
    const uint64_t n = g->numVertices, m = g->numEdges;
    if (g->rowPtr[n] != m) return WTC_ERR_CSR;
    CsrHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.n = n;
    h.m = m;
    h.tag = tag;
    h.checksum = payload_checksum(n, m, g->rowPtr, g->colInd);
    // write to a sibling temporary and rename, so readers never see a partial file
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return WTC_ERR_IO;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
              std::fwrite(g->rowPtr, 8, (size_t)n + 1, f) == (size_t)n + 1 &&
              (m == 0 || std::fwrite(g->colInd, 4, (size_t)m, f) == (size_t)m);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return WTC_ERR_IO;
    }
    return WTC_OK;
}

int wtc_read_csr(const char *path, uint64_t *tag, wtc_big_graph **out) {
    if (!path || !out) return WTC_ERR_ARG;
    *out = nullptr;
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    CsrHeader h;
    if (f.size < sizeof h) return WTC_ERR_FORMAT;
    std::memcpy(&h, f.p, sizeof h);
    // the writer zeroes the reserved words; anything else is a damaged header
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.n == 0 || h.n > 0xffffffffULL ||
        (h.reserved[0] | h.reserved[1] | h.reserved[2]) != 0)
        return WTC_ERR_FORMAT;
    if (h.m > (f.size - sizeof h) / 4 || (f.size - sizeof h - 4 * h.m) / 8 != h.n + 1 ||
        sizeof h + 8 * (h.n + 1) + 4 * h.m != f.size)
        return WTC_ERR_FORMAT;
    wtc_big_graph *g = static_cast<wtc_big_graph *>(std::calloc(1, sizeof(wtc_big_graph)));
    uint64_t *rp = alloc_n<uint64_t>(h.n + 1);
    uint32_t *ci = alloc_n<uint32_t>(std::max<uint64_t>(h.m, 1));
    if (!g || !rp || !ci) {
        std::free(g);
        std::free(rp);
        std::free(ci);
        return WTC_ERR_NOMEM;
    }
    const char *prow = f.p + sizeof h, *pcol = prow + 8 * (h.n + 1);
    par_chunks(h.n + 1, 1u << 20, [&](uint64_t a, uint64_t b, int) { std::memcpy(rp + a, prow + 8 * a, 8 * (b - a)); });
    par_chunks(h.m, 1u << 21, [&](uint64_t a, uint64_t b, int) { std::memcpy(ci + a, pcol + 4 * a, 4 * (b - a)); });
    // checksum, then structure: offsets non-decreasing from 0 to m, ids < n
    std::atomic<int> bad{rp[0] != 0 || rp[h.n] != h.m || payload_checksum(h.n, h.m, rp, ci) != h.checksum};
    if (!bad) {
        par_chunks(h.n, 1u << 20, [&](uint64_t a, uint64_t b, int) {
            for (uint64_t v = a; v < b; v++)
                if (rp[v] > rp[v + 1]) bad = 1;
        });
        par_chunks(h.m, 1u << 21, [&](uint64_t a, uint64_t b, int) {
            for (uint64_t j = a; j < b; j++)
                if (ci[j] >= h.n) bad = 1;
        });
    }
    if (bad) {
        std::free(g);
        std::free(rp);
        std::free(ci);
        return WTC_ERR_FORMAT;
    }
    g->numVertices = (uint32_t)h.n;
    g->numEdges = h.m;
    g->rowPtr = rp;
    g->colInd = ci;
    if (tag) *tag = h.tag;
    *out = g;
    return WTC_OK;
}

}  // extern "C"
