// tc_main.cpp -- the `tc` command line (and `tc_big_graphs` when built with
// WTC_BIG=1).  Same flags, usage exit and results table as the reference's
// main (tc.cu:816-918; 64-bit table of experiments/tc_big_graphs.cu:874-891);
// the counting itself goes through the library's C ABI.  Built with
// WTC_KERNEL=1/2/3 it is the experiments' `tc_arrow` (experiments/tc_arrow.cu,
// same flags and table as tc), `tc_vertex` or `tc_edge` (experiments/
// tc_vertex.cu:733-850, tc_edge.cu: no -s, -w <wedge style>, the style in
// the table's fifth column).
#include <sys/stat.h>
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>

#include "wtc.h"

#ifndef WTC_BIG
#define WTC_BIG 0
#endif
#ifndef WTC_KERNEL
#define WTC_KERNEL 0  // 0 wedge-parallel, 1 arrow wedges, 2 vertex-parallel, 3 edge-parallel
#endif
#if WTC_BIG && WTC_KERNEL
#error "the experiment command lines are built with 32-bit offsets only"
#endif
constexpr bool kStyled = WTC_KERNEL >= 2;  // -w instead of -s

static const char *title() {
    switch (WTC_KERNEL) {
        case 1: return "Arrow Wedge Parallel Triangle Counting";
        case 2: return "Vertex-parallel Triangle Counting";
        case 3: return "Edge-parallel Triangle Counting";
        default: return WTC_BIG ? "Wedge Parallel Triangle Counting for big graphs" : "Wedge Parallel Triangle Counting";
    }
}

static const char *binary() {
    switch (WTC_KERNEL) {
        case 1: return "tc_arrow";
        case 2: return "tc_vertex";
        case 3: return "tc_edge";
        default: return WTC_BIG ? "tc_big_graphs" : "tc";
    }
}

static double wall_seconds() {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + (double)tv.tv_usec / 1e6;
}

static uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// Key of the oriented CSR a run builds: the text file's size and mtime and
// the -m / -z flags (the -p style does not change the CSR).  0 if the file
// cannot be stat'ed.
static uint64_t source_tag(const char *file, int mm, int zero) {
    struct stat st;
    if (stat(file, &st) != 0) return 0;
    const uint64_t ns = (uint64_t)st.st_mtim.tv_sec * 1000000000ull + (uint64_t)st.st_mtim.tv_nsec;
    return mix(mix((uint64_t)st.st_size) ^ ns) ^ (uint64_t)(mm * 2 + zero + 1);
}

[[noreturn]] static void usage() {
    std::printf("%s\n\n", title());
    std::printf("Usage:\n\n");
    std::printf("Either one of these must be selected:\n");
    std::printf(" -m <filename>        [Input graph in Matrix Market format]\n");
    std::printf(" -e <filename>        [Input graph in edge list format]\n");
    std::printf("Required arguments:\n");
    if (!kStyled) std::printf(" -s <num>             [Spread, a.k.a. wedges/thread]\n");
    std::printf(" -a <num>             [Adjacency matrix length] (must be divisble by 32)\n");
    std::printf("Optional arguments:\n");
    if (kStyled) std::printf(" -w <num>             [Wedge style, 0:outgoing (default), 1:arrow, 2:mixed]\n");
    std::printf(" -l <num>             [Loop count]\n");
    std::printf(" -z                   [Input graph is zero-indexed]\n");
    std::printf(" -p                   [Preprocessing style, 0:CPU, 1:GPU, 2:GPU low-memory (default)]\n");
    std::printf("\n");
    std::printf("Example:\n");
    std::printf(kStyled ? "./%s -m Amazon0302.mtx -a 8192 -l 10\n" : "./%s -m Amazon0302.mtx -s 5 -a 8192 -l 10\n",
                binary());
    std::exit(EXIT_FAILURE);
}

[[noreturn]] static void die(int rc) {
    const char *detail = wtc_last_error_detail();
    std::fprintf(stderr, "ERROR: %s%s%s\n", wtc_strerror(rc), (detail && *detail) ? ": " : "", detail ? detail : "");
    std::exit(EXIT_FAILURE);
}

int main(int argc, char **argv) {
    char *graph_filename = nullptr;
    bool graph_mm = false, graph_zero_indexed = false;
    uint32_t loop_cnt = 1, spread = 0, adjacency_matrix_len = 0;
    int preprocess_style = WTC_PREPROCESS_GPU_CONSTRAINED;
    int style = WTC_WEDGE_OUTGOING;

    while (argc > 1 && argv[1][0] == '-') {
        switch (argv[1][1]) {
            case 'm':
                graph_mm = true;
                [[fallthrough]];
            case 'e':
                if (argc < 3) usage();
                graph_filename = argv[2];
                argv += 2;
                argc -= 2;
                break;
            case 'z':
                graph_zero_indexed = true;
                argv++;
                argc--;
                break;
            case 'a':
                if (argc < 3) usage();
                adjacency_matrix_len = (uint32_t)std::atoi(argv[2]);
                if (adjacency_matrix_len % 32 != 0) usage();
                argv += 2;
                argc -= 2;
                break;
            case 'w':
                if (!kStyled || argc < 3) usage();
                if (std::atoi(argv[2]) < WTC_WEDGE_OUTGOING || std::atoi(argv[2]) > WTC_WEDGE_MIXED) usage();
                style = std::atoi(argv[2]);
                argv += 2;
                argc -= 2;
                break;
            case 's':
                if (kStyled || argc < 3) usage();
                spread = (uint32_t)std::atoi(argv[2]);
                if (spread == 0) usage();
                argv += 2;
                argc -= 2;
                break;
            case 'l':
                if (argc < 3) usage();
                loop_cnt = (uint32_t)std::atoi(argv[2]);
                argv += 2;
                argc -= 2;
                break;
            case 'p':
                if (argc < 3) usage();
                if (std::atoi(argv[2]) < WTC_PREPROCESS_CPU || std::atoi(argv[2]) > WTC_PREPROCESS_GPU_CONSTRAINED)
                    usage();
                preprocess_style = std::atoi(argv[2]);
                argv += 2;
                argc -= 2;
                break;
            default:
                usage();
        }
    }
    if (graph_filename == nullptr || (!kStyled && spread == 0)) usage();
    (void)style;  // read only by the vertex-/edge-parallel command lines

    // WTC_CSR_CACHE=<file>: the oriented CSR is loaded from that binary file
    // when its key matches the input (the prepro column is then the load
    // time); otherwise it is built as usual and written there.
    const char *cache = std::getenv("WTC_CSR_CACHE");
    const uint64_t tag = cache && *cache ? source_tag(graph_filename, graph_mm, graph_zero_indexed) : 0;
    wtc_big_graph *graph = nullptr;
    double t_preprocessing = wall_seconds();
    int rc = WTC_OK;
    if (tag) {
        uint64_t got = 0;
        if (wtc_read_csr(cache, &got, &graph) == WTC_OK && got != tag) {
            wtc_free_graph(graph);
            graph = nullptr;
        }
    }
    if (graph) {
        t_preprocessing = wall_seconds() - t_preprocessing;
    } else {
        wtc_big_graph *original = nullptr;
        rc = wtc_read_graph(graph_filename, graph_mm, graph_zero_indexed, preprocess_style, &original);
        if (rc == WTC_ERR_IO) {
            std::fprintf(stderr, "ERROR: unable to open graph file.\n");
            usage();
        }
        if (rc == WTC_ERR_ZERO_ID) {
            std::fprintf(stderr, "ERROR: zero vertex id detected but -z was not set.\n");
            usage();
        }
        if (rc == WTC_ERR_FORMAT) usage();
        if (rc) die(rc);

        t_preprocessing = wall_seconds();
        rc = wtc_preprocess(original, preprocess_style, &graph);
        t_preprocessing = wall_seconds() - t_preprocessing;
        wtc_free_graph(original);
        if (rc) die(rc);
        if (tag && wtc_write_csr(cache, graph, tag) != WTC_OK)
            std::fprintf(stderr, "WARNING: could not write the CSR cache %s\n", cache);
    }

#if !WTC_BIG
    wtc_graph *graph32 = nullptr;
    if ((rc = wtc_narrow_graph(graph, &graph32))) die(rc);
    if (kStyled)
        std::printf("%-60s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s\n", "graph", "n", "m", "a", "wedge style",
                    "triangles", "prepro (s)", "GPU copy (s)", "GPU exec (s)", "GPU total (s)", "CPU+GPU (s)");
    else
        std::printf("%-60s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s\n", "graph", "n", "m", "s", "a",
                    "triangles", "prepro (s)", "GPU copy (s)", "GPU exec (s)", "GPU total (s)", "CPU+GPU (s)");
#else
    std::printf("%-60s %16s %22s %16s %16s %22s %16s %16s %16s %16s %16s\n", "graph", "n", "m", "s", "a", "triangles",
                "prepro (s)", "GPU copy (s)", "GPU exec (s)", "GPU total (s)", "CPU+GPU (s)");
#endif

    bool warmed_up = false;
    for (uint32_t i = 0; i < loop_cnt + 1; i++) {
        double t_cpu = wall_seconds();
        wtc_gpu_time t_gpu = {0.0, 0.0};
        uint64_t triangles = 0;
#if !WTC_BIG
        switch (WTC_KERNEL) {
            case 1: rc = wtc_tc_arrow_gpu(graph32, spread, adjacency_matrix_len, &t_gpu, &triangles); break;
            case 2: rc = wtc_tc_vertex_gpu(graph32, adjacency_matrix_len, style, &t_gpu, &triangles); break;
            case 3: rc = wtc_tc_edge_gpu(graph32, adjacency_matrix_len, style, &t_gpu, &triangles); break;
            default: rc = wtc_tc_gpu(graph32, spread, adjacency_matrix_len, &t_gpu, &triangles);
        }
#else
        rc = wtc_tc_gpu_big(graph, spread, adjacency_matrix_len, &t_gpu, &triangles);
#endif
        if (rc) die(rc);
        t_cpu = wall_seconds() - t_cpu;
        t_gpu.copy /= 1000.0;
        t_gpu.exec /= 1000.0;
        if (warmed_up) {
#if !WTC_BIG
            if (kStyled) {
                const char *style_str = style == WTC_WEDGE_ARROW ? "arrow" : style == WTC_WEDGE_MIXED ? "mixed" : "outgoing";
                std::printf("%-60s %16u %16u %16u %16s %16lu %16.6f %16.6f %16.6f %16.6f %16.6f\n", graph_filename,
                            graph32->numVertices, graph32->numEdges, adjacency_matrix_len, style_str,
                            (unsigned long)triangles, t_preprocessing, t_gpu.copy, t_gpu.exec,
                            t_gpu.copy + t_gpu.exec, t_cpu);
            } else {
                std::printf("%-60s %16u %16u %16u %16u %16lu %16.6f %16.6f %16.6f %16.6f %16.6f\n", graph_filename,
                            graph32->numVertices, graph32->numEdges, spread, adjacency_matrix_len,
                            (unsigned long)triangles, t_preprocessing, t_gpu.copy, t_gpu.exec,
                            t_gpu.copy + t_gpu.exec, t_cpu);
            }
#else
            std::printf("%-60s %16u %22lu %16u %16u %22lu %16.6f %16.6f %16.6f %16.6f %16.6f\n", graph_filename,
                        graph->numVertices, (unsigned long)graph->numEdges, spread, adjacency_matrix_len,
                        (unsigned long)triangles, t_preprocessing, t_gpu.copy, t_gpu.exec, t_gpu.copy + t_gpu.exec,
                        t_cpu);
#endif
        } else {
            warmed_up = true;
        }
    }
#if !WTC_BIG
    wtc_free_graph32(graph32);
#endif
    wtc_free_graph(graph);
    return EXIT_SUCCESS;
}
