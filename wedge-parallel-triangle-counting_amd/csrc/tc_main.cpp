// tc_main.cpp -- the `tc` command line (and `tc_big_graphs` when built with
// WTC_BIG=1).  Same flags, usage exit and results table as the reference's
// main (tc.cu:816-918; 64-bit table of experiments/tc_big_graphs.cu:874-891);
// the counting itself goes through the library's C ABI.
#include <sys/stat.h>
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>

#include "wtc.h"

#ifndef WTC_BIG
#define WTC_BIG 0
#endif

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
    std::printf(WTC_BIG ? "Wedge Parallel Triangle Counting for big graphs\n\n" : "Wedge Parallel Triangle Counting\n\n");
    std::printf("Usage:\n\n");
    std::printf("Either one of these must be selected:\n");
    std::printf(" -m <filename>        [Input graph in Matrix Market format]\n");
    std::printf(" -e <filename>        [Input graph in edge list format]\n");
    std::printf("Required arguments:\n");
    std::printf(" -s <num>             [Spread, a.k.a. wedges/thread]\n");
    std::printf(" -a <num>             [Adjacency matrix length] (must be divisble by 32)\n");
    std::printf("Optional arguments:\n");
    std::printf(" -l <num>             [Loop count]\n");
    std::printf(" -z                   [Input graph is zero-indexed]\n");
    std::printf(" -p                   [Preprocessing style, 0:CPU, 1:GPU, 2:GPU low-memory (default)]\n");
    std::printf("\n");
    std::printf("Example:\n");
    std::printf(WTC_BIG ? "./tc_big_graphs -m Amazon0302.mtx -s 5 -a 8192 -l 10\n"
                        : "./tc -m Amazon0302.mtx -s 5 -a 8192 -l 10\n");
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
            case 's':
                if (argc < 3) usage();
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
    if (graph_filename == nullptr || spread == 0) usage();

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
    std::printf("%-60s %16s %16s %16s %16s %16s %16s %16s %16s %16s %16s\n", "graph", "n", "m", "s", "a", "triangles",
                "prepro (s)", "GPU copy (s)", "GPU exec (s)", "GPU total (s)", "CPU+GPU (s)");
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
        rc = wtc_tc_gpu(graph32, spread, adjacency_matrix_len, &t_gpu, &triangles);
#else
        rc = wtc_tc_gpu_big(graph, spread, adjacency_matrix_len, &t_gpu, &triangles);
#endif
        if (rc) die(rc);
        t_cpu = wall_seconds() - t_cpu;
        t_gpu.copy /= 1000.0;
        t_gpu.exec /= 1000.0;
        if (warmed_up) {
#if !WTC_BIG
            std::printf("%-60s %16u %16u %16u %16u %16lu %16.6f %16.6f %16.6f %16.6f %16.6f\n", graph_filename,
                        graph32->numVertices, graph32->numEdges, spread, adjacency_matrix_len,
                        (unsigned long)triangles, t_preprocessing, t_gpu.copy, t_gpu.exec, t_gpu.copy + t_gpu.exec,
                        t_cpu);
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
