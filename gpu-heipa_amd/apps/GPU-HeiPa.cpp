// GPU-HeiPa: k-way edge-cut partitioner (reference: app/GPU-HeiPa.cpp:36-162).
//   GPU-HeiPa --graph g.graph --k 32 [--imbalance 0.03] --config default|ultra
//             [--mapping out.txt] [--seed S] [--verbose-level L]
// Writes one block id per line to --mapping when given (util.h:288-329).
#include "cli.hpp"

int main(int argc, char **argv) {
    auto t_start = std::chrono::steady_clock::now();
    cli::Options opt({
        {"--help", "", "Produces the help message", "", ""},
        {"--graph", "-g", "Filepath to the graph.", "", ""},
        {"--mapping", "-m", "Output filepath to the generated mapping.", "GPU-HeiPa_par.txt", ""},
        {"--k", "-k", "Number of blocks k", "", ""},
        {"--imbalance", "-e", "Allowed imbalance (for example 0.03).", "0.03", ""},
        {"--config", "-c", "Algorithm Config {default, ultra}.", "", ""},
        {"--verbose-level", "", "Whether to print.", "1", ""},
        {"--seed", "", "Seed of the initial partitioner (default: random).", "", ""},
        {"--arena-gb", "", "Device arena size in GiB (0 = a quarter of free HBM).", "0", ""},
    });
    if (argc == 1) {
        opt.print_help();
        return 0;
    }
    opt.parse(argc, argv);
    const std::string graph_in = opt.get("--graph");
    const uint32_t k = (uint32_t) std::stoul(opt.get("--k"));
    const double imbalance = std::stod(opt.get("--imbalance"));
    const std::string config = opt.get("--config");
    const int verbose = std::stoi(opt.get("--verbose-level"));
    const uint64_t seed = opt.is_set("--seed") ? std::stoull(opt.get("--seed")) : (uint64_t) std::random_device{}();
    const uint64_t arena = (uint64_t) (std::stod(opt.get("--arena-gb")) * (1ull << 30));
    if (config != "default" && config != "ultra") {
        std::cerr << "Error: Invalid config: " << config << std::endl;
        return EXIT_FAILURE;
    }
    if (k == 0) {
        std::cerr << "Error: k must be positive" << std::endl;
        return EXIT_FAILURE;
    }

    heipa_host_graph *hg = nullptr;
    int rc = heipa_read_metis(graph_in.c_str(), &hg);
    if (rc != HEIPA_OK) cli::die("reading the graph", rc);
    uint32_t n, m;
    int64_t gw;
    heipa_host_graph_view(hg, &n, &m, &gw, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (verbose >= 1) std::cout << "Read graph in     : " << cli::ms_since(t_start) << std::endl;

    heipa_ctx *ctx = nullptr;
    rc = heipa_ctx_create(0, arena, &ctx);
    if (rc != HEIPA_OK) cli::die("heipa_ctx_create", rc);
    auto t_solve = std::chrono::steady_clock::now();
    std::vector<uint32_t> part(n);
    heipa_solve_stats st{};
    {
        cli::DeviceGraph dg(hg);
        uint32_t *dpart = nullptr;
        CLI_HIP(hipMalloc(&dpart, 4ull * (n ? n : 1)));
        rc = heipa_partition(ctx, &dg.g, k, imbalance, seed, config == "ultra", dpart, &st, nullptr, 0);
        if (rc != HEIPA_OK) cli::die("heipa_partition", rc, ctx);
        CLI_HIP(hipMemcpy(part.data(), dpart, 4ull * n, hipMemcpyDeviceToHost));
        CLI_HIP(hipFree(dpart));
    }
    const double solve_ms = cli::ms_since(t_solve);
    if (verbose >= 1) {
        int64_t empty = 0, oload = 0, too_much = 0;
        std::vector<int64_t> bw(k, 0);
        const int32_t *vw;
        heipa_host_graph_view(hg, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &vw);
        for (uint32_t u = 0; u < n; ++u) bw[part[u]] += vw ? vw[u] : 1;
        for (uint32_t b = 0; b < k; ++b) {
            empty += bw[b] == 0;
            oload += bw[b] > st.lmax;
            too_much += bw[b] > st.lmax ? bw[b] - st.lmax : 0;
        }
        std::cout << "------- Info -------" << std::endl;
        std::cout << "Graph             : " << graph_in << std::endl;
        std::cout << "#Vertices         : " << n << std::endl;
        std::cout << "#Edges            : " << m << std::endl;
        std::cout << "k                 : " << k << std::endl;
        std::cout << "imbalance         : " << imbalance << std::endl;
        std::cout << "Lmax              : " << st.lmax << std::endl;
        std::cout << "------- Stat -------" << std::endl;
        std::cout << "Init. edge-cut    : " << st.initial_cut << std::endl;
        std::cout << "Final edge-cut    : " << st.cut << std::endl;
        std::cout << "Final max block w : " << st.maxw << std::endl;
        std::cout << "#empty partitions : " << empty << std::endl;
        std::cout << "#oload partitions : " << oload << std::endl;
        std::cout << "Sum oload weights : " << too_much << std::endl;
        std::cout << "Levels            : " << st.levels << std::endl;
        std::cout << "------- Time -------" << std::endl;
        std::cout << "Coarsening        : " << st.ms_coarsening << std::endl;
        std::cout << "Contraction       : " << st.ms_contraction << std::endl;
        std::cout << "Init. Part.       : " << st.ms_initial << std::endl;
        std::cout << "Uncontraction     : " << st.ms_uncontraction << std::endl;
        std::cout << "Refinement        : " << st.ms_refinement << std::endl;
        std::cout << "Solved in         : " << solve_ms << std::endl;
    }
    if (opt.is_set("--mapping")) {
        auto t_w = std::chrono::steady_clock::now();
        rc = heipa_write_partition(opt.get("--mapping").c_str(), part.data(), n);
        if (rc != HEIPA_OK) cli::die("writing the partition", rc);
        if (verbose >= 1) std::cout << "Write partition in: " << cli::ms_since(t_w) << std::endl;
    }
    heipa_ctx_destroy(ctx);
    heipa_host_graph_free(hg);
    if (verbose >= 1)
        std::cout << "Total Time in GPU-HeiPa.cpp : " << cli::ms_since(t_start) / 1000.0 << " seconds." << std::endl;
    return 0;
}
