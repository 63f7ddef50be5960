// GPU-HeiProMap: hierarchical process mapping (reference: app/GPU-HeiProMap.cpp:36-165).
//   GPU-HeiProMap --graph g.graph --hierarchy 4:8:6 --distance 1:10:100
//                 --config IM|HM|HM-ultra [--imbalance 0.03] [--mapping out.txt] [--seed S]
// k = product of the hierarchy (promap_configuration.h:111-112).  IM runs the
// multilevel process-mapping solver (promap_solver.h:146-347: coarsening,
// global multisection, Jet over the comm cost J); HM runs the hierarchical
// multisection (promap_solver.h:238-241).  J(C,D,Pi) of the result is
// reported with 64-bit accumulation.
#include "cli.hpp"

int main(int argc, char **argv) {
    auto t_start = std::chrono::steady_clock::now();
    cli::Options opt({
        {"--help", "", "Produces the help message", "", ""},
        {"--graph", "-g", "Filepath to the graph.", "", ""},
        {"--mapping", "-m", "Output filepath to the generated mapping.", "GPU-HeiProMap_par.txt", ""},
        {"--hierarchy", "-h", "Hierarchy in the form a1:a2:...:al .", "", ""},
        {"--distance", "-d", "Distance in the form d1:d2:...:dl .", "", ""},
        {"--imbalance", "-e", "Allowed imbalance (for example 0.03).", "0.03", ""},
        {"--config", "-c", "Algorithm Config {IM, HM, HM-ultra}.", "", ""},
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
    const auto hier = cli::split_ints<uint32_t>(opt.get("--hierarchy"), ':');
    const auto dist = cli::split_ints<int32_t>(opt.get("--distance"), ':');
    const double imbalance = std::stod(opt.get("--imbalance"));
    const std::string config = opt.get("--config");
    const int verbose = std::stoi(opt.get("--verbose-level"));
    const uint64_t seed = opt.is_set("--seed") ? std::stoull(opt.get("--seed")) : (uint64_t) std::random_device{}();
    const uint64_t arena = (uint64_t) (std::stod(opt.get("--arena-gb")) * (1ull << 30));
    if (config != "IM" && config != "HM" && config != "HM-ultra") {
        std::cerr << "Error: Invalid config: " << config << " (supported: IM, HM, HM-ultra)" << std::endl;
        return EXIT_FAILURE;
    }
    if (hier.empty() || hier.size() != dist.size()) {
        std::cerr << "Error: --hierarchy and --distance need the same number of levels" << std::endl;
        return EXIT_FAILURE;
    }
    uint64_t k = 1;
    for (auto h: hier) k *= h;

    heipa_host_graph *hg = nullptr;
    int rc = heipa_read_metis(graph_in.c_str(), &hg);
    if (rc != HEIPA_OK) cli::die("reading the graph", rc);
    uint32_t n, m;
    heipa_host_graph_view(hg, &n, &m, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (verbose >= 1) std::cout << "Read graph in     : " << cli::ms_since(t_start) << std::endl;

    heipa_ctx *ctx = nullptr;
    rc = heipa_ctx_create(0, arena, &ctx);
    if (rc != HEIPA_OK) cli::die("heipa_ctx_create", rc);
    auto t_solve = std::chrono::steady_clock::now();
    std::vector<uint32_t> part(n);
    int64_t J = 0, cut = 0;
    {
        cli::DeviceGraph dg(hg);
        uint32_t *dpart = nullptr;
        int32_t *dD = nullptr;
        CLI_HIP(hipMalloc(&dpart, 4ull * (n ? n : 1)));
        if (config == "IM") {
            heipa_solve_stats st{};
            rc = heipa_im_partition(ctx, &dg.g, hier.data(), dist.data(), (uint32_t) hier.size(), imbalance, seed,
                                    dpart, &st, nullptr, 0);
            if (rc != HEIPA_OK) cli::die("heipa_im_partition", rc, ctx);
            if (verbose >= 1) {
                std::cout << "Init. comm-cost   : " << st.initial_cut << std::endl;
                std::cout << "Levels            : " << st.levels << std::endl;
            }
        } else {
            rc = heipa_hm_partition(ctx, &dg.g, hier.data(), (uint32_t) hier.size(), imbalance, seed,
                                    config == "HM-ultra", dpart);
            if (rc != HEIPA_OK) cli::die("heipa_hm_partition", rc, ctx);
        }
        std::vector<int32_t> D(k * k);
        heipa_distance_matrix(hier.data(), dist.data(), (uint32_t) hier.size(), D.data());
        CLI_HIP(hipMalloc(&dD, 4ull * k * k));
        CLI_HIP(hipMemcpy(dD, D.data(), 4ull * k * k, hipMemcpyHostToDevice));
        rc = heipa_comm_cost(ctx, &dg.g, dpart, dD, (uint32_t) k, &J);
        if (rc != HEIPA_OK) cli::die("heipa_comm_cost", rc, ctx);
        rc = heipa_edge_cut(ctx, &dg.g, dpart, &cut);
        if (rc != HEIPA_OK) cli::die("heipa_edge_cut", rc, ctx);
        CLI_HIP(hipMemcpy(part.data(), dpart, 4ull * n, hipMemcpyDeviceToHost));
        CLI_HIP(hipFree(dpart));
        CLI_HIP(hipFree(dD));
    }
    const double solve_ms = cli::ms_since(t_solve);
    if (verbose >= 1) {
        std::cout << "------- Info -------" << std::endl;
        std::cout << "#Vertices         : " << n << std::endl;
        std::cout << "#Edges            : " << m << std::endl;
        std::cout << "k                 : " << k << std::endl;
        std::cout << "imbalance         : " << imbalance << std::endl;
        std::cout << "------- Stat -------" << std::endl;
        std::cout << "Final comm-cost   : " << J << std::endl;
        std::cout << "Final edge-cut    : " << cut << std::endl;
        std::cout << "Solved in         : " << solve_ms << std::endl;
    }
    if (opt.is_set("--mapping")) {
        rc = heipa_write_partition(opt.get("--mapping").c_str(), part.data(), n);
        if (rc != HEIPA_OK) cli::die("writing the mapping", rc);
    }
    heipa_ctx_destroy(ctx);
    heipa_host_graph_free(hg);
    if (verbose >= 1)
        std::cout << "Total Time spent in GPU-HeiProMap.cpp: " << cli::ms_since(t_start) / 1000.0 << " seconds."
                  << std::endl;
    return 0;
}
