// `pdhg` command line: a drop-in for the reference CLI (gpu/pdhg.cu:937-992). Same flags, stdout status line,
// JSON keys, .sol/.dual files and exit codes (0 after any solve, 1 on a parse error, 2 on usage errors and
// unsupported QP/MILP input, which is refused before any device work or output file). This build runs on the
// GPU only: `--device cpu` is accepted by the parser (so scripts keep working) but ends with exit 2 and a
// message, after the input gate, instead of running a CPU solver.
//
// Additional flags: --fast (the fast path: FP32 reflected Halpern PDHG iterates inside FP64 iterative
// refinement = --fp32 --halpern --refine), --halpern, --reflection R, --refine,
// --max-refinements N, --profile-kernels, --warm-sol FILE / --warm-dual FILE (start point in the .sol/.dual
// formats this program writes).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "driver.hpp"

using namespace tpdhg;
using Clock = std::chrono::steady_clock;

static double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// name value lines (the .sol / .dual formats) into a dense vector by name; missing names stay 0.
static bool read_named(const char* path, const std::vector<std::string>& names, std::vector<double>& out) {
    std::ifstream in(path);
    if (!in) return false;
    std::unordered_map<std::string, size_t> pos;
    for (size_t i = 0; i < names.size(); ++i) pos.emplace(names[i], i);
    out.assign(names.size(), 0.0);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string nm, v;
        if (!(ls >> nm >> v)) continue;
        auto it = pos.find(nm);
        if (it != pos.end()) out[it->second] = std::strtod(v.c_str(), nullptr);
    }
    return true;
}

int main(int argc, char** argv) {
    const char *model = nullptr, *json = nullptr, *sol = nullptr, *dual = nullptr;
    const char *warm_sol = nullptr, *warm_dual = nullptr;
    std::string device = "gpu";
    int threads = 0;  // accepted for compatibility; the GPU path has no host threads to size
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&] {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--device") device = next();
        else if (a == "--threads") threads = std::atoi(next());
        else if (a == "--tol") o.tol = std::atof(next());
        else if (a == "--time-limit") o.time_limit = std::atof(next());
        else if (a == "--max-iter") o.max_iter = std::atol(next());
        else if (a == "--row-rel") o.row_rel = std::atof(next());
        else if (a == "--eval-freq") o.eval_freq = std::max(1, std::atoi(next()));
        else if (a == "--json") json = next();
        else if (a == "--sol") sol = next();
        else if (a == "--dual") dual = next();
        else if (a == "--fp32") o.precision_bits = 32;
        else if (a == "--no-graph") o.graph = false;
        else if (a == "--verbose") o.verbose = true;
        else if (a == "--refine") o.refine = true, o.halpern = true;
        else if (a == "--halpern") o.halpern = true;
        else if (a == "--reflection") o.rho = std::atof(next());
        else if (a == "--fast") o.halpern = true, o.refine = true, o.precision_bits = 32;
        else if (a == "--max-refinements") o.max_refinements = std::atoi(next());
        else if (a == "--profile-kernels") o.profile_kernels = true;
        else if (a == "--warm-sol") warm_sol = next();
        else if (a == "--warm-dual") warm_dual = next();
        else model = argv[i];
    }
    if (!model || (device != "gpu" && device != "cpu")) {
        std::fprintf(stderr,
                     "usage: pdhg MODEL.mps --device gpu|cpu [--threads N] [--tol 1e-4] [--time-limit S] "
                     "[--max-iter N] [--json OUT.json] [--sol OUT.sol] [--dual OUT.dual] [--fp32] [--eval-freq 64] "
                     "[--row-rel 0] [--no-graph] [--fast] [--halpern] [--reflection 1.9] [--refine] "
                     "[--max-refinements 8] "
                     "[--profile-kernels] [--warm-sol X.sol] [--warm-dual X.dual]\n");
        return 2;
    }
    auto t_start = Clock::now();
    Timings tm;
    LP lp;
    try {
        lp = flatten(read_mps(model));
    } catch (const ParseError& e) {
        std::printf("status parse_error: %s\n", e.what());
        return 1;
    }
    if (const char* why = input_error(lp)) {  // gpu/pdhg_input.hpp:20-26
        std::printf("status unsupported_model\n");
        std::fflush(stdout);
        std::fprintf(stderr, "%s\n", why);
        return 2;
    }
    tm.parse = since(t_start);
    (void)threads;
    if (o.refine && !o.halpern) o.halpern = true;
    if (device == "cpu") {
        std::fprintf(stderr, "this build solves on the GPU only; use --device gpu (there is no CPU backend)\n");
        return 2;
    }
    std::vector<double> x0, y0;
    if (warm_sol && !read_named(warm_sol, lp.col_names, x0)) {
        std::fprintf(stderr, "cannot read %s\n", warm_sol);
        return 2;
    }
    if (warm_dual && !read_named(warm_dual, lp.row_names, y0)) {
        std::fprintf(stderr, "cannot read %s\n", warm_dual);
        return 2;
    }
    Report r;
    try {
        r = solve_lp(lp, o, tm, t_start, warm_sol ? x0.data() : nullptr, warm_dual ? y0.data() : nullptr);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "HIP error: %s\n", e.what());
        return 1;
    }
    std::printf("status %s objective %.12g iterations %ld restarts %d rel_p %.2e rel_d %.2e rel_gap %.2e setup %.3fs "
                "solve %.3fs total %.3fs device %s\n",
                r.out.status.c_str(), r.objective, long(r.out.iterations), r.out.restarts, r.out.rel_primal,
                r.out.rel_dual, r.out.rel_gap, r.setup_s, r.solve_s, r.total_s, r.device_name.c_str());
    if (json)
        if (std::FILE* f = std::fopen(json, "w")) {
            std::fputs(report_json(r, lp, o).c_str(), f);
            std::fclose(f);
        }
    if (sol)
        if (std::FILE* f = std::fopen(sol, "w")) {
            for (int j = 0; j < lp.n; ++j) std::fprintf(f, "%s %.17g\n", lp.col_names[size_t(j)].c_str(), r.x[size_t(j)]);
            std::fclose(f);
        }
    if (dual)  // row multipliers y = Dr z in minimisation form (gpu/pdhg.cu:903-909)
        if (std::FILE* f = std::fopen(dual, "w")) {
            for (int i = 0; i < lp.m; ++i) std::fprintf(f, "%s %.17g\n", lp.row_names[size_t(i)].c_str(), r.y[size_t(i)]);
            std::fclose(f);
        }
    return 0;
}
