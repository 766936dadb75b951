// `pdhg` command line: a drop-in for the reference CLI (gpu/pdhg.cu:937-992). Same flags, stdout status line,
// JSON keys, .sol/.dual files and exit codes (0 after any solve, 1 on a parse error, 2 on usage errors and
// unsupported QP/MILP input, which is refused before any device work or output file). `--device gpu` runs the
// HIP backend, `--device cpu [--threads N]` the reference scheme on host threads (the reference's CPU backend;
// N = 0 or absent: one thread below 200,000 nonzeros, else every hardware thread).
//
// Additional flags: --fast (the fast path: FP32 reflected Halpern PDHG iterates inside FP64 iterative
// refinement = --fp32 --halpern --refine), --halpern, --reflection R, --refine,
// --max-refinements N, --profile-kernels, --warm-sol FILE / --warm-dual FILE (start point in the .sol/.dual
// formats this program writes), --recheck (original-model recheck: chk_* keys in the JSON), --batch LIST (many
// models in one batched solve).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "driver.hpp"
#include "recheck.hpp"

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

// Appends the original-model recheck (gpu/bench_pdhg.py's chk_* keys) to a JSON object, and prints it.
void add_recheck(std::string& js, const LP& lp, const Report& r, bool print) {
    const Recheck c = recheck(lp, r.x.data(), r.y.data());
    char num[64];
    auto fmt = [&](double v) {
        std::snprintf(num, sizeof num, "%.17g", v);
        return std::string(std::isfinite(v) ? num : "null");
    };
    if (print)
        std::printf("recheck objective %.12g row_viol_rel %.2e bound_viol %.2e dual_bound %s dual_gap_rel %.2e "
                    "dual_infeas_rel %.2e\n",
                    c.objective, c.row_viol_rel, c.bound_viol, c.dual_bound_finite ? fmt(c.dual_bound).c_str() : "none",
                    c.dual_gap_rel, c.dual_infeas_rel);
    const std::string extra = ", \"chk_objective\": " + fmt(c.objective) + ", \"chk_row_viol_rel\": " +
                              fmt(c.row_viol_rel) + ", \"chk_bound_viol\": " + fmt(c.bound_viol) +
                              ", \"chk_dual_bound\": " + (c.dual_bound_finite ? fmt(c.dual_bound) : "null") +
                              ", \"chk_dual_obj_loose\": " + fmt(c.dual_obj_loose) + ", \"chk_dual_infeas_rel\": " +
                              fmt(c.dual_infeas_rel) + ", \"chk_dual_gap_rel\": " + fmt(c.dual_gap_rel) +
                              ", \"chk_dual_sign\": " + std::to_string(c.dual_sign);
    const size_t brace = js.rfind('}');
    if (brace != std::string::npos) js.insert(brace, extra);
}

void status_line(const Report& r) {
    std::printf("status %s objective %.12g iterations %ld restarts %d rel_p %.2e rel_d %.2e rel_gap %.2e setup %.3fs "
                "solve %.3fs total %.3fs device %s\n",
                r.out.status.c_str(), r.objective, long(r.out.iterations), r.out.restarts, r.out.rel_primal,
                r.out.rel_dual, r.out.rel_gap, r.setup_s, r.solve_s, r.total_s, r.device_name.c_str());
}

// --batch LIST: every MPS path listed (one per line; blank lines and lines starting with # skipped) is read, gated
// and solved in one batch by the fast path; one "case PATH" + status line per LP on stdout and, with --json, one
// JSON object per LP and line (the single-solve keys plus "case"). Exit codes as for one model.
int run_batch_list(const char* list, Options o, const char* json, bool do_recheck) {
    std::ifstream in(list);
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", list);
        return 2;
    }
    std::vector<std::string> paths;
    for (std::string line; std::getline(in, line);) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        size_t a = line.find_first_not_of(" \t");
        if (a == std::string::npos || line[a] == '#') continue;
        paths.push_back(line.substr(a));
    }
    if (paths.empty()) {
        std::fprintf(stderr, "%s lists no models\n", list);
        return 2;
    }
    std::vector<LP> lps(paths.size());
    for (size_t k = 0; k < paths.size(); ++k) {
        try {
            lps[k] = flatten(read_mps(paths[k]));
        } catch (const ParseError& e) {
            std::printf("status parse_error: %s: %s\n", paths[k].c_str(), e.what());
            return 1;
        }
        if (const char* why = input_error(lps[k])) {
            std::printf("status unsupported_model\n");
            std::fflush(stdout);
            std::fprintf(stderr, "%s: %s\n", paths[k].c_str(), why);
            return 2;
        }
    }
    o.halpern = true;  // batches run the fast path
    std::vector<Report> reps;
    try {
        std::vector<const LP*> ptrs;
        for (const LP& lp : lps) ptrs.push_back(&lp);
        BatchSession bs(ptrs, o);
        reps = bs.solve();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "HIP error: %s\n", e.what());
        return 1;
    }
    std::FILE* f = json ? std::fopen(json, "w") : nullptr;
    for (size_t k = 0; k < paths.size(); ++k) {
        std::printf("case %s\n", paths[k].c_str());
        status_line(reps[k]);
        std::string js = report_json(reps[k], lps[k], o);
        if (do_recheck) add_recheck(js, lps[k], reps[k], true);
        while (!js.empty() && (js.back() == '\n' || js.back() == ' ')) js.pop_back();
        if (!js.empty() && js.front() == '{') js.insert(1, "\"case\": \"" + paths[k] + "\", ");
        if (f) std::fprintf(f, "%s\n", js.c_str());
    }
    if (f) std::fclose(f);
    return 0;
}

int main(int argc, char** argv) {
    const char *model = nullptr, *json = nullptr, *sol = nullptr, *dual = nullptr;
    const char *warm_sol = nullptr, *warm_dual = nullptr, *batch_list = nullptr;
    bool do_recheck = false;
    std::string device = "gpu";
    int threads = 0;  // CPU backend threads (0: the reference's automatic choice)
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
        else if (a == "--restart-metric") o.restart_metric = std::atoi(next());
        else if (a == "--omega-rule") o.omega_rule = std::atoi(next());
        else if (a == "--fast") o.halpern = true, o.refine = true, o.precision_bits = 32;
        else if (a == "--max-refinements") o.max_refinements = std::atoi(next());
        else if (a == "--profile-kernels") o.profile_kernels = true;
        else if (a == "--warm-sol") warm_sol = next();
        else if (a == "--recheck") do_recheck = true;
        else if (a == "--batch") batch_list = next();
        else if (a == "--warm-dual") warm_dual = next();
        else model = argv[i];
    }
    if (o.restart_metric < 0 || o.restart_metric > 1 || o.omega_rule < 0 || o.omega_rule > 4) {
        std::fprintf(stderr, "--restart-metric must be 0 or 1 and --omega-rule 0, 1, 2, 3 or 4\n");
        return 2;
    }
    if (batch_list && !model && device == "gpu") return run_batch_list(batch_list, o, json, do_recheck);
    if (!model || (device != "gpu" && device != "cpu")) {
        std::fprintf(stderr,
                     "usage: pdhg MODEL.mps --device gpu|cpu [--threads N] [--tol 1e-4] [--time-limit S] "
                     "[--max-iter N] [--json OUT.json] [--sol OUT.sol] [--dual OUT.dual] [--fp32] [--eval-freq 64] "
                     "[--row-rel 0] [--no-graph] [--fast] [--halpern] [--reflection 1.9] [--restart-metric 1] [--omega-rule 3] [--refine] "
                     "[--max-refinements 8] "
                     "[--profile-kernels] [--warm-sol X.sol] [--warm-dual X.dual] [--recheck]\n"
                     "       pdhg --batch LIST.txt [--json OUT.jsonl] [solver flags]\n");
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
    if (o.refine && !o.halpern) o.halpern = true;
    if (device == "cpu") {  // the reference's CPU backend (gpu/pdhg.cu:914-934): the reference scheme on host threads
        if (o.halpern) {
            std::fprintf(stderr, "the fast path (--fast, --halpern, --refine) runs on the GPU only; use --device gpu\n");
            return 2;
        }
        o.cpu_threads = std::max(0, threads);
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
    status_line(r);
    std::string js = report_json(r, lp, o);
    if (do_recheck) add_recheck(js, lp, r, true);
    if (json)
        if (std::FILE* f = std::fopen(json, "w")) {
            std::fputs(js.c_str(), f);
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
