// gpupsat -- command-line drop-in for the reference solver's executable
// (src/SATSolver/main.cu:109-329, options src/FileManager/ParametersManager.cpp:31-192),
// built on the C ABI of libgpsat.so only.
//
// Options (defaults as in ParametersManager.cpp:36-46):
//   -i/--input-file FILE (also positional)   -o/--output-file FILE (accepted, unused)
//   -t/--number-of-threads N (32)            -b/--number-of-blocks N (32)
//   -v/--verbosity-level N (1)               -s/--strategy distributed|uniform
//   -p/--sequential-as-parallel              -u/--preprocess-unary-clauses
//   -l/--write-log                           --help  --version
// The README's spellings --input, --threads, --blocks, --verbosity, --log are aliases.
// Extra options of this engine: --max-conflicts N, --timeout-ms MS, --decision vsids|ref,
// --learnt-cap N, --max-iterations N, --trajectory (a sequential run follows the reference's
// own search step for step, MAX_ITERATIONS 1000 unless --max-iterations is given).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "gpsat.h"

namespace {

struct Options {
    std::string input = "task.cnf";
    std::string output = "solution.txt";
    int threads = 32, blocks = 32, verbosity = 1;
    std::string strategy = "distributed";
    bool seq_as_parallel = false, write_log = false;
    long long max_conflicts = 0, max_iterations = 0;
    bool trajectory = false;
    double timeout_ms = 0;
    int decision = 0, learnt_cap = 0;
    bool any = false;
};

void usage()
{
    printf("runsat: ./runsat <file> [options]:\n"
           "  --help                                Displays this help message\n"
           "  --version                             Displays version number\n"
           "  -i [ --input-file ] arg               Input file\n"
           "  -o [ --output-file ] arg              Output file\n"
           "  -t [ --number-of-threads ] arg        Number of threads\n"
           "  -b [ --number-of-blocks ] arg         Number of blocks\n"
           "  -v [ --verbosity-level ] arg          Set verbosity level\n"
           "  -s [ --strategy ] arg                 Strategy for generating the jobs: distributed or uniform\n"
           "  -p [ --sequential-as-parallel ]       Forces the execution of the parallel strategy, with jobs\n"
           "                                        creation, but running with 1 thread and 1 block\n"
           "  -u [ --preprocess-unary-clauses ]     Turns on the pre-processing of unary clauses\n"
           "  -l [ --write-log ]                    Prints file,threads,blocks,ms to autolog.txt\n"
           "  --max-conflicts arg                   Conflict budget per job (0: none)\n"
           "  --max-iterations arg                  Decision budget per job (0: none)\n"
           "  --timeout-ms arg                      Wall-clock budget per job (0: none)\n"
           "  --decision arg                        vsids (default) or ref (last free var, positive)\n"
           "  --trajectory                          sequential run: the reference's own search, step for step\n"
           "  --learnt-cap arg                      Learned clauses held per job\n");
}

bool parse_args(int argc, char** argv, Options& o)
{
    auto need = [&](int& i) -> const char* {
        if (i + 1 >= argc) {
            fprintf(stderr, "the required argument for option '%s' is missing\n", argv[i]);
            exit(0);
        }
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string val;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
        }
        auto value = [&](int& idx) -> std::string { return val.empty() ? std::string(need(idx)) : val; };
        o.any = true;
        if (a == "--help") {
            usage();
            exit(0);
        } else if (a == "--version") {
            printf("%s\n", gpsat_version());
            exit(0);
        } else if (a == "-i" || a == "--input-file" || a == "--input") {
            o.input = value(i);
        } else if (a == "-o" || a == "--output-file") {
            o.output = value(i);
        } else if (a == "-t" || a == "--number-of-threads" || a == "--threads") {
            o.threads = atoi(value(i).c_str());
        } else if (a == "-b" || a == "--number-of-blocks" || a == "--blocks") {
            o.blocks = atoi(value(i).c_str());
        } else if (a == "-v" || a == "--verbosity-level" || a == "--verbosity") {
            o.verbosity = atoi(value(i).c_str());
        } else if (a == "-s" || a == "--strategy") {
            o.strategy = value(i);
        } else if (a == "-p" || a == "--sequential-as-parallel") {
            o.seq_as_parallel = true;
        } else if (a == "-u" || a == "--preprocess-unary-clauses" || a == "--preprocess") {
            // the reference's -u sets sequential-as-parallel (ParametersManager.cpp:116-117);
            // unary preprocessing is always on (ParametersManager.cpp:176-182)
            o.seq_as_parallel = true;
        } else if (a == "-l" || a == "--write-log" || a == "--log") {
            o.write_log = true;
        } else if (a == "--max-conflicts") {
            o.max_conflicts = atoll(value(i).c_str());
        } else if (a == "--max-iterations") {
            o.max_iterations = atoll(value(i).c_str());
        } else if (a == "--timeout-ms") {
            o.timeout_ms = atof(value(i).c_str());
        } else if (a == "--trajectory") {
            o.trajectory = true;
        } else if (a == "--decision") {
            const std::string d = value(i);
            o.decision = d == "ref" ? 1 : 0;
        } else if (a == "--learnt-cap") {
            o.learnt_cap = atoi(value(i).c_str());
        } else if (!a.empty() && a[0] == '-') {
            fprintf(stderr, "unrecognised option '%s'\n", a.c_str());
            exit(0);
        } else {
            o.input = a;   // positional input file
        }
    }
    if (!o.any) {
        usage();
        exit(0);
    }
    if (o.strategy != "distributed" && o.strategy != "uniform") {
        fprintf(stderr, "Strategy must be either distributed or uniform!\n");
        exit(0);
    }
    printf("input file:\t\t\t%s\n", o.input.c_str());
    return true;
}

bool file_exists(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

int check(int rc)
{
    if (rc != GPSAT_OK) {
        fprintf(stderr, "gpsat error %d: %s\n", rc, gpsat_last_error());
        exit(1);
    }
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    parse_args(argc, argv, o);
    if (!file_exists(o.input)) {
        printf("The specified CNF file (%s) was not found!\n", o.input.c_str());
        return -1;
    }
    gpsat_formula* f = nullptr;
    if (gpsat_dimacs_read(o.input.c_str(), &f) != GPSAT_OK) {
        printf("Error parsing inputs.\n");
        return -1;
    }
    int32_t n = 0, m = 0, hv = 0, hc = 0;
    int64_t L = 0;
    gpsat_formula_info(f, &n, &m, &L, &hv, &hc);
    std::vector<int32_t> lits(std::max<int64_t>(L, 1));
    std::vector<int64_t> off(m + 1);
    gpsat_formula_arrays(f, lits.data(), off.data());
    gpsat_formula_free(f);
    if (n > hv) printf("header claims %d vars, but highest var found is %d. Using %d...", hv, n, n);
    if (m != hc) printf("header claims %d clauses, but found %d clauses. Using %d...", hc, m, m);

    // most frequent var (CnfManager::add_lit, CnfReader.cpp:73-83)
    std::vector<int32_t> occ(std::max(n, 1), 0);
    int32_t best_var = -1, best_occ = -1;
    for (int64_t k = 0; k < L; k++) {
        const int32_t v = lits[k] >> 1;
        if (++occ[v] > best_occ) {
            best_occ = occ[v];
            best_var = v;
        }
    }

    // preprocessing (always on, ParametersManager.cpp:176-182)
    std::vector<int32_t> plits(std::max<int64_t>(L, 1)), solved(m + 1);
    std::vector<int64_t> poff(m + 1);
    int32_t pm = 0, ns = 0, pstatus = 2;
    check(gpsat_preprocess(lits.data(), off.data(), m, n, plits.data(), poff.data(), &pm, solved.data(), &ns, &pstatus));
    solved.resize(ns);

    if (n < 3 && (o.blocks > 1 || o.threads > 1 || o.seq_as_parallel)) {
        printf("Warning: There are %d vars in the formula and at least %d are necessary to parallelize. "
               "Forcing sequential execution!\n", n, 3);
        o.blocks = o.threads = 1;
        o.seq_as_parallel = false;
    }
    const bool sequential = o.threads == 1 && o.blocks == 1 && !o.seq_as_parallel;

    gpsat_solve_opts so;
    gpsat_solve_opts_default(&so);
    so.decision = o.decision;
    so.max_conflicts = o.max_conflicts;
    so.max_decisions = o.max_iterations;
    so.max_ms = o.timeout_ms;
    if (o.learnt_cap > 0) so.learnt_cap = o.learnt_cap;

    if (o.verbosity >= 1) {
        printf("Solver configuration:\n");
        printf("Input file: %s\n", o.input.c_str());
        printf("Formula has %d vars and %d clauses\n", n, pm);
        printf("Variable '%d', the most frequent, has been found %d times.\n", best_var + 1, best_occ);
        if (sequential) {
            printf("Parallelization strategy: SEQUENTIAL RUN\n");
        } else {
            printf("Parallelization strategy: Divide and Conquer\n");
            printf("Number of blocks: %d\n", o.blocks);
            printf("Number of threads: %d\n", o.threads);
            printf("Job creation strategy: %s\n", o.strategy == "uniform" ? "uniform" : "distribution per thread");
        }
        printf("Conflict analysis: ON (%s)\n", sequential && o.trajectory ? "decision scheme" : "1-UIP");
        printf("Unary clauses pre-processing is ON\n");
        printf("Conflict analysis is two wached literals\n");
    }
    printf("VSIDS is %s\n", so.decision == 0 && !(sequential && o.trajectory) ? "ON" : "OFF");
    printf("Restart is ON\n");
    printf("Clause learning is ON with learnt clause capacity of %d\n", sequential && o.trajectory ? 20 : so.learnt_cap);
    printf("Simple jobs generation is OFF\n");

    std::vector<int8_t> value(std::max(n, 1), -1);   // -1 unassigned, 0 false, 1 true
    int status = 2;
    double gpu_ms = 0.0;
    if (pstatus != 2) {
        printf("Solved in pre-processing.\n");
        status = pstatus;
    } else {
        // dead vars: the solved literals' vars (main.cu:165-174)
        std::vector<int32_t> dead;
        for (int32_t l : solved) dead.push_back(l >> 1);
        std::vector<int64_t> icoff = {0, pm};
        gpsat_batch* b = nullptr;
        check(gpsat_batch_create(plits.data(), poff.data(), icoff.data(), &n, 1, &b));
        std::vector<int32_t> ji, co{0}, cl;
        int32_t n_jobs = 0;
        if (sequential) {
            printf("About to call sequential kernel!\n");
        } else {
            int32_t vars[16], d = 0;
            check(gpsat_cube_vars(plits.data(), poff.data(), pm, n, dead.data(), (int32_t)dead.size(), o.threads,
                                  o.blocks, o.strategy == "uniform" ? 1 : 0, vars, &d));
            n_jobs = 1 << d;
            std::vector<int32_t> cubes((size_t)n_jobs * d);
            check(gpsat_cubes(vars, d, cubes.data()));
            for (int32_t j = 0; j < n_jobs; j++) {
                ji.push_back(0);
                cl.insert(cl.end(), cubes.begin() + (size_t)j * d, cubes.begin() + (size_t)(j + 1) * d);
                co.push_back((int32_t)cl.size());
            }
            printf("Number of jobs = %d\n", n_jobs);
            printf("About to invoke kernel...\n");
        }
        int32_t st = 2, verified = 0;
        std::vector<uint32_t> rows((2 * (size_t)n + 31) / 32 + 1);
        std::vector<int64_t> js((size_t)std::max(n_jobs, 1) * GPSAT_JOB_STATS);
        gpsat_solve_info info;
        if (sequential && o.trajectory) {
            // the reference's sequential search (SATSolver.cu:67-218), MAX_ITERATIONS 1000
            std::vector<uint8_t> deadv(std::max(n, 1), 0);
            for (int32_t v : dead) deadv[v] = 1;
            int64_t stats[5];
            const auto t0 = std::chrono::steady_clock::now();
            check(gpsat_solve_ref(b, deadv.data(), o.max_iterations > 0 ? (int32_t)o.max_iterations : 1000, 0, 0, &st,
                                  rows.data(), stats));
            gpu_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (st == 4) {
                fprintf(stderr, "trajectory mode ran out of learned-clause capacity\n");
                gpsat_batch_destroy(b);
                return 1;
            }
            if (o.verbosity >= 1)
                printf("Decisions %lld, conflicts %lld, learned %lld, iterations %lld, restarts %lld\n",
                       (long long)stats[0], (long long)stats[1], (long long)stats[2], (long long)stats[3],
                       (long long)stats[4]);
        } else {
            check(gpsat_solve(b, &so, sequential ? nullptr : ji.data(), sequential ? nullptr : co.data(),
                              sequential ? nullptr : cl.data(), sequential ? 0 : n_jobs, nullptr, nullptr, &st,
                              rows.data(), &verified, js.data(), nullptr, &info));
            gpu_ms = info.kernel_ms + info.verify_ms;
        }
        gpsat_batch_destroy(b);
        if (!sequential) {
            int solved_jobs = 0;
            for (int32_t j = 0; j < n_jobs; j++)
                solved_jobs += js[(size_t)j * GPSAT_JOB_STATS] == 0 || js[(size_t)j * GPSAT_JOB_STATS] == 1;
            printf("Kernel was invoked 1 times\n");
            printf("Jobs size = %d\n", n_jobs);
            printf("There were %d jobs created.\nThere were %d solved jobs\n", n_jobs, solved_jobs);
        }
        status = st;
        if (status == 0)
            for (int32_t v = 0; v < n; v++) {
                const bool t = (rows[(2 * v + 1) >> 5] >> ((2 * v + 1) & 31)) & 1;
                const bool fl = (rows[(2 * v) >> 5] >> ((2 * v) & 31)) & 1;
                value[v] = t ? 1 : fl ? 0 : -1;
            }
    }
    printf("Total time on GPU: %f ms\n", gpu_ms);

    // Results::print_results (src/SATSolver/Results.cu:62-141)
    if (status == 2) {
        printf("UNDEFINED\n");
    } else if (status == 1) {
        printf("UNSATISFIABLE\n");
    } else {
        printf("SATISFIABLE\n");
        for (int32_t l : solved) value[l >> 1] = (int8_t)(l & 1);
        std::string line;
        for (int32_t v = 0; v < n; v++) {
            line += value[v] == 0 ? "-" : "";
            line += std::to_string(v + 1) + " ";
        }
        printf("%s0\n", line.c_str());
        // every clause of the input formula must hold a true literal (unassigned vars
        // count as true, as printed)
        bool ok = true;
        for (int32_t c = 0; c < m && ok; c++) {
            bool sat = false;
            for (int64_t k = off[c]; k < off[c + 1] && !sat; k++) {
                const int32_t l = lits[k];
                const int8_t x = value[l >> 1];
                sat = (l & 1) ? x != 0 : x == 0;
            }
            ok = sat;
        }
        printf("Solution %s\n", ok ? "was verified" : "was not verified");
    }
    if (o.write_log) {
        FILE* lf = fopen("autolog.txt", "a");
        if (lf) {
            fprintf(lf, "%s,%d,%d,%f\n", o.input.c_str(), o.threads, o.blocks, gpu_ms);
            fclose(lf);
        }
    }
    return 0;
}
