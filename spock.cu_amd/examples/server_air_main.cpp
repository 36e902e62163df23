// The reference's server air-conditioning timing driver (examples/serverAirCondition/main.cu:1-56)
// on the drop-in header: the problem folder ./data/ (or argv[1]; generate it with
// spock_amd.api.configs.generate("server", folder, horizon=h), the example's main.py) is solved
// from its initialState five times, three warm-up runs and two timed ones, each with tol(1e-3)
// and a 5-minute cap, the cache reset between runs; the average of the timed runs is appended to
// time.csv.  A run that does not converge ends the loop with "SPOCK failed! : <status>" and an
// average of 0, as in the reference.
//
// Build: hipcc -std=c++17 -O2 -I../../include server_air_main.cpp -L../lib -lspock_amd \
//        -Wl,-rpath,'$ORIGIN/../lib' -o server_air_main
#include <fstream>
#include <iostream>
#include <numeric>
#include <vector>

#include "spock_amd.hpp"

using namespace spock_amd;
typedef double real_t;

int main(int argc, char **argv) {
    const real_t minute = 60.;
    real_t avgTime = 0.;
    try {
        std::cout << "Reading tree files...\n";
        ScenarioTree<real_t> tree(argc > 1 ? argv[1] : "./data/");
        std::cout << "Reading problem files...\n";
        ProblemData<real_t> problem(tree);
        const real_t tol = 1e-3, maxTime = 5 * minute;
        std::cout << "Allocating cache...\n";
        CacheBuilder<real_t> builder(tree, problem);
        Cache<real_t> cache = builder.tol(tol).maxTimeSecs(maxTime).build();
        DTensor<real_t> d_initState = DTensor<real_t>::parseFromFile(tree.path() + "initialState" + tree.fpFileExt());
        std::vector<real_t> initState(tree.numStates());
        d_initState.download(initState);
        const size_t runs = 2, warm = 3, totalRuns = runs + warm;
        std::vector<real_t> runTimes(totalRuns, 0.);
        std::cout << "Computing average solve time over (" << runs << ") runs with (" << warm << ") warm up runs...\n";
        for (size_t i = 0; i < totalRuns; i++) {
            const int status = cache.runSpock(initState);
            if (status != converged) throw std::runtime_error(toString(status));
            runTimes[i] = cache.solveTime();
            if (i < totalRuns - 1) cache.reset();
            std::cout << "Run (" << i << ") : " << runTimes[i] << " s.\n";
        }
        avgTime = std::accumulate(runTimes.begin() + warm, runTimes.end(), real_t(0)) / runs;
    } catch (const std::exception &e) {
        std::cout << "SPOCK failed! : " << e.what() << std::endl;
    } catch (...) {
        std::cout << "SPOCK failed! : No error info.\n";
    }
    std::ofstream file("time.csv", std::ios::app);
    file << avgTime << std::endl;
    std::cout << "Saved (avgTime = " << avgTime << " s).\n";
    return 0;
}
