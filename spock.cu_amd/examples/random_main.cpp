// The reference's benchmark driver (examples/random/main.cu:24-85) on the drop-in header: solve the
// problem in ./data/ (or argv[1]) from its initialState file a few times and append the average
// solve time to time.csv.
#include <fstream>
#include <iostream>
#include <numeric>
#include <vector>

#include "spock_amd.hpp"

using namespace spock_amd;
typedef double real_t;

int main(int argc, char **argv) {
    real_t avgTime = 0.;
    try {
        std::cout << "Reading tree files...\n";
        ScenarioTree<real_t> tree(argc > 1 ? argv[1] : "./data/");
        std::cout << "Reading problem files...\n";
        ProblemData<real_t> problem(tree);
        const real_t tol = 1e-3, maxTime = 5 * 60.;
        std::cout << "Allocating cache...\n";
        CacheBuilder<real_t> builder(tree, problem);
        Cache<real_t> cache = builder.tol(tol).maxTimeSecs(maxTime).build();
        std::vector<real_t> initState = parseFromFile<real_t>(tree.path() + "initialState" + tree.fpFileExt());
        const size_t runs = 2, warm = 1;
        std::vector<real_t> runTimes(runs + warm, 0.);
        for (size_t i = 0; i < runs + warm; i++) {
            int status = cache.runSpock(initState);
            if (status != converged) throw std::runtime_error(toString(status));
            runTimes[i] = cache.solveTime();
            cache.reset();
            std::cout << "Run (" << i << ") : " << runTimes[i] << " s, " << cache.solveIter() << " iterations.\n";
        }
        avgTime = std::accumulate(runTimes.begin() + warm, runTimes.end(), 0.) / runs;
    } catch (const std::exception &e) {
        std::cout << "SPOCK failed! : " << e.what() << std::endl;
        return 1;
    }
    std::ofstream timeScaling("time.csv", std::ios::app);
    timeScaling << avgTime << std::endl;
    std::cout << "Saved (avgTime = " << avgTime << " s).\n";
    return 0;
}
