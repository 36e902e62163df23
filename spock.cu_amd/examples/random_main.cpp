// The reference's timing driver (examples/random/main.cu:24-85) on the drop-in header: solve the
// problem folder ./data/ (or argv[1]) from its initialState file once as a warm-up and twice
// timed, then append "average solve time, host memory MB, device memory MB" to time.csv.
// The only changes a reference driver needs are the include, the namespace, and
// parseFromFile<T>(...) returning a host vector where the reference downloads a DTensor.
//
// Build: hipcc -std=c++17 -O2 -I../../include random_main.cpp -L../lib -lspock_amd \
//        -Wl,-rpath,'$ORIGIN/../lib' -o random_main
#include <hip/hip_runtime_api.h>
#include <sys/resource.h>

#include <fstream>
#include <iostream>
#include <numeric>
#include <vector>

#include "spock_amd.hpp"

using namespace spock_amd;
typedef double real_t;

static real_t deviceMemoryMB() {
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) return 0.;
    return (real_t) (totalB - freeB) / (1024. * 1024.);
}

static real_t peakHostMB() {
    struct rusage usage;
    getrusage(RUSAGE_SELF, &usage);
    return usage.ru_maxrss / 1024.;
}

int main(int argc, char **argv) {
    const real_t hostBefore = peakHostMB(), devBefore = deviceMemoryMB();
    real_t avgTime = 0., ram = 0., ramGPU = 0.;
    try {
        std::cout << "Reading tree files...\n";
        ScenarioTree<real_t> tree(argc > 1 ? argv[1] : "./data/");
        std::cout << "Reading problem files...\n";
        ProblemData<real_t> problem(tree);
        const real_t tol = 1e-3, maxTime = 5 * 60.;
        std::cout << "Allocating cache...\n";
        CacheBuilder builder(tree, problem);
        Cache cache = builder.tol(tol).maxTimeSecs(maxTime).build();
        std::vector<real_t> initState = parseFromFile<real_t>(tree.path() + "initialState" + tree.fpFileExt());
        const size_t runs = 2, warm = 1;
        std::vector<real_t> runTimes(runs + warm, 0.);
        std::cout << "Computing average solve time over (" << runs << ") runs with (" << warm << ") warm up runs...\n";
        for (size_t i = 0; i < runs + warm; i++) {
            const int status = cache.runSpock(initState);
            if (status != converged) throw std::runtime_error(toString(status));
            runTimes[i] = cache.solveTime();
            if (i == 0) {
                ram = peakHostMB() - hostBefore;
                ramGPU = deviceMemoryMB() - devBefore;
            }
            std::cout << "Run (" << i << ") : " << runTimes[i] << " s, " << cache.solveIter() << " iterations.\n";
            cache.reset();
        }
        avgTime = std::accumulate(runTimes.begin() + warm, runTimes.end(), 0.) / runs;
    } catch (const std::exception &e) {
        std::cout << "SPOCK failed! : " << e.what() << std::endl;
        return 1;
    }
    std::ofstream timeScaling("time.csv", std::ios::app);
    timeScaling << avgTime << ", " << ram << ", " << ramGPU << std::endl;
    std::cout << "Saved (avgTime = " << avgTime << " s).\n";
    return 0;
}
