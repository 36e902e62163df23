// The reference's Anderson-buffer study (examples/andersonBuffer/main.cu) on the drop-in header:
// three caches of the problem folder ./data/ (or argv[1]) with Anderson buffers 3, 5 and 7, each
// solved from the folder's initialState once as a warm-up and once timed, then the line
// "nodes x (states + inputs), nodes, states + inputs, time(3), time(5), time(7)" appended to
// time.csv, as the reference writes it.
//
// Build: hipcc -std=c++17 -O2 -I../../include anderson_buffer_main.cpp -L../lib -lspock_amd \
//        -Wl,-rpath,'$ORIGIN/../lib' -o anderson_buffer_main
#include <array>
#include <fstream>
#include <iostream>
#include <vector>

#include "spock_amd.hpp"

using namespace spock_amd;
typedef double real_t;

int main(int argc, char **argv) {
    const real_t minute = 60.;
    const std::array<size_t, 3> buffers = {3, 5, 7};
    std::array<real_t, 3> times = {0., 0., 0.};
    std::cout << "Reading tree files...\n";
    ScenarioTree<real_t> tree(argc > 1 ? argv[1] : "./data/");
    std::cout << "Reading problem files...\n";
    ProblemData<real_t> problem(tree);
    int rc = 0;
    try {
        const real_t tol = 1e-3, maxTime = 5 * minute;
        std::cout << "Allocating cache...\n";
        CacheBuilder builder(tree, problem);
        std::vector<std::unique_ptr<Cache<real_t>>> caches;
        for (size_t m : buffers) caches.push_back(builder.tol(tol).maxTimeSecs(maxTime).andersonBuffer(m).make_unique());
        DTensor<real_t> d_initState = DTensor<real_t>::parseFromFile(tree.path() + "initialState" + tree.fpFileExt());
        std::vector<real_t> initState(tree.numStates());
        d_initState.download(initState);
        const size_t runs = 1, warm = 1;
        std::cout << "Computing average solve time over (" << runs << ") runs with (" << warm << ") warm up runs...\n";
        for (size_t i = 0; i < runs + warm; i++) {
            for (size_t k = 0; k < buffers.size(); k++) {
                const int status = caches[k]->runSpock(initState);
                if (status != converged) throw std::runtime_error(toString(status));
                times[k] = caches[k]->solveTime();
                caches[k]->reset();
                std::cout << "Run (" << i << "), buff (" << buffers[k] << ") : " << times[k] << " s.\n";
            }
        }
    } catch (const std::exception &e) {
        std::cout << "SPOCK failed! : " << e.what() << std::endl;
        rc = 1;
    }
    std::ofstream timeScaling("time.csv", std::ios::app);
    timeScaling << tree.numNodes() * tree.numStatesAndInputs() << ", " << tree.numNodes() << ", "
                << tree.numStatesAndInputs() << ", " << times[0] << ", " << times[1] << ", " << times[2] << std::endl;
    return rc;
}
