"""Handle lifetimes of the C ABI (spock.cu_amd/csrc/handles.hpp), on the CPU under AddressSanitizer,
LeakSanitizer and UndefinedBehaviorSanitizer.

The library's context, problem, batch and network handles follow one protocol: a batch holds a
reference to its context and to each distinct problem of its instances, a problem and a network
hold one to their context, and spk_*_destroy drops the caller's reference.  This harness builds
handles with the same ownership on handles.hpp (the context's stream is a heap block that its last
release frees, and each handle's free reads its context, as the library's do) and destroys them
in all 120 orders of five handles.  A use after free, a double free or a leak fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r'''
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "handles.hpp"

struct Ctx : spk::Counted { int *stream = new int(7); };
void freeCtx(Ctx *c) { delete c->stream; delete c; }
void releaseCtx(Ctx *c) { spk::release(c, freeCtx); }

struct Problem : spk::Counted {
    Ctx *ctx; std::vector<double> tables = std::vector<double>(64, 1.0);
    explicit Problem(Ctx *c) : ctx(spk::retain(c)) {}
    ~Problem() { if (*ctx->stream != 7) std::abort(); releaseCtx(ctx); }
};
struct Batch {
    Ctx *ctx; std::vector<Problem *> held;
    Batch(Ctx *c, std::vector<Problem *> probs) : ctx(spk::retain(c)) {
        for (Problem *p : probs)
            if (std::find(held.begin(), held.end(), p) == held.end()) held.push_back(spk::retain(p));
    }
    ~Batch() {
        if (*ctx->stream != 7) std::abort();  // waits on the context's stream
        for (Problem *p : held) { if (p->tables[3] != 1.0) std::abort(); spk::release(p, [](Problem *q) { delete q; }); }
        releaseCtx(ctx);
    }
};
struct Net {
    Ctx *ctx;
    explicit Net(Ctx *c) : ctx(spk::retain(c)) {}
    ~Net() { if (*ctx->stream != 7) std::abort(); releaseCtx(ctx); }
};

int main() {
    int orders = 0;
    std::vector<int> order = {0, 1, 2, 3, 4};
    do {
        Ctx *c = new Ctx();
        Problem *p = new Problem(c), *q = new Problem(c);
        Batch *b = new Batch(c, {p, p, q, p});  // several instances of one problem
        Net *m = new Net(c);
        for (int h : order) {
            switch (h) {
                case 0: releaseCtx(c); break;                                          // spk_ctx_destroy
                case 1: spk::release(p, [](Problem *x) { delete x; }); break;          // spk_problem_destroy
                case 2: spk::release(q, [](Problem *x) { delete x; }); break;
                case 3: delete b; break;                                              // spk_batch_destroy
                case 4: delete m; break;                                              // spk_mlp_destroy
            }
        }
        orders++;
    } while (std::next_permutation(order.begin(), order.end()));
    std::printf("%d orders\n", orders);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_handles_survive_every_destroy_order(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{os.path.join(ROOT, 'spock.cu_amd', 'csrc')}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    assert run.returncode == 0, run.stderr[-2000:]
    assert run.stdout.strip() == "120 orders"
