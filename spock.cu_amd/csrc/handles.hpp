// Reference counts of the C ABI's handles.
//
// A context stays alive while the caller's handle or any problem, batch or network created on it
// does; a problem while the caller's handle or any batch holding an instance of it does.
// spk_*_destroy drops the caller's reference and frees whatever reaches zero, so the handles may
// be destroyed in any order (host-language finalizers, thread-local holders and static objects
// end in orders the caller does not control).  The counter is atomic; using one handle from
// several threads at once is still the caller's business.
#pragma once
#include <atomic>

namespace spk {

struct Counted {
    std::atomic<int> refs{1};  // the caller's handle
};

template <class H>
H *retain(H *h) {
    if (h) h->refs.fetch_add(1, std::memory_order_relaxed);
    return h;
}

// Drops one reference; the last one calls free(h).
template <class H, class F>
void release(H *h, F &&free) {
    if (h && h->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) free(h);
}

}  // namespace spk
