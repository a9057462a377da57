// Ordering of a buffer's free against the streams that used it (host logic only, no HIP calls,
// so the rule can be exercised on the CPU against a model of streams).
//
// A buffer is freed on the stream it was allocated on. Work that used it may sit on other
// streams: the releasing thread's (the buffer's last readers there) and any stream on which
// another thread queued reads or writes of it (recorded at use). The free must not run before
// any of that work, so the allocation stream waits for an event recorded on each such stream
// just before the free is queued.
#pragma once

#include <algorithm>
#include <vector>

namespace ckh {

// The streams whose work so far must complete before a free queued on `alloc`: `releasing` and
// each of `users`, without `alloc` itself (stream order covers it) and without repeats.
template <class S>
std::vector<S> free_order(S alloc, S releasing, const std::vector<S>& users) {
    std::vector<S> r;
    auto add = [&](S s) {
        if (s != alloc && std::find(r.begin(), r.end(), s) == r.end()) r.push_back(s);
    };
    add(releasing);
    for (S u : users) add(u);
    return r;
}

}  // namespace ckh
