// A CPU model of streams and events for the free-ordering rule of ckks-train_amd/csrc/lifetime.h.
//
// Streams are FIFO queues of operations; an event wait is ready once its record has run. A buffer
// allocated on one stream is used by work on several streams (each use noted in the buffer's
// ledger, as the library does at every operation), then released by a thread working on some
// stream: the release queues, for every stream the rule returns, a record there and a wait on the
// allocation stream, then the free on the allocation stream. A random scheduler runs the queues in
// any order they allow; a free that runs while a use of the buffer is still queued is a violation.
//
// Prints "<rule> <scenarios> <violations>" for the library's rule and for the rule that orders
// only the releasing thread's stream.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <vector>

#include "../ckks-train_amd/csrc/lifetime.h"

namespace {

enum Kind { USE, NOP, RECORD, WAIT, FREE };
struct Op {
    Kind k;
    int ev;
};

struct Model {
    std::vector<std::deque<Op>> q;
    std::vector<bool> recorded;
    int uses_left = 0;
    bool violated = false, freed = false;

    explicit Model(int streams) : q(streams) {}
    int event() {
        recorded.push_back(false);
        return (int)recorded.size() - 1;
    }
    bool ready(const Op& o) const { return o.k != WAIT || recorded[o.ev]; }
    void run(const Op& o) {
        if (o.k == USE) --uses_left;
        if (o.k == RECORD) recorded[o.ev] = true;
        if (o.k == FREE) {
            freed = true;
            if (uses_left > 0) violated = true;
        }
    }
    // Runs the queues to completion in a random order they allow; false on a deadlock.
    bool drain(std::mt19937& g) {
        for (;;) {
            std::vector<int> can;
            bool any = false;
            for (int s = 0; s < (int)q.size(); ++s) {
                if (q[s].empty()) continue;
                any = true;
                if (ready(q[s].front())) can.push_back(s);
            }
            if (!any) return true;
            if (can.empty()) return false;
            const int s = can[g() % can.size()];
            Op o = q[s].front();
            q[s].pop_front();
            run(o);
        }
    }
};

// One scenario: returns 1 for a violation, 0 otherwise (-1 on a deadlock, which no rule may cause).
int scenario(unsigned seed, bool library_rule) {
    std::mt19937 g(seed);
    const int S = 2 + (int)(g() % 4);
    Model m(S);
    const int alloc = (int)(g() % S), releasing = (int)(g() % S);
    std::vector<int> users;  // the buffer's ledger (note_use: streams other than alloc)
    const int nuse = 1 + (int)(g() % 6);
    for (int i = 0; i < nuse; ++i) {
        const int s = (int)(g() % S);
        // unrelated work before the use, and sometimes a dependency between streams
        for (int k = (int)(g() % 3); k > 0; --k) m.q[s].push_back({NOP, -1});
        if (g() % 4 == 0) {
            const int t = (int)(g() % S);
            const int e = m.event();
            m.q[t].push_back({RECORD, e});
            m.q[s].push_back({WAIT, e});
        }
        m.q[s].push_back({USE, -1});
        ++m.uses_left;
        if (s != alloc) {
            bool seen = false;
            for (int u : users) seen = seen || u == s;
            if (!seen) users.push_back(s);
        }
    }
    // the releasing thread's last use of the buffer is on its own stream (it held a reference)
    m.q[releasing].push_back({USE, -1});
    ++m.uses_left;
    const std::vector<int> order =
        library_rule ? ckh::free_order(alloc, releasing, users) : ckh::free_order(alloc, releasing, std::vector<int>{});
    for (int r : order) {
        const int e = m.event();
        m.q[r].push_back({RECORD, e});
        m.q[alloc].push_back({WAIT, e});
    }
    m.q[alloc].push_back({FREE, -1});
    if (!m.drain(g)) return -1;
    return m.violated ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    const unsigned n = argc > 1 ? (unsigned)atoi(argv[1]) : 20000;
    for (int lib = 1; lib >= 0; --lib) {
        unsigned viol = 0, dead = 0;
        for (unsigned s = 0; s < n; ++s) {
            const int r = scenario(s * 2654435761u + 17u, lib == 1);
            viol += r == 1;
            dead += r == -1;
        }
        printf("%s %u %u %u\n", lib ? "ledger" : "releasing_stream_only", n, viol, dead);
    }
    return 0;
}
