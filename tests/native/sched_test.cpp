// Admission policy checks (host only): FIFO order, no overtaking of a large
// head request, budget and slot limits, rejection of impossible requests;
// shortest-first order, ties, the age limit and an aged request that blocks.
#include <cassert>
#include <cstdio>

#include "ddfan_sched.h"

int main() {
    using ddf::Admission;
    {
        Admission a(100, 3);
        assert(!a.submit(1, 101));  // never fits
        assert(a.submit(2, 40) && a.submit(3, 40) && a.submit(4, 40) && a.submit(5, 10));
        auto v = a.admit();
        assert(v.size() == 2 && v[0] == 2 && v[1] == 3);  // 4 (40) does not fit: 80 used
        assert(a.used() == 80 && a.active() == 2 && a.waiting() == 2);
        assert(a.admit().empty());  // 5 (10) would fit but must not overtake 4
        a.release(40);
        v = a.admit();
        assert(v.size() == 2 && v[0] == 4 && v[1] == 5 && a.used() == 90 && a.active() == 3);
        a.release(40);
        a.release(40);
        a.release(10);
        assert(a.used() == 0 && a.active() == 0);
    }
    {
        Admission a(1000, 2);  // slot limit binds before memory
        for (int i = 0; i < 5; ++i) assert(a.submit(i, 10));
        auto v = a.admit();
        assert(v.size() == 2 && v[0] == 0 && v[1] == 1);
        a.release(10);
        v = a.admit();
        assert(v.size() == 1 && v[0] == 2);
    }
    {
        Admission a(100, 2);  // immediate admission only with nobody waiting
        assert(!a.admit_now(101));
        assert(a.admit_now(60) && a.used() == 60 && a.active() == 1);
        assert(!a.admit_now(50) && a.used() == 60);  // does not fit
        assert(a.submit(7, 50));                      // queued instead
        assert(!a.admit_now(10));                     // 7 is waiting: no overtaking
        assert(a.admit().empty());
        a.release(60);
        auto v = a.admit();
        assert(v.size() == 1 && v[0] == 7);
        assert(a.admit_now(40) && a.active() == 2);
        assert(!a.admit_now(1));                      // both slots taken
    }
    {
        // shortest first: least work that fits, earliest on ties
        Admission a(100, 2);
        a.set_policy(Admission::SHORTEST, 10.0);
        assert(a.submit(1, 30, 500.0, 0.0) && a.submit(2, 30, 100.0, 1.0) &&
               a.submit(3, 30, 100.0, 2.0) && a.submit(4, 90, 10.0, 3.0));
        auto v = a.admit(4.0);
        // 4 has the least work but 90 + 30 > 100 once 2 runs: 4 first (fits), then nothing
        assert(v.size() == 1 && v[0] == 4 && a.used() == 90);
        a.release(90);
        v = a.admit(5.0);
        assert(v.size() == 2 && v[0] == 2 && v[1] == 3);  // tie on work: earlier first
        a.release(30);
        a.release(30);
        v = a.admit(6.0);
        assert(v.size() == 1 && v[0] == 1);
        a.release(30);
        assert(a.used() == 0 && a.active() == 0 && a.waiting() == 0);
    }
    {
        // age limit: a request waiting max_wait or longer goes first and is not overtaken
        Admission a(100, 1);
        a.set_policy(Admission::SHORTEST, 5.0);
        assert(a.submit(1, 60, 900.0, 0.0) && a.submit(2, 10, 1.0, 1.0));
        auto v = a.admit(2.0);
        assert(v.size() == 1 && v[0] == 2);  // 1 has waited 2 s < 5 s: shortest first
        assert(a.submit(3, 10, 1.0, 2.5));
        a.release(10);
        v = a.admit(6.0);
        assert(v.size() == 1 && v[0] == 1);  // 1 has waited 6 s: it goes before 3
        a.release(60);
        v = a.admit(6.5);
        assert(v.size() == 1 && v[0] == 3);
        a.release(10);
        // an aged request that does not fit blocks the others until it does
        Admission b(100, 4);
        b.set_policy(Admission::SHORTEST, 1.0);
        assert(b.admit_now(50));
        assert(b.submit(5, 80, 900.0, 0.0) && b.submit(6, 10, 1.0, 0.5));
        v = b.admit(0.6);
        assert(v.size() == 1 && v[0] == 6);  // 5 not yet aged, does not fit; 6 fits
        assert(b.submit(7, 10, 1.0, 0.7));
        v = b.admit(2.0);
        assert(v.empty());  // 5 aged: 7 must not overtake it
        b.release(50);
        b.release(10);
        v = b.admit(2.1);
        assert(v.size() == 2 && v[0] == 5 && v[1] == 7);
    }
    std::printf("sched ok\n");
    return 0;
}
