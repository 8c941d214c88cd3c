// Admission policy checks (host only): FIFO order, no overtaking of a large
// head request, budget and slot limits, rejection of impossible requests.
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
    std::printf("sched ok\n");
    return 0;
}
