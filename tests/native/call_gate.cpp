// Thread schedules over csrc/call_gate.hpp, run by tests/test_call_gate.py.
// A "resource" (an SRS and the owner state it uses) is released only after
// the gate drains; every access a call makes checks that it is still alive.
// Thread A holds the resource's lock, fails and retires; thread B calls on
// the same resource before A locks (ran), while A holds the lock (bailed:
// sees the retirement after taking the lock) or after the retirement
// (refused at the gate).  Prints "<ran> <bailed> <refused> <violations>".
#include <atomic>
#include <chrono>
#include <cstdio>
#include <mutex>
#include <random>
#include <thread>

#include "call_gate.hpp"

using namespace dvgpu;

struct Resource {
    std::atomic<bool> alive{true};
    std::mutex mu;
    int users = 0;  // under the gate lock
};

struct Counts {
    std::atomic<int> ran{0}, bailed{0}, refused{0}, violations{0};
};

static void pause_us(int us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); }

// One call on `r` (the shape of a KZG entry point): register, lock, re-check.
static void call(CallGate& gate, Resource& r, std::atomic<bool>& retired, Counts& c, int work_us) {
    bool entered = gate.enter([&] { ++r.users; });
    if (!entered) {
        ++c.refused;
        return;
    }
    {
        std::unique_lock<std::mutex> lock(r.mu);
        if (retired) {
            ++c.bailed;
        } else {
            if (!r.alive) ++c.violations;
            pause_us(work_us);
            if (!r.alive) ++c.violations;
            ++c.ran;
        }
    }
    gate.leave([&] { --r.users; });
}

// A holds the lock for `hold_us`, fails, leaves, retires; B starts after `b_delay_us`.
static void schedule(int b_delay_us, int hold_us, int work_us, Counts& c) {
    CallGate gate;
    Resource r;
    std::atomic<bool> retired{false};
    std::atomic<bool> a_locked{false};
    std::thread a([&] {
        if (!gate.enter([&] { ++r.users; })) return;
        {
            std::unique_lock<std::mutex> lock(r.mu);
            a_locked = true;
            pause_us(hold_us);
        }  // the failure unwinds: lock released, then the registration
        gate.leave([&] { --r.users; });
        retired.exchange(true);   // Owner::retire
        gate.close_and_drain();
        if (r.users) ++c.violations;
        r.alive = false;          // release_device
    });
    std::thread b([&] {
        pause_us(b_delay_us);
        call(gate, r, retired, c, work_us);
    });
    a.join();
    b.join();
    (void)a_locked;
}

int main() {
    Counts c;
    schedule(0, 20000, 100, c);        // B waits on the lock A holds: bailed (or ran first)
    schedule(60000, 1000, 100, c);     // B long after the retirement: refused
    std::mt19937 rng(7);
    for (int i = 0; i < 300; ++i) schedule(int(rng() % 3000), int(rng() % 2000), int(rng() % 500), c);
    // B first with a long call, A later: B runs, A's retirement waits for it.
    {
        CallGate gate;
        Resource r;
        std::atomic<bool> retired{false};
        std::thread b([&] { call(gate, r, retired, c, 30000); });
        pause_us(5000);
        std::thread a([&] {
            retired.exchange(true);
            gate.close_and_drain();
            if (r.users) ++c.violations;
            r.alive = false;
        });
        a.join();
        b.join();
    }
    std::printf("%d %d %d %d\n", c.ran.load(), c.bailed.load(), c.refused.load(), c.violations.load());
    return 0;
}
