// Registration of calls that may run concurrently on one owner (the KZG
// entry points allow several threads per SRS).  Retirement closes the gate
// and waits until every registered call has left, so nothing such a call
// uses (its SRS, the owner's stream and buffers) is released under it; a
// call arriving after the gate closed is refused.  Host-only (no HIP), so
// tests build it with g++.
#pragma once
#include <condition_variable>
#include <mutex>

namespace dvgpu {

class CallGate {
  public:
    // Register a call, running `f` under the gate's lock first (lookups of
    // objects that retirement or destruction remove).  False, with `f` not
    // run, once the gate is closed.
    template <class F>
    bool enter(F&& f) {
        std::lock_guard<std::mutex> g(m_);
        if (closed_) return false;
        f();
        ++inside_;
        return true;
    }
    bool enter() {
        return enter([] {});
    }

    // Leave, running `f` under the gate's lock first (release of what the
    // call's enter() lookup took).
    template <class F>
    void leave(F&& f) {
        {
            std::lock_guard<std::mutex> g(m_);
            f();
            --inside_;
        }
        idle_.notify_all();
    }
    void leave() {
        leave([] {});
    }

    // Wait, under the gate's lock, until `ready()` holds; re-checked after
    // every leave().
    template <class P>
    void wait(P&& ready) {
        std::unique_lock<std::mutex> g(m_);
        idle_.wait(g, ready);
    }

    // Run `f` under the gate's lock (updates of what enter() looks up).
    template <class F>
    void locked(F&& f) {
        std::lock_guard<std::mutex> g(m_);
        f();
    }

    // Close the gate and wait until no call is inside.  The caller must not
    // be inside itself.
    void close_and_drain() {
        std::unique_lock<std::mutex> g(m_);
        closed_ = true;
        idle_.wait(g, [this] { return inside_ == 0; });
    }

    bool closed() const {
        std::lock_guard<std::mutex> g(m_);
        return closed_;
    }

  private:
    mutable std::mutex m_;
    std::condition_variable idle_;
    int inside_ = 0;
    bool closed_ = false;
};


}  // namespace dvgpu
