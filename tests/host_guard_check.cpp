// Host-only check of the exception containment of the library's C ABI
// (elfel.jl_amd/csrc/common.h), built and run by tests/test_host_guard.py:
//   - host_guard turns std::bad_alloc into ELFEL_ENOMEM and any other
//     exception into ELFEL_EINVAL, with the message on the context, and passes
//     ordinary return codes through;
//   - join_on_exit joins a thread pool when its scope is left by an exception,
//     after waking threads that wait for a stop flag (the batch builder's
//     worker pool), so no joinable std::thread is destroyed.
// Prints one line per failed check and exits with the number of failures.
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <stdexcept>

#include "common.h"

// api.hip's definition, restated for this host-only program
int set_err(elfel_hip_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

static int fails = 0;
static void check(bool ok, const char *what)
{
    if (!ok) {
        printf("FAIL %s\n", what);
        fails++;
    }
}

int main()
{
    elfel_hip_ctx *c = new elfel_hip_ctx();
    int rc = host_guard(c, [&]() -> int { throw std::bad_alloc(); });
    check(rc == ELFEL_ENOMEM && c->err == "host allocation failed", "bad_alloc -> ENOMEM");
    rc = host_guard(c, [&]() -> int { throw std::runtime_error("bad table"); });
    check(rc == ELFEL_EINVAL && c->err.find("bad table") != std::string::npos, "runtime_error -> EINVAL");
    rc = host_guard(c, [&]() -> int { throw 3; });
    check(rc == ELFEL_EINVAL && c->err == "host exception", "non-std exception -> EINVAL");
    rc = host_guard(nullptr, [&]() -> int { throw std::bad_alloc(); });
    check(rc == ELFEL_ENOMEM, "no context");
    rc = host_guard(c, [&]() -> int { return ELFEL_ESTATE; });
    check(rc == ELFEL_ESTATE, "return code passed through");
    rc = host_guard(c, [&]() -> int { return ELFEL_OK; });
    check(rc == ELFEL_OK, "success passed through");

    // a pool whose threads wait for work or a stop flag, left by an exception
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    std::atomic<int> exited{0};
    bool caught = false;
    try {
        std::vector<std::thread> pool;
        auto joiner = join_on_exit(pool, [&] {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv.notify_all();
        });
        for (int t = 0; t < 4; t++)
            pool.emplace_back([&] {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop; });
                exited++;
            });
        throw std::runtime_error("builder failed");
    } catch (const std::runtime_error &) {
        caught = true;
    }
    check(caught && exited == 4, "waiting pool stopped and joined on an exception");
    delete c;
    if (!fails) printf("OK\n");
    return fails;
}
