// gm_drive.cpp: open-loop load generator for the request server.
//
// Submits request r (sources src[rp[r] .. rp[r+1])) at time arrive_us[r]
// after the start, from one host thread that spins on the steady clock, so
// arrivals follow the given schedule (e.g. Poisson gaps) without the
// overhead of an interpreter between submissions.  After the last arrival
// it waits for every request and fills out[r] (latencies measured by the
// server from submission to completion).  *wall_us is the time from the
// start to the last completion.

#include <chrono>
#include <vector>
#include "gm_kernels.h"

extern "C" int gm_server_drive(gm_server *s, int64_t nreq, const int64_t *rp, const int64_t *src,
                               const double *arrive_us, gm_server_result *out, double *wall_us)
{
    using clk = std::chrono::steady_clock;
    if (!s || nreq < 0 || (nreq > 0 && (!rp || !src || !out))) return GM_INVALID_VALUE;
    std::vector<int64_t> ids((size_t) nreq);
    std::vector<double> sub((size_t) nreq);
    const clk::time_point t0 = clk::now();
    auto since = [&t0]() {
        return std::chrono::duration<double, std::micro>(clk::now() - t0).count();
    };
    for (int64_t r = 0; r < nreq; r++)
    {
        const double due = arrive_us ? arrive_us[r] : 0.0;
        while (since() < due) { }
        sub[(size_t) r] = since();
        ids[(size_t) r] = gm_server_submit(s, src + rp[r], rp[r + 1] - rp[r], nullptr, nullptr);
        if (ids[(size_t) r] <= 0) return (int) ids[(size_t) r];
    }
    double last = 0.0;
    for (int64_t r = 0; r < nreq; r++)
    {
        const int rc = gm_server_wait(s, ids[(size_t) r], &out[r]);
        if (rc != GM_SUCCESS) return rc;
        const double done = sub[(size_t) r] + out[r].latency_us;
        if (done > last) last = done;
    }
    for (int64_t r = 0; r < nreq; r++) gm_server_release(s, ids[(size_t) r]);
    if (wall_us) *wall_us = last;
    return GM_SUCCESS;
}
