// Admission rule of the serving loop (host code only, no HIP): whether a queued
// request may start now, given the device footprint already counted.
//
//  * `resident` counts the requests that run plus the per-worker reservations of
//    the 3D handles, which stay allocated while a worker is idle;
//  * a request starts when its footprint fits beside them, or when no request is
//    running at all, so idle reservations never block the queue's head for ever;
//  * a request that could not fit the budget even on an empty device is rejected
//    when it is submitted (never_fits), so the second clause cannot admit it.
#pragma once

#include <cstdint>

namespace sdot_amd {

inline bool srv_admit(double resident, double need, double budget, int64_t running) {
    return resident + need <= budget || running == 0;
}

// bytes: the request's own footprint; reserve: the reservation of the worker
// handle that may have to be created for it (0 in the 2D lane)
inline bool srv_never_fits(double bytes, double reserve, double budget) { return bytes + reserve > budget; }

}  // namespace sdot_amd
