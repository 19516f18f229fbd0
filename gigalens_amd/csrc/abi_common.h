// Error reporting shared by the translation units of the C ABI: every entry point
// returns a gl status code and leaves a thread-local message for gl_last_error().
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "gigalens_amd.h"

namespace glc {
int fail(int code, const std::string& msg);
}  // namespace glc

#define GL_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return glc::fail(GL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace glc {

// Orders the calls on one object that owns scratch memory (a plan's workspace or
// launch-order buffer, a batch's workspace, a network's activations) across streams:
// when a call arrives on another stream than the previous call, the new stream waits
// on an event recorded on the previous one, so the scratch is never used by two
// streams at once.  Calls on one stream cost nothing extra.  A stream under graph
// capture is neither waited on nor recorded (the caller orders work before a capture).
struct StreamGuard {
    std::mutex mu;
    hipStream_t last = nullptr;
    bool used = false;
    hipEvent_t ev = nullptr;

    hipError_t init() { return hipEventCreateWithFlags(&ev, hipEventDisableTiming); }
    void release() {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    hipError_t enter(hipStream_t s) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipError_t e = hipStreamIsCapturing(s, &cs);
        if (e != hipSuccess) return e;
        if (cs != hipStreamCaptureStatusNone) return hipSuccess;
        std::lock_guard<std::mutex> lock(mu);
        if (used && s != last && ev) {
            hipStreamCaptureStatus ls = hipStreamCaptureStatusNone;
            if ((e = hipStreamIsCapturing(last, &ls)) != hipSuccess) return e;
            if (ls == hipStreamCaptureStatusNone) {
                if ((e = hipEventRecord(ev, last)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(s, ev, 0)) != hipSuccess) return e;
            }
        }
        last = s;
        used = true;
        return hipSuccess;
    }
};

}  // namespace glc
