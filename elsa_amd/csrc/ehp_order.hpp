// ehp_order.hpp -- ordering of the calls on one operator (or batch) across streams and threads.
//
// An operator owns device scratch that every call rewrites: the forward's plane-major volume
// views, the back projection's fixed-point scale, the bfloat16 accumulation volume. Two calls
// on one operator must therefore not run at the same time on the device. elsa's
// LinearOperator::apply is const and may be called from several host threads, and HIP callers
// may pass any stream, so the library orders the calls itself: a call holds the operator's
// host mutex while it enqueues, makes its stream wait for the event recorded after the
// previous call (when that call went to another stream), and records the event again after
// its own launches. Calls on one stream stay in stream order at no extra cost.
//
// Inside a stream capture (hipGraph) the wait and the record are left out: the capture
// orders its own nodes, and the stream that captures was ordered after the caller's stream
// before the capture began. Replays of a captured graph are ordered by the caller.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>

namespace ehp
{
    struct SharedEvent {
        hipEvent_t ev = nullptr;
        SharedEvent() = default;
        SharedEvent(const SharedEvent&) = delete;
        SharedEvent& operator=(const SharedEvent&) = delete;
        ~SharedEvent()
        {
            if (ev)
                (void) hipEventDestroy(ev);
        }
    };

    struct OrderGuard {
        std::mutex mu;
        std::shared_ptr<SharedEvent> own;  // recorded after this guard's own calls
        std::shared_ptr<SharedEvent> last; // the event after the latest call (own, or a batch's)
        hipStream_t last_stream = nullptr;
    };

    inline bool stream_capturing(hipStream_t s)
    {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess)
            return false;
        return st != hipStreamCaptureStatusNone;
    }

    // Before the launches of a call on stream s (mutex held): wait for the previous call when
    // it went to another stream.
    inline hipError_t order_begin(const OrderGuard& g, hipStream_t s, bool capturing)
    {
        if (capturing || !g.last || g.last_stream == s)
            return hipSuccess;
        return hipStreamWaitEvent(s, g.last->ev, 0);
    }

    // The guard's event, created up front (an operator does this at create time, so that its
    // calls create nothing).
    inline hipError_t order_init(OrderGuard& g)
    {
        if (g.own)
            return hipSuccess;
        auto ev = std::make_shared<SharedEvent>();
        const hipError_t e = hipEventCreateWithFlags(&ev->ev, hipEventDisableTiming);
        if (e != hipSuccess)
            return e;
        g.own = ev;
        return hipSuccess;
    }

    // After the launches (mutex held): record the guard's event on s.
    inline hipError_t order_end(OrderGuard& g, hipStream_t s, bool capturing)
    {
        if (capturing)
            return hipSuccess;
        hipError_t e = order_init(g);
        if (e != hipSuccess)
            return e;
        e = hipEventRecord(g.own->ev, s);
        if (e != hipSuccess)
            return e;
        g.last = g.own;
        g.last_stream = s;
        return hipSuccess;
    }
} // namespace ehp
