// Request server for the primitive core (extension: the reference has no
// counterpart).
//
// Requests (Sort_By_Key, Sort, Inclusive_Scan, Reduce) name device-resident
// inputs; the server owns their outputs.  A dispatcher thread admits queued
// requests in arrival order while their HBM footprint (outputs plus
// temporaries) fits the budget, coalesces everything admitted into one
// segmented launch per operation group on one of `slots` streams, and keeps
// launching new batches while earlier ones are still running (continuous
// batching: work admitted while the GPU is busy overlaps with it).  Each
// request's result is exactly what the single-vector call would produce:
// sorts and integer scans bit-exact, reductions in an order fixed by the
// request's length alone, so a result is the same alone and in any mix.
//
// With batching off, the same dispatcher serves one request at a time through
// the single-call functions and waits for each (the per-call baseline).
//
// Callers: services that receive many independent sort / scan / reduce
// requests of mixed sizes.
#ifndef _TRTC_server_h
#define _TRTC_server_h

#include <string>

#include "TRTC_api.h"
#include "TRTCContext.h"

enum TRTC_Request_Op
{
    TRTC_REQ_SORT_BY_KEY = 0,   // in_a keys, in_b values -> outputs 0 keys, 1 values (stable, Less)
    TRTC_REQ_SORT = 1,          // in_a keys -> output 0 keys (Less)
    TRTC_REQ_INCLUSIVE_SCAN = 2, // in_a -> output 0 (Plus)
    TRTC_REQ_REDUCE = 3          // in_a -> output 0: one element (Plus, no init)
};

class TRTC_Server_Impl;
class THRUST_RTC_API TRTC_Server
{
public:
    // hbm_budget: bytes of outputs and temporaries the server may hold at
    // once, reserved in device memory when the server starts (its buffer
    // pool; grown on demand if that reservation fails); slots: batches in
    // flight at once (streams); batching false: one request at a time through
    // the single-call path.
    TRTC_Server(size_t hbm_budget, int slots, bool batching);
    ~TRTC_Server();
    TRTC_Server(const TRTC_Server&) = delete;
    TRTC_Server& operator=(const TRTC_Server&) = delete;

    // Batch size limits of the batching server: a batch closes at
    // max_requests requests or before its elements pass max_elements (it
    // always takes at least one request).  Defaults 4096 and 2^28.
    void Set_Batch_Limits(size_t max_requests, size_t max_elements);

    // Queue a request; returns a ticket (>= 0) or -1 for an invalid request.
    // key_elem / val_elem are element type names ("uint32_t", "float", ...;
    // val_elem only for Sort_By_Key).  Inputs must stay unchanged until the
    // request is done.
    long long Submit(int op, const char* key_elem, const char* val_elem, const void* in_a, const void* in_b,
                     size_t n);
    // Blocks until the request is done; false if it failed.
    bool Wait(long long ticket);
    bool Done(long long ticket);
    // Device address of output `which` (see TRTC_Request_Op); valid until Release.
    void* Output(long long ticket, int which);
    // Microseconds from Submit to the moment the dispatcher saw the request's
    // batch complete.
    double Latency_Us(long long ticket);
    // Returns the request's outputs to the server's pool.
    void Release(long long ticket);
    // JSON: requests served, batches, mean batch size, peak HBM held.
    std::string Stats();

private:
    TRTC_Server_Impl* m_impl;
};

// Load replay: request i is submitted arrival_us[i] microseconds after the
// start (op, element types, inputs and length per request); a consumer thread
// waits for the requests in submission order and releases each one's outputs
// once it is done.  latency_us[i] receives each request's latency (-1 if it
// failed).  Returns the seconds from the start to the last completion.
double THRUST_RTC_API TRTC_Server_Replay(TRTC_Server& server, size_t count, const double* arrival_us, const int* ops,
                                         const char* const* key_elem, const char* const* val_elem,
                                         const void* const* in_a, const void* const* in_b, const size_t* n,
                                         double* latency_us);

#endif
