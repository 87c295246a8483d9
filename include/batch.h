// Batched segmented primitives (extension: the reference has no counterpart).
//
// A segment is a range [offsets[s], offsets[s+1]) of a vector (offsets: host
// array of num_segments + 1 entries).  Every segment ends up exactly as a
// separate call of the single-vector function on that range would leave it
// (sorts stable and bit-exact, integer scans bit-exact, reductions in a fixed
// order that depends only on the segment's length).
//
// The one-shot functions below (TRTC_Segmented_*) run all segments with one
// launch per kernel stage.  TRTC_Batch records several such operations,
// tunes the tile schedule of its sorts, allocates everything once, captures
// all of it into one HIP graph and replays it with one hipGraphLaunch per
// Launch().  Callers: workloads with many independent sorts / scans of
// uneven sizes (per-row sorts, per-request work in a server).
#ifndef _TRTC_batch_h
#define _TRTC_batch_h

#include <string>

#include "TRTC_api.h"
#include "TRTCContext.h"
#include "DVVector.h"
#include "functor.h"

bool THRUST_RTC_API TRTC_Segmented_Sort(DVVectorLike& keys, const size_t* offsets, size_t num_segments,
                                        const Functor& comp);
// (TRTC_Segmented_Sort_By_Key is declared in sort.h, TRTC_Segmented_Inclusive_Scan in scan.h.)
bool THRUST_RTC_API TRTC_Segmented_Exclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out,
                                                  const size_t* offsets, size_t num_segments);
// results[s] = reduction of segment s with op (no init); every segment must be non-empty.
bool THRUST_RTC_API TRTC_Segmented_Reduce(const DVVectorLike& vec_in, const size_t* offsets, size_t num_segments,
                                          DVVectorLike& results, const Functor& op);

// Autotuning of the batched sort's tile schedule (on by default): the first
// batch of a new shape times a few schedules on a copy of its data.
void THRUST_RTC_API TRTC_Set_Batch_Autotune(bool on);
// Use one fixed tile schedule for every batched sort, e.g.
// "<=2048:256x8 rest:512x16" (segments of up to 2048 keys in 256-thread x 8
// tiles, the rest in 512 x 16); NULL or "" returns to automatic.  False if
// the text is not a valid schedule.
bool THRUST_RTC_API TRTC_Set_Batch_Schedule(const char* schedule);
// JSON list of the tuning runs so far (signature, candidate times, choice).
std::string THRUST_RTC_API TRTC_Batch_Tuning_Report();

class TRTC_Batch_Impl;
class THRUST_RTC_API TRTC_Batch
{
public:
    TRTC_Batch();
    ~TRTC_Batch();
    TRTC_Batch(const TRTC_Batch&) = delete;
    TRTC_Batch& operator=(const TRTC_Batch&) = delete;

    // Record operations (vectors must stay alive and in place until the
    // batch is destroyed; their contents may change between launches).
    bool Add_Sort(DVVectorLike& keys, const size_t* offsets, size_t num_segments, const Functor& comp);
    bool Add_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const size_t* offsets, size_t num_segments,
                         const Functor& comp);
    bool Add_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const size_t* offsets,
                            size_t num_segments, const Functor& op);
    bool Add_Reduce(const DVVectorLike& vec_in, const size_t* offsets, size_t num_segments, DVVectorLike& results,
                    const Functor& op);
    // Tune, allocate, upload the descriptor tables and capture the graph.
    bool Instantiate();
    // Replay the whole batch: one hipGraphLaunch on the calling thread's stream.
    bool Launch();
    // Schedules chosen for the recorded sorts.
    std::string Describe() const;

private:
    TRTC_Batch_Impl* m_impl;
};

#endif
