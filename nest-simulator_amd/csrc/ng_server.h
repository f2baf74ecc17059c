// ng_server.h -- served form: a request loop with HBM-budget admission and continuous batching.
//
// Modelled on nest-server (pynest/nest/server/hl_api_server.py), which runs one simulation
// kernel per request and serialises calls. Here every admitted request's network stays
// resident in HBM across its Simulate calls, and all running requests advance together in
// the same slice launches (Runner::run_chunk). New requests join at chunk boundaries as
// soon as the HBM budget has room; finished ones leave and release their device memory.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ng_engine.h"
#include "ng_nest.h"

namespace nestgpu
{

enum RequestState
{
  REQ_QUEUED = 0,
  REQ_ADMITTED = 1,
  REQ_RUNNING = 2,
  REQ_DONE = 3,
  REQ_FAILED = -1
};

struct RequestInfo
{
  int state = REQ_QUEUED;
  int64_t calls_done = 0;  // Simulate calls completed
  double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;  // seconds since the server started
  uint64_t bytes = 0;                                         // device bytes reserved
  uint64_t events = 0;                                        // recurrent synaptic events delivered
  std::string error;
};

class Server
{
public:
  static constexpr int PREPARERS = 4;  // admitted requests whose device copies build at once

  Server( int device, uint64_t hbm_budget_bytes, int slices_per_chunk );
  ~Server();
  Server( const Server& ) = delete;
  Server& operator=( const Server& ) = delete;

  // k: a network built on the host (Create/Connect), not yet simulated; the caller must not
  // touch it until the request is done. call_steps: the Simulate calls, in steps.
  int64_t submit( Kernel* k, const std::vector< int64_t >& call_steps );
  // An open request (a session): its network stays resident after its calls are done, until
  // close(); extend() appends a Simulate call, which joins the batch at the next chunk
  // boundary. Between calls the caller may read the kernel's state (no device work runs on
  // it then).
  int64_t submit_open( Kernel* k, const std::vector< int64_t >& call_steps );
  void extend( int64_t id, int64_t steps );
  void close( int64_t id );
  // Waits up to timeout_s (negative: forever) until the request has completed n calls, or
  // is done or failed; returns its info.
  RequestInfo wait_calls( int64_t id, int64_t n, double timeout_s );
  // time base of a request's kernel (fixed once the kernel is submitted)
  const TimeBase& timebase_of( int64_t id );
  // Waits up to timeout_s (negative: forever) for the request to finish; returns its info.
  RequestInfo wait( int64_t id, double timeout_s );
  std::string stats_json();

private:
  struct Req
  {
    int64_t id;
    Kernel* k;
    std::vector< int64_t > calls;
    size_t call = 0;
    int64_t remaining = 0;
    RequestInfo info;
    bool open = false;    // stays resident between calls until closed
    bool parked = false;  // open, on the device, all its calls done
    bool closing = false;
  };
  double now() const;
  void admit_loop();
  void run_loop();
  void finish( Req* r, const std::string& error );

  int device_;
  uint64_t budget_, used_ = 0;
  int S_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_;
  bool stop_ = false;
  int64_t next_id_ = 1;
  std::map< int64_t, std::unique_ptr< Req > > reqs_;
  std::deque< Req* > queue_;  // waiting for admission, FIFO
  std::vector< Req* > ready_;  // on the device, joining at the next chunk boundary
  int64_t submit_req( Kernel* k, const std::vector< int64_t >& call_steps, bool open );
  // statistics
  uint64_t chunks_ = 0, events_done_ = 0;
  double busy_s_ = 0;
  size_t max_concurrent_ = 0;
  std::vector< std::thread > admit_threads_;  // preparers: admitted requests build concurrently
  std::thread run_thread_;
};

}  // namespace nestgpu
