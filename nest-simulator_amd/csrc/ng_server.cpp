// ng_server.cpp -- served form (see ng_server.h).
#include "ng_server.h"

#include <algorithm>
#include <chrono>
#include <sstream>
#include <stdexcept>

namespace nestgpu
{

Server::Server( int device, uint64_t hbm_budget_bytes, int slices_per_chunk )
  : device_( device )
  , budget_( hbm_budget_bytes )
  , S_( std::max( 1, slices_per_chunk ) )
  , t0_( std::chrono::steady_clock::now() )
{
  // admission stays FIFO against the budget; the admitted requests' device copies are built
  // by several preparers at once, so one large network does not hold up the ones behind it
  for ( int i = 0; i < PREPARERS; ++i )
    admit_threads_.emplace_back( [ this ] { admit_loop(); } );
  run_thread_ = std::thread( [ this ] { run_loop(); } );
}

Server::~Server()
{
  {
    std::lock_guard< std::mutex > g( mu_ );
    stop_ = true;
  }
  cv_.notify_all();
  for ( auto& th : admit_threads_ )
    th.join();
  run_thread_.join();
}

double
Server::now() const
{
  return std::chrono::duration< double >( std::chrono::steady_clock::now() - t0_ ).count();
}

int64_t
Server::submit( Kernel* k, const std::vector< int64_t >& call_steps )
{
  return submit_req( k, call_steps, false );
}

int64_t
Server::submit_open( Kernel* k, const std::vector< int64_t >& call_steps )
{
  return submit_req( k, call_steps, true );
}

void
Server::extend( int64_t id, int64_t steps )
{
  if ( steps < 0 )
    throw std::runtime_error( "BadParameter: The simulation time cannot be negative." );
  std::lock_guard< std::mutex > g( mu_ );
  auto it = reqs_.find( id );
  if ( it == reqs_.end() )
    throw std::runtime_error( "unknown request id" );
  Req* r = it->second.get();
  if ( not r->open or r->closing or r->info.state == REQ_DONE or r->info.state == REQ_FAILED )
    throw std::runtime_error( "the request is not open" );
  r->calls.push_back( steps );
  if ( r->parked )
  {
    // rejoin the batch at the next chunk boundary with the new call (the first call of a
    // request that was prepared without one, or the one after its last done call)
    r->parked = false;
    r->call = r->calls.size() == 1 ? 0 : r->call + 1;
    r->remaining = r->calls[ r->call ];
    ready_.push_back( r );
  }
  cv_.notify_all();
}

void
Server::close( int64_t id )
{
  Req* r = nullptr;
  {
    std::lock_guard< std::mutex > g( mu_ );
    auto it = reqs_.find( id );
    if ( it == reqs_.end() )
      throw std::runtime_error( "unknown request id" );
    r = it->second.get();
    if ( not r->open or r->closing )
      return;
    r->closing = true;
    if ( not r->parked )
    {
      cv_.notify_all();
      return;  // the run loop finishes it after its remaining calls
    }
    r->parked = false;
  }
  finish( r, "" );
}

const TimeBase&
Server::timebase_of( int64_t id )
{
  std::lock_guard< std::mutex > g( mu_ );
  auto it = reqs_.find( id );
  if ( it == reqs_.end() )
    throw std::runtime_error( "unknown request id" );
  return it->second->k->timebase();
}

RequestInfo
Server::wait_calls( int64_t id, int64_t n, double timeout_s )
{
  std::unique_lock< std::mutex > g( mu_ );
  auto it = reqs_.find( id );
  if ( it == reqs_.end() )
    throw std::runtime_error( "unknown request id" );
  Req* r = it->second.get();
  auto done = [ r, n ] { return r->info.calls_done >= n or r->info.state == REQ_DONE or r->info.state == REQ_FAILED; };
  if ( timeout_s < 0 )
    cv_.wait( g, done );
  else
    cv_.wait_for( g, std::chrono::duration< double >( timeout_s ), done );
  return r->info;
}

int64_t
Server::submit_req( Kernel* k, const std::vector< int64_t >& call_steps, bool open )
{
  if ( k == nullptr )
    throw std::runtime_error( "null kernel" );
  if ( k->prepared() or k->simulated() )
    throw std::runtime_error( "a served network must not have been prepared or simulated" );
  if ( k->device() != device_ )
    throw std::runtime_error( "the network is bound to another device than the server" );
  auto r = std::make_unique< Req >();
  r->k = k;
  r->calls = call_steps;
  r->open = open;
  r->info.t_submit = now();
  r->info.bytes = k->estimate_device_bytes();
  if ( r->info.bytes > budget_ )
    throw std::runtime_error( "request needs " + std::to_string( r->info.bytes ) + " device bytes, more than the budget of "
      + std::to_string( budget_ ) );
  std::lock_guard< std::mutex > g( mu_ );
  r->id = next_id_++;
  Req* p = r.get();
  reqs_[ p->id ] = std::move( r );
  queue_.push_back( p );
  cv_.notify_all();
  return p->id;
}

RequestInfo
Server::wait( int64_t id, double timeout_s )
{
  std::unique_lock< std::mutex > g( mu_ );
  auto it = reqs_.find( id );
  if ( it == reqs_.end() )
    throw std::runtime_error( "unknown request id" );
  Req* r = it->second.get();
  auto done = [ r ] { return r->info.state == REQ_DONE or r->info.state == REQ_FAILED; };
  if ( timeout_s < 0 )
    cv_.wait( g, done );
  else
    cv_.wait_for( g, std::chrono::duration< double >( timeout_s ), done );
  return r->info;
}

void
Server::admit_loop()
{
  // FIFO admission: the head request is prepared on the device (host store build + upload)
  // once the budget has room for it; this overlaps with the chunks the run loop executes
  for ( ;; )
  {
    Req* r = nullptr;
    {
      std::unique_lock< std::mutex > g( mu_ );
      cv_.wait( g, [ this ] { return stop_ or ( not queue_.empty() and used_ + queue_.front()->info.bytes <= budget_ ); } );
      if ( stop_ )
        return;
      r = queue_.front();
      queue_.pop_front();
      used_ += r->info.bytes;
      r->info.state = REQ_ADMITTED;
      r->info.t_admit = now();
    }
    std::string err;
    try
    {
      r->k->prepare_for_batch();
    }
    catch ( const std::exception& e )
    {
      err = e.what();
    }
    std::lock_guard< std::mutex > g( mu_ );
    if ( not err.empty() )
    {
      used_ -= r->info.bytes;
      r->info.state = REQ_FAILED;
      r->info.error = err;
      r->info.t_done = now();
    }
    else
    {
      r->call = 0;
      if ( r->calls.empty() and r->open )
      {
        r->info.state = REQ_RUNNING;
        r->info.t_start = now();
        r->parked = true;  // prepared, waiting for its first call
      }
      else
      {
        r->remaining = r->calls.empty() ? 0 : r->calls[ 0 ];
        ready_.push_back( r );
      }
    }
    cv_.notify_all();
  }
}

void
Server::finish( Req* r, const std::string& error )
{
  // called with mu_ not held; device work of r is complete
  uint64_t ev = 0;
  std::string err = error;
  try
  {
    r->k->after_batch_run( 0 );
    ev = r->k->synaptic_events();
    r->k->release_device();
  }
  catch ( const std::exception& e )
  {
    if ( err.empty() )
      err = e.what();
  }
  std::lock_guard< std::mutex > g( mu_ );
  used_ -= r->info.bytes;
  r->info.events = ev;
  r->info.t_done = now();
  r->info.state = err.empty() ? REQ_DONE : REQ_FAILED;
  r->info.error = err;
  events_done_ += ev;
  cv_.notify_all();
}

void
Server::run_loop()
{
  Runner runner( device_, S_ );
  // slices are launched directly, not replayed from a captured graph: the preparers allocate
  // and upload on the device while this loop runs, and HIP fails a stream capture when other
  // threads make such calls during it (the batch changes at most chunk boundaries, so a
  // graph would be captured often anyway)
  runner.use_graph = false;
  std::vector< Req* > active;
  for ( ;; )
  {
    {
      std::unique_lock< std::mutex > g( mu_ );
      cv_.wait( g, [ & ] { return stop_ or not ready_.empty() or not active.empty(); } );
      if ( stop_ )
        return;
      // continuous batching: prepared requests join at this chunk boundary
      const double t = now();
      for ( Req* r : ready_ )
      {
        r->info.state = REQ_RUNNING;
        r->info.t_start = t;
        active.push_back( r );
      }
      ready_.clear();
      max_concurrent_ = std::max( max_concurrent_, active.size() );
    }
    std::vector< Engine* > engines;
    std::vector< int64_t > remaining;
    for ( Req* r : active )
    {
      engines.push_back( &r->k->engine() );
      remaining.push_back( r->remaining );
    }
    std::string err;
    const double t_chunk = now();
    try
    {
      runner.run_chunk( engines, remaining );
    }
    catch ( const std::exception& e )
    {
      err = e.what();
    }
    {
      std::lock_guard< std::mutex > g( mu_ );
      busy_s_ += now() - t_chunk;
      ++chunks_;
    }
    std::vector< Req* > still;
    for ( size_t i = 0; i < active.size(); ++i )
    {
      Req* r = active[ i ];
      if ( not err.empty() )
      {
        finish( r, err );
        continue;
      }
      r->remaining = remaining[ i ];
      // a Simulate call ends: the next one of the same request continues from its state
      bool park = false;
      {
        std::lock_guard< std::mutex > g( mu_ );
        while ( r->remaining == 0 and r->call + 1 < r->calls.size() )
        {
          r->k->after_batch_run( 0 );
          ++r->info.calls_done;
          ++r->call;
          r->remaining = r->calls[ r->call ];
        }
        if ( r->remaining == 0 and r->open and not r->closing )
        {
          // an open request keeps its network resident and waits for its next call
          r->k->after_batch_run( 0 );
          ++r->info.calls_done;
          r->parked = true;
          park = true;
          cv_.notify_all();
        }
      }
      if ( park )
        continue;
      if ( r->remaining == 0 )
      {
        {
          std::lock_guard< std::mutex > g( mu_ );
          ++r->info.calls_done;
        }
        finish( r, "" );
      }
      else
        still.push_back( r );
    }
    active.swap( still );
  }
}

std::string
Server::stats_json()
{
  std::lock_guard< std::mutex > g( mu_ );
  size_t queued = queue_.size(), done = 0, failed = 0;
  for ( const auto& kv : reqs_ )
  {
    done += kv.second->info.state == REQ_DONE;
    failed += kv.second->info.state == REQ_FAILED;
  }
  std::ostringstream o;
  o.precision( 17 );
  o << "{\"requests\": " << reqs_.size() << ", \"queued\": " << queued << ", \"done\": " << done
    << ", \"failed\": " << failed << ", \"budget_bytes\": " << budget_ << ", \"used_bytes\": " << used_
    << ", \"chunks\": " << chunks_ << ", \"busy_s\": " << busy_s_ << ", \"events\": " << events_done_
    << ", \"max_concurrent\": " << max_concurrent_ << ", \"uptime_s\": " << now() << "}";
  return o.str();
}

}  // namespace nestgpu
