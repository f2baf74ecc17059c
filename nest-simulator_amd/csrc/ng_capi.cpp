// ng_capi.cpp -- extern "C" boundary (include/nestgpu.h) over the host kernel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <chrono>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nestgpu.h"
#include "ng_nest.h"
#include "ng_server.h"

struct nestgpu_kernel
{
  std::unique_ptr< nestgpu::Kernel > k;
  std::string err;
  int device;
};

namespace
{
thread_local std::string g_err;

template < typename F >
int
guarded( nestgpu_kernel* h, F&& f )
{
  try
  {
    if ( h == nullptr )
      throw std::runtime_error( "null kernel handle" );
    f();
    return 0;
  }
  catch ( const std::exception& e )
  {
    if ( h )
      h->err = e.what();
    g_err = e.what();
    return 1;
  }
  catch ( ... )
  {
    if ( h )
      h->err = "unknown error";
    g_err = "unknown error";
    return 1;
  }
}

std::vector< int64_t >
ids_of( const int64_t* ids, int64_t n )
{
  if ( n < 0 or ( n > 0 and ids == nullptr ) )
    throw std::runtime_error( "invalid id array" );
  return std::vector< int64_t >( ids, ids + n );
}
}  // namespace

extern "C" {

int
nestgpu_kernel_new( int device, nestgpu_kernel** out )
{
  try
  {
    if ( out == nullptr )
      throw std::runtime_error( "out is NULL" );
    auto* h = new nestgpu_kernel;
    h->device = device;
    h->k = std::make_unique< nestgpu::Kernel >( device );
    *out = h;
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

void
nestgpu_kernel_free( nestgpu_kernel* k )
{
  delete k;
}

const char*
nestgpu_last_error( const nestgpu_kernel* k )
{
  return k ? k->err.c_str() : g_err.c_str();
}

int
nestgpu_reset_kernel( nestgpu_kernel* h )
{
  return guarded( h, [ & ] { h->k = std::make_unique< nestgpu::Kernel >( h->device ); } );
}

int
nestgpu_set_kernel_status( nestgpu_kernel* h, const char* json )
{
  return guarded( h, [ & ] { h->k->set_kernel_status( nestgpu::Json::parse( json ) ); } );
}

int
nestgpu_get_kernel_status( nestgpu_kernel* h, char* buf, size_t cap, size_t* len )
{
  return guarded( h,
    [ & ]
    {
      const std::string s = h->k->get_kernel_status_json();
      if ( len )
        *len = s.size();
      if ( buf and cap > 0 )
      {
        const size_t n = std::min( cap - 1, s.size() );
        std::memcpy( buf, s.data(), n );
        buf[ n ] = 0;
      }
    } );
}

int
nestgpu_create( nestgpu_kernel* h, const char* model, int64_t n, const char* params_json, int64_t* first_id )
{
  return guarded( h,
    [ & ]
    {
      const int64_t f = h->k->create( model ? model : "", n, nestgpu::Json::parse( params_json ) );
      if ( first_id )
        *first_id = f;
    } );
}

int
nestgpu_set_status( nestgpu_kernel* h, const int64_t* ids, int64_t n, const char* params_json )
{
  return guarded( h, [ & ] { h->k->set_status( ids_of( ids, n ), nestgpu::Json::parse( params_json ) ); } );
}

int
nestgpu_set_values( nestgpu_kernel* h, const int64_t* ids, int64_t n, const char* key, const double* values )
{
  return guarded( h, [ & ] { h->k->set_values( ids_of( ids, n ), key ? key : "", values ); } );
}

int
nestgpu_get_values( nestgpu_kernel* h, const int64_t* ids, int64_t n, const char* key, double* out )
{
  return guarded( h, [ & ] { h->k->get_values( ids_of( ids, n ), key ? key : "", out ); } );
}

int
nestgpu_connect( nestgpu_kernel* h, const int64_t* pre, int64_t n_pre, const int64_t* post, int64_t n_post,
  const char* conn_spec_json, const char* syn_spec_json )
{
  return guarded( h,
    [ & ]
    {
      nestgpu::Json cs = nestgpu::Json::parse( conn_spec_json );
      h->k->connect( ids_of( pre, n_pre ), ids_of( post, n_post ), cs, nestgpu::Json::parse( syn_spec_json ) );
    } );
}

int
nestgpu_connect_arrays( nestgpu_kernel* h, const int64_t* sources, const int64_t* targets, const double* weights,
  const double* delays_ms, int64_t n )
{
  return guarded( h, [ & ] { h->k->connect_arrays( sources, targets, weights, delays_ms, n ); } );
}

int
nestgpu_set_spike_counts( nestgpu_kernel* h, int64_t first_id, int64_t n, int64_t first_step, int64_t n_steps,
  const uint8_t* counts )
{
  return guarded( h, [ & ] { h->k->set_spike_counts( first_id, n, first_step, n_steps, counts ); } );
}

int64_t
nestgpu_num_connections( nestgpu_kernel* h )
{
  return h ? h->k->num_connections() : -1;
}

int
nestgpu_get_connections( nestgpu_kernel* h, int64_t* sources, int64_t* targets, double* weights, int64_t* delay_steps,
  int32_t* vps )
{
  return guarded( h, [ & ] { h->k->get_connections( sources, targets, weights, delay_steps, vps ); } );
}

extern "C" int
nestgpu_get_connection_models( nestgpu_kernel* h, int32_t* models )
{
  return guarded( h, [ & ] { h->k->get_connection_models( models ); } );
}

int
nestgpu_prepare( nestgpu_kernel* h )
{
  return guarded( h, [ & ] { h->k->prepare(); } );
}

int
nestgpu_run( nestgpu_kernel* h, double t_ms )
{
  return guarded( h, [ & ] { h->k->run( t_ms ); } );
}

int
nestgpu_cleanup( nestgpu_kernel* h )
{
  return guarded( h, [ & ] { h->k->cleanup(); } );
}

int
nestgpu_simulate( nestgpu_kernel* h, double t_ms )
{
  return guarded( h, [ & ] { h->k->simulate( t_ms ); } );
}

int
nestgpu_recorder_events( nestgpu_kernel* h, int64_t recorder_id, int64_t* senders, double* times, int64_t cap,
  int64_t* n )
{
  return guarded( h,
    [ & ]
    {
      std::vector< int64_t > s, st;
      h->k->recorder_events( recorder_id, s, st );
      if ( n )
        *n = static_cast< int64_t >( s.size() );
      if ( senders )
      {
        const double ms = h->k->timebase().ms_per_tic * h->k->timebase().tics_per_step;
        const int64_t m = std::min< int64_t >( cap, static_cast< int64_t >( s.size() ) );
        for ( int64_t i = 0; i < m; ++i )
        {
          senders[ i ] = s[ i ];
          if ( times )
            times[ i ] = ms * static_cast< double >( st[ i ] );  // Time::step(stamp).get_ms()
        }
      }
    } );
}

int
nestgpu_multimeter_events( nestgpu_kernel* h, int64_t multimeter_id, int64_t* senders, double* times, double* values,
  int64_t cap, int64_t* n )
{
  return guarded( h,
    [ & ]
    {
      std::vector< int64_t > s, st;
      std::vector< double > v;
      h->k->multimeter_events( multimeter_id, s, st, v );
      const size_t nf = h->k->multimeter_record_from( multimeter_id ).size();
      if ( n )
        *n = static_cast< int64_t >( s.size() );
      if ( senders )
      {
        const double ms = h->k->timebase().ms_per_tic * h->k->timebase().tics_per_step;
        const int64_t m = std::min< int64_t >( cap, static_cast< int64_t >( s.size() ) );
        for ( int64_t i = 0; i < m; ++i )
        {
          senders[ i ] = s[ i ];
          if ( times )
            times[ i ] = ms * static_cast< double >( st[ i ] );
          if ( values )
            for ( size_t j = 0; j < nf; ++j )
              values[ i * nf + j ] = v[ i * nf + j ];
        }
      }
    } );
}

int
nestgpu_node_status( nestgpu_kernel* h, int64_t id, char* buf, size_t cap, size_t* len )
{
  return guarded( h,
    [ & ]
    {
      const std::string js = h->k->node_status_json( id );
      if ( len )
        *len = js.size();
      if ( buf and cap > 0 )
      {
        const size_t m = std::min( cap - 1, js.size() );
        std::memcpy( buf, js.data(), m );
        buf[ m ] = 0;
      }
    } );
}

int
nestgpu_batch_simulate( nestgpu_kernel** ks, int n, const double* t_ms )
{
  try
  {
    if ( n <= 0 )
      return 0;
    std::vector< nestgpu::Engine* > engines;
    std::vector< int64_t > steps;
    const int dev = ks[ 0 ]->device;
    for ( int i = 0; i < n; ++i )
    {
      if ( ks[ i ]->device != dev )
        throw std::runtime_error( "batched kernels must share one device" );
      ks[ i ]->k->prepare_for_batch();
      engines.push_back( &ks[ i ]->k->engine() );
      steps.push_back( ks[ i ]->k->steps_for( t_ms[ i ] ) );
    }
    static thread_local std::unique_ptr< nestgpu::Runner > runner;
    if ( not runner )
      runner = std::make_unique< nestgpu::Runner >( dev, 64 );
    runner->run( engines, steps );
    for ( int i = 0; i < n; ++i )
      ks[ i ]->k->after_batch_run( steps[ i ] );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_batch_simulate_noalloc( nestgpu_kernel** ks, int n, const double* t_ms )
{
  try
  {
    if ( n <= 0 )
      return 0;
    if ( ks == nullptr or t_ms == nullptr )
      throw std::runtime_error( "null argument" );
    if ( n > nestgpu::NOALLOC_MAX_NETS )
      throw std::runtime_error( "batch_simulate_noalloc: " + std::to_string( n ) + " kernels, at most "
        + std::to_string( nestgpu::NOALLOC_MAX_NETS ) + " per call" );
    std::vector< nestgpu::Engine* > engines;
    std::vector< int64_t > steps;
    const int dev = ks[ 0 ]->device;
    for ( int i = 0; i < n; ++i )
    {
      if ( ks[ i ] == nullptr )
        throw std::runtime_error( "null kernel handle" );
      if ( ks[ i ]->device != dev )
        throw std::runtime_error( "batched kernels must share one device" );
      if ( not ks[ i ]->k->device_copy_current() )
        throw std::runtime_error( "KernelException: batch_simulate_noalloc needs every kernel prepared, with no "
                                  "change since its Prepare (Prepare allocates the network's device memory)" );
      steps.push_back( ks[ i ]->k->checked_steps( t_ms[ i ] ) );
    }
    for ( int i = 0; i < n; ++i )
    {
      ks[ i ]->k->prepare_for_batch();  // the device copy is current: only the Prepare flags change
      engines.push_back( &ks[ i ]->k->engine() );
    }
    nestgpu::run_noalloc( dev, engines, steps, 64 );
    for ( int i = 0; i < n; ++i )
      ks[ i ]->k->after_batch_run( steps[ i ] );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_batch_autotune( nestgpu_kernel** ks, int n, double probe_ms, const int* candidates, int n_cand,
  const int* window_caps, int n_win, int* chosen, int* chosen_win )
{
  try
  {
    if ( n <= 0 )
      return 0;
    if ( n_cand <= 0 or candidates == nullptr )
      throw std::runtime_error( "batch_autotune needs at least one candidate targets_per_block" );
    const int dev = ks[ 0 ]->device;
    // classes: one per set of neuron models and octave of the neuron count
    std::map< std::pair< int, int >, std::vector< int > > classes;
    for ( int i = 0; i < n; ++i )
    {
      if ( ks[ i ]->device != dev )
        throw std::runtime_error( "batched kernels must share one device" );
      const uint64_t nn = std::max< uint64_t >( 1, ks[ i ]->k->num_neurons() );
      int c = 0;
      while ( ( uint64_t( 2 ) << c ) <= nn )
        ++c;
      classes[ { ks[ i ]->k->model_mask(), c } ].push_back( i );
    }
    // A probe of a few instances may not fill the device, so its wall time measures one
    // block's latency rather than throughput. The score is the probe's block-time,
    // ms * min(blocks, slots) with two workgroups per compute unit: what the class costs
    // inside a batch that fills the device.
    int n_cu = 0;
    if ( hipDeviceGetAttribute( &n_cu, hipDeviceAttributeMultiprocessorCount, dev ) != hipSuccess or n_cu <= 0 )
      n_cu = 256;
    const double slots = 2.0 * n_cu;
    nestgpu::Runner runner( dev, 64 );
    for ( auto& kv : classes )
    {
      const std::vector< int >& members = kv.second;
      std::vector< int > sample;
      const size_t ns = std::min< size_t >( 4, members.size() );
      for ( size_t q = 0; q < ns; ++q )
        sample.push_back( members[ q * members.size() / ns ] );
      int best = -1, best_win = 0;
      double best_ms = 0;
      const int nw = window_caps and n_win > 0 ? n_win : 1;
      for ( int c = 0; c < n_cand * nw; ++c )
      {
        const int tb = candidates[ c / nw ];
        const int wcap = window_caps and n_win > 0 ? window_caps[ c % nw ] : 0;
        if ( wcap < 0 )
          throw std::runtime_error( "batch_autotune: window slot caps must be >= 0 (0: as many as fit)" );
        bool ok = tb >= 1 and tb <= static_cast< int >( nestgpu::MAX_TB );
        for ( int i : sample )
          ok = ok and ( ks[ i ]->k->num_neurons() + tb - 1 ) / tb <= nestgpu::PMAX_GLOB;
        if ( not ok )
          continue;
        std::vector< nestgpu::Engine* > engines;
        std::vector< int64_t > steps;
        std::vector< bool > keep;
        for ( int i : sample )
        {
          keep.push_back( ks[ i ]->k->keep_host_connections() );
          ks[ i ]->k->set_keep_host_connections( true );  // probes rebuild the device copy
          ks[ i ]->k->reset_device();
          ks[ i ]->k->set_layout( tb, ks[ i ]->k->delivery_mode(), wcap );
          ks[ i ]->k->prepare_for_batch();
          engines.push_back( &ks[ i ]->k->engine() );
          steps.push_back( ks[ i ]->k->steps_for( probe_ms ) );
        }
        runner.run( engines, steps );  // warm-up: graph capture, caches
        const auto t0 = std::chrono::steady_clock::now();
        runner.run( engines, steps );
        double blocks = 0;
        for ( int i : sample )
          blocks += static_cast< double >( ( ks[ i ]->k->num_neurons() + tb - 1 ) / tb );
        const double ms = std::chrono::duration< double, std::milli >( std::chrono::steady_clock::now() - t0 ).count()
          * std::min( blocks, slots );
        if ( best < 0 or ms < best_ms )
        {
          best = tb;
          best_win = wcap;
          best_ms = ms;
        }
        for ( size_t q = 0; q < sample.size(); ++q )
        {
          ks[ sample[ q ] ]->k->reset_device();
          ks[ sample[ q ] ]->k->set_keep_host_connections( keep[ q ] );
        }
      }
      if ( best < 0 )
        throw std::runtime_error( "batch_autotune: no candidate fits the instances of one size class" );
      for ( int i : members )
      {
        ks[ i ]->k->reset_device();
        ks[ i ]->k->set_layout( best, ks[ i ]->k->delivery_mode(), best_win );
        if ( chosen )
          chosen[ i ] = best;
        if ( chosen_win )
          chosen_win[ i ] = best_win;
      }
    }
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

struct nestgpu_server
{
  std::unique_ptr< nestgpu::Server > s;
};

int
nestgpu_server_new( int device, uint64_t hbm_budget_bytes, int slices_per_chunk, nestgpu_server** out )
{
  try
  {
    if ( out == nullptr )
      throw std::runtime_error( "out is NULL" );
    auto* h = new nestgpu_server;
    h->s = std::make_unique< nestgpu::Server >( device, hbm_budget_bytes, slices_per_chunk );
    *out = h;
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

void
nestgpu_server_free( nestgpu_server* s )
{
  delete s;
}

int
nestgpu_server_submit( nestgpu_server* s, nestgpu_kernel* k, const double* t_ms, int n_calls, int64_t* id )
{
  try
  {
    if ( s == nullptr or k == nullptr or id == nullptr or ( n_calls > 0 and t_ms == nullptr ) )
      throw std::runtime_error( "null argument" );
    std::vector< int64_t > steps;
    for ( int i = 0; i < n_calls; ++i )
    {
      if ( t_ms[ i ] < 0 )
        throw std::runtime_error( "BadParameter: The simulation time cannot be negative." );
      steps.push_back( k->k->steps_for( t_ms[ i ] ) );
    }
    *id = s->s->submit( k->k.get(), steps );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_submit_open( nestgpu_server* s, nestgpu_kernel* k, const double* t_ms, int n_calls, int64_t* id )
{
  try
  {
    if ( s == nullptr or k == nullptr or id == nullptr or ( n_calls > 0 and t_ms == nullptr ) )
      throw std::runtime_error( "null argument" );
    std::vector< int64_t > steps;
    for ( int i = 0; i < n_calls; ++i )
      steps.push_back( k->k->checked_steps( t_ms[ i ] ) );
    *id = s->s->submit_open( k->k.get(), steps );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_extend( nestgpu_server* s, int64_t id, double t_ms )
{
  try
  {
    if ( s == nullptr )
      throw std::runtime_error( "null server" );
    if ( t_ms < 0 )
      throw std::runtime_error( "BadParameter: The simulation time cannot be negative." );
    // the kernel's time base converts the duration; the request's kernel is not touched here
    const nestgpu::TimeBase& tb = s->s->timebase_of( id );
    if ( tb.ms_to_tics( t_ms ) % tb.tics_per_step != 0 )
      throw std::runtime_error( "BadParameter: The simulation time must be a multiple of the simulation resolution." );
    s->s->extend( id, tb.ms_to_steps( t_ms ) );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_close( nestgpu_server* s, int64_t id )
{
  try
  {
    if ( s == nullptr )
      throw std::runtime_error( "null server" );
    s->s->close( id );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_wait_calls( nestgpu_server* s, int64_t id, int64_t n, double timeout_s, int* state, int64_t* calls_done )
{
  try
  {
    if ( s == nullptr )
      throw std::runtime_error( "null server" );
    const nestgpu::RequestInfo r = s->s->wait_calls( id, n, timeout_s );
    if ( state )
      *state = r.state;
    if ( calls_done )
      *calls_done = r.calls_done;
    if ( r.state == nestgpu::REQ_FAILED )
      throw std::runtime_error( r.error );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_wait( nestgpu_server* s, int64_t id, double timeout_s, int* state, double* times, uint64_t* events )
{
  try
  {
    if ( s == nullptr )
      throw std::runtime_error( "null server" );
    const nestgpu::RequestInfo r = s->s->wait( id, timeout_s );
    if ( state )
      *state = r.state;
    if ( times )
    {
      times[ 0 ] = r.t_submit;
      times[ 1 ] = r.t_admit;
      times[ 2 ] = r.t_start;
      times[ 3 ] = r.t_done;
    }
    if ( events )
      *events = r.events;
    if ( r.state == nestgpu::REQ_FAILED )
      g_err = r.error;
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_server_stats( nestgpu_server* s, char* buf, size_t cap, size_t* len )
{
  try
  {
    if ( s == nullptr )
      throw std::runtime_error( "null server" );
    const std::string j = s->s->stats_json();
    if ( len )
      *len = j.size();
    if ( buf and cap > 0 )
    {
      const size_t n = std::min( cap - 1, j.size() );
      std::memcpy( buf, j.data(), n );
      buf[ n ] = 0;
    }
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_poisson_counts( uint64_t seed, double lambda, uint64_t stream0, int64_t n, int64_t step0, int64_t n_steps,
  uint8_t* out )
{
  try
  {
    nestgpu::poisson_counts( seed, lambda, stream0, n, step0, n_steps, out );
    return 0;
  }
  catch ( const std::exception& e )
  {
    g_err = e.what();
    return 1;
  }
}

int
nestgpu_device_count( int* n )
{
  int c = 0;
  if ( hipGetDeviceCount( &c ) != hipSuccess )
    c = 0;
  if ( n )
    *n = c;
  return 0;
}

}  // extern "C"
