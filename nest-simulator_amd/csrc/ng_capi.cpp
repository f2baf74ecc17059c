// ng_capi.cpp -- extern "C" boundary (include/nestgpu.h) over the host kernel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nestgpu.h"
#include "ng_nest.h"

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
