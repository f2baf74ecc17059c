// ng_engine.cpp -- HBM placement of a network instance and the chunked slice loop.
#include "ng_engine.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace nestgpu
{

void
hip_check( hipError_t e, const char* what )
{
  if ( e != hipSuccess )
    throw std::runtime_error( std::string( "HIP error in " ) + what + ": " + hipGetErrorString( e ) );
}

Engine::Engine( int device )
  : device_( device )
{
}

Engine::~Engine()
{
  if ( built_ )
  {
    hipSetDevice( device_ );
    for ( void* p : store_allocs_ )
      hipFree( p );
    for ( void* p : state_allocs_ )
      hipFree( p );
  }
}

void*
Engine::alloc( size_t bytes )
{
  void* p = nullptr;
  if ( bytes == 0 )
    bytes = 8;
  bytes = ( bytes + 255 ) & ~size_t( 255 );
  hip_check( hipMalloc( &p, bytes ), "hipMalloc" );
  hip_check( hipMemset( p, 0, bytes ), "hipMemset" );
  bytes_ += bytes;
  return p;
}

template < typename T >
T*
Engine::upload( const std::vector< T >& v )
{
  T* p = static_cast< T* >( alloc( v.size() * sizeof( T ) ) );
  if ( not v.empty() )
    hip_check( hipMemcpy( p, v.data(), v.size() * sizeof( T ), hipMemcpyHostToDevice ), "hipMemcpy H2D" );
  return p;
}

void
Engine::free_store()
{
  for ( void* p : store_allocs_ )
    hipFree( p );
  store_allocs_.clear();
}

void
Engine::build( const BuildSpec& b, bool keep_state )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  if ( keep_state and not built_ )
    keep_state = false;
  // ---- synapse store and delivery scratch (always rebuilt)
  free_store();
  auto S = [ this ]( void* p )
  {
    store_allocs_.push_back( p );
    return p;
  };
  net_.n = b.n;
  for ( int m = 0; m <= N_MODELS; ++m )
    net_.model_begin[ m ] = b.model_begin[ m ];
  net_.min_delay = b.min_delay;
  net_.max_delay = b.max_delay;
  net_.L = b.min_delay + b.max_delay;
  net_.nw = ( b.min_delay + 63 ) / 64;
  if ( net_.nw > 4 )
    throw std::runtime_error( "min_delay above 256 steps is not supported on the device path" );
  net_.tgt_bits = b.tgt_bits;
  int sb = 1;
  while ( ( 1 << sb ) < b.max_delay )
    ++sb;
  net_.slot_bits = sb;
  net_.n_vp = b.n_vp;
  net_.small_max = 32;
  net_.s_cap = b.s_cap;
  uint64_t max_len = 0;
  for ( uint32_t i = 0; i < b.n; ++i )
    max_len = std::max( max_len, b.row_ptr[ i + 1 ] - b.row_ptr[ i ] );
  if ( max_len >= ( uint64_t( 1 ) << ( 32 - sb ) ) )
    throw std::runtime_error( "out-degree too large for the record packing" );
  net_.cmax = static_cast< uint32_t >( std::max< uint64_t >( 1, ( max_len + 255 ) / 256 ) );
  net_.row_ptr = static_cast< const uint64_t* >( S( upload( b.row_ptr ) ) );
  net_.tgt_d = static_cast< const uint32_t* >( S( upload( b.tgt_d ) ) );
  net_.wgt = static_cast< const double* >( S( upload( b.w ) ) );
  std::vector< uint64_t > inp = b.in_ptr;
  uint64_t max_cap = 0;
  for ( uint32_t i = 0; i < b.n; ++i )
    max_cap = std::max( max_cap, ( inp[ i + 1 ] - inp[ i ] ) * static_cast< uint64_t >( b.s_cap ) );
  net_.max_cap = max_cap;
  net_.in_ptr = static_cast< const uint64_t* >( S( upload( inp ) ) );
  net_.cnt = static_cast< uint32_t* >( S( alloc( sizeof( uint32_t ) * std::max< uint32_t >( b.n, 1 ) ) ) );
  const uint64_t n_rec = std::max< uint64_t >( 1, inp[ b.n ] * static_cast< uint64_t >( b.s_cap ) );
  net_.rec = static_cast< Rec* >( S( alloc( sizeof( Rec ) * n_rec ) ) );
  // merge scratch only needed when some target can exceed the LDS sort size
  const uint64_t tmp_regions = max_cap > 2048 ? NG_GRID_BIG : 0;
  net_.rec_tmp = static_cast< Rec* >( S( alloc( sizeof( Rec ) * std::max< uint64_t >( 1, tmp_regions * max_cap ) ) ) );
  net_.big = static_cast< uint32_t* >( S( alloc( sizeof( uint32_t ) * std::max< uint32_t >( b.n, 1 ) ) ) );
  // ---- drive (inputs; replaced on rebuild)
  net_.drive_counts = static_cast< const uint8_t* >( S( upload( b.drive_counts ) ) );
  net_.drive_cols = b.drive_cols;
  net_.drive_idx = static_cast< const uint32_t* >( S( upload( b.drive_idx ) ) );
  net_.ptables = static_cast< const PoissonTable* >( S( upload( b.ptables ) ) );
  net_.psets = static_cast< const ParamSet* >( S( upload( b.psets ) ) );
  for ( int m = 0; m < N_MODELS; ++m )
    net_.pidx[ m ] = static_cast< uint16_t* >( S( upload( b.pidx[ m ] ) ) );
  if ( not b.dev2ord.empty() )
  {
    std::vector< uint32_t > o2d( b.n );
    for ( uint32_t i = 0; i < b.n; ++i )
      o2d[ b.dev2ord[ i ] ] = i;
    net_.dev2ord = static_cast< const uint32_t* >( S( upload( b.dev2ord ) ) );
    net_.ord2dev = static_cast< const uint32_t* >( S( upload( o2d ) ) );
  }
  else
  {
    net_.dev2ord = nullptr;
    net_.ord2dev = nullptr;
  }
  net_.vp_begin = static_cast< const uint32_t* >( S( upload( b.vp_begin ) ) );

  if ( not keep_state )
  {
    for ( void* p : state_allocs_ )
      hipFree( p );
    state_allocs_.clear();
    auto T = [ this ]( void* p )
    {
      state_allocs_.push_back( p );
      return p;
    };
    for ( int m = 0; m < N_MODELS; ++m )
    {
      const uint32_t nm = b.model_begin[ m + 1 ] - b.model_begin[ m ];
      for ( int f = 0; f < 5; ++f )
        net_.st[ m ][ f ] = b.st[ m ][ f ].empty() ? nullptr : static_cast< double* >( T( upload( b.st[ m ][ f ] ) ) );
      net_.r[ m ] = static_cast< int32_t* >( T( upload( b.r[ m ] ) ) );
      net_.rb[ m ] = static_cast< double* >(
        T( alloc( sizeof( double ) * static_cast< size_t >( net_.L ) * model_channels( m ) * std::max< uint32_t >( nm, 1 ) ) ) );
    }
    const uint64_t spk_cap = static_cast< uint64_t >( std::max< uint32_t >( b.n, 1 ) ) * b.min_delay;
    net_.mask = static_cast< uint64_t* >( T( alloc( sizeof( uint64_t ) * std::max< uint32_t >( b.n, 1 ) * net_.nw ) ) );
    net_.spk = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * spk_cap ) ) );
    net_.spk_lag = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * spk_cap ) ) );
    net_.spk_tmp = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * 2 * spk_cap ) ) );
    net_.ctr = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * 16 ) ) );
    net_.scan_state = static_cast< uint32_t* >( T( alloc( sizeof( uint64_t ) * ( ( b.n + 255 ) / 256 + 1 ) ) ) );
    net_.log_cap = b.log_cap;
    net_.log_src = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * b.log_cap ) ) );
    net_.log_step = static_cast< int64_t* >( T( alloc( sizeof( int64_t ) * b.log_cap ) ) );
    net_.log_n = static_cast< unsigned long long* >( T( alloc( sizeof( unsigned long long ) ) ) );
    clock = 0;
    from_step = 0;
  }
  net_.clock = clock;
  net_.from0 = from_step;
  net_.todo = 0;
  net_.active = 0;
  built_ = true;
}

void
Engine::set_state( int model, int field, uint32_t j, double v )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  hip_check( hipMemcpy( net_.st[ model ][ field ] + j, &v, sizeof v, hipMemcpyHostToDevice ), "set_state" );
}

void
Engine::set_r( int model, uint32_t j, int32_t v )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  hip_check( hipMemcpy( net_.r[ model ] + j, &v, sizeof v, hipMemcpyHostToDevice ), "set_r" );
}

void
Engine::get_state( int model, int field, std::vector< double >& out )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  const uint32_t nm = net_.model_begin[ model + 1 ] - net_.model_begin[ model ];
  out.resize( nm );
  if ( nm )
    hip_check( hipMemcpy( out.data(), net_.st[ model ][ field ], nm * sizeof( double ), hipMemcpyDeviceToHost ),
      "get_state" );
}

void
Engine::get_r( int model, std::vector< int32_t >& out )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  const uint32_t nm = net_.model_begin[ model + 1 ] - net_.model_begin[ model ];
  out.resize( nm );
  if ( nm )
    hip_check( hipMemcpy( out.data(), net_.r[ model ], nm * sizeof( int32_t ), hipMemcpyDeviceToHost ), "get_r" );
}

void
Engine::drain_log()
{
  std::vector< uint32_t >& src = log_src;
  std::vector< int64_t >& step = log_step;
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  unsigned long long n = 0;
  hip_check( hipMemcpy( &n, net_.log_n, sizeof n, hipMemcpyDeviceToHost ), "log_n" );
  if ( n > net_.log_cap )
    throw std::runtime_error( "spike log overflow: more spikes in one chunk than the log holds" );
  const size_t old = src.size();
  src.resize( old + n );
  step.resize( old + n );
  if ( n )
  {
    hip_check( hipMemcpy( src.data() + old, net_.log_src, n * sizeof( uint32_t ), hipMemcpyDeviceToHost ), "log" );
    hip_check( hipMemcpy( step.data() + old, net_.log_step, n * sizeof( int64_t ), hipMemcpyDeviceToHost ), "log" );
  }
  unsigned long long z = 0;
  hip_check( hipMemcpy( net_.log_n, &z, sizeof z, hipMemcpyHostToDevice ), "log reset" );
}

// ------------------------------------------------------------------ Runner
Runner::Runner( int device, int slices_per_chunk )
  : device_( device )
  , S_( slices_per_chunk )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  hip_check( hipStreamCreateWithFlags( &stream_, hipStreamNonBlocking ), "hipStreamCreate" );
  hipDeviceProp_t prop;
  hip_check( hipGetDeviceProperties( &prop, device_ ), "hipGetDeviceProperties" );
  const int cus = prop.multiProcessorCount;
  grid_expand = cus * 8;
  grid_big = NG_GRID_BIG;  // k_big indexes one merge-scratch region per workgroup
  grid_update = cus * 8;
  grid_compact = cus;  // one resident workgroup per CU: the look-back never waits on a non-resident tile
}

Runner::~Runner()
{
  hipSetDevice( device_ );
  if ( exec_ )
    hipGraphExecDestroy( exec_ );
  if ( d_nets_ )
    hipFree( d_nets_ );
  if ( h_nets_ )
    hipHostFree( h_nets_ );
  if ( stream_ )
    hipStreamDestroy( stream_ );
}

void
Runner::ensure_capacity( size_t n )
{
  if ( n <= cap_ )
    return;
  if ( exec_ )
  {
    hipGraphExecDestroy( exec_ );
    exec_ = nullptr;
    exec_n_ = -1;
  }
  if ( d_nets_ )
    hipFree( d_nets_ );
  if ( h_nets_ )
    hipHostFree( h_nets_ );
  hip_check( hipMalloc( &d_nets_, n * sizeof( NetDev ) ), "hipMalloc nets" );
  hip_check( hipHostMalloc( &h_nets_, n * sizeof( NetDev ), hipHostMallocDefault ), "hipHostMalloc nets" );
  cap_ = n;
}

void
Runner::capture( int n_nets, const bool* has_model )
{
  if ( exec_ and exec_n_ == n_nets and std::equal( has_model, has_model + N_MODELS, exec_models_ ) )
    return;
  if ( exec_ )
  {
    hipGraphExecDestroy( exec_ );
    exec_ = nullptr;
  }
  LaunchCfg cfg { grid_expand, grid_big, grid_update, grid_compact };
  hipGraph_t g;
  hip_check( hipStreamBeginCapture( stream_, hipStreamCaptureModeThreadLocal ), "hipStreamBeginCapture" );
  for ( int s = 0; s < S_; ++s )
    launch_slice( d_nets_, n_nets, s, has_model, cfg, stream_ );
  hip_check( hipStreamEndCapture( stream_, &g ), "hipStreamEndCapture" );
  hip_check( hipGraphInstantiate( &exec_, g, nullptr, nullptr, 0 ), "hipGraphInstantiate" );
  hip_check( hipGraphDestroy( g ), "hipGraphDestroy" );
  exec_n_ = n_nets;
  std::copy( has_model, has_model + N_MODELS, exec_models_ );
}

void
Runner::run( const std::vector< Engine* >& engines, const std::vector< int64_t >& steps )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  const int n = static_cast< int >( engines.size() );
  if ( n == 0 )
    return;
  ensure_capacity( n );
  bool has_model[ N_MODELS ] = { false, false, false };
  for ( auto* e : engines )
    for ( int m = 0; m < N_MODELS; ++m )
      has_model[ m ] = has_model[ m ] or e->has_model( m );
  std::vector< int64_t > remaining = steps;
  LaunchCfg cfg { grid_expand, grid_big, grid_update, grid_compact };
  hipEvent_t e0, e1;
  hip_check( hipEventCreate( &e0 ), "hipEventCreate" );
  hip_check( hipEventCreate( &e1 ), "hipEventCreate" );
  last_ms_ = 0;
  for ( ;; )
  {
    bool any = false;
    for ( int i = 0; i < n; ++i )
    {
      Engine& E = *engines[ i ];
      NetDev d = E.net();
      d.clock = E.clock;
      d.from0 = E.from_step;
      d.chunk_seq = E.chunk_seq;
      const int64_t span = static_cast< int64_t >( d.min_delay - E.from_step ) + static_cast< int64_t >( S_ - 1 ) * d.min_delay;
      d.todo = std::min( remaining[ i ], span );
      d.active = d.todo > 0 ? 1 : 0;
      any = any or d.active;
      h_nets_[ i ] = d;
    }
    if ( not any )
      break;
    if ( use_graph )
      capture( n, has_model );
    hip_check( hipMemcpyAsync( d_nets_, h_nets_, n * sizeof( NetDev ), hipMemcpyHostToDevice, stream_ ), "nets H2D" );
    hip_check( hipEventRecord( e0, stream_ ), "hipEventRecord" );
    if ( use_graph )
      hip_check( hipGraphLaunch( exec_, stream_ ), "hipGraphLaunch" );
    else
      for ( int s = 0; s < S_; ++s )
        hip_check( launch_slice( d_nets_, n, s, has_model, cfg, stream_ ), "launch_slice" );
    hip_check( hipEventRecord( e1, stream_ ), "hipEventRecord" );
    hip_check( hipStreamSynchronize( stream_ ), "hipStreamSynchronize" );
    float ms = 0;
    hipEventElapsedTime( &ms, e0, e1 );
    last_ms_ += ms;
    for ( int i = 0; i < n; ++i )
    {
      Engine& E = *engines[ i ];
      const NetDev& d = h_nets_[ i ];
      if ( not d.active )
        continue;
      uint32_t ctr[ 8 ];
      hip_check( hipMemcpy( ctr, E.net().ctr, sizeof ctr, hipMemcpyDeviceToHost ), "ctr" );
      if ( ctr[ 2 ] )
      {
        std::string why;
        if ( ctr[ 2 ] & 1u )
          why += " event records exceeded their capacity (more spikes per source and slice than reserved);";
        if ( ctr[ 2 ] & 2u )
          why += " NumericalInstability in aeif_cond_exp;";
        if ( ctr[ 2 ] & 4u )
          why += " aeif_cond_exp emitted two spikes within one step;";
        if ( ctr[ 2 ] & 8u )
          why += " spike compaction did not converge;";
        throw std::runtime_error( "device run failed:" + why );
      }
      const int64_t total = E.from_step + d.todo;
      E.clock += ( total / d.min_delay ) * d.min_delay;
      E.from_step = static_cast< int32_t >( total % d.min_delay );
      E.chunk_seq += 1;
      remaining[ i ] -= d.todo;
      unsigned long long ln = 0;
      hip_check( hipMemcpy( &ln, E.net().log_n, sizeof ln, hipMemcpyDeviceToHost ), "log_n" );
      if ( ln > E.net().log_cap )
        throw std::runtime_error( "spike log overflow within one chunk" );
      if ( ln > E.net().log_cap / 2 )
        E.drain_log();
    }
  }
  hipEventDestroy( e0 );
  hipEventDestroy( e1 );
}

}  // namespace nestgpu
