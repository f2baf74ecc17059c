// ng_engine.cpp -- HBM placement of a network instance and the chunked slice loop.
#include "ng_engine.h"

#include <algorithm>
#include <cstring>
#include <mutex>
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
    free_list( store_allocs_ );
    free_list( state_allocs_ );
    free_list( log_allocs_ );
    free_list( mm_allocs_ );
  }
}

void
Engine::free_list( std::vector< void* >& l )
{
  for ( void* p : l )
  {
    hipFree( p );
    const auto it = sizes_.find( p );
    if ( it != sizes_.end() )
    {
      bytes_ -= it->second;
      sizes_.erase( it );
    }
  }
  l.clear();
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
  sizes_[ p ] = bytes;
  return p;
}

template < typename T, typename A >
T*
Engine::upload( const std::vector< T, A >& v )
{
  T* p = static_cast< T* >( alloc( v.size() * sizeof( T ) ) );
  if ( not v.empty() )
    hip_check( hipMemcpy( p, v.data(), v.size() * sizeof( T ), hipMemcpyHostToDevice ), "hipMemcpy H2D" );
  return p;
}

void
Engine::build( const BuildSpec& b, bool keep_state )
{
  build_layout( b, keep_state, false );
}

void
Engine::build_layout( const BuildSpec& b, bool keep_state, bool relayout )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  if ( keep_state
       and ( not built_ or net_.n != b.n or net_.min_delay != b.min_delay or net_.max_delay != b.max_delay
             or net_.nch != b.nch or ( not relayout and ( net_.tb != b.tb or bstart_ != b.bstart ) ) ) )
    throw std::runtime_error( "device store rebuild changed the network shape after Simulate" );
  if ( b.min_delay > 256 )
    throw std::runtime_error( "NotImplemented: min_delay above 256 steps on the device path" );
  if ( b.tb == 0 or b.tb > MAX_TB )
    throw std::runtime_error( "targets_per_block must be in [1, " + std::to_string( MAX_TB ) + "]" );
  if ( b.n >= ( 1u << 24 ) )
    throw std::runtime_error( "NotImplemented: 2^24 or more neurons in one instance" );
  if ( b.bstart.size() != b.n_blocks() + 1 or b.bstart.front() != 0 or b.bstart.back() != b.n )
    throw std::runtime_error( "target blocks do not cover the neurons" );
  for ( size_t q = 0; q + 1 < b.bstart.size(); ++q )
    if ( b.bstart[ q + 1 ] <= b.bstart[ q ] or b.bstart[ q + 1 ] - b.bstart[ q ] > b.tb )
      throw std::runtime_error( "target blocks must be non-empty and at most targets_per_block large" );
  if ( b.nch != 1 and b.nch != 2 )
    throw std::runtime_error( "ring-buffer channels must be 1 or 2" );
  const uint32_t P = b.n_blocks();
  if ( P > ( b.win ? PMAX_WIN : PMAX_GLOB ) )
    throw std::runtime_error( "NotImplemented: too many target blocks in one instance; raise targets_per_block" );
  if ( b.cache and ( b.win or P > PMAX_WIN or b.np != 1 ) )
    throw std::runtime_error( "cache delivery: at most " + std::to_string( PMAX_WIN )
      + " blocks, one pass and no window" );
  if ( b.win
       and ( b.wn < 1 or b.wa < 0 or b.wa + b.wn > b.max_delay
             or static_cast< uint64_t >( b.wn ) * b.nch * b.tb > WIN_CAP ) )
    throw std::runtime_error( "window delivery: the window does not fit the LDS or the ring" );
  if ( b.syn.size() >= ( uint64_t( 1 ) << 32 ) )
    throw std::runtime_error( "NotImplemented: 2^32 or more synapses in one instance" );
  if ( static_cast< uint64_t >( b.min_delay + b.max_delay ) * b.nch * b.tb * sizeof( double ) >= ( uint64_t( 1 ) << 31 ) )
    throw std::runtime_error( "NotImplemented: ring buffer of one block at or above 2 GiB (delays too long)" );
  free_list( store_allocs_ );
  auto S = [ this ]( void* p )
  {
    store_allocs_.push_back( p );
    return p;
  };
  net_.n = b.n;
  net_.tb = b.tb;
  net_.n_blocks = P;
  net_.min_delay = b.min_delay;
  net_.max_delay = b.max_delay;
  net_.L = b.min_delay + b.max_delay;
  net_.n_vp = b.n_vp;
  net_.uniform_model = b.uniform_model;
  net_.spk_cap = b.tb * static_cast< uint32_t >( b.min_delay );
  net_.nch = b.nch;
  net_.win = b.win ? 1 : 0;
  net_.cache = b.cache ? 1 : 0;
  net_.wa = b.win ? b.wa : 0;
  net_.wn = b.win ? b.wn : 0;
  if ( b.np < 1 or b.np > MAX_PASSES or static_cast< int64_t >( b.np ) * b.ws < b.max_delay or ( b.win and b.np != 1 ) )
    throw std::runtime_error( "invalid delivery passes" );
  net_.np = b.np;
  net_.ws = b.ws;
  net_.log_enabled = b.log_enabled ? 1 : 0;
  mask_ = 0;
  for ( const auto& ps : b.psets )
    mask_ |= 1 << ps.model;
  if ( not b.stdp_prm.empty() )
    mask_ |= MASK_STDP;  // the launch needs the stdp_synapse instantiation of the slice kernel
  // ---- store and inputs (replaced on every build)
  net_.boff = static_cast< const uint32_t* >( S( upload( b.boff ) ) );
  net_.syn = static_cast< const SynRec* >( S( upload( b.syn ) ) );
  net_.psets = static_cast< const ParamSet* >( S( upload( b.psets ) ) );
  net_.pidx = static_cast< const uint16_t* >( S( upload( b.pidx ) ) );
  net_.model = static_cast< const uint8_t* >( S( upload( b.model ) ) );
  net_.vp_begin = static_cast< const uint32_t* >( S( upload( b.vp_begin ) ) );
  net_.bstart = static_cast< const uint32_t* >( S( upload( b.bstart ) ) );
  bstart_ = b.bstart;
  n_syn_ = b.syn.size();
  // stdp_synapse: parameter sets, record state and neuron flags come with the store; the
  // post-spike histories are state
  net_.stdp = b.stdp_prm.empty() ? 0u : 1u;
  net_.ms_per_tic = b.ms_per_tic;
  net_.ms_per_step = b.ms_per_step;
  net_.tics_per_step = b.tics_per_step;
  net_.spre = nullptr;
  net_.pre_row0 = net_.srow = nullptr;
  net_.stdp_in = nullptr;
  net_.tau_minus_inv = nullptr;
  if ( net_.stdp )
  {
    if ( keep_state and not relayout )
      throw std::runtime_error( "NotImplemented: rebuilding the device copy of a network with stdp_synapse after Simulate" );
    if ( b.stdp_srow.size() != b.syn.size() or b.stdp_in.size() != b.n or b.stdp_pre_row0.size() != b.n + 1
         or b.stdp_pre.size() != b.stdp_pre_row0.back() or b.tau_minus_inv.size() != b.n )
      throw std::runtime_error( "stdp_synapse: trace rows, per-record rows and per-neuron flags expected" );
    net_.spre = static_cast< StdpPre* >( S( upload( b.stdp_pre ) ) );
    net_.pre_row0 = static_cast< const uint32_t* >( S( upload( b.stdp_pre_row0 ) ) );
    net_.srow = static_cast< const uint32_t* >( S( upload( b.stdp_srow ) ) );
    net_.stdp_in = static_cast< const uint8_t* >( S( upload( b.stdp_in ) ) );
    net_.tau_minus_inv = static_cast< const double* >( S( upload( b.tau_minus_inv ) ) );
  }
  net_.drive_counts = static_cast< const uint8_t* >( S( upload( b.drive_counts ) ) );
  net_.drive_cols = b.drive_cols;
  net_.drive_idx = static_cast< const uint32_t* >( S( upload( b.drive_idx ) ) );
  net_.ptables = static_cast< const PoissonTable* >( S( upload( b.ptables ) ) );
  if ( not keep_state )
  {
    free_list( state_allocs_ );
    auto T = [ this ]( void* p )
    {
      state_allocs_.push_back( p );
      return p;
    };
    for ( int f = 0; f < 5; ++f )
      net_.v[ f ] = static_cast< double* >( T( upload( b.v[ f ] ) ) );
    net_.r = static_cast< int32_t* >( T( upload( b.r ) ) );
    net_.rb = static_cast< double* >( T( alloc( sizeof( double ) * static_cast< size_t >( P ) * net_.L * b.nch * b.tb ) ) );
    net_.blk_cnt = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * 2 * std::max< uint32_t >( P, 1 ) ) ) );
    net_.blk_spk = static_cast< uint32_t* >(
      T( alloc( sizeof( uint32_t ) * 2 * static_cast< size_t >( std::max< uint32_t >( P, 1 ) ) * net_.spk_cap ) ) );
    net_.blk_part = static_cast< uint8_t* >(
      T( alloc( sizeof( uint8_t ) * 2 * static_cast< size_t >( std::max< uint32_t >( P, 1 ) ) * net_.spk_cap ) ) );
    const size_t flat_cap = static_cast< size_t >( std::max< uint32_t >( b.n, 1 ) ) * b.min_delay;
    net_.flat_src = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * flat_cap ) ) );
    net_.flat_lag = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * flat_cap ) ) );
    net_.flat_meta = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * 4 ) ) );
    net_.log_cap = 0;
    net_.log_n = static_cast< unsigned long long* >( T( alloc( sizeof( unsigned long long ) ) ) );
    net_.counters = static_cast< unsigned long long* >( T( alloc( sizeof( unsigned long long ) * N_COUNTERS ) ) );
    rx_ = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * std::max< uint32_t >( b.n, 1 ) ) ) );
    net_.rx = nullptr;
    net_.hist_t = net_.hist_k = net_.kminus = net_.last_spike = nullptr;
    net_.hist_n = nullptr;
    if ( net_.stdp )
    {
      // ArchivingNode state: no history, Kminus 0, last spike -1 ms (archiving_node.cpp:38-48)
      if ( b.stdp_hcap < 2 or b.stdp_hcap > STDP_HIST_MAX or ( b.stdp_hcap & ( b.stdp_hcap - 1 ) ) )
        throw std::runtime_error( "stdp_synapse: history capacity must be a power of two in [2, 65536]" );
      net_.stdp_hcap = b.stdp_hcap;
      const size_t nh = static_cast< size_t >( b.n ) * b.stdp_hcap;
      net_.hist_t = static_cast< double* >( T( alloc( sizeof( double ) * nh ) ) );
      net_.hist_k = static_cast< double* >( T( alloc( sizeof( double ) * nh ) ) );
      net_.hist_n = static_cast< uint32_t* >( T( alloc( sizeof( uint32_t ) * b.n ) ) );
      net_.kminus = static_cast< double* >( T( alloc( sizeof( double ) * b.n ) ) );
      net_.last_spike = static_cast< double* >( T( upload( std::vector< double >( b.n, -1.0 ) ) ) );
    }
    clock = 0;
    from_step = 0;
    part = 0;
  }
  if ( keep_state and relayout )
  {
    // new ring buffers and spike lists for the new blocks (the caller moves their contents)
    auto replace = [ this ]( void* old, size_t bytes )
    {
      void* p = alloc( bytes );
      for ( auto& q : state_allocs_ )
        if ( q == old )
          q = p;
      std::vector< void* > one( 1, old );
      free_list( one );
      return p;
    };
    const size_t P1 = std::max< uint32_t >( P, 1 );
    net_.rb = static_cast< double* >( replace( net_.rb, sizeof( double ) * static_cast< size_t >( P ) * net_.L * b.nch * b.tb ) );
    net_.blk_cnt = static_cast< uint32_t* >( replace( net_.blk_cnt, sizeof( uint32_t ) * 2 * P1 ) );
    net_.blk_spk = static_cast< uint32_t* >( replace( net_.blk_spk, sizeof( uint32_t ) * 2 * P1 * net_.spk_cap ) );
    net_.blk_part = static_cast< uint8_t* >( replace( net_.blk_part, sizeof( uint8_t ) * 2 * P1 * net_.spk_cap ) );
  }
  if ( b.log_cap != net_.log_cap )
  {
    // the caller drains the log before a rebuild, so the buffers can be replaced
    free_list( log_allocs_ );
    auto Lg = [ this ]( void* p )
    {
      log_allocs_.push_back( p );
      return p;
    };
    net_.log_cap = b.log_cap;
    net_.log_src = static_cast< uint32_t* >( Lg( alloc( sizeof( uint32_t ) * b.log_cap ) ) );
    net_.log_step = static_cast< int64_t* >( Lg( alloc( sizeof( int64_t ) * b.log_cap ) ) );
    unsigned long long z = 0;
    hip_check( hipMemcpy( net_.log_n, &z, sizeof z, hipMemcpyHostToDevice ), "log reset" );
  }
  // multimeter sampling: columns and a sample buffer for the rows of one chunk (at most
  // MM_ROWS rows; the runner shortens chunks that would sample more)
  free_list( mm_allocs_ );
  net_.mm_int = 0;
  net_.mm_col = nullptr;
  net_.mm_buf = nullptr;
  net_.mm_cols = net_.mm_rows = 0;
  net_.mm_off = b.mm_off;
  if ( b.mm_int > 0 and b.mm_cols > 0 )
  {
    if ( b.mm_col.size() != b.n )
      throw std::runtime_error( "multimeter: one sample column entry per neuron expected" );
    auto M = [ this ]( void* p )
    {
      mm_allocs_.push_back( p );
      return p;
    };
    net_.mm_int = b.mm_int;
    net_.mm_cols = b.mm_cols;
    net_.mm_rows = MM_ROWS;
    net_.mm_col = static_cast< const uint32_t* >( M( upload( b.mm_col ) ) );
    net_.mm_buf = static_cast< double* >( M( alloc( sizeof( double ) * 4 * MM_ROWS * static_cast< size_t >( b.mm_cols ) ) ) );
  }
  net_.clock = clock;
  net_.from0 = from_step;
  net_.part0 = part;
  net_.todo = 0;
  net_.active = 0;
  built_ = true;
}

bool
Engine::relayout( const BuildSpec& b )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  if ( not built_ or part != 0 or from_step != 0 or b.n != net_.n or b.nch != net_.nch or b.min_delay != net_.min_delay
       or b.max_delay != net_.max_delay or b.bstart.size() < 2 )
    return false;
  hip_check( hipDeviceSynchronize(), "relayout: synchronize" );
  const uint32_t P0 = net_.n_blocks, tb0 = net_.tb, cap0 = net_.spk_cap;
  const size_t cells = static_cast< size_t >( net_.L ) * net_.nch;  // ring cells per target
  const std::vector< uint32_t > start0 = bstart_;
  std::vector< uint32_t > cnt0( 2 * static_cast< size_t >( P0 ) ), spk0( cnt0.size() * cap0 );
  std::vector< uint8_t > part0( spk0.size() );
  std::vector< double > rb0( static_cast< size_t >( P0 ) * cells * tb0 );
  auto d2h = []( void* dst, const void* src, size_t bytes )
  {
    if ( bytes )
      hip_check( hipMemcpy( dst, src, bytes, hipMemcpyDeviceToHost ), "relayout: read" );
  };
  d2h( cnt0.data(), net_.blk_cnt, cnt0.size() * sizeof( uint32_t ) );
  d2h( spk0.data(), net_.blk_spk, spk0.size() * sizeof( uint32_t ) );
  d2h( part0.data(), net_.blk_part, part0.size() );
  d2h( rb0.data(), net_.rb, rb0.size() * sizeof( double ) );
  build_layout( b, true, true );
  const uint32_t P1 = net_.n_blocks, tb1 = net_.tb, cap1 = net_.spk_cap;
  auto block_of = [ this ]( uint32_t i )
  { return static_cast< uint32_t >( std::upper_bound( bstart_.begin(), bstart_.end(), i ) - bstart_.begin() ) - 1; };
  // ring cells: target i keeps its cells, at its place in its new block
  std::vector< double > rb1( static_cast< size_t >( P1 ) * cells * tb1, 0.0 );
  for ( uint32_t q0 = 0; q0 < P0; ++q0 )
    for ( uint32_t i = start0[ q0 ]; i < start0[ q0 + 1 ]; ++i )
    {
      const uint32_t q1 = block_of( i );
      const size_t l0 = i - start0[ q0 ], l1 = i - bstart_[ q1 ];
      for ( size_t c = 0; c < cells; ++c )
        rb1[ ( q1 * cells + c ) * tb1 + l1 ] = rb0[ ( q0 * cells + c ) * tb0 + l0 ];
    }
  // spike lists: the lists of consecutive blocks, read in block order, hold their entries in
  // device-index order; cutting that sequence at the new bounds keeps it
  std::vector< uint32_t > cnt1( 2 * static_cast< size_t >( P1 ), 0 ), spk1( cnt1.size() * cap1, 0 );
  std::vector< uint8_t > part1( spk1.size(), 0 );
  for ( size_t par = 0; par < 2; ++par )
    for ( uint32_t q0 = 0; q0 < P0; ++q0 )
    {
      const size_t l0 = par * P0 + q0;
      for ( uint32_t k = 0; k < std::min( cnt0[ l0 ], cap0 ); ++k )
      {
        const uint32_t e = spk0[ l0 * cap0 + k ];
        const size_t l1 = par * P1 + block_of( e >> 8 );
        const uint32_t o = cnt1[ l1 ]++;
        if ( o >= cap1 )
          throw std::runtime_error( "relayout: spike list of a new block overflows" );
        spk1[ l1 * cap1 + o ] = e;
        part1[ l1 * cap1 + o ] = part0[ l0 * cap0 + k ];
      }
    }
  auto h2d = []( void* dst, const void* src, size_t bytes )
  {
    if ( bytes )
      hip_check( hipMemcpy( dst, src, bytes, hipMemcpyHostToDevice ), "relayout: write" );
  };
  h2d( net_.rb, rb1.data(), rb1.size() * sizeof( double ) );
  h2d( net_.blk_cnt, cnt1.data(), cnt1.size() * sizeof( uint32_t ) );
  h2d( net_.blk_spk, spk1.data(), spk1.size() * sizeof( uint32_t ) );
  h2d( net_.blk_part, part1.data(), part1.size() );
  return true;
}

void
Engine::read_weights( std::vector< double >& w ) const
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  std::vector< SynRec > recs( n_syn_ );
  if ( n_syn_ )
    hip_check( hipMemcpy( recs.data(), net_.syn, n_syn_ * sizeof( SynRec ), hipMemcpyDeviceToHost ), "read weights" );
  w.resize( n_syn_ );
  for ( size_t i = 0; i < n_syn_; ++i )
  {
    const uint64_t bits = ( static_cast< uint64_t >( recs[ i ].w_hi ) << 32 ) | recs[ i ].w_lo;
    std::memcpy( &w[ i ], &bits, sizeof bits );
  }
}

void
Engine::count_received( bool on )
{
  net_.rx = on ? rx_ : nullptr;
}

void
Engine::read_received( std::vector< uint32_t >& out ) const
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  out.resize( net_.n );
  if ( net_.n )
    hip_check( hipMemcpy( out.data(), rx_, net_.n * sizeof( uint32_t ), hipMemcpyDeviceToHost ), "read received" );
}

void
Engine::clear_received()
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  if ( net_.n )
    hip_check( hipMemset( rx_, 0, net_.n * sizeof( uint32_t ) ), "clear received" );
}

void
Engine::set_state( int field, uint32_t dev, double v )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  hip_check( hipMemcpy( net_.v[ field ] + dev, &v, sizeof v, hipMemcpyHostToDevice ), "set_state" );
}

void
Engine::get_state( int field, std::vector< double >& out )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  out.resize( net_.n );
  if ( net_.n )
    hip_check( hipMemcpy( out.data(), net_.v[ field ], net_.n * sizeof( double ), hipMemcpyDeviceToHost ), "get_state" );
}

void
Engine::get_r( std::vector< int32_t >& out )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  out.resize( net_.n );
  if ( net_.n )
    hip_check( hipMemcpy( out.data(), net_.r, net_.n * sizeof( int32_t ), hipMemcpyDeviceToHost ), "get_r" );
}

unsigned long long
Engine::read_counter( int i ) const
{
  unsigned long long v = 0;
  hip_check( hipMemcpy( &v, net_.counters + i, sizeof v, hipMemcpyDeviceToHost ), "read counter" );
  return v;
}

void
Engine::read_counters( unsigned long long* out ) const
{
  hip_check( hipMemcpy( out, net_.counters, sizeof( unsigned long long ) * N_COUNTERS, hipMemcpyDeviceToHost ),
    "read counters" );
}

void
Engine::clear_phase_ticks()
{
  const unsigned long long z[ 3 ] = { 0, 0, 0 };
  hip_check( hipMemcpy( net_.counters + 4, z, sizeof z, hipMemcpyHostToDevice ), "clear phase ticks" );
}

void
Engine::clear_error_flags()
{
  unsigned long long z = 0;
  hip_check( hipMemcpy( net_.counters, &z, sizeof z, hipMemcpyHostToDevice ), "clear flags" );
}

void
Engine::drain_log()
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  unsigned long long n = 0;
  hip_check( hipMemcpy( &n, net_.log_n, sizeof n, hipMemcpyDeviceToHost ), "log_n" );
  if ( n > net_.log_cap )
    throw std::runtime_error( "spike log overflow: more spikes in one chunk than the log holds" );
  if ( not net_.log_enabled )
    return;
  const size_t old = log_src.size();
  log_src.resize( old + n );
  log_step.resize( old + n );
  if ( n )
  {
    hip_check( hipMemcpy( log_src.data() + old, net_.log_src, n * sizeof( uint32_t ), hipMemcpyDeviceToHost ), "log" );
    hip_check( hipMemcpy( log_step.data() + old, net_.log_step, n * sizeof( int64_t ), hipMemcpyDeviceToHost ), "log" );
  }
  unsigned long long z = 0;
  hip_check( hipMemcpy( net_.log_n, &z, sizeof z, hipMemcpyHostToDevice ), "log reset" );
}

void
Engine::drain_samples()
{
  if ( mm_chunk_rows <= 0 )
    return;
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  const size_t per_row = 4 * static_cast< size_t >( net_.mm_cols );
  const size_t old = mm_data.size();
  mm_data.resize( old + per_row * mm_chunk_rows );
  hip_check( hipMemcpy( mm_data.data() + old, net_.mm_buf, per_row * mm_chunk_rows * sizeof( double ),
               hipMemcpyDeviceToHost ),
    "multimeter samples" );
  for ( int64_t r = 0; r < mm_chunk_rows; ++r )
    mm_stamps.push_back( net_.mm_off + ( mm_chunk_row0 + r ) * net_.mm_int );
  mm_chunk_rows = 0;
}

// ------------------------------------------------------------------ Runner
Runner::Runner( int device, int slices_per_chunk )
  : device_( device )
  , S_( slices_per_chunk )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  hip_check( hipStreamCreateWithFlags( &stream_, hipStreamNonBlocking ), "hipStreamCreate" );
  for ( int q = 0; q < 2; ++q )
  {
    hip_check( hipEventCreate( &ev_chunk_[ q ] ), "hipEventCreate" );
    hip_check( hipEventCreate( &ev_kernel_[ q ] ), "hipEventCreate" );
  }
}

Runner::~Runner()
{
  hipSetDevice( device_ );
  if ( exec_ )
    hipGraphExecDestroy( exec_ );
  if ( d_nets_ )
    hipFree( d_nets_ );
  if ( d_prefix_ )
    hipFree( d_prefix_ );
  if ( h_nets_ )
    hipHostFree( h_nets_ );
  if ( d_status_ )
    hipFree( d_status_ );
  if ( h_status_ )
    hipHostFree( h_status_ );
  for ( int q = 0; q < 2; ++q )
  {
    if ( ev_chunk_[ q ] )
      hipEventDestroy( ev_chunk_[ q ] );
    if ( ev_kernel_[ q ] )
      hipEventDestroy( ev_kernel_[ q ] );
  }
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
  }
  exec_n_ = -1;
  exec_failed_ = false;
  if ( d_nets_ )
    hipFree( d_nets_ );
  if ( d_prefix_ )
    hipFree( d_prefix_ );
  if ( h_nets_ )
    hipHostFree( h_nets_ );
  if ( d_status_ )
    hipFree( d_status_ );
  if ( h_status_ )
    hipHostFree( h_status_ );
  hip_check( hipMalloc( &d_nets_, n * sizeof( NetDev ) ), "hipMalloc nets" );
  hip_check( hipMalloc( &d_status_, 2 * n * sizeof( unsigned long long ) ), "hipMalloc status" );
  hip_check( hipHostMalloc( &h_status_, 2 * n * sizeof( unsigned long long ), hipHostMallocDefault ),
    "hipHostMalloc status" );
  hip_check( hipMalloc( &d_prefix_, ( n + 1 ) * sizeof( uint32_t ) ), "hipMalloc prefix" );
  hip_check( hipHostMalloc( &h_nets_, n * sizeof( NetDev ), hipHostMallocDefault ), "hipHostMalloc nets" );
  cap_ = n;
  prefix_.clear();
}

bool
Runner::capture( int n_nets, uint32_t total_blocks, int mask, bool win )
{
  const bool same_key = exec_n_ == n_nets and exec_total_ == total_blocks and exec_mask_ == mask and exec_win_ == win
    and exec_wg_ == wg_per_cu;
  if ( same_key and ( exec_ or exec_failed_ ) )
    return exec_ != nullptr;
  if ( exec_ )
  {
    hipGraphExecDestroy( exec_ );
    exec_ = nullptr;
  }
  // A capture fails when another host thread makes a call HIP does not allow during it (an
  // allocation or a synchronous copy, e.g. while other networks are prepared): the capture is
  // then dropped and the caller launches this launch shape's slices directly; a new launch
  // shape is captured again. Failures are counted (graph_capture_failures in kernel status).
  exec_n_ = n_nets;
  exec_total_ = total_blocks;
  exec_mask_ = mask;
  exec_win_ = win;
  exec_wg_ = wg_per_cu;
  exec_failed_ = true;
  hipGraph_t g = nullptr;
  if ( hipStreamBeginCapture( stream_, hipStreamCaptureModeThreadLocal ) != hipSuccess )
  {
    (void)hipGetLastError();
    ++capture_failures;
    return false;
  }
  for ( int s = 0; s < S_; ++s )
    launch_slice( d_nets_, d_prefix_, n_nets, total_blocks, mask, win, s, slice_lds_for_occupancy( wg_per_cu, win ),
      stream_ );
  const hipError_t ended = hipStreamEndCapture( stream_, &g );
  if ( ended != hipSuccess or g == nullptr or hipGraphInstantiate( &exec_, g, nullptr, nullptr, 0 ) != hipSuccess )
  {
    (void)hipGetLastError();
    if ( g )
      hipGraphDestroy( g );
    exec_ = nullptr;
    ++capture_failures;
    return false;
  }
  hip_check( hipGraphDestroy( g ), "hipGraphDestroy" );
  exec_failed_ = false;
  return true;
}

void
Runner::launch_one( int n, uint32_t total, int mask, bool win, int s, hipEvent_t* ev )
{
  if ( ev )
    hip_check( hipEventRecord( ev[ 0 ], stream_ ), "hipEventRecord" );
  hip_check( launch_slice( d_nets_, d_prefix_, n, total, mask, win, s, slice_lds_for_occupancy( wg_per_cu, win ), stream_ ),
    "launch_slice" );
  if ( ev )
  {
    hip_check( hipEventRecord( ev[ 1 ], stream_ ), "hipEventRecord" );
    hip_check( hipStreamSynchronize( stream_ ), "hipStreamSynchronize" );
    float ms = 0;
    hip_check( hipEventElapsedTime( &ms, ev[ 0 ], ev[ 1 ] ), "hipEventElapsedTime" );
    kernel_ms[ K_SLICE ] += ms;
    kernel_launches[ K_SLICE ] += 1;
  }
}

void
Runner::run( const std::vector< Engine* >& engines, const std::vector< int64_t >& steps )
{
  std::vector< int64_t > remaining = steps;
  last_ms_ = 0;
  while ( run_chunk( engines, remaining ) )
  {
  }
}

// ---- chunk pieces shared by Runner::run_chunk and run_noalloc

namespace
{
// Launch records of one chunk: every instance's clock, first step and steps to do. Returns
// false when no instance has work; prefix, mask, win and split describe the launch.
bool
stage_chunk( const std::vector< Engine* >& engines, const std::vector< int64_t >& remaining, int S, NetDev* h_nets,
  std::vector< uint32_t >& prefix, int& mask, bool& win, bool& split )
{
  const int n = static_cast< int >( engines.size() );
  mask = 0;
  win = false;
  split = false;
  prefix.assign( n + 1, 0 );
  for ( int i = 0; i < n; ++i )
  {
    mask |= engines[ i ]->model_mask();
    // the window kernel also carries the cell cache of cache delivery
    win = win or engines[ i ]->window() or engines[ i ]->cache();
    prefix[ i + 1 ] = prefix[ i ] + engines[ i ]->net().n_blocks;
  }
  if ( prefix[ n ] == 0 )
    return false;
  bool any = false;
  for ( int i = 0; i < n; ++i )
  {
    Engine& E = *engines[ i ];
    NetDev d = E.net();
    d.clock = E.clock;
    d.from0 = E.from_step;
    d.part0 = E.part;
    const int64_t span =
      static_cast< int64_t >( d.min_delay - E.from_step ) + static_cast< int64_t >( S - 1 ) * d.min_delay;
    d.todo = std::min( remaining[ i ], span );
    E.mm_chunk_rows = 0;
    if ( d.mm_int > 0 and d.todo > 0 )
    {
      // sampling stamps off + k * int within the chunk's stamps [first, first + todo); a chunk
      // holds at most mm_rows of them
      const int64_t first = E.clock + E.from_step + 1;
      int64_t s0 = first <= d.mm_off ? d.mm_off : first + ( d.mm_int - ( first - d.mm_off ) % d.mm_int ) % d.mm_int;
      const int64_t last_allowed = s0 + static_cast< int64_t >( d.mm_rows - 1 ) * d.mm_int;
      if ( first + d.todo - 1 > last_allowed )
        d.todo = last_allowed - first + 1;
      d.mm_row0 = ( s0 - d.mm_off ) / d.mm_int;
      const int64_t last = first + d.todo - 1;
      E.mm_chunk_row0 = d.mm_row0;
      E.mm_chunk_rows = s0 > last ? 0 : ( last - s0 ) / d.mm_int + 1;
    }
    d.active = d.todo > 0 ? 1 : 0;
    any = any or d.active;
    split = split or ( d.active and E.from_step > 0 );
    if ( d.part0 > 255 )
      throw std::runtime_error( "NotImplemented: a slice split into more than 256 Run parts" );
    h_nets[ i ] = d;
  }
  return any;
}

// Direct launches of a chunk's slices (no graph); launch(s) enqueues slice s. A slice that
// completes a slice split across Run calls is followed by the receive-order rebuild.
template < typename F >
void
launch_chunk_direct( const NetDev* h_nets, NetDev* d_nets, int n, int S, bool split, hipStream_t st, F&& launch )
{
  for ( int s = 0; s < S; ++s )
  {
    // launches where no instance has work are skipped (their workgroups would exit at once)
    bool active = false;
    for ( int i = 0; i < n and not active; ++i )
      active = slice_phase( h_nets[ i ], s ).active;
    if ( not active )
      break;
    launch( s );
    if ( s == 0 and split )
      for ( int i = 0; i < n; ++i )
      {
        // the slice that continued a previous part completes here: restore NEST's order
        const NetDev& d = h_nets[ i ];
        const SlicePhase ph = slice_phase( d, 0 );
        if ( ph.active and d.from0 > 0 and ph.complete )
          hip_check( launch_flatten( d_nets + i, ph.parity, d.part0 + 1, st ), "launch_flatten" );
      }
  }
}

// After the status read-back: raise device errors, advance every instance's clock and run
// part, and drain spike logs that are half full.
void
finish_chunk( const std::vector< Engine* >& engines, const NetDev* h_nets, const unsigned long long* h_status,
  std::vector< int64_t >& remaining )
{
  const int n = static_cast< int >( engines.size() );
  for ( int i = 0; i < n; ++i )
  {
    Engine& E = *engines[ i ];
    const NetDev& d = h_nets[ i ];
    if ( not d.active )
      continue;
    const unsigned long long flags = h_status[ 2 * i ];
    if ( flags )
    {
      E.clear_error_flags();
      std::string why;
      if ( flags & ERR_AEIF_INSTABILITY )
        why += " NumericalInstability in aeif_cond_exp;";
      if ( flags & ERR_AEIF_STEPS )
        why += " aeif_cond_exp needed more than " + std::to_string( AEIF_MAX_STEPS )
          + " integration steps within one time step;";
      if ( flags & ERR_AEIF_DOUBLE_SPIKE )
        why += " aeif_cond_exp emitted two spikes within one step;";
      if ( flags & ERR_SPIKE_LIST )
        why += " spike list capacity exceeded;";
      if ( flags & ERR_LOG )
        why += " spike log capacity exceeded;";
      if ( flags & ERR_SAMPLES )
        why += " multimeter sample buffer exceeded;";
      if ( flags & ERR_STDP_HISTORY )
        why += " a neuron spiked more than " + std::to_string( d.stdp_hcap )
          + " times between two spikes of a presynaptic stdp_synapse neuron (post-spike history;"
            " raise stdp_history_capacity);";
      if ( flags & ERR_STDP_TWICE )
        why += " a neuron with outgoing stdp_synapse connections spiked twice within one min_delay slice;";
      throw std::runtime_error( "device run failed:" + why );
    }
    const int64_t total_steps = E.from_step + d.todo;
    const int64_t slices_done = total_steps / d.min_delay;
    E.clock += slices_done * d.min_delay;
    const int32_t nf = static_cast< int32_t >( total_steps % d.min_delay );
    if ( nf == 0 )
      E.part = 0;
    else if ( slices_done == 0 )
      E.part += 1;  // another part of the same slice
    else
      E.part = 1;
    E.from_step = nf;
    remaining[ i ] -= d.todo;
    if ( h_status[ 2 * i + 1 ] > E.net().log_cap / 2 )
      E.drain_log();
    E.drain_samples();
  }
}
}  // namespace

bool
Runner::run_chunk( const std::vector< Engine* >& engines, std::vector< int64_t >& remaining )
{
  hip_check( hipSetDevice( device_ ), "hipSetDevice" );
  const int n = static_cast< int >( engines.size() );
  if ( n == 0 )
    return false;
  ensure_capacity( n );
  int mask = 0;
  bool win = false, split = false;
  std::vector< uint32_t > prefix;
  // the window kernel runs instances with a window (at most PMAX_WIN blocks, checked at
  // build) and without one (their block prefix lives in the unused window)
  if ( not stage_chunk( engines, remaining, S_, h_nets_, prefix, mask, win, split ) )
    return false;
  const uint32_t total = prefix[ n ];
  if ( prefix != prefix_ )
  {
    hip_check( hipMemcpy( d_prefix_, prefix.data(), ( n + 1 ) * sizeof( uint32_t ), hipMemcpyHostToDevice ), "prefix" );
    prefix_ = prefix;
  }
  const bool graph = use_graph and not time_kernels and not split and capture( n, total, mask, win );
  ++( graph ? graph_chunks : direct_chunks );
  hip_check( hipMemcpyAsync( d_nets_, h_nets_, n * sizeof( NetDev ), hipMemcpyHostToDevice, stream_ ), "nets H2D" );
  hip_check( hipEventRecord( ev_chunk_[ 0 ], stream_ ), "hipEventRecord" );
  if ( not graph )
    launch_chunk_direct( h_nets_, d_nets_, n, S_, split, stream_,
      [ & ]( int s ) { launch_one( n, total, mask, win, s, time_kernels ? ev_kernel_ : nullptr ); } );
  else
    hip_check( hipGraphLaunch( exec_, stream_ ), "hipGraphLaunch" );
  hip_check( hipEventRecord( ev_chunk_[ 1 ], stream_ ), "hipEventRecord" );
  // error flags and spike-log fill of every instance in one read-back
  hip_check( launch_status( d_nets_, n, d_status_, stream_ ), "launch_status" );
  hip_check( hipMemcpyAsync( h_status_, d_status_, 2 * n * sizeof( unsigned long long ), hipMemcpyDeviceToHost, stream_ ),
    "status D2H" );
  hip_check( hipStreamSynchronize( stream_ ), "hipStreamSynchronize" );
  float ms = 0;
  hipEventElapsedTime( &ms, ev_chunk_[ 0 ], ev_chunk_[ 1 ] );
  last_ms_ += ms;
  finish_chunk( engines, h_nets_, h_status_, remaining );
  return true;
}

// ------------------------------------------------------------------ run without allocation
// The launch records, block prefix and status words of run_noalloc live in this fixed
// device array (each device holds its own copy until the process ends). Host code never
// uses the array's name as an address: it takes the device address from hipGetSymbolAddress
// once per device (noalloc_scratch) and passes that to kernels and copies.
namespace
{
struct NoAllocArea
{
  NetDev nets[ NOALLOC_MAX_NETS ];
  uint32_t prefix[ NOALLOC_MAX_NETS + 1 ];
  uint32_t pad_;
  unsigned long long status[ 2 * NOALLOC_MAX_NETS ];
};
}  // namespace

static __device__ __align__( 256 ) unsigned char scratch[ sizeof( NoAllocArea ) ];

namespace
{
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
NoAllocArea* g_noalloc_ptr[ NOALLOC_MAX_DEVICES ] = {};

// Device address of this device's scratch array; the device must be current.
NoAllocArea*
noalloc_scratch( int device )
{
  if ( device < 0 or device >= NOALLOC_MAX_DEVICES )
    throw std::runtime_error( "run without allocation: device index out of range" );
  std::lock_guard< std::mutex > lock( g_noalloc_mu );
  if ( g_noalloc_ptr[ device ] == nullptr )
  {
    void* p = nullptr;
    hip_check( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ), "hipGetSymbolAddress(scratch)" );
    if ( p == nullptr )
      throw std::runtime_error( "run without allocation: no device address for the scratch array" );
    g_noalloc_ptr[ device ] = static_cast< NoAllocArea* >( p );
  }
  return g_noalloc_ptr[ device ];
}
}  // namespace

void
run_noalloc( int device, const std::vector< Engine* >& engines, const std::vector< int64_t >& steps, int S )
{
  const int n = static_cast< int >( engines.size() );
  if ( n > NOALLOC_MAX_NETS )
    throw std::runtime_error( "run without allocation: " + std::to_string( n ) + " instances, at most "
      + std::to_string( NOALLOC_MAX_NETS ) + " fit the fixed scratch array" );
  if ( S < 1 )
    throw std::runtime_error( "run without allocation: slices per chunk must be at least 1" );
  for ( Engine* E : engines )
    if ( not E->built() )
      throw std::runtime_error( "run without allocation: every kernel must be prepared first (Prepare allocates "
                                "the network's device memory)" );
  if ( n == 0 )
    return;
  hip_check( hipSetDevice( device ), "hipSetDevice" );
  NoAllocArea* area = noalloc_scratch( device );
  NetDev* d_nets = area->nets;  // device addresses inside the scratch array
  uint32_t* d_prefix = area->prefix;
  unsigned long long* d_status = area->status;
  const hipStream_t st = nullptr;  // the device's null stream: nothing to create
  NetDev h_nets[ NOALLOC_MAX_NETS ];
  unsigned long long h_status[ 2 * NOALLOC_MAX_NETS ];
  std::vector< int64_t > remaining = steps;
  std::vector< uint32_t > prefix;
  for ( ;; )
  {
    int mask = 0;
    bool win = false, split = false;
    if ( not stage_chunk( engines, remaining, S, h_nets, prefix, mask, win, split ) )
      break;
    const uint32_t total = prefix[ n ];
    // every word the kernels read is written here for each chunk, so nothing left in the
    // array by an earlier call is used
    hip_check( hipMemcpyAsync( d_prefix, prefix.data(), ( n + 1 ) * sizeof( uint32_t ), hipMemcpyHostToDevice, st ),
      "prefix" );
    hip_check( hipMemcpyAsync( d_nets, h_nets, n * sizeof( NetDev ), hipMemcpyHostToDevice, st ), "nets H2D" );
    launch_chunk_direct( h_nets, d_nets, n, S, split, st,
      [ & ]( int s ) { hip_check( launch_slice( d_nets, d_prefix, n, total, mask, win, s, 0, st ), "launch_slice" ); } );
    hip_check( launch_status( d_nets, n, d_status, st ), "launch_status" );
    hip_check( hipMemcpyAsync( h_status, d_status, 2 * n * sizeof( unsigned long long ), hipMemcpyDeviceToHost, st ),
      "status D2H" );
    hip_check( hipStreamSynchronize( st ), "hipStreamSynchronize" );
    finish_chunk( engines, h_nets, h_status, remaining );
  }
}

}  // namespace nestgpu
