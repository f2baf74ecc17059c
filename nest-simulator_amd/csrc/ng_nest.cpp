// ng_nest.cpp -- host kernel: NEST's Create/Connect/Simulate surface over the device engine.
#include "ng_nest.h"
#include "ng_conngen.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace nestgpu
{

// ------------------------------------------------------------------ time base
void
TimeBase::set_resolution( double ms )
{
  // Time::set_resolution (nest_time.cpp:88-103)
  tics_per_step = static_cast< long >( std::floor( tics_per_ms * ms + 0.5 ) );
  tics_per_step_inv = 1. / static_cast< double >( tics_per_step );
  tics_per_step_rnd = tics_per_step - 1;
  ms_per_step = tics_per_step / tics_per_ms;
  steps_per_ms = 1 / ms_per_step;
}

long
TimeBase::delay_ms_to_steps( double ms ) const
{
  // Time::delay_ms_to_steps = ld_round( ms * STEPS_PER_MS ) (nest_time.h:539-542, numerics.cpp:41-44)
  return static_cast< long >( std::floor( ms * steps_per_ms + 0.5 ) );
}

// ------------------------------------------------------------------ RNG
SeedSeqFe128::SeedSeqFe128( const std::vector< uint32_t >& entropy )
{
  // hash each entropy word into the mixer, cross-mix all ordered pairs, then fold in any
  // remaining words (randutils seed_seq_fe::mix_entropy, count = 4, one round)
  uint32_t hash_const = 0x43b0d7e5u;
  auto hash = [ &hash_const ]( uint32_t value )
  {
    value ^= hash_const;
    hash_const *= 0x931e8875u;
    value *= hash_const;
    value ^= value >> 16;
    return value;
  };
  auto mix = []( uint32_t x, uint32_t y )
  {
    uint32_t result = 0xca01f9ddu * x - 0x4973f715u * y;
    result ^= result >> 16;
    return result;
  };
  size_t next = 0;
  for ( uint32_t& e : mixer )
    e = hash( next < entropy.size() ? entropy[ next++ ] : 0u );
  for ( int a = 0; a < 4; ++a )
    for ( int b = 0; b < 4; ++b )
      if ( a != b )
        mixer[ b ] = mix( mixer[ b ], hash( mixer[ a ] ) );
  for ( ; next < entropy.size(); ++next )
    for ( uint32_t& e : mixer )
      e = mix( e, hash( entropy[ next ] ) );
}

VpRng::VpRng( const std::vector< uint32_t >& seed, int rng_type )
  : type( rng_type )
{
  SeedSeqFe128 sseq( seed );
  if ( type == RNG_PHILOX_32 )
  {
    // r123::Engine's seeding: the key is the first two words of the seed sequence, the
    // counter zero (Random123/array.h seed(), conventional/Engine.hpp)
    uint32_t k[ 2 ];
    sseq.generate( k, k + 2 );
    philox.k0 = k[ 0 ];
    philox.k1 = k[ 1 ];
  }
  else
    engine.seed( sseq );
}

// ------------------------------------------------------------------ parameters
// The per-VP distributions of a normal or lognormal leaf: those of the parameter object its
// "id" names when the kernel has seen it before (same VP count), else new ones.
static std::shared_ptr< std::vector< VpNormal > >
shared_normals( const Json& v, int n_vp, double mean, double sd, ParamStates* states )
{
  std::shared_ptr< std::vector< VpNormal > >* slot = nullptr;
  if ( states and v.has( "id" ) )
  {
    slot = &states->normals[ static_cast< int64_t >( v.at( "id" ).number() ) ];
    if ( *slot and static_cast< int >( ( *slot )->size() ) == n_vp and ( **slot )[ 0 ].mean == mean
         and ( **slot )[ 0 ].sd == sd )
      return *slot;
  }
  auto d = std::make_shared< std::vector< VpNormal > >();
  d->reserve( n_vp );
  for ( int i = 0; i < n_vp; ++i )
    d->emplace_back( mean, sd );
  if ( slot )
    *slot = d;
  return d;
}

std::unique_ptr< Param >
Param::from_json( const Json& j, int n_vp, bool top_level, ParamStates* states )
{
  auto p = std::make_unique< Param >();
  if ( j.kind == Json::NUM or j.kind == Json::BOOL )
  {
    p->op_ = CONST;
    p->a_ = j.number();
    return p;
  }
  if ( j.kind == Json::ARR and top_level )
  {
    p->op_ = ARRAY;
    p->values_.reserve( j.arr.size() );
    for ( const Json& x : j.arr )
    {
      if ( x.kind != Json::NUM )
        throw std::runtime_error( "BadProperty: an array parameter holds numbers only" );
      p->values_.push_back( x.num );
    }
    p->cursor_.resize( n_vp );
    return p;
  }
  if ( j.kind != Json::OBJ or j.obj.size() != 1 )
    throw std::runtime_error( "BadProperty: a parameter is a number or a one-key object" );
  const std::string& k = j.obj[ 0 ].first;
  const Json& v = j.obj[ 0 ].second;
  if ( k == "uniform" )
  {
    // UniformParameter (parameter.h:180-221)
    p->op_ = UNIFORM;
    p->a_ = v.has( "min" ) ? v.at( "min" ).number() : 0.0;
    double mx = v.has( "max" ) ? v.at( "max" ).number() : 1.0;
    if ( p->a_ >= mx )
      throw std::runtime_error( "BadProperty: nest::UniformParameter: min < max required." );
    p->b_ = mx - p->a_;
  }
  else if ( k == "uniform_int" )
  {
    p->op_ = UNIFORM_INT;
    p->a_ = v.has( "max" ) ? v.at( "max" ).number() : 1.0;
    if ( p->a_ <= 0 )
      throw std::runtime_error( "BadProperty: nest::UniformIntParameter: max > 0 required." );
  }
  else if ( k == "normal" )
  {
    p->op_ = NORMAL;
    p->a_ = v.has( "mean" ) ? v.at( "mean" ).number() : 0.0;
    p->b_ = v.has( "std" ) ? v.at( "std" ).number() : 1.0;
    if ( p->b_ < 0 )
      throw std::runtime_error( "BadProperty: nest::NormalParameter: std >= 0 required." );
    p->normals_ = shared_normals( v, n_vp, p->a_, p->b_, states );
  }
  else if ( k == "lognormal" )
  {
    // LognormalParameter (parameter.cpp:120-157): std::lognormal_distribution, i.e.
    // exp(std * z + mean) with z from its own standard normal draw
    p->op_ = LOGNORMAL;
    p->a_ = v.has( "mean" ) ? v.at( "mean" ).number() : 0.0;
    p->b_ = v.has( "std" ) ? v.at( "std" ).number() : 1.0;
    if ( p->b_ < 0 )
      throw std::runtime_error( "BadProperty: nest::LognormalParameter: std ≥ 0 required." );
    p->normals_ = shared_normals( v, n_vp, 0.0, 1.0, states );
  }
  else if ( k == "exponential" )
  {
    // ExponentialParameter (parameter.h:313-346)
    p->op_ = EXPONENTIAL;
    p->a_ = v.has( "beta" ) ? v.at( "beta" ).number() : 1.0;
    if ( p->a_ < 0 )
      throw std::runtime_error( "BadProperty: nest::ExponentialParameter: beta ≥ 0 required." );
  }
  else if ( k == "exp" or k == "sin" or k == "cos" )
  {
    // ExpParameter, SinParameter, CosParameter (parameter.h:991,1037,1082)
    p->op_ = k == "exp" ? EXP : k == "sin" ? SIN : COS;
    p->l_ = from_json( v, n_vp, false, states );
  }
  else if ( k == "compare" )
  {
    // ComparingParameter (parameter.h:680-752): 1 or 0 for a < b, <=, ==, !=, >=, > (0..5)
    if ( v.kind != Json::ARR or v.arr.size() != 3 )
      throw std::runtime_error( "BadProperty: compare takes two parameters and a comparator" );
    p->op_ = COMPARE;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->r_ = from_json( v.arr[ 1 ], n_vp, false, states );
    p->a_ = v.arr[ 2 ].number();
    if ( p->a_ < 0 or p->a_ > 5 or p->a_ != std::floor( p->a_ ) )
      throw std::runtime_error( "BadProperty: Comparator specification has to be in the range 0-5." );
  }
  else if ( k == "conditional" )
  {
    // ConditionalParameter (parameter.h:760-820): only the chosen branch is evaluated
    if ( v.kind != Json::ARR or v.arr.size() != 3 )
      throw std::runtime_error( "BadProperty: conditional takes a condition and two parameters" );
    p->op_ = CONDITIONAL;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->r_ = from_json( v.arr[ 1 ], n_vp, false, states );
    p->f_ = from_json( v.arr[ 2 ], n_vp, false, states );
  }
  else if ( k == "pow" )
  {
    // PowParameter (parameter.h:1130): a parameter to a fixed exponent
    if ( v.kind != Json::ARR or v.arr.size() != 2 )
      throw std::runtime_error( "BadProperty: pow takes a parameter and an exponent" );
    p->op_ = POW;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->a_ = v.arr[ 1 ].number();
  }
  else if ( k == "sum" or k == "product" or k == "quotient" )
  {
    // SumParameter, ProductParameter, QuotientParameter (parameter.h:496,546 and the like)
    if ( v.kind != Json::ARR or v.arr.size() != 2 )
      throw std::runtime_error( "BadProperty: sum/product/quotient take two operands" );
    p->op_ = k == "sum" ? SUM : k == "product" ? PRODUCT : QUOTIENT;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->r_ = from_json( v.arr[ 1 ], n_vp, false, states );
  }
  else if ( k == "max" or k == "min" )
  {
    if ( v.kind != Json::ARR or v.arr.size() != 2 )
      throw std::runtime_error( "BadProperty: max/min take a parameter and a value" );
    p->op_ = k == "max" ? MAX : MIN;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->a_ = v.arr[ 1 ].number();
  }
  else if ( k == "redraw" )
  {
    if ( v.kind != Json::ARR or v.arr.size() != 3 )
      throw std::runtime_error( "BadProperty: redraw takes a parameter, min and max" );
    p->op_ = REDRAW;
    p->l_ = from_json( v.arr[ 0 ], n_vp, false, states );
    p->a_ = v.arr[ 1 ].number();
    p->b_ = v.arr[ 2 ].number();
  }
  else
    throw std::runtime_error( "BadProperty: unknown parameter type '" + k + "'" );
  return p;
}

double
Param::value( VpRng& rng, int vp )
{
  switch ( op_ )
  {
  case CONST:
    return a_;
  case UNIFORM:
    return a_ + rng.drand() * b_;
  case UNIFORM_INT:
    return static_cast< double >( rng.ulrand( static_cast< unsigned long >( a_ ) ) );
  case NORMAL:
    if ( b_ == 0 )
      return a_;
    return ( *normals_ )[ vp ]( rng );
  case SUM:
  {
    const double x = l_->value( rng, vp );
    return x + r_->value( rng, vp );
  }
  case PRODUCT:
  {
    const double x = l_->value( rng, vp );
    return x * r_->value( rng, vp );
  }
  case MAX:
    return std::max( l_->value( rng, vp ), a_ );
  case MIN:
    return std::min( l_->value( rng, vp ), a_ );
  case EXP:
    return std::exp( l_->value( rng, vp ) );
  case SIN:
    return std::sin( l_->value( rng, vp ) );
  case COS:
    return std::cos( l_->value( rng, vp ) );
  case QUOTIENT:
  {
    const double x = l_->value( rng, vp );
    return x / r_->value( rng, vp );
  }
  case POW:
    return std::pow( l_->value( rng, vp ), a_ );
  case COMPARE:
  {
    // ComparingParameter::value (parameter.h:712-715) passes both operand values as the
    // arguments of one call, compare_( p1->value(), p2->value() ). g++ on x86-64 evaluates
    // call arguments right to left, so the reference draws the right operand first.
    const double y = r_->value( rng, vp );
    const double x = l_->value( rng, vp );
    switch ( static_cast< int >( a_ ) )
    {
    case 0:
      return x < y;
    case 1:
      return x <= y;
    case 2:
      return x == y;
    case 3:
      return x != y;
    case 4:
      return x >= y;
    default:
      return x > y;
    }
  }
  case CONDITIONAL:
    return l_->value( rng, vp ) != 0 ? r_->value( rng, vp ) : f_->value( rng, vp );
  case LOGNORMAL:
    if ( b_ == 0 )
      return std::exp( a_ );
    return std::exp( b_ * ( *normals_ )[ vp ]( rng ) + a_ );
  case EXPONENTIAL:
    return a_ * ( -std::log( 1 - rng.drand() ) );
  case ARRAY:
  {
    size_t& i = cursor_[ vp ].next;
    if ( i >= values_.size() )
      throw std::runtime_error( "KernelException: Parameter values exhausted." );
    return values_[ i++ ];
  }
  case REDRAW:
    for ( int i = 0; i < 1000; ++i )
    {
      const double x = l_->value( rng, vp );
      if ( a_ <= x and x <= b_ )
        return x;
    }
    throw std::runtime_error( "KernelException: Number of redraws exceeded limit." );
  }
  return 0;
}

void
Param::skip( int vp, size_t n )
{
  if ( op_ != ARRAY )
    return;
  size_t& i = cursor_[ vp ].next;
  if ( i >= values_.size() )
    throw std::runtime_error( "KernelException: Parameter values exhausted." );
  i += n;
}

// ------------------------------------------------------------------ Poisson drive
PoissonTable
make_poisson_table( double lambda, uint64_t seed )
{
  PoissonTable t {};
  t.seed = seed;
  double p = std::exp( -lambda );
  double c = 0;
  int k = 0;
  for ( ; k < 63; ++k )
  {
    c += p;
    t.cdf[ k ] = c;
    if ( 1.0 - c < 1e-17 )
      break;
    p *= lambda / ( k + 1 );
  }
  t.kmax = k;
  for ( int q = k + 1; q < 64; ++q )
    t.cdf[ q ] = 2.0;
  if ( lambda > 20.0 )
    throw std::runtime_error( "BadProperty: Poisson drive rate per step must be <= 20" );
  return t;
}

static inline uint32_t
poisson_count_host( const PoissonTable& T, uint64_t stream, int64_t step )
{
  uint64_t z = T.seed + stream * 0x9E3779B97F4A7C15ull + static_cast< uint64_t >( step ) * 0xD1B54A32D192ED03ull;
  z = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
  z = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
  z = z ^ ( z >> 31 );
  const double u = static_cast< double >( z >> 11 ) * ( 1.0 / 9007199254740992.0 );
  uint32_t k = 0;
  while ( static_cast< int32_t >( k ) < T.kmax and u >= T.cdf[ k ] )
    ++k;
  return k;
}

void
poisson_counts( uint64_t seed, double lambda, uint64_t stream0, int64_t n, int64_t step0, int64_t n_steps, uint8_t* out )
{
  const PoissonTable T = make_poisson_table( lambda, seed );
  const int nt = static_cast< int >( std::max< int64_t >( 1, std::min< int64_t >( 16, n_steps / 64 ) ) );
  std::vector< std::thread > th;
  for ( int t = 0; t < nt; ++t )
    th.emplace_back(
      [ & ]( int tid )
      {
        for ( int64_t s = tid; s < n_steps; s += nt )
          for ( int64_t i = 0; i < n; ++i )
            out[ s * n + i ] = static_cast< uint8_t >( poisson_count_host( T, stream0 + i, step0 + s ) );
      },
      t );
  for ( auto& x : th )
    x.join();
}

// ------------------------------------------------------------------ kernel
Kernel::Kernel( int device )
  : device_( device )
{
  reset();
}

Kernel::~Kernel() = default;

void
Kernel::reset()
{
  tb_ = TimeBase();
  n_vp_ = 1;
  seed_ = 143202461;
  rng_type_ = RNG_MT19937_64;
  rngs_.clear();
  rng_dirty_ = true;
  nodes_.clear();
  params_.clear();
  state_.clear();
  neuron_gid_.clear();
  gens_.clear();
  gen_gid_.clear();
  dense_.clear();
  recorders_.clear();
  rec_window_.clear();
  rec_gid_.clear();
  meters_.clear();
  meter_gid_.clear();
  mm_col_gid_.clear();
  rec_events_.clear();
  conns_.assign( 1, ConnList() );
  dev_conns_.clear();
  d_min_seen_ = -1;
  d_max_seen_ = 1;
  min_delay_ = max_delay_ = 1;
  simulated_ = prepared_ = false;
  store_dirty_ = true;
  state_dirty_ = false;
  engine_.reset();
  gid_to_dev_.clear();
  dev_to_gid_.clear();
  rec_events_.clear();
  spike_count_ = 0;
  targets_per_block_ = 0;
  balanced_blocks_ = true;
  connect_on_device_ = false;
  device_connect_calls_ = 0;
  block_start_.clear();
  rebalanced_ = false;
  rx_steps_ = 0;
  block_relayouts_ = 0;
  stdp_sets_.clear();
  syn_pos_.clear();
  delivery_mode_ = 0;
  wg_per_cu_ = 0;
  param_states_.normals.clear();
  graph_launch_ = true;
  phase_timing_ = false;
  graph_chunks_ = direct_chunks_ = capture_failures_ = 0;
  window_slots_ = 0;
  stdp_history_ = STDP_HIST_DEFAULT;
  delivery_passes_ = 0;
  keep_host_connections_ = true;
  host_connections_freed_ = false;
  released_ = false;
  released_events_ = released_spikes_ = 0;
}

void
Kernel::init_rngs()
{
  // RandomManager::initialize (random_manager.cpp:63-100): vp-specific engines seeded with
  // seed_seq_fe128{ base_seed, THREAD_SPECIFIC_SEEDER_, vp }
  rngs_.clear();
  for ( int vp = 0; vp < n_vp_; ++vp )
    rngs_.emplace_back( std::vector< uint32_t > { seed_, 0xb84c9baeu, static_cast< uint32_t >( vp ) }, rng_type_ );
  rng_dirty_ = false;
}

void
Kernel::set_kernel_status( const Json& d )
{
  for ( const auto& kv : d.obj )
  {
    const std::string& k = kv.first;
    if ( k == "resolution" )
    {
      // SimulationManager::set_status (simulation_manager.cpp:183-266)
      if ( not nodes_.empty() )
        throw std::runtime_error(
          "KernelException: Nodes have already been created: time representation cannot be changed." );
      const double r = kv.second.number();
      if ( not( r >= tb_.ms_per_tic ) )
        throw std::runtime_error( "KernelException: Resolution must be greater than or equal to one tic. Value unchanged." );
      // is_integer (libnestutil/numerics.cpp:62-76): within two bits of rounding of an integer
      const double n_tics = r / tb_.ms_per_tic;
      double int_part;
      double frac_part = std::modf( n_tics, &int_part );
      if ( frac_part > 0.5 )
        frac_part = 1 - frac_part;
      if ( not( frac_part < 4 * n_tics * std::numeric_limits< double >::epsilon() ) )
        throw std::runtime_error( "KernelException: Resolution must be a multiple of the tic length. Value unchanged." );
      tb_.set_resolution( r );
    }
    else if ( k == "rng_seed" )
    {
      const double s = kv.second.number();
      if ( s < 1 or s > 4294967295.0 )
        throw std::runtime_error( "BadProperty: RNG seed must be in (0, 2^32-1]." );
      seed_ = static_cast< uint32_t >( s );
      rng_dirty_ = true;
    }
    else if ( k == "total_num_virtual_procs" or k == "local_num_threads" )
    {
      if ( not nodes_.empty() )
        throw std::runtime_error( "KernelException: Nodes exist: number of threads cannot be changed." );
      const int n = static_cast< int >( kv.second.number() );
      if ( n < 1 )
        throw std::runtime_error( "BadProperty: number of virtual processes must be positive." );
      n_vp_ = n;
      conns_.assign( n_vp_, ConnList() );
      rng_dirty_ = true;
    }
    else if ( k == "slices_per_chunk" )
    {
      slices_per_chunk_ = std::max( 1, static_cast< int >( kv.second.number() ) );
      own_runner_.reset();
    }
    else if ( k == "targets_per_block" )
    {
      const int v = static_cast< int >( kv.second.number() );
      if ( v < 0 or v > static_cast< int >( MAX_TB ) )
        throw std::runtime_error(
          "BadProperty: targets_per_block must be in [0, " + std::to_string( MAX_TB ) + "] (0 = automatic)" );
      if ( simulated_ )
        throw std::runtime_error( "KernelException: targets_per_block cannot change after Simulate" );
      targets_per_block_ = v;
      store_dirty_ = true;
    }
    else if ( k == "connect_on_device" )
      connect_on_device_ = kv.second.number() != 0;
    else if ( k == "balanced_blocks" )
    {
      const bool v = kv.second.number() != 0;
      if ( simulated_ and v != balanced_blocks_ )
        throw std::runtime_error( "KernelException: balanced_blocks cannot change after Simulate" );
      balanced_blocks_ = v;
      store_dirty_ = true;
    }
    else if ( k == "delivery_mode" )
    {
      const std::string v = kv.second.string();
      const int m = v == "auto" ? 0 : v == "window" ? 1 : v == "global" ? 2 : v == "cache" ? 3 : -1;
      if ( m < 0 )
        throw std::runtime_error( "BadProperty: delivery_mode must be 'auto', 'window', 'global' or 'cache'" );
      if ( simulated_ and m != delivery_mode_ )
        throw std::runtime_error( "KernelException: delivery_mode cannot change after Simulate" );
      delivery_mode_ = m;
      store_dirty_ = true;
    }
    else if ( k == "delivery_passes" )
    {
      const int v = static_cast< int >( kv.second.number() );
      if ( v < 0 or v > MAX_PASSES )
        throw std::runtime_error(
          "BadProperty: delivery_passes must be in [0, " + std::to_string( MAX_PASSES ) + "] (0 or 1 = one pass)" );
      if ( simulated_ and v != delivery_passes_ )
        throw std::runtime_error( "KernelException: delivery_passes cannot change after Simulate" );
      delivery_passes_ = v;
      store_dirty_ = true;
    }
    else if ( k == "window_slots" )
    {
      const int v = static_cast< int >( kv.second.number() );
      if ( v < 0 )
        throw std::runtime_error( "BadProperty: window_slots must be >= 0 (0 = as many as fit)" );
      if ( simulated_ and v != window_slots_ )
        throw std::runtime_error( "KernelException: window_slots cannot change after Simulate" );
      window_slots_ = v;
      store_dirty_ = true;
    }
    else if ( k == "workgroups_per_cu" )
    {
      const int v = static_cast< int >( kv.second.number() );
      if ( v < 0 or v > 8 )
        throw std::runtime_error( "BadProperty: workgroups_per_cu must be in [0, 8] (0 = as many as fit)" );
      wg_per_cu_ = v;
    }
    else if ( k == "stdp_history_capacity" )
    {
      // entries of the post-spike history ring kept per neuron; NEST keeps an unbounded
      // history (ArchivingNode::history_), a run that needs an entry the ring dropped fails
      const double v = kv.second.number();
      const uint32_t u = static_cast< uint32_t >( v );
      if ( not( v >= 2 and v <= STDP_HIST_MAX ) or static_cast< double >( u ) != v or ( u & ( u - 1 ) ) )
        throw std::runtime_error( "BadProperty: stdp_history_capacity must be a power of two in [2, 65536]" );
      if ( simulated_ and u != stdp_history_ )
        throw std::runtime_error( "KernelException: stdp_history_capacity cannot change after Simulate" );
      stdp_history_ = u;
      store_dirty_ = true;
    }
    else if ( k == "keep_host_connections" )
      keep_host_connections_ = kv.second.number() != 0;
    else if ( k == "kernel_timing" )
    {
      time_kernels_ = kv.second.number() != 0;
      if ( own_runner_ )
        own_runner_->time_kernels = time_kernels_;
    }
    else if ( k == "graph_launch" )
      graph_launch_ = kv.second.number() != 0;
    else if ( k == "phase_timing" )
    {
      // slice workgroups time their delivery phase and their whole work (100 MHz clock);
      // switching it on clears the sums
      phase_timing_ = kv.second.number() != 0;
      if ( engine_ and engine_->built() )
      {
        engine_->net().phase_timing = phase_timing_ ? 1 : 0;
        if ( phase_timing_ )
          engine_->clear_phase_ticks();
      }
    }
    else if ( k == "rng_type" )
    {
      // RandomManager::set_status (random_manager.cpp:160-185): the streams restart
      const std::string v = kv.second.string();
      if ( v != "mt19937_64" and v != "Philox_32" )
        throw std::runtime_error( "BadProperty: rng_type must be 'mt19937_64' or 'Philox_32'." );
      rng_type_ = v == "Philox_32" ? RNG_PHILOX_32 : RNG_MT19937_64;
      rng_dirty_ = true;
    }
    else if ( k == "print_time" or k == "overwrite_files" or k == "data_path" )
      continue;
    else if ( k == "biological_time" )
    {
      // SimulationManager::set_status (simulation_manager.cpp:151-158)
      if ( kv.second.number() != 0.0 )
        throw std::runtime_error( "BadProperty: The simulation time can only be set to 0.0." );
      if ( simulated_ and engine_ and engine_->clock + engine_->from_step > 0 )
        throw std::runtime_error( "NotImplemented: resetting the simulation time after Simulate on the device path." );
    }
    else
      throw std::runtime_error( "UnaccessedDictionaryEntry: kernel status key '" + k + "'" );
  }
}

std::string
Kernel::get_kernel_status_json() const
{
  std::ostringstream o;
  o.precision( 17 );
  const int64_t clk = engine_ ? engine_->clock : 0;
  const int32_t fr = engine_ ? engine_->from_step : 0;
  // until the first simulation the delay extrema are those of the connections made so far,
  // the resolution when there are none (update_delay_extrema_, connection_manager.cpp:485-527)
  const long dmin = simulated_ ? min_delay_ : ( d_min_seen_ < 0 ? 1 : d_min_seen_ );
  const long dmax = simulated_ ? max_delay_ : std::max( d_max_seen_, 1L );
  o << "{\"resolution\": " << tb_.resolution_ms() << ", \"tics_per_step\": " << tb_.tics_per_step
    << ", \"rng_seed\": " << seed_ << ", \"rng_type\": \"" << ( rng_type_ == RNG_PHILOX_32 ? "Philox_32" : "mt19937_64" )
    << "\", \"rng_types\": [\"mt19937_64\", \"Philox_32\"]";
  // ModelManager::get_status (model_manager.cpp:116-131): the registered neuron models, then
  // the devices; the synapse models
  o << ", \"node_models\": [";
  for ( const std::string& nm : ModelManager::instance().names() )
    o << "\"" << nm << "\", ";
  o << "\"spike_generator\", \"spike_recorder\", \"multimeter\", \"voltmeter\"]"
    << ", \"synapse_models\": [\"static_synapse\", \"stdp_synapse\"]"
    << ", \"total_num_virtual_procs\": " << n_vp_
    << ", \"local_num_threads\": " << n_vp_ << ", \"min_delay\": " << dmin * tb_.ms_per_step
    << ", \"max_delay\": " << dmax * tb_.ms_per_step << ", \"min_delay_steps\": " << dmin
    << ", \"max_delay_steps\": " << dmax << ", \"num_connections\": " << num_connections()
    << ", \"network_size\": " << nodes_.size() << ", \"biological_time\": " << ( clk + fr ) * tb_.ms_per_step
    << ", \"clock_steps\": " << clk << ", \"from_step\": " << fr << ", \"local_spike_counter\": " << local_spike_counter()
    << ", \"time_simulate_kernel_ms\": " << last_kernel_ms_ << ", \"time_last_run_ms\": " << last_wall_ms_
    << ", \"time_simulate\": " << time_simulate_ms_ * 1e-3
    // delivery, update and spike gathering run fused in one slice kernel: its time is
    // reported as time_update and included in time_deliver_spike_data's place
    << ", \"time_update\": " << time_kernel_ms_[ K_SLICE ] * 1e-3
    << ", \"time_gather_spike_data\": " << time_kernel_ms_[ K_FLATTEN ] * 1e-3
    << ", \"time_deliver_spike_data\": 0, \"fused_deliver_update\": true"
    << ", \"targets_per_block\": " << ( engine_ and engine_->built() ? static_cast< int64_t >( engine_->net().tb ) : targets_per_block_ )
    << ", \"balanced_blocks\": " << ( balanced_blocks_ ? "true" : "false" )
    << ", \"connect_on_device\": " << ( connect_on_device_ ? "true" : "false" )
    << ", \"device_connect_calls\": " << device_connect_calls_
    << ", \"block_relayouts\": " << block_relayouts_
    << ", \"target_blocks\": " << ( engine_ and engine_->built() ? static_cast< int64_t >( engine_->net().n_blocks ) : 0 )
    << ", \"delivery_mode\": \""
    << ( delivery_mode_ == 0 ? "auto" : delivery_mode_ == 1 ? "window" : delivery_mode_ == 2 ? "global" : "cache" ) << "\""
    << ", \"window_delivery\": " << ( engine_ and engine_->built() and engine_->window() ? "true" : "false" )
    << ", \"cache_delivery\": " << ( engine_ and engine_->built() and engine_->cache() ? "true" : "false" )
    << ", \"delivery_passes\": " << ( engine_ and engine_->built() ? engine_->net().np : delivery_passes_ )
    << ", \"window_slots\": " << ( engine_ and engine_->built() ? engine_->net().wn : window_slots_ )
    << ", \"stdp_history_capacity\": " << stdp_history_
    << ", \"workgroups_per_cu\": " << wg_per_cu_ << ", \"kernel_timing\": " << ( time_kernels_ ? "true" : "false" )
    << ", \"graph_launch\": " << ( graph_launch_ ? "true" : "false" ) << ", \"graph_chunks\": " << graph_chunks_
    << ", \"direct_chunks\": " << direct_chunks_ << ", \"graph_capture_failures\": " << capture_failures_
    << ", \"kernel_ms\": {";
  static const char* kn[] = { "slice", "flatten" };
  for ( int q = 0; q < N_KINDS; ++q )
    o << ( q ? ", " : "" ) << "\"" << kn[ q ] << "\": " << time_kernel_ms_[ q ];
  o << "}, \"kernel_launches\": {";
  for ( int q = 0; q < N_KINDS; ++q )
    o << ( q ? ", " : "" ) << "\"" << kn[ q ] << "\": " << kernel_launches_[ q ];
  unsigned long long cn[ N_COUNTERS ] = {};
  if ( engine_ and engine_->built() and not released_ )
    engine_->read_counters( cn );
  o << "}, \"phase_timing\": " << ( phase_timing_ ? "true" : "false" ) << ", \"phase_ticks\": {\"deliver\": " << cn[ 4 ]
    << ", \"total\": " << cn[ 5 ] << ", \"workgroups\": " << cn[ 6 ] << ", \"clock_hz\": 100000000}";
  o << ", \"synaptic_events\": " << synaptic_events()
    << ", \"device_bytes\": " << ( engine_ ? engine_->device_bytes() : 0 )
    << ", \"device_bytes_estimate\": " << estimate_device_bytes() << "}";
  return o.str();
}

Kernel::NodeRec&
Kernel::node( int64_t gid )
{
  if ( gid < 1 or gid > static_cast< int64_t >( nodes_.size() ) )
    throw std::runtime_error( "UnknownNode: node id " + std::to_string( gid ) );
  return nodes_[ gid - 1 ];
}

const Kernel::NodeRec&
Kernel::node( int64_t gid ) const
{
  if ( gid < 1 or gid > static_cast< int64_t >( nodes_.size() ) )
    throw std::runtime_error( "UnknownNode: node id " + std::to_string( gid ) );
  return nodes_[ gid - 1 ];
}

std::string
Kernel::model_name( int64_t gid ) const
{
  const NodeRec& r = node( gid );
  if ( r.kind == 1 )
    return "spike_generator";
  if ( r.kind == 2 )
    return "spike_recorder";
  if ( r.kind == 3 )
    return meters_[ r.idx ].voltmeter ? "voltmeter" : "multimeter";
  return ModelManager::instance().get( params_[ r.idx ].model ).name();
}

// ArchivingNode::set_status (archiving_node.cpp:238-257) for every neuron, then the model's
// Parameters_::set and State_::set; a key that neither reads fails the call
void
Kernel::apply_params( NeuronParams& p, NeuronState& s, const Json& d, bool )
{
  std::set< std::string > used;
  // the trace time constant of stdp synapses onto this neuron, checked first and set once the
  // model's own checks pass
  double tau_minus = p.tau_minus;
  if ( d.has( "tau_minus" ) )
  {
    used.insert( "tau_minus" );
    tau_minus = d.at( "tau_minus" ).number();
    if ( tau_minus <= 0.0 )
      throw std::runtime_error( "BadProperty: All time constants must be strictly positive." );
  }
  ModelManager::instance().get( p.model ).set_status( p, s, d, used );
  for ( const auto& kv : d.obj )
    if ( not used.count( kv.first ) )
      throw std::runtime_error( "UnaccessedDictionaryEntry: '" + kv.first + "' is not a parameter of this model" );
  p.tau_minus = tau_minus;
}

int64_t
Kernel::create( const std::string& model, int64_t n, const Json& params )
{
  if ( params.kind != Json::OBJ and params.kind != Json::NUL )
    throw std::runtime_error( "BadProperty: node parameters must be a dictionary" );
  if ( n < 1 )
    throw std::runtime_error( "BadProperty: Create requires n >= 1." );
  // recording devices may join after Simulate (they change what is read back, not the network
  // on the device); neurons and generators may not
  if ( simulated_ and model != "spike_recorder" and model != "multimeter" and model != "voltmeter" )
    throw std::runtime_error( "NotImplemented: creating nodes after Simulate is not supported on the device path." );
  const int64_t first = static_cast< int64_t >( nodes_.size() ) + 1;
  if ( const NodeModel* M = ModelManager::instance().find( model ) )
  {
    NeuronParams p {};
    NeuronState s {};
    p.model = M->id();
    p.tau_minus = 20.0;  // ArchivingNode (archiving_node.cpp:40)
    M->defaults( p, s );
    apply_params( p, s, params, true );
    for ( int64_t i = 0; i < n; ++i )
    {
      nodes_.push_back( NodeRec { 0, static_cast< uint32_t >( params_.size() ) } );
      params_.push_back( p );
      state_.push_back( s );
      neuron_gid_.push_back( first + i );
    }
  }
  else if ( model == "spike_generator" )
  {
    for ( int64_t i = 0; i < n; ++i )
    {
      nodes_.push_back( NodeRec { 1, static_cast< uint32_t >( gens_.size() ) } );
      gens_.emplace_back();
      gen_gid_.push_back( first + i );
    }
    if ( not params.obj.empty() )
    {
      std::vector< int64_t > g( n );
      for ( int64_t i = 0; i < n; ++i )
        g[ i ] = first + i;
      set_status( g, params );
    }
  }
  else if ( model == "spike_recorder" )
  {
    for ( int64_t i = 0; i < n; ++i )
    {
      nodes_.push_back( NodeRec { 2, static_cast< uint32_t >( recorders_.size() ) } );
      recorders_.emplace_back();
      rec_window_.emplace_back();
      rec_events_.emplace_back();
      rec_gid_.push_back( first + i );
    }
    if ( params.kind == Json::OBJ and not params.obj.empty() )
    {
      std::vector< int64_t > g( n );
      for ( int64_t i = 0; i < n; ++i )
        g[ i ] = first + i;
      set_status( g, params );
    }
  }
  else if ( model == "multimeter" or model == "voltmeter" )
  {
    // multimeter::Parameters_ (models/multimeter.cpp:37-60): interval 1 ms, offset 0; a
    // voltmeter records V_m (:227-233)
    for ( int64_t i = 0; i < n; ++i )
    {
      nodes_.push_back( NodeRec { 3, static_cast< uint32_t >( meters_.size() ) } );
      Meter m;
      m.interval = tb_.ms_to_steps( 1.0 );
      m.voltmeter = model == "voltmeter";
      if ( m.voltmeter )
        m.record_from = { "V_m" };
      meters_.push_back( m );
      meter_gid_.push_back( first + i );
    }
    if ( params.kind == Json::OBJ and not params.obj.empty() )
    {
      std::vector< int64_t > g( n );
      for ( int64_t i = 0; i < n; ++i )
        g[ i ] = first + i;
      set_status( g, params );
    }
  }
  else
    throw std::runtime_error( "UnknownModelName: '" + model + "' is not available on the device path." );
  store_dirty_ = true;
  return first;
}

void
Kernel::set_status( const std::vector< int64_t >& gids, const Json& d )
{
  if ( d.kind != Json::OBJ )
    throw std::runtime_error( "BadProperty: node parameters must be a dictionary" );
  for ( int64_t g : gids )
  {
    NodeRec& r = node( g );
    if ( r.kind == 0 )
    {
      if ( simulated_ )
      {
        // after Simulate only the membrane potential can be set (it lives on the device)
        if ( d.obj.size() != 1 or not d.has( "V_m" ) )
          throw std::runtime_error( "NotImplemented: only V_m can be changed after Simulate on the device path." );
        const double v = d.at( "V_m" ).number();
        set_values( { g }, "V_m", &v );
        continue;
      }
      NeuronParams p = params_[ r.idx ];
      NeuronState s = state_[ r.idx ];
      apply_params( p, s, d, false );
      params_[ r.idx ] = p;
      state_[ r.idx ] = s;
      state_dirty_ = true;
      store_dirty_ = true;
    }
    else if ( r.kind == 1 )
    {
      // set on a copy, kept only when every check passes (Node::set_status)
      SpikeGen G = gens_[ r.idx ];
      for ( const auto& kv : d.obj )
      {
        const std::string& k = kv.first;
        if ( update_window( G.window, k, kv.second ) )
          continue;
        if ( k == "spike_times" )
        {
          G.stamps.clear();
          for ( const Json& t : kv.second.arr )
          {
            // spike_generator::Parameters_::assert_valid_spike_time_and_insert_ (spike_generator.cpp:85-140)
            const double tm = t.number();
            if ( tm == 0.0 )
              throw std::runtime_error( "BadProperty: spike time cannot be set to 0." );
            const long tics = tb_.ms_to_tics( tm );
            if ( tics % tb_.tics_per_step != 0 )
              throw std::runtime_error(
                "BadProperty: spike_generator: Time point " + std::to_string( tm ) + " is not representable in current resolution." );
            G.stamps.push_back( tics / tb_.tics_per_step );
          }
          for ( size_t i = 1; i < G.stamps.size(); ++i )
            if ( G.stamps[ i ] < G.stamps[ i - 1 ] )
              throw std::runtime_error( "BadProperty: spike times must be sorted in non-descending order." );
        }
        else if ( k == "spike_multiplicities" )
        {
          G.mults.clear();
          for ( const Json& m : kv.second.arr )
            G.mults.push_back( static_cast< int64_t >( m.number() ) );
        }
        else if ( k == "poisson_rate" or k == "poisson_seed" or k == "poisson_start" or k == "poisson_stop" )
          continue;
        else if ( k == "precise_times" or k == "allow_offgrid_times" or k == "shift_now_spikes" )
        {
          if ( kv.second.number() != 0 )
            throw std::runtime_error( "NotImplemented: off-grid spike_generator options are not supported." );
        }
        else
          throw std::runtime_error( "UnaccessedDictionaryEntry: '" + k + "' is not a spike_generator parameter" );
      }
      if ( d.has( "poisson_rate" ) )
      {
        // seeded Poisson spike train: counts per step ~ Poisson(rate * h), pre-generated
        const double rate = d.at( "poisson_rate" ).number();
        const double start = d.has( "poisson_start" ) ? d.at( "poisson_start" ).number() : 0.0;
        const double stop = d.has( "poisson_stop" ) ? d.at( "poisson_stop" ).number() : 1e9;
        G.poisson_lambda = rate * tb_.resolution_ms() * 1e-3;
        G.poisson_seed = d.has( "poisson_seed" ) ? static_cast< uint64_t >( d.at( "poisson_seed" ).number() ) : 1;
        G.poisson_t0 = tb_.ms_to_steps( start );
        G.poisson_steps = tb_.ms_to_steps( stop ) - G.poisson_t0;
      }
      if ( not G.mults.empty() and G.mults.size() != G.stamps.size() )
        throw std::runtime_error( "BadProperty: spike_multiplicities must have the same length as spike_times." );
      if ( not G.window.stop_inf and G.window.stop < G.window.start )
        throw std::runtime_error( "BadProperty: stop >= start required." );
      gens_[ r.idx ] = std::move( G );
      store_dirty_ = true;
    }
    else if ( r.kind == 3 )
    {
      Meter m = meters_[ r.idx ];
      set_meter_status( m, d );
      meters_[ r.idx ] = m;
    }
    else
    {
      DeviceWindow w = rec_window_[ r.idx ];
      for ( const auto& kv : d.obj )
      {
        const std::string& k = kv.first;
        if ( update_window( w, k, kv.second ) )
          continue;
        if ( k == "record_to" )
        {
          if ( kv.second.string() != "memory" )
            throw std::runtime_error( "BadProperty: spike_recorder records to memory only on the device path." );
        }
        else
          throw std::runtime_error( "UnaccessedDictionaryEntry: spike_recorder parameter '" + k + "'" );
      }
      if ( not w.stop_inf and w.stop < w.start )
        throw std::runtime_error( "BadProperty: stop >= start required." );
      rec_window_[ r.idx ] = w;
    }
  }
}

void
Kernel::set_meter_status( Meter& m, const Json& d )
{
  // multimeter::Parameters_::set (models/multimeter.cpp:115-162), then the device window
  // (nestkernel/device.cpp:83-118); a failed set changes nothing (the caller keeps a copy)
  const bool fixed = d.has( "interval" ) or d.has( "offset" ) or d.has( "record_from" );
  if ( not m.targets.empty() and fixed )
    throw std::runtime_error( "BadProperty: The recording interval, the interval offset and the list of properties "
                              "to record cannot be changed after the multimeter has been connected to nodes." );
  for ( const auto& kv : d.obj )
  {
    const std::string& k = kv.first;
    if ( k == "interval" )
    {
      const int64_t tics = tb_.ms_to_tics( kv.second.number() );
      if ( tics < tb_.tics_per_step )
        throw std::runtime_error( "BadProperty: The sampling interval must be at least as long as the simulation "
                                  "resolution." );
      if ( tics % tb_.tics_per_step != 0 )
        throw std::runtime_error( "BadProperty: The sampling interval must be a multiple of the simulation "
                                  "resolution." );
      m.interval = tics / tb_.tics_per_step;
    }
    else if ( k == "offset" )
    {
      const double ms = kv.second.number();
      if ( ms != 0 and tb_.ms_to_tics( ms ) < tb_.tics_per_step )
        throw std::runtime_error( "BadProperty: The offset for the sampling interval must be at least as long as the "
                                  "simulation resolution." );
      m.offset = tb_.ms_to_steps( ms );  // Time::step( Time( Time::ms( offset ) ).get_steps() )
    }
    else if ( k == "record_from" )
    {
      if ( m.voltmeter )
        throw std::runtime_error( "BadProperty: a voltmeter records V_m only" );
      std::vector< std::string > rf;
      if ( kv.second.kind != Json::ARR )
        throw std::runtime_error( "BadProperty: record_from must be a list of names" );
      for ( const Json& v : kv.second.arr )
        rf.push_back( v.string() );
      m.record_from = rf;
    }
    else if ( update_window( m.window, k, kv.second ) )
      continue;
    else if ( k == "record_to" )
    {
      if ( kv.second.string() != "memory" )
        throw std::runtime_error( "BadProperty: multimeter records to memory only on the device path." );
    }
    else
      throw std::runtime_error( "UnaccessedDictionaryEntry: multimeter parameter '" + k + "'" );
  }
  if ( not m.window.stop_inf and m.window.stop < m.window.start )
    throw std::runtime_error( "BadProperty: stop >= start required." );
}

bool
Kernel::update_window( DeviceWindow& w, const std::string& k, const Json& v ) const
{
  // Device::Parameters_::update_ (device.cpp:83-101): finite values on the time grid, or an
  // infinite stop
  if ( k != "origin" and k != "start" and k != "stop" )
    return false;
  const double ms = v.number();
  const bool infinite = not std::isfinite( ms ) or std::fabs( ms ) >= 1e15;
  long steps = 0;
  if ( not infinite )
  {
    const long tics = tb_.ms_to_tics( ms );
    if ( tics % tb_.tics_per_step != 0 )
      throw std::runtime_error( "BadProperty: " + k + " must be a multiple of the simulation resolution." );
    steps = tics / tb_.tics_per_step;
  }
  else if ( k != "stop" or ms < 0 )
    throw std::runtime_error( "BadProperty: " + k + " must be finite on the device path." );
  if ( k == "origin" )
    w.origin = steps;
  else if ( k == "start" )
    w.start = steps;
  else
  {
    w.stop = steps;
    w.stop_inf = infinite;
  }
  w.set = true;
  return true;
}

void
Kernel::collect_log()
{
  if ( not engine_ or not engine_->built() )
    return;
  engine_->drain_log();
  auto& src = engine_->log_src;
  auto& st = engine_->log_step;
  // spike_recorder::handle (models/spike_recorder.cpp:100-112): a recorder keeps the spikes of
  // the senders it observes whose stamps lie in its window at this time
  for ( size_t r = 0; r < recorders_.size(); ++r )
  {
    const auto& rec = recorders_[ r ];
    if ( rec.empty() )
      continue;
    const DeviceWindow& w = rec_window_[ r ];
    for ( size_t i = 0; i < src.size(); ++i )
    {
      const int64_t g = dev_to_gid_[ src[ i ] ];
      if ( std::binary_search( rec.begin(), rec.end(), g ) and ( not w.set or w.active( st[ i ] ) ) )
        rec_events_[ r ].emplace_back( st[ i ], g );
    }
  }
  spike_count_ += src.size();
  src.clear();
  st.clear();
  collect_samples();
}

void
Kernel::collect_samples()
{
  if ( not engine_ or not engine_->built() )
    return;
  // multimeter::handle (models/multimeter.cpp:190-216): each multimeter keeps the samples of
  // its targets whose stamps lie in its window at this time, the values in record_from order
  auto& stamps = engine_->mm_stamps;
  auto& data = engine_->mm_data;
  const size_t cols = mm_col_gid_.size();
  if ( stamps.empty() or cols == 0 )
  {
    stamps.clear();
    data.clear();
    return;
  }
  for ( Meter& m : meters_ )
  {
    if ( m.targets.empty() or m.record_from.empty() )
      continue;
    // per target: the device field of each recordable and the offset its model adds (the
    // iaf models keep V_m relative to E_L)
    const size_t nf = m.record_from.size();
    std::vector< size_t > tcol;  // columns of the targets
    std::vector< int > tfield;
    std::vector< double > toff;
    for ( int64_t g : m.targets )
    {
      const auto it = std::lower_bound( mm_col_gid_.begin(), mm_col_gid_.end(), g );
      if ( it == mm_col_gid_.end() or *it != g )
        continue;
      tcol.push_back( static_cast< size_t >( it - mm_col_gid_.begin() ) );
      const NeuronParams& p = params_[ node( g ).idx ];
      const NodeModel& M = ModelManager::instance().get( p.model );
      for ( const std::string& v : m.record_from )
      {
        const int f = M.state_field( v );
        tfield.push_back( f < 0 ? 0 : f );
        toff.push_back( M.state_offset( p, f ) );
      }
    }
    for ( size_t row = 0; row < stamps.size(); ++row )
    {
      const int64_t stamp = stamps[ row ];
      if ( m.window.set and not m.window.active( stamp ) )
        continue;
      for ( size_t k = 0; k < tcol.size(); ++k )
      {
        const size_t c = tcol[ k ];
        m.ev_gid.push_back( mm_col_gid_[ c ] );
        m.ev_stamp.push_back( stamp );
        for ( size_t j = 0; j < nf; ++j )
          m.ev_vals.push_back( data[ ( row * cols + c ) * 4 + tfield[ k * nf + j ] ] + toff[ k * nf + j ] );
      }
    }
  }
  stamps.clear();
  data.clear();
}

const std::vector< std::string >&
Kernel::multimeter_record_from( int64_t mm_gid ) const
{
  const NodeRec& r = node( mm_gid );
  if ( r.kind != 3 )
    throw std::runtime_error( "BadProperty: node is not a multimeter" );
  return meters_[ r.idx ].record_from;
}

void
Kernel::multimeter_events( int64_t mm_gid, std::vector< int64_t >& senders, std::vector< int64_t >& steps,
  std::vector< double >& values )
{
  const NodeRec& r = node( mm_gid );
  if ( r.kind != 3 )
    throw std::runtime_error( "BadProperty: node is not a multimeter" );
  collect_log();
  const Meter& m = meters_[ r.idx ];
  const size_t nf = m.record_from.size();
  std::vector< size_t > order( m.ev_gid.size() );
  for ( size_t i = 0; i < order.size(); ++i )
    order[ i ] = i;
  std::stable_sort( order.begin(), order.end(), [ &m ]( size_t a, size_t b ) {
    return m.ev_stamp[ a ] != m.ev_stamp[ b ] ? m.ev_stamp[ a ] < m.ev_stamp[ b ] : m.ev_gid[ a ] < m.ev_gid[ b ];
  } );
  senders.resize( order.size() );
  steps.resize( order.size() );
  values.resize( order.size() * nf );
  for ( size_t e = 0; e < order.size(); ++e )
  {
    const size_t i = order[ e ];
    senders[ e ] = m.ev_gid[ i ];
    steps[ e ] = m.ev_stamp[ i ];
    for ( size_t j = 0; j < nf; ++j )
      values[ e * nf + j ] = m.ev_vals[ i * nf + j ];
  }
}

std::string
Kernel::node_status_json( int64_t gid ) const
{
  const NodeRec& r = node( gid );
  std::ostringstream o;
  o.precision( 17 );
  o << "{\"model\": \"" << model_name( gid ) << "\", \"global_id\": " << gid;
  if ( r.kind == 3 )
  {
    const Meter& m = meters_[ r.idx ];
    o << ", \"interval\": " << static_cast< double >( m.interval ) * tb_.ms_per_step
      << ", \"offset\": " << static_cast< double >( m.offset ) * tb_.ms_per_step << ", \"record_from\": [";
    for ( size_t j = 0; j < m.record_from.size(); ++j )
      o << ( j ? ", " : "" ) << "\"" << m.record_from[ j ] << "\"";
    o << "]";
  }
  o << "}";
  return o.str();
}

void
Kernel::set_values( const std::vector< int64_t >& gids, const std::string& key, const double* v )
{
  // per-node values (NodeCollection.set(key=array)); live state goes to the device directly
  const bool live = engine_ and engine_->built() and simulated_;
  for ( size_t i = 0; i < gids.size(); ++i )
  {
    NodeRec& r = node( gids[ i ] );
    if ( r.kind != 0 )
      throw std::runtime_error( "BadProperty: per-node values are only supported for neurons" );
    NeuronParams& p = params_[ r.idx ];
    NeuronState& s = state_[ r.idx ];
    if ( key == "V_m" and live )
    {
      const int dev = gid_to_dev_[ gids[ i ] ];
      engine_->set_state( 0, static_cast< uint32_t >( dev ), v[ i ] - ModelManager::instance().get( p.model ).state_offset( p, 0 ) );
      continue;
    }
    Json d;
    d.kind = Json::OBJ;
    Json x;
    x.kind = Json::NUM;
    x.num = v[ i ];
    d.obj.emplace_back( key, x );
    if ( live )
      throw std::runtime_error( "NotImplemented: only V_m can be changed after Simulate on the device path." );
    apply_params( p, s, d, false );
    store_dirty_ = true;
  }
}

void
Kernel::get_values( const std::vector< int64_t >& gids, const std::string& key, double* out )
{
  const bool live = engine_ and engine_->built() and simulated_;
  std::vector< double > cache[ 5 ];
  std::vector< int32_t > rcache;
  for ( size_t i = 0; i < gids.size(); ++i )
  {
    const NodeRec& r = node( gids[ i ] );
    if ( r.kind == 3 and ( key == "interval" or key == "offset" ) )
    {
      const Meter& m = meters_[ r.idx ];
      out[ i ] = static_cast< double >( key == "interval" ? m.interval : m.offset ) * tb_.ms_per_step;
      continue;
    }
    if ( ( ( r.kind == 2 or r.kind == 3 ) and key == "n_events" )
         or ( r.kind != 0 and ( key == "origin" or key == "start" or key == "stop" ) ) )
    {
      const DeviceWindow& w = r.kind == 2 ? rec_window_[ r.idx ]
        : r.kind == 3                     ? meters_[ r.idx ].window
                                          : gens_[ r.idx ].window;
      if ( key == "n_events" )
      {
        std::vector< int64_t > senders, steps;
        if ( r.kind == 2 )
          recorder_events( gids[ i ], senders, steps );
        else
        {
          std::vector< double > vals;
          multimeter_events( gids[ i ], senders, steps, vals );
        }
        out[ i ] = static_cast< double >( senders.size() );
      }
      else if ( key == "stop" and w.stop_inf )
        out[ i ] = std::numeric_limits< double >::infinity();
      else
        out[ i ] = static_cast< double >( key == "origin" ? w.origin : key == "start" ? w.start : w.stop )
          * tb_.ms_per_tic * static_cast< double >( tb_.tics_per_step );
      continue;
    }
    if ( r.kind != 0 )
      throw std::runtime_error( "BadProperty: '" + key + "' is not a property of " + model_name( gids[ i ] ) );
    const NeuronParams& p = params_[ r.idx ];
    const NeuronState& s = state_[ r.idx ];
    const NodeModel& M = ModelManager::instance().get( p.model );
    const int field = M.state_field( key );
    if ( field >= 0 or key == "refractory_steps" )
    {
      double v;
      int32_t rr;
      if ( live )
      {
        const uint32_t dev = static_cast< uint32_t >( gid_to_dev_[ gids[ i ] ] );
        if ( field >= 0 )
        {
          if ( cache[ field ].empty() )
            engine_->get_state( field, cache[ field ] );
          v = cache[ field ][ dev ];
        }
        else
        {
          if ( rcache.empty() )
            engine_->get_r( rcache );
          rr = rcache[ dev ];
        }
      }
      else
      {
        const double vals[] = { s.v0, s.v1, s.v2, s.v3, s.v4 };
        v = field >= 0 ? vals[ field ] : 0;
        rr = s.r;
      }
      if ( field < 0 )
        out[ i ] = rr;
      else
        out[ i ] = v + M.state_offset( p, field );
      continue;
    }
    double pv;
    if ( M.get_param( p, key, pv ) )
      out[ i ] = pv;
    else if ( key == "tau_minus" )
      out[ i ] = p.tau_minus;
    else if ( key == "global_id" )
      out[ i ] = static_cast< double >( gids[ i ] );
    else if ( key == "vp" or key == "thread" )
      out[ i ] = vp_of( gids[ i ] );  // one process: the node's thread is its VP (vp_manager_impl.h:43-52)
    else
      throw std::runtime_error( "BadProperty: unknown property '" + key + "'" );
  }
}

void
Kernel::check_delay( long steps )
{
  // DelayChecker::assert_valid_delay_ms (delay_checker.cpp:146-171)
  if ( steps < 1 )
    throw std::runtime_error( "BadDelay: Delay must be greater than or equal to resolution" );
  if ( simulated_ and ( steps < min_delay_ or steps > max_delay_ ) )
    throw std::runtime_error( "BadDelay: Minimum and maximum delay cannot be changed after Simulate has been called." );
}

void
Kernel::single_connect( int64_t s, int64_t t, int vp, VpRng& rng, Param* wp, Param* dp, uint16_t syn )
{
  // BipartiteConnBuilder::single_connect_ (conn_builder.cpp:379-437): delay is drawn before
  // the weight when both are given; defaults of static_synapse are weight 1, delay 1 ms
  double delay = 1.0, weight = 1.0;
  if ( wp and dp )
  {
    delay = dp->value( rng, vp );
    weight = wp->value( rng, vp );
  }
  else if ( dp )
    delay = dp->value( rng, vp );
  else if ( wp )
    weight = wp->value( rng, vp );
  const long ds = tb_.delay_ms_to_steps( delay );
  check_delay( ds );
  if ( syn )
  {
    // stdp_synapse::set_status (stdp_synapse.h:303-314): weight and Wmax of the same sign
    const double Wmax = stdp_sets_[ syn - 1 ].Wmax;
    if ( not( ( ( weight >= 0 ) - ( weight < 0 ) ) == ( ( Wmax >= 0 ) - ( Wmax < 0 ) ) ) )
      throw std::runtime_error( "BadProperty: Weight and Wmax must have same sign." );
  }
  ConnList& C = conns_[ vp ];
  C.src.push_back( static_cast< uint32_t >( s ) );
  C.tgt.push_back( static_cast< uint32_t >( t ) );
  C.w.push_back( weight );
  if ( ds > 65535 )
    throw std::runtime_error( "BadDelay: delays above 65535 steps are not supported" );
  C.d.push_back( static_cast< uint16_t >( ds ) );
  if ( syn )
  {
    C.syn.resize( C.size(), 0 );
    C.syn.back() = syn;
  }
}

void
Kernel::connect( const std::vector< int64_t >& pre, const std::vector< int64_t >& post, const Json& conn_spec,
  const Json& syn_spec )
{
  if ( host_connections_freed_ )
    throw std::runtime_error( "KernelException: the host connection table was released; Connect is not possible" );
  // ConnectionManager::connect (connection_manager.cpp:440-447)
  if ( pre.empty() )
    throw std::runtime_error( "IllegalConnection: Presynaptic nodes cannot be an empty NodeCollection" );
  if ( post.empty() )
    throw std::runtime_error( "IllegalConnection: Postsynaptic nodes cannot be an empty NodeCollection" );
  if ( rng_dirty_ )
    init_rngs();
  std::string rule = "all_to_all";
  bool autapses = true, multapses = true;
  double indegree = 0, p = 0;
  bool has_indegree = false, has_p = false;
  if ( conn_spec.kind == Json::STR )
    rule = conn_spec.str;
  else if ( conn_spec.kind == Json::OBJ )
  {
    for ( const auto& kv : conn_spec.obj )
    {
      if ( kv.first == "rule" )
        rule = kv.second.string();
      else if ( kv.first == "allow_autapses" )
        autapses = kv.second.number() != 0;
      else if ( kv.first == "allow_multapses" )
        multapses = kv.second.number() != 0;
      else if ( kv.first == "indegree" )
      {
        indegree = kv.second.number();
        has_indegree = true;
      }
      else if ( kv.first == "p" )
      {
        p = kv.second.number();
        has_p = true;
      }
      else
        throw std::runtime_error( "UnaccessedDictionaryEntry: conn_spec key '" + kv.first + "'" );
    }
  }
  std::unique_ptr< Param > wp, dp;
  // stdp_synapse (models/stdp_synapse.h:200-215): its parameters, constant per Connect call
  bool stdp = false;
  if ( syn_spec.has( "synapse_model" ) )
  {
    const std::string& sm = syn_spec.at( "synapse_model" ).string();
    if ( sm == "stdp_synapse" )
      stdp = true;
    else if ( sm != "static_synapse" and sm != "static_synapse_hpc" )
      throw std::runtime_error( "UnknownSynapseType: '" + sm + "' is not available on the device path." );
  }
  StdpSet sset;
  for ( const auto& kv : syn_spec.obj )
  {
    if ( kv.first == "synapse_model" )
      continue;
    double* sdst = kv.first == "tau_plus" ? &sset.tau_plus
      : kv.first == "lambda"              ? &sset.lambda
      : kv.first == "alpha"               ? &sset.alpha
      : kv.first == "mu_plus"             ? &sset.mu_plus
      : kv.first == "mu_minus"            ? &sset.mu_minus
      : kv.first == "Wmax"                ? &sset.Wmax
      : kv.first == "Kplus"               ? &sset.Kplus
                                          : nullptr;
    if ( sdst )
    {
      if ( not stdp )
        throw std::runtime_error( "UnaccessedDictionaryEntry: syn_spec key '" + kv.first + "'" );
      if ( kv.second.kind != Json::NUM )
        throw std::runtime_error( "NotImplemented: stdp_synapse parameter '" + kv.first + "' must be a number" );
      *sdst = kv.second.number();
    }
    else if ( kv.first == "weight" )
      wp = Param::from_json( kv.second, n_vp_, true, &param_states_ );
    else if ( kv.first == "delay" )
      dp = Param::from_json( kv.second, n_vp_, true, &param_states_ );
    else if ( kv.first == "receptor_type" )
    {
      // the models here have the one receptor 0 (handles_test_event, e.g. iaf_psc_exp.h:466-474)
      const auto check = [ & ]( const Json& v )
      {
        const long rt = static_cast< long >( v.number() );
        if ( rt != 0 )
          throw std::runtime_error( "UnknownReceptorType: Receptor type " + std::to_string( rt )
            + " is not available in " + model_name( post[ 0 ] ) + "." );
      };
      if ( kv.second.kind == Json::ARR )
        for ( const Json& v : kv.second.arr )
          check( v );
      else
        check( kv.second );
    }
    else if ( not sdst )
      throw std::runtime_error( "UnaccessedDictionaryEntry: syn_spec key '" + kv.first + "'" );
  }
  uint16_t syn = 0;
  if ( stdp )
  {
    if ( sset.Kplus < 0 )
      throw std::runtime_error( "BadProperty: Kplus must be non-negative." );
    // facilitate_/depress_ raise (1 - w/Wmax) and w/Wmax to mu_plus and mu_minus: 0 and 1 (the
    // defaults) are exact without pow, which device and host would round differently
    if ( ( sset.mu_plus != 0.0 and sset.mu_plus != 1.0 ) or ( sset.mu_minus != 0.0 and sset.mu_minus != 1.0 ) )
      throw std::runtime_error( "NotImplemented: stdp_synapse with mu_plus or mu_minus other than 0 or 1" );
    if ( not( sset.tau_plus > 0 ) or sset.Wmax == 0 )
      throw std::runtime_error( "BadProperty: stdp_synapse needs tau_plus > 0 and Wmax != 0" );
    if ( simulated_ )
      throw std::runtime_error( "NotImplemented: stdp_synapse connections after Simulate" );
    if ( stdp_sets_.size() >= 65534 )
      throw std::runtime_error( "NotImplemented: more than 65534 stdp_synapse Connect calls" );
  }
  if ( pre.empty() or post.empty() )
    return;
  const int pre_kind = node( pre[ 0 ] ).kind;
  const int post_kind = node( post[ 0 ] ).kind;
  for ( int64_t g : pre )
    if ( node( g ).kind != pre_kind )
      throw std::runtime_error( "IllegalConnection: mixed node kinds in source collection" );
  for ( int64_t g : post )
    if ( node( g ).kind != post_kind )
      throw std::runtime_error( "IllegalConnection: mixed node kinds in target collection" );
  if ( rule == "one_to_one" and pre.size() != post.size() )
    throw std::runtime_error( "DimensionMismatch: one_to_one requires equal numbers of sources and targets." );

  if ( post_kind == 2 )
  {
    // connections to a spike_recorder: the recorder observes these senders
    if ( pre_kind != 0 )
      throw std::runtime_error( "IllegalConnection: spike_recorder records neurons only" );
    for ( int64_t r : post )
    {
      auto& v = recorders_[ node( r ).idx ];
      v.insert( v.end(), pre.begin(), pre.end() );
      std::sort( v.begin(), v.end() );
      v.erase( std::unique( v.begin(), v.end() ), v.end() );
    }
    return;
  }
  if ( pre_kind == 3 or post_kind == 3 )
  {
    // a multimeter samples the neurons it is connected to (multimeter -> neuron); each
    // recordable must exist in the target model (UniversalDataLogger::connect_logging_device,
    // nestkernel/universal_data_logger.h:280-295)
    if ( post_kind == 3 or pre_kind != 3 or post_kind != 0 )
      throw std::runtime_error( "IllegalConnection: a multimeter connects to neurons as their source" );
    if ( rule != "all_to_all" and rule != "one_to_one" )
      throw std::runtime_error( "IllegalConnection: multimeter connections use one_to_one or all_to_all" );
    for ( size_t ti = 0; ti < post.size(); ++ti )
    {
      const std::vector< std::string > rec =
        ModelManager::instance().get( params_[ node( post[ ti ] ).idx ].model ).recordables();
      const std::vector< int64_t > meters = rule == "one_to_one" ? std::vector< int64_t > { pre[ ti ] } : pre;
      for ( int64_t mg : meters )
        for ( const std::string& v : meters_[ node( mg ).idx ].record_from )
        {
          const bool ok = std::find( rec.begin(), rec.end(), v ) != rec.end();
          if ( not ok )
            throw std::runtime_error( "IllegalConnection: Cannot connect with unknown recordable " + v );
        }
    }
    for ( size_t ti = 0; ti < post.size(); ++ti )
    {
      const std::vector< int64_t > meters = rule == "one_to_one" ? std::vector< int64_t > { pre[ ti ] } : pre;
      for ( int64_t mg : meters )
        meters_[ node( mg ).idx ].targets.push_back( post[ ti ] );
    }
    for ( int64_t mg : pre )
    {
      auto& v = meters_[ node( mg ).idx ].targets;
      std::sort( v.begin(), v.end() );
      v.erase( std::unique( v.begin(), v.end() ), v.end() );
    }
    store_dirty_ = true;
    return;
  }
  if ( stdp and ( pre_kind != 0 or post_kind != 0 ) )
    throw std::runtime_error( "NotImplemented: stdp_synapse connects neurons only on the device path" );
  if ( stdp )
  {
    stdp_sets_.push_back( sset );
    syn = static_cast< uint16_t >( stdp_sets_.size() );
  }
  if ( pre_kind == 2 )
    throw std::runtime_error( "IllegalConnection: spike_recorder cannot be a source" );
  if ( post_kind == 1 )
    throw std::runtime_error( "IllegalConnection: spike_generator cannot be a target" );

  if ( rule == "fixed_indegree" and pre_kind == 0 )
  {
    if ( indegree < 0 )
      throw std::runtime_error( "BadProperty: Indegree cannot be less than zero." );
    if ( not multapses and static_cast< size_t >( indegree ) > pre.size() )
      throw std::runtime_error( "BadProperty: Indegree cannot be larger than population size." );
  }
  if ( rule == "pairwise_bernoulli" and ( p < 0 or p > 1 ) )
    throw std::runtime_error( "BadProperty: Connection probability 0 <= p <= 1 required." );
  if ( rule != "one_to_one" and rule != "all_to_all" and rule != "fixed_indegree" and rule != "pairwise_bernoulli" )
    throw std::runtime_error( "BadProperty: unknown or unsupported connection rule '" + rule + "'" );
  // a key the rule does not read is an error (conn_spec.all_entries_accessed,
  // connection_manager.cpp:470); the rule's own key must be there (conn_builder.cpp:1403,1861)
  if ( has_indegree and rule != "fixed_indegree" )
    throw std::runtime_error( "UnaccessedDictionaryEntry: conn_spec key 'indegree'" );
  if ( has_p and rule != "pairwise_bernoulli" )
    throw std::runtime_error( "UnaccessedDictionaryEntry: conn_spec key 'p'" );
  if ( rule == "fixed_indegree" and not has_indegree )
    throw std::runtime_error( "BadProperty: rule fixed_indegree needs 'indegree'" );
  if ( rule == "pairwise_bernoulli" and not has_p )
    throw std::runtime_error( "BadProperty: rule pairwise_bernoulli needs 'p'" );
  // array parameters hold one value per connection in PyNEST's layout: one_to_one [pair],
  // all_to_all [target][source], fixed_indegree [target][draw]
  // (pynest/nest/lib/hl_api_connection_helpers.py:48-130)
  const bool arrays = ( wp and wp->is_array() ) or ( dp and dp->is_array() );
  if ( arrays )
  {
    size_t expect = 0;
    if ( rule == "one_to_one" )
      expect = pre.size();
    else if ( rule == "all_to_all" )
      expect = pre.size() * post.size();
    else if ( rule == "fixed_indegree" and pre_kind == 0 )
      expect = post.size() * static_cast< size_t >( std::max( 0L, std::lround( indegree ) ) );
    else
      throw std::runtime_error( "BadProperty: array parameters need the rule one_to_one, all_to_all or fixed_indegree" );
    for ( const Param* q : { wp.get(), dp.get() } )
      if ( q and q->is_array() and q->array_size() != expect )
        throw std::runtime_error(
          "BadProperty: an array parameter of this connection needs " + std::to_string( expect ) + " values" );
  }
  // an array's values of connections that another VP creates: each VP's cursor passes over them
  auto skip = [ &wp, &dp ]( int vp, size_t n )
  {
    if ( n == 0 )
      return;
    if ( wp )
      wp->skip( vp, n );
    if ( dp )
      dp->skip( vp, n );
  };

  if ( pre_kind == 1 )
  {
    // spike_generator -> neurons (connect_from_device_): generators exist on every thread;
    // each target's VP runs the builder with its own stream
    if ( rule != "one_to_one" and rule != "all_to_all" )
      throw std::runtime_error( "NotImplemented: spike_generator connections support one_to_one and all_to_all" );
    for ( int vp = 0; vp < n_vp_; ++vp )
      for ( size_t ti = 0; ti < post.size(); ++ti )
      {
        const int64_t t = post[ ti ];
        if ( vp_of( t ) != vp )
        {
          if ( arrays )
            skip( vp, rule == "one_to_one" ? 1 : pre.size() );
          continue;
        }
        auto one = [ & ]( int64_t s )
        {
          double delay = 1.0, weight = 1.0;
          if ( wp and dp )
          {
            delay = dp->value( rngs_[ vp ], vp );
            weight = wp->value( rngs_[ vp ], vp );
          }
          else if ( dp )
            delay = dp->value( rngs_[ vp ], vp );
          else if ( wp )
            weight = wp->value( rngs_[ vp ], vp );
          const long ds = tb_.delay_ms_to_steps( delay );
          check_delay( ds );
          if ( d_min_seen_ < 0 or ds < d_min_seen_ )
            d_min_seen_ = ds;
          d_max_seen_ = std::max( d_max_seen_, ds );
          dev_conns_.push_back( DevConn { s, t, weight, ds } );
        };
        if ( rule == "one_to_one" )
          one( pre[ ti ] );
        else
          for ( int64_t s : pre )
            one( s );
      }
    store_dirty_ = true;
    return;
  }

  // neuron -> neuron on the device: the builders whose draws take a known number of words of
  // a counter-based stream, with constant weight and delay
  const bool const_params = ( not wp or wp->is_const() ) and ( not dp or dp->is_const() );
  if ( connect_on_device_ and const_params and not stdp
       and ( rule == "pairwise_bernoulli" or ( rule == "fixed_indegree" and autapses and multapses ) ) )
  {
    if ( rng_type_ != RNG_PHILOX_32 )
      throw std::runtime_error( "BadProperty: connect_on_device draws from counter-based streams: set rng_type "
                                "'Philox_32' first" );
    // single_connect_ with constant parameters: delay, then weight (conn_builder.cpp:379-437)
    const double delay = dp ? dp->const_value() : 1.0;
    const double weight = wp ? wp->const_value() : 1.0;
    const long ds = tb_.delay_ms_to_steps( delay );
    check_delay( ds );
    if ( ds > 65535 )
      throw std::runtime_error( "BadDelay: delays above 65535 steps are not supported" );
    device_connect( pre, post, rule, rule == "fixed_indegree" ? std::round( indegree ) : p, autapses, weight,
      static_cast< uint16_t >( ds ) );
    store_dirty_ = true;
    return;
  }

  // neuron -> neuron: every VP runs the builder for its local targets with its own stream
  // (FixedInDegreeBuilder::inner_connect_ conn_builder.cpp:1508-1545, BernoulliBuilder
  // :1937-1965, OneToOne :960-1031, AllToAll :1170-1255)
  std::vector< std::string > errors( n_vp_ );
  std::vector< long > dmin( n_vp_, -1 ), dmax( n_vp_, 1 );
  std::vector< size_t > start( n_vp_ );
  for ( int vp = 0; vp < n_vp_; ++vp )
    start[ vp ] = conns_[ vp ].size();
  auto worker = [ & ]( int vp )
  {
    try
    {
      VpRng& rng = rngs_[ vp ];
      const long K = static_cast< long >( std::round( indegree ) );
      std::vector< unsigned long > chosen;
      size_t local = 0;
      for ( const int64_t t : post )
        local += vp_of( t ) == vp;
      if ( rule == "one_to_one" )
        conns_[ vp ].reserve_more( local );
      else if ( rule == "all_to_all" )
        conns_[ vp ].reserve_more( local * pre.size() );
      else if ( rule == "fixed_indegree" )
        conns_[ vp ].reserve_more( local * static_cast< size_t >( K ) );
      else
        conns_[ vp ].reserve_more( static_cast< size_t >( p * static_cast< double >( local * pre.size() ) * 1.02 ) + 1024 );
      for ( size_t ti = 0; ti < post.size(); ++ti )
      {
        const int64_t t = post[ ti ];
        if ( vp_of( t ) != vp )
        {
          // conn_builder.cpp:985-996 (one_to_one checks autapses first and skips no value for
          // one), 1186-1192 (all_to_all), 1466-1474 (fixed_indegree)
          if ( arrays )
          {
            if ( rule == "one_to_one" )
            {
              if ( not( pre[ ti ] == t and not autapses ) )
                skip( vp, 1 );
            }
            else if ( rule == "all_to_all" )
              skip( vp, pre.size() );
            else if ( rule == "fixed_indegree" )
              skip( vp, static_cast< size_t >( K ) );
          }
          continue;
        }
        if ( rule == "one_to_one" )
        {
          const int64_t s = pre[ ti ];
          if ( s == t and not autapses )
            continue;
          single_connect( s, t, vp, rng, wp.get(), dp.get(), syn );
        }
        else if ( rule == "all_to_all" )
        {
          for ( int64_t s : pre )
          {
            if ( s == t and not autapses )
            {
              if ( arrays )
                skip( vp, 1 );  // conn_builder.cpp:1244-1251
              continue;
            }
            single_connect( s, t, vp, rng, wp.get(), dp.get(), syn );
          }
        }
        else if ( rule == "fixed_indegree" )
        {
          chosen.clear();
          for ( long j = 0; j < K; ++j )
          {
            unsigned long s_id;
            int64_t s;
            bool skip_autapse, skip_multapse;
            do
            {
              s_id = rng.ulrand( pre.size() );
              s = pre[ s_id ];
              skip_autapse = not autapses and s == t;
              skip_multapse = not multapses and std::find( chosen.begin(), chosen.end(), s_id ) != chosen.end();
            } while ( skip_autapse or skip_multapse );
            if ( not multapses )
              chosen.push_back( s_id );
            single_connect( s, t, vp, rng, wp.get(), dp.get(), syn );
          }
        }
        else
        {
          for ( int64_t s : pre )
          {
            if ( not autapses and s == t )
              continue;
            if ( rng.drand() >= p )
              continue;
            single_connect( s, t, vp, rng, wp.get(), dp.get(), syn );
          }
        }
      }
    }
    catch ( const std::exception& e )
    {
      errors[ vp ] = e.what();
    }
  };
  if ( n_vp_ == 1 )
    worker( 0 );
  else
  {
    std::vector< std::thread > th;
    for ( int vp = 0; vp < n_vp_; ++vp )
      th.emplace_back( worker, vp );
    for ( auto& x : th )
      x.join();
  }
  for ( auto& e : errors )
    if ( not e.empty() )
      throw std::runtime_error( e );
  for ( int vp = 0; vp < n_vp_; ++vp )
  {
    const ConnList& C = conns_[ vp ];
    if ( C.size() > static_cast< size_t >( ( 1u << 27 ) - 2 ) )
      throw std::runtime_error( "KernelException: Too many connections: at most 134217726 connections supported per "
                                "virtual process and synapse model." );
    for ( size_t i = start[ vp ]; i < C.size(); ++i )
    {
      const long ds = C.d[ i ];
      if ( d_min_seen_ < 0 or ds < d_min_seen_ )
        d_min_seen_ = ds;
      if ( ds > d_max_seen_ )
        d_max_seen_ = ds;
    }
  }
  store_dirty_ = true;
}

void
Kernel::device_connect( const std::vector< int64_t >& pre, const std::vector< int64_t >& post, const std::string& rule,
  double param, bool autapses, double weight, uint16_t delay_steps )
{
  // each VP's local targets in builder order, as in the host worker of connect()
  std::vector< uint32_t > pre32( pre.size() );
  for ( size_t i = 0; i < pre.size(); ++i )
    pre32[ i ] = static_cast< uint32_t >( pre[ i ] );
  std::vector< GenStream > streams( n_vp_ );
  for ( int vp = 0; vp < n_vp_; ++vp )
    streams[ vp ] = GenStream { rngs_[ vp ].philox.k0, rngs_[ vp ].philox.k1, rngs_[ vp ].philox.pos };
  std::vector< std::vector< uint32_t > > src( n_vp_ ), tgt( n_vp_ );
  if ( rule == "pairwise_bernoulli" )
  {
    // the source index of each target when it is also a source (its autapse is skipped)
    std::map< int64_t, int32_t > index_in_pre;
    if ( not autapses )
      for ( size_t i = 0; i < pre.size(); ++i )
        index_in_pre.emplace( pre[ i ], static_cast< int32_t >( i ) );
    std::vector< GenRow > rows;
    std::vector< uint64_t > draws( n_vp_, 0 );
    for ( int vp = 0; vp < n_vp_; ++vp )
      for ( const int64_t t : post )
      {
        if ( vp_of( t ) != vp )
          continue;
        int32_t aut = -1;
        if ( not autapses )
        {
          const auto it = index_in_pre.find( t );
          if ( it != index_in_pre.end() )
            aut = it->second;
        }
        rows.push_back( GenRow { static_cast< uint32_t >( t ), aut, draws[ vp ], static_cast< uint32_t >( vp ) } );
        draws[ vp ] += pre.size() - ( aut >= 0 ? 1 : 0 );
      }
    gen_pairwise_bernoulli( device_, pre32, rows, streams, param, src, tgt );
    for ( int vp = 0; vp < n_vp_; ++vp )
      rngs_[ vp ].philox.pos += 2 * draws[ vp ];  // one drand, two words, per visited pair
  }
  else
  {
    const uint32_t K = static_cast< uint32_t >( std::max( 0.0, param ) );
    std::vector< std::vector< uint32_t > > rows( n_vp_ );
    for ( const int64_t t : post )
      rows[ vp_of( t ) ].push_back( static_cast< uint32_t >( t ) );
    gen_fixed_indegree( device_, pre32, rows, K, streams, src, tgt );
    for ( int vp = 0; vp < n_vp_; ++vp )
      rngs_[ vp ].philox.pos = streams[ vp ].pos;
  }
  for ( int vp = 0; vp < n_vp_; ++vp )
  {
    ConnList& C = conns_[ vp ];
    const size_t n = src[ vp ].size();
    C.reserve_more( n );
    C.src.insert( C.src.end(), src[ vp ].begin(), src[ vp ].end() );
    C.tgt.insert( C.tgt.end(), tgt[ vp ].begin(), tgt[ vp ].end() );
    C.w.insert( C.w.end(), n, weight );
    C.d.insert( C.d.end(), n, delay_steps );
    if ( C.size() > static_cast< size_t >( ( 1u << 27 ) - 2 ) )
      throw std::runtime_error( "KernelException: Too many connections: at most 134217726 connections supported per "
                                "virtual process and synapse model." );
    if ( n )
    {
      if ( d_min_seen_ < 0 or delay_steps < d_min_seen_ )
        d_min_seen_ = delay_steps;
      d_max_seen_ = std::max< long >( d_max_seen_, delay_steps );
    }
  }
  ++device_connect_calls_;
}

void
Kernel::connect_arrays( const int64_t* src, const int64_t* tgt, const double* w, const double* d_ms, int64_t n )
{
  if ( host_connections_freed_ )
    throw std::runtime_error( "KernelException: the host connection table was released; Connect is not possible" );
  // ConnectionManager::connect_arrays: each VP creates the connections of its targets in
  // array order (connection_manager.cpp:624-760). Every element is checked before any
  // connection is added, so a call that fails adds none.
  for ( int64_t i = 0; i < n; ++i )
  {
    if ( node( src[ i ] ).kind != 0 or node( tgt[ i ] ).kind != 0 )
      throw std::runtime_error( "IllegalConnection: connect_arrays connects neurons only" );
  }
  std::vector< uint16_t > steps( static_cast< size_t >( n ) );
  std::vector< size_t > per_vp( n_vp_, 0 );
  long dmin = -1, dmax = 1;
  for ( int64_t i = 0; i < n; ++i )
  {
    const long ds = tb_.delay_ms_to_steps( d_ms ? d_ms[ i ] : 1.0 );
    check_delay( ds );
    if ( ds > 65535 )
      throw std::runtime_error( "BadDelay: delays above 65535 steps are not supported" );
    steps[ i ] = static_cast< uint16_t >( ds );
    if ( dmin < 0 or ds < dmin )
      dmin = ds;
    dmax = std::max( dmax, ds );
    ++per_vp[ vp_of( tgt[ i ] ) ];
  }
  std::vector< std::string > errors( n_vp_ );
  auto fill = [ & ]( int vp )
  {
    try
    {
      ConnList& C = conns_[ vp ];
      C.reserve_more( per_vp[ vp ] );
      for ( int64_t i = 0; i < n; ++i )
      {
        if ( vp_of( tgt[ i ] ) != vp )
          continue;
        C.src.push_back( static_cast< uint32_t >( src[ i ] ) );
        C.tgt.push_back( static_cast< uint32_t >( tgt[ i ] ) );
        C.w.push_back( w ? w[ i ] : 1.0 );
        C.d.push_back( steps[ i ] );
      }
    }
    catch ( const std::exception& e )
    {
      errors[ vp ] = e.what();
    }
  };
  if ( n_vp_ == 1 or n < 65536 )
    for ( int vp = 0; vp < n_vp_; ++vp )
      fill( vp );
  else
  {
    std::vector< std::thread > th;
    for ( int vp = 0; vp < n_vp_; ++vp )
      th.emplace_back( fill, vp );
    for ( auto& x : th )
      x.join();
  }
  for ( const auto& e : errors )
    if ( not e.empty() )
      throw std::runtime_error( e );
  if ( n > 0 )
  {
    if ( d_min_seen_ < 0 or dmin < d_min_seen_ )
      d_min_seen_ = dmin;
    d_max_seen_ = std::max( d_max_seen_, dmax );
  }
  store_dirty_ = true;
}

void
Kernel::set_spike_counts( int64_t gen_first, int64_t n, int64_t t0_step, int64_t n_steps, const uint8_t* counts )
{
  for ( int64_t i = 0; i < n; ++i )
    if ( node( gen_first + i ).kind != 1 )
      throw std::runtime_error( "BadProperty: spike counts can only be set on spike_generator nodes" );
  DenseBlock b;
  b.t0 = t0_step;
  b.n_steps = n_steps;
  b.cols = n;
  b.counts.assign( counts, counts + n * n_steps );
  dense_.push_back( std::move( b ) );
  for ( int64_t i = 0; i < n; ++i )
  {
    SpikeGen& G = gens_[ node( gen_first + i ).idx ];
    G.dense_block = static_cast< int >( dense_.size() - 1 );
    G.dense_col = i;
  }
  store_dirty_ = true;
}

int64_t
Kernel::num_connections() const
{
  int64_t n = host_connections_freed_ ? freed_connections_ : 0;
  for ( const auto& c : conns_ )
    n += static_cast< int64_t >( c.size() );
  return n + static_cast< int64_t >( dev_conns_.size() );
}

void
Kernel::get_connections( int64_t* src, int64_t* tgt, double* w, int64_t* d, int32_t* vp ) const
{
  if ( host_connections_freed_ )
    throw std::runtime_error( "KernelException: the host connection table was released (keep_host_connections false)" );
  // neuron connections per VP in creation order, then device connections; stdp_synapse
  // weights as the device holds them after a run
  std::vector< double > dev_w;
  if ( not stdp_sets_.empty() and engine_ and engine_->built() and simulated_ and not released_
       and syn_pos_.size() == static_cast< size_t >( num_connections() - static_cast< int64_t >( dev_conns_.size() ) ) )
    engine_->read_weights( dev_w );
  int64_t k = 0;
  for ( int v = 0; v < n_vp_; ++v )
  {
    const ConnList& C = conns_[ v ];
    for ( size_t i = 0; i < C.size(); ++i, ++k )
    {
      src[ k ] = C.src[ i ];
      tgt[ k ] = C.tgt[ i ];
      w[ k ] = not dev_w.empty() and C.syn_at( i ) ? dev_w[ syn_pos_[ k ] ] : C.w[ i ];
      d[ k ] = C.d[ i ];
      if ( vp )
        vp[ k ] = v;
    }
  }
  for ( const auto& c : dev_conns_ )
  {
    src[ k ] = c.gen;
    tgt[ k ] = c.tgt;
    w[ k ] = c.w;
    d[ k ] = c.d;
    if ( vp )
      vp[ k ] = vp_of( c.tgt );
    ++k;
  }
}

void
Kernel::get_connection_models( int32_t* model ) const
{
  if ( host_connections_freed_ )
    throw std::runtime_error( "KernelException: the host connection table was released (keep_host_connections false)" );
  int64_t k = 0;
  for ( int v = 0; v < n_vp_; ++v )
  {
    const ConnList& C = conns_[ v ];
    for ( size_t i = 0; i < C.size(); ++i )
      model[ k++ ] = C.syn_at( i ) ? 1 : 0;
  }
  for ( size_t i = 0; i < dev_conns_.size(); ++i )
    model[ k++ ] = 0;
}

void
Kernel::prepare()
{
  // SimulationManager::prepare (simulation_manager.cpp:507-570)
  if ( released_ )
    throw std::runtime_error( "KernelException: this network was released from the device after its last run" );
  if ( not simulated_ )
  {
    min_delay_ = d_min_seen_ < 0 ? 1 : d_min_seen_;
    max_delay_ = std::max( d_max_seen_, 1L );
  }
  if ( params_.empty() )
  {
    prepared_ = true;
    return;
  }
  if ( not engine_ )
    engine_ = std::make_unique< Engine >( device_ );
  if ( store_dirty_ or not engine_->built() )
  {
    if ( simulated_ and state_dirty_ )
      throw std::runtime_error( "NotImplemented: parameter changes after Simulate are not supported on the device path." );
    collect_log();
    if ( host_connections_freed_ )
      throw std::runtime_error( "KernelException: the host connection table was released; cannot rebuild the device copy" );
    BuildSpec b;
    build_spec( b );
    engine_->build( b, simulated_ );
    engine_->net().phase_timing = phase_timing_ ? 1 : 0;
    store_dirty_ = false;
    state_dirty_ = false;
    if ( not keep_host_connections_ )
    {
      freed_connections_ = 0;
      for ( auto& C : conns_ )
      {
        freed_connections_ += static_cast< int64_t >( C.size() );
        C = ConnList();
      }
      host_connections_freed_ = true;
    }
  }
  prepared_ = true;
}

void
Kernel::reset_device()
{
  if ( host_connections_freed_ )
    throw std::runtime_error( "KernelException: the host connection table was released; cannot rebuild the device copy" );
  engine_.reset();
  prepared_ = false;
  simulated_ = false;
  store_dirty_ = true;
  rebalanced_ = false;
  rx_steps_ = 0;
  for ( auto& e : rec_events_ )
    e.clear();
  for ( Meter& m : meters_ )
  {
    m.ev_gid.clear();
    m.ev_stamp.clear();
    m.ev_vals.clear();
  }
  spike_count_ = 0;
  last_kernel_ms_ = last_wall_ms_ = time_simulate_ms_ = 0;
  for ( int q = 0; q < N_KINDS; ++q )
  {
    time_kernel_ms_[ q ] = 0;
    kernel_launches_[ q ] = 0;
  }
}

uint64_t
Kernel::estimate_device_bytes() const
{
  const uint64_t N = params_.size();
  if ( N == 0 )
    return 0;
  int64_t E = 0;
  for ( const auto& C : conns_ )
    E += static_cast< int64_t >( C.size() );
  E += host_connections_freed_ ? freed_connections_ : 0;
  const long dmin = simulated_ ? min_delay_ : ( d_min_seen_ < 0 ? 1 : d_min_seen_ );
  const long dmax = simulated_ ? max_delay_ : std::max( d_max_seen_, 1L );
  int nch = 1;
  for ( const auto& Pm : params_ )
    nch = std::max( nch, ModelManager::instance().get( Pm.model ).channels() );
  uint32_t tb = 0;
  bool win = false;
  try
  {
    choose_layout( static_cast< uint32_t >( N ), static_cast< int >( dmax ), nch, tb, win );
  }
  catch ( const std::exception& )
  {
    // a forced window layout that does not fit fails at Prepare; estimate the global one
    tb = std::max< uint32_t >( 32, std::min< uint32_t >( MAX_TB, static_cast< uint32_t >( ( N + 511 ) / 512 ) ) );
  }
  const uint64_t P = ( N + tb - 1 ) / tb;
  // balanced blocks: as many blocks at most, each up to twice as large
  if ( targets_per_block_ == 0 and balanced_blocks_ )
    tb = block_cap( tb );
  const uint64_t L = dmin + dmax;
  uint64_t b = static_cast< uint64_t >( E ) * sizeof( SynRec ) + N * ( P + 1 ) * 4  // store
    + P * L * nch * tb * 8                                                          // ring buffers
    + N * ( 5 * 8 + 4 + 2 + 1 + 4 )                                                 // state, indices
    + 2 * P * tb * dmin * 5 + 2 * N * dmin * 4                                     // spike lists
    + ( std::any_of( recorders_.begin(), recorders_.end(), []( const auto& r ) { return not r.empty(); } )
          ? std::max< uint64_t >( 1u << 20, N * 16 ) * 12
          : 0 );
  for ( const auto& D : dense_ )
    b += D.counts.size();
  return b + ( 1u << 20 );
}

void
Kernel::release_device()
{
  if ( engine_ and engine_->built() )
  {
    collect_log();
    std::vector< double > f[ 5 ];
    for ( int q = 0; q < 5; ++q )
      engine_->get_state( q, f[ q ] );
    std::vector< int32_t > r;
    engine_->get_r( r );
    for ( size_t i = 0; i < state_.size(); ++i )
    {
      const uint32_t dev = static_cast< uint32_t >( gid_to_dev_[ neuron_gid_[ i ] ] );
      NeuronState& S = state_[ i ];
      S.v0 = f[ 0 ][ dev ];
      S.v1 = f[ 1 ][ dev ];
      S.v2 = f[ 2 ][ dev ];
      S.v3 = f[ 3 ][ dev ];
      S.v4 = f[ 4 ][ dev ];
      S.r = r[ dev ];
    }
    released_events_ = engine_->read_counter( 1 );
    released_spikes_ = engine_->read_counter( 2 );
    // stdp_synapse weights live on the device: keep them in the host connection table
    if ( not stdp_sets_.empty() and simulated_ and not host_connections_freed_
         and syn_pos_.size() == static_cast< size_t >( num_connections() - static_cast< int64_t >( dev_conns_.size() ) ) )
    {
      std::vector< double > dev_w;
      engine_->read_weights( dev_w );
      int64_t k = 0;
      for ( auto& C : conns_ )
        for ( size_t i = 0; i < C.size(); ++i, ++k )
          if ( C.syn_at( i ) )
            C.w[ i ] = dev_w[ syn_pos_[ k ] ];
    }
  }
  engine_.reset();
  prepared_ = false;
  released_ = true;
}

void
Kernel::set_layout( int targets_per_block, int delivery_mode, int window_slots )
{
  if ( simulated_ )
    throw std::runtime_error( "KernelException: the device layout cannot change after Simulate" );
  targets_per_block_ = targets_per_block;
  delivery_mode_ = delivery_mode;
  window_slots_ = window_slots;
  store_dirty_ = true;
}

void
Kernel::prepare_for_batch()
{
  if ( not prepared_ )
    prepare();
}

void
Kernel::after_batch_run( int64_t )
{
  simulated_ = true;
  collect_log();
}

int64_t
Kernel::checked_steps( double t_ms ) const
{
  // nest::run (nestkernel/nest.cpp:506-525)
  if ( t_ms < 0 )
    throw std::runtime_error( "BadParameter: The simulation time cannot be negative." );
  if ( tb_.ms_to_tics( t_ms ) % tb_.tics_per_step != 0 )
    throw std::runtime_error( "BadParameter: The simulation time must be a multiple of the simulation resolution." );
  return tb_.ms_to_steps( t_ms );
}

void
Kernel::run( double t_ms )
{
  if ( not prepared_ )
    throw std::runtime_error( "KernelException: Run called without calling Prepare." );
  const long steps = checked_steps( t_ms );
  simulated_ = true;
  if ( steps == 0 or params_.empty() )
    return;
  if ( not runner_ )
  {
    if ( not own_runner_ )
      own_runner_ = std::make_unique< Runner >( device_, slices_per_chunk_ );
    runner_ = own_runner_.get();
  }
  runner_->time_kernels = time_kernels_;
  runner_->wg_per_cu = wg_per_cu_;
  if ( runner_ == own_runner_.get() )
    runner_->use_graph = graph_launch_;
  double km0[ N_KINDS ];
  int64_t kl0[ N_KINDS ];
  for ( int q = 0; q < N_KINDS; ++q )
  {
    km0[ q ] = runner_->kernel_ms[ q ];
    kl0[ q ] = runner_->kernel_launches[ q ];
  }
  const int64_t gc0 = runner_->graph_chunks, dc0 = runner_->direct_chunks, cf0 = runner_->capture_failures;
  // received events are counted for the measured block balance until it has been applied
  const bool measure = runner_ == own_runner_.get() and balanced_blocks_ and targets_per_block_ == 0
    and not rebalanced_ and not host_connections_freed_;
  engine_->count_received( measure );
  const auto t0 = std::chrono::steady_clock::now();
  runner_->run( { engine_.get() }, { steps } );
  const auto t1 = std::chrono::steady_clock::now();
  engine_->count_received( false );
  if ( measure )
    rx_steps_ += steps;
  last_wall_ms_ = std::chrono::duration< double, std::milli >( t1 - t0 ).count();
  time_simulate_ms_ += last_wall_ms_;
  last_kernel_ms_ = runner_->last_kernel_ms();
  for ( int q = 0; q < N_KINDS; ++q )
  {
    time_kernel_ms_[ q ] += runner_->kernel_ms[ q ] - km0[ q ];
    kernel_launches_[ q ] += runner_->kernel_launches[ q ] - kl0[ q ];
  }
  graph_chunks_ += runner_->graph_chunks - gc0;
  direct_chunks_ += runner_->direct_chunks - dc0;
  capture_failures_ += runner_->capture_failures - cf0;
  collect_log();
  if ( measure )
  {
    const auto t2 = std::chrono::steady_clock::now();
    rebalance_blocks();
    const double ms = std::chrono::duration< double, std::milli >( std::chrono::steady_clock::now() - t2 ).count();
    last_wall_ms_ += ms;
    time_simulate_ms_ += ms;
  }
}

void
Kernel::cleanup()
{
  prepared_ = false;
}

void
Kernel::simulate( double t_ms )
{
  prepare();
  run( t_ms );
  cleanup();
}

uint64_t
Kernel::synaptic_events() const
{
  if ( released_ )
    return released_events_;
  if ( not engine_ or not engine_->built() )
    return 0;
  return engine_->read_counter( 1 );
}

void
Kernel::recorder_events( int64_t rec_gid, std::vector< int64_t >& senders, std::vector< int64_t >& steps )
{
  const NodeRec& r = node( rec_gid );
  if ( r.kind != 2 )
    throw std::runtime_error( "BadProperty: node is not a spike_recorder" );
  collect_log();
  std::vector< std::pair< int64_t, int64_t > > ev = rec_events_[ r.idx ];
  std::sort( ev.begin(), ev.end() );
  senders.resize( ev.size() );
  steps.resize( ev.size() );
  for ( size_t i = 0; i < ev.size(); ++i )
  {
    steps[ i ] = ev[ i ].first;
    senders[ i ] = ev[ i ].second;
  }
}

}  // namespace nestgpu
