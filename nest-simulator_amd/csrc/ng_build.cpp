// ng_build.cpp -- Prepare: from NEST's node/connection tables to the device layout.
//
// Counterpart of SimulationManager::prepare -> update_connection_infrastructure
// (nestkernel/simulation_manager.cpp:507-570,750-833: sort connections by source, build
// the presynaptic target tables) and NodeManager::prepare_nodes (node_manager.cpp:590-640:
// init buffers, pre_run_hook propagators).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <thread>

#include "ng_nest.h"

namespace nestgpu
{

void
Kernel::choose_layout( uint32_t n, int max_delay, int nch, uint32_t& tb, bool& win ) const
{
  // Block size: at most two blocks per compute unit of an MI355X (256 CUs, two workgroups
  // each), multiple of 8, at least 32, few enough blocks for the per-instance limit. The
  // smallest such block leaves the most ring slots in the LDS window.
  const uint32_t pmax = delivery_mode_ == 1 or delivery_mode_ == 3 ? PMAX_WIN : PMAX_GLOB;
  if ( targets_per_block_ > 0 )
    tb = static_cast< uint32_t >( targets_per_block_ );
  else
  {
    tb = ( n + 511 ) / 512;
    tb = ( ( tb + 7 ) / 8 ) * 8;
    tb = std::max< uint32_t >( tb, 32 );
    tb = std::max< uint32_t >( tb, ( n + pmax - 1 ) / pmax );
    tb = std::min< uint32_t >( tb, MAX_TB );
  }
  const uint32_t P = ( n + tb - 1 ) / tb;
  // auto: window delivery wherever it fits (it keeps the densest slots' read-modify-writes in
  // LDS), global delivery for slot-range passes and for instances beyond the window's limits
  const bool fits = window_slots( max_delay, nch, tb ) >= 1 and P <= PMAX_WIN;
  // (cache delivery starts from the window layout; build_spec turns it into the cell cache)
  win = fits and ( delivery_mode_ == 1 or delivery_mode_ == 3 or ( delivery_mode_ == 0 and delivery_passes_ <= 1 ) );
  if ( ( delivery_mode_ == 1 or delivery_mode_ == 3 ) and not win )
    throw std::runtime_error( "BadProperty: delivery_mode 'window' does not fit: channels * targets_per_block must "
                              "be <= "
      + std::to_string( WIN_CAP ) + " and the instance at most " + std::to_string( PMAX_WIN ) + " blocks" );
}

int
Kernel::window_slots( int max_delay, int nch, uint32_t tb )
{
  return static_cast< int >(
    std::min< uint64_t >( max_delay, WIN_CAP / ( static_cast< uint64_t >( nch ) * std::max< uint32_t >( tb, 1 ) ) ) );
}

// Contiguous blocks of device indices balancing the work of their targets: at most p_max
// blocks of at most cap targets each, the largest block work as small as greedy filling under
// a bisected bound makes it (a block closes when the next target would exceed the bound or
// the block is full). start has one entry per block and the end.
static std::vector< uint32_t >
partition_by_work( const std::vector< double >& work, uint32_t p_max, uint32_t cap )
{
  const uint32_t N = static_cast< uint32_t >( work.size() );
  double total = 0, heaviest = 0;
  for ( double x : work )
  {
    total += x;
    heaviest = std::max( heaviest, x );
  }
  auto fill = [ & ]( double bound, std::vector< uint32_t >* out )
  {
    uint32_t blocks = 0, size = 0;
    double sum = 0;
    if ( out )
      out->assign( 1, 0 );
    for ( uint32_t i = 0; i < N; ++i )
    {
      if ( size > 0 and ( size == cap or sum + work[ i ] > bound ) )
      {
        ++blocks;
        if ( out )
          out->push_back( i );
        size = 0;
        sum = 0;
      }
      ++size;
      sum += work[ i ];
    }
    ++blocks;
    if ( out )
      out->push_back( N );
    return blocks;
  };
  // at the total every block is cut by size only: ceil(N / cap) <= p_max blocks
  double lo = std::max( total / p_max, heaviest ), hi = total;
  for ( int it = 0; it < 50 and hi - lo > 1e-9 * hi; ++it )
  {
    const double mid = 0.5 * ( lo + hi );
    if ( fill( mid, nullptr ) <= p_max )
      hi = mid;
    else
      lo = mid;
  }
  std::vector< uint32_t > start;
  fill( hi, &start );
  return start;
}

static double
largest_block_work( const std::vector< double >& work, const std::vector< uint32_t >& start )
{
  double m = 0;
  for ( size_t q = 0; q + 1 < start.size(); ++q )
  {
    double sum = 0;
    for ( uint32_t i = start[ q ]; i < start[ q + 1 ]; ++i )
      sum += work[ i ];
    m = std::max( m, sum );
  }
  return m;
}

std::vector< uint32_t >
Kernel::balanced_partition( uint32_t N, uint32_t tb ) const
{
  // in-degree of every device index: the events it receives per spike of all its sources
  const int nw = static_cast< int >( std::max< size_t >( 1, std::min< size_t >( conns_.size(), 16 ) ) );
  std::vector< std::vector< uint32_t > > part( nw, std::vector< uint32_t >( N, 0 ) );
  {
    std::vector< std::thread > th;
    for ( int w = 0; w < nw; ++w )
      th.emplace_back(
        [ &, w ]
        {
          for ( size_t v = w; v < conns_.size(); v += nw )
            for ( uint32_t g : conns_[ v ].tgt )
              ++part[ w ][ gid_to_dev_[ g ] ];
        } );
    for ( auto& x : th )
      x.join();
  }
  std::vector< double > work( N, 0.0 );
  double total = 0;
  for ( uint32_t i = 0; i < N; ++i )
  {
    uint64_t k = 0;
    for ( int w = 0; w < nw; ++w )
      k += part[ w ][ i ];
    work[ i ] = static_cast< double >( k );
    total += work[ i ];
  }
  // a target's update costs about as much as delivering a fifth of the mean in-degree
  const double upd = 0.2 * total / N + 1.0;
  for ( double& x : work )
    x += upd;
  return partition_by_work( work, ( N + tb - 1 ) / tb, block_cap( tb ) );
}

uint32_t
Kernel::block_cap( uint32_t tb )
{
  return std::min< uint32_t >( MAX_TB, ( 2 * tb + 7 ) / 8 * 8 );
}

void
Kernel::rebalance_blocks()
{
  // Target blocks balanced by the events their targets received so far (measured by the slice
  // kernel), once, at the first Run boundary that ends a slice after at least 64 slices
  // counted. A target's update costs about as much as delivering 4 events per slice (the
  // update and the delivery phases of an iaf_psc_exp circuit on the MI355X). The layout
  // changes only if the largest block work drops by 5 % or more; results do not depend on it.
  Engine& E = *engine_;
  if ( not stdp_sets_.empty() )
  {
    // the move rebuilds the store from the host table, which holds no stdp_synapse state
    rebalanced_ = true;
    return;
  }
  const int64_t slices = rx_steps_ / std::max< long >( min_delay_, 1 );
  if ( slices < 64 or E.part != 0 or E.from_step != 0 )
    return;
  rebalanced_ = true;
  std::vector< uint32_t > rx;
  E.read_received( rx );
  const uint32_t N = static_cast< uint32_t >( rx.size() );
  std::vector< double > work( N );
  double events = 0;
  for ( uint32_t i = 0; i < N; ++i )
  {
    work[ i ] = rx[ i ] + 4.0 * static_cast< double >( slices );
    events += rx[ i ];
  }
  if ( events < static_cast< double >( N ) * slices )  // under one event per target and slice
    return;
  uint32_t tb_nom = 0;
  bool win = false;
  choose_layout( N, static_cast< int >( max_delay_ ), E.net().nch, tb_nom, win );
  std::vector< uint32_t > start = partition_by_work( work, ( N + tb_nom - 1 ) / tb_nom, block_cap( tb_nom ) );
  if ( largest_block_work( work, start ) > 0.95 * largest_block_work( work, E.block_start() ) )
    return;
  const std::vector< uint32_t > before = block_start_;
  block_start_ = start;
  BuildSpec b;
  build_spec( b );
  if ( E.relayout( b ) )
    ++block_relayouts_;
  else
    block_start_ = before;
}

void
Kernel::build_spec( BuildSpec& b )
{
  const uint32_t N = static_cast< uint32_t >( params_.size() );
  const double h = tb_.resolution_ms();
  b.n_vp = n_vp_;
  b.min_delay = static_cast< int >( min_delay_ );
  b.max_delay = static_cast< int >( max_delay_ );
  b.n = N;
  // ring-buffer channels: the most any model here uses
  b.nch = 1;
  for ( const auto& Pm : params_ )
    b.nch = std::max( b.nch, ModelManager::instance().get( Pm.model ).channels() );
  choose_layout( N, b.max_delay, b.nch, b.tb, b.win );
  // ---- device index = position in NEST's receive order (VP, then node id)
  std::vector< uint32_t > order( N );
  for ( uint32_t i = 0; i < N; ++i )
    order[ i ] = i;
  std::stable_sort( order.begin(), order.end(),
    [ this ]( uint32_t a, uint32_t c ) { return vp_of( neuron_gid_[ a ] ) < vp_of( neuron_gid_[ c ] ); } );
  std::vector< uint32_t > dev_of( N );
  for ( uint32_t o = 0; o < N; ++o )
    dev_of[ order[ o ] ] = o;
  b.vp_begin.assign( n_vp_ + 1, 0 );
  for ( uint32_t o = 0; o < N; ++o )
    b.vp_begin[ vp_of( neuron_gid_[ order[ o ] ] ) + 1 ]++;
  for ( int v = 0; v < n_vp_; ++v )
    b.vp_begin[ v + 1 ] += b.vp_begin[ v ];
  gid_to_dev_.assign( nodes_.size() + 1, -1 );
  dev_to_gid_.assign( N, 0 );
  for ( uint32_t i = 0; i < N; ++i )
  {
    gid_to_dev_[ neuron_gid_[ i ] ] = static_cast< int32_t >( dev_of[ i ] );
    dev_to_gid_[ dev_of[ i ] ] = neuron_gid_[ i ];
  }

  // ---- target blocks: contiguous device-index ranges. With the block size left to the
  // library, the blocks are balanced by in-degree (no more blocks than uniform ones of the
  // chosen size, each at most twice as large); a rebuild after Simulate keeps the partition,
  // which the ring buffers and pending spike lists are laid out by (the measured balance
  // moves them explicitly).
  if ( simulated_ and not block_start_.empty() and block_start_.back() == N )
    b.bstart = block_start_;
  else if ( targets_per_block_ == 0 and balanced_blocks_ and N > 0 )
    b.bstart = balanced_partition( N, b.tb );
  else
  {
    b.bstart.resize( ( N + b.tb - 1 ) / b.tb + 1 );
    for ( size_t q = 0; q < b.bstart.size(); ++q )
      b.bstart[ q ] = static_cast< uint32_t >( std::min< uint64_t >( static_cast< uint64_t >( q ) * b.tb, N ) );
  }
  {
    uint32_t largest = 0;
    for ( size_t q = 0; q + 1 < b.bstart.size(); ++q )
      largest = std::max( largest, b.bstart[ q + 1 ] - b.bstart[ q ] );
    if ( targets_per_block_ == 0 )
      b.tb = std::max< uint32_t >( 8, ( largest + 7 ) / 8 * 8 );
  }
  block_start_ = b.bstart;
  const uint32_t P = b.n_blocks();
  if ( b.win )
  {
    // The window covers the wn consecutive slots that expect the most events: with the
    // emission lag uniform over the slice, a synapse of delay d feeds slots
    // [d - min_delay, d - 1] equally.
    // Staging costs the same for every slice while a slice's events grow with its length:
    // by default the window holds at most 6 slots per step of min_delay (the measured best on
    // a one-step min_delay circuit; ten-step slices keep every slot that fits).
    b.wn = window_slots( b.max_delay, b.nch, b.tb );
    b.wn = std::min( b.wn, window_slots_ > 0 ? window_slots_ : 6 * b.min_delay );
    std::vector< double > per_delay( b.max_delay + 1, 0.0 ), density( b.max_delay, 0.0 );
    for ( const auto& C : conns_ )
      for ( uint16_t dd : C.d )
        per_delay[ dd ] += 1.0;
    for ( int dd = b.min_delay; dd <= b.max_delay; ++dd )
      for ( int l = 0; l < b.min_delay; ++l )
        density[ l + dd - b.min_delay ] += per_delay[ dd ];
    double run = 0, best = -1, total = 0;
    b.wa = 0;
    for ( int s = 0; s < b.max_delay; ++s )
    {
      total += density[ s ];
      run += density[ s ];
      if ( s >= b.wn )
        run -= density[ s - b.wn ];
      if ( s >= b.wn - 1 and run > best )
      {
        best = run;
        b.wa = s - b.wn + 1;
      }
    }
    // cache delivery: the cells each slice touches are cached in LDS instead of a window,
    // loaded once and written back once per slice
    (void)total;
    if ( delivery_mode_ == 3 )
    {
      b.win = false;
      b.cache = true;
      b.wa = b.wn = 0;
    }
  }
  else if ( delivery_mode_ == 3 )
    throw std::runtime_error( "BadProperty: delivery_mode 'cache' needs at most " + std::to_string( PMAX_WIN )
      + " target blocks" );
  // Global delivery can run in slot-range passes (delivery_passes), each touching a smaller
  // part of the ring. Finding each spike's split points costs dependent loads that measured
  // more than the smaller working set saves on the MI355X, so one pass is the default.
  b.np = 1;
  b.ws = b.max_delay;
  if ( not b.win and delivery_passes_ > 1 )
  {
    b.np = std::max( 1, std::min( delivery_passes_, b.max_delay ) );
    b.ws = ( b.max_delay + b.np - 1 ) / b.np;
  }
  // ---- drive: at most one spike_generator per neuron
  std::vector< int64_t > gen_of( N, -1 );
  std::vector< double > drive_w( N, 0 );
  std::vector< long > drive_d( N, 0 );
  for ( const auto& c : dev_conns_ )
  {
    const uint32_t i = node( c.tgt ).idx;
    if ( gen_of[ i ] >= 0 )
      throw std::runtime_error( "NotImplemented: at most one spike_generator per neuron on the device path" );
    gen_of[ i ] = c.gen;
    drive_w[ i ] = c.w;
    drive_d[ i ] = c.d;
  }
  // dense count blocks: user blocks plus one for generators given by spike_times
  std::vector< DenseBlock > blocks = dense_;
  {
    int64_t tmax = 0;
    bool any = false;
    for ( const auto& G : gens_ )
      if ( G.dense_block < 0 and G.poisson_lambda == 0 and not G.stamps.empty() )
      {
        any = true;
        tmax = std::max( tmax, G.window.origin + G.stamps.back() );
      }
    if ( any )
    {
      DenseBlock sb;
      sb.t0 = 0;
      sb.n_steps = tmax;
      sb.cols = static_cast< int64_t >( gens_.size() );
      sb.counts.assign( static_cast< size_t >( sb.n_steps * sb.cols ), 0 );
      for ( size_t g = 0; g < gens_.size(); ++g )
      {
        const SpikeGen& G = gens_[ g ];
        if ( G.dense_block >= 0 or G.poisson_lambda != 0 )
          continue;
        for ( size_t k = 0; k < G.stamps.size(); ++k )
        {
          if ( k > 0 and G.stamps[ k ] == G.stamps[ k - 1 ] )
            throw std::runtime_error(
              "NotImplemented: repeated spike times in one spike_generator; use spike_multiplicities" );
          const int64_t m = G.mults.empty() ? 1 : G.mults[ k ];
          if ( m < 0 or m > 255 )
            throw std::runtime_error( "NotImplemented: spike multiplicities above 255" );
          // emitted at origin + stamp while the window is active; stamps before the first step
          // are never sent (spike_generator.cpp:334-352)
          const int64_t stamp = G.window.origin + G.stamps[ k ];
          if ( stamp < 1 or ( G.window.set and not G.window.active( stamp ) ) )
            continue;
          sb.counts[ static_cast< size_t >( ( stamp - 1 ) * sb.cols + g ) ] = static_cast< uint8_t >( m );
        }
      }
      blocks.push_back( std::move( sb ) );
    }
  }
  std::vector< int64_t > block_col0( blocks.size(), 0 );
  int64_t t0 = 0, t1 = 0, cols = 0;
  for ( size_t k = 0; k < blocks.size(); ++k )
  {
    if ( k == 0 )
    {
      t0 = blocks[ k ].t0;
      t1 = blocks[ k ].t0 + blocks[ k ].n_steps;
    }
    t0 = std::min( t0, blocks[ k ].t0 );
    t1 = std::max( t1, blocks[ k ].t0 + blocks[ k ].n_steps );
    block_col0[ k ] = cols;
    cols += blocks[ k ].cols;
  }
  if ( blocks.size() == 1 )
  {
    b.drive_counts = blocks[ 0 ].counts;
    b.drive_cols = cols;
  }
  else if ( blocks.size() > 1 )
  {
    b.drive_cols = cols;
    b.drive_counts.assign( static_cast< size_t >( ( t1 - t0 ) * cols ), 0 );
    for ( size_t k = 0; k < blocks.size(); ++k )
      for ( int64_t s = 0; s < blocks[ k ].n_steps; ++s )
        std::memcpy( &b.drive_counts[ static_cast< size_t >( ( blocks[ k ].t0 - t0 + s ) * cols + block_col0[ k ] ) ],
          &blocks[ k ].counts[ static_cast< size_t >( s * blocks[ k ].cols ) ], static_cast< size_t >( blocks[ k ].cols ) );
  }
  const size_t user_blocks = dense_.size();

  // ---- parameter sets (pre_run_hook of each model) and initial state
  std::map< std::string, uint16_t > pset_ids;
  std::map< std::pair< double, uint64_t >, int32_t > tables;
  b.drive_idx.assign( N, 0xffffffffu );
  b.pidx.assign( N, 0 );
  b.model.assign( N, 0 );
  b.r.assign( N, 0 );
  for ( int f = 0; f < 5; ++f )
    b.v[ f ].assign( N, 0.0 );
  int first_model = -1;
  bool uniform = true;
  for ( uint32_t i = 0; i < N; ++i )
  {
    const NeuronParams& Pm = params_[ i ];
    const NeuronState& S = state_[ i ];
    ParamSet ps;
    std::memset( &ps, 0, sizeof ps );
    ps.model = Pm.model;
    ps.refractory_counts = static_cast< int32_t >( tb_.ms_to_steps( Pm.t_ref ) );  // Time( Time::ms( t_ref ) ).get_steps()
    const NodeModel& M = ModelManager::instance().get( Pm.model );
    M.pre_run_hook( Pm, h, ps );
    if ( gen_of[ i ] >= 0 )
    {
      const SpikeGen& G = gens_[ node( gen_of[ i ] ).idx ];
      ps.drive_w = drive_w[ i ];
      ps.drive_d = static_cast< int32_t >( drive_d[ i ] );
      if ( G.poisson_lambda > 0 )
      {
        ps.drive_mode = 2;
        auto key = std::make_pair( G.poisson_lambda, G.poisson_seed );
        auto it = tables.find( key );
        if ( it == tables.end() )
        {
          it = tables.emplace( key, static_cast< int32_t >( b.ptables.size() ) ).first;
          b.ptables.push_back( make_poisson_table( G.poisson_lambda, G.poisson_seed ) );
        }
        ps.drive_table = it->second;
        ps.drive_t0 = G.poisson_t0;
        ps.drive_steps = G.poisson_steps;
        b.drive_idx[ dev_of[ i ] ] = static_cast< uint32_t >( gen_of[ i ] );
      }
      else if ( G.dense_block >= 0 or not G.stamps.empty() )
      {
        ps.drive_mode = 1;
        ps.drive_t0 = blocks.size() > 1 ? t0 : blocks[ 0 ].t0;
        ps.drive_steps = blocks.size() > 1 ? t1 - t0 : blocks[ 0 ].n_steps;
        const int64_t col = G.dense_block >= 0 ? block_col0[ G.dense_block ] + G.dense_col
                                               : block_col0[ user_blocks ] + node( gen_of[ i ] ).idx;
        b.drive_idx[ dev_of[ i ] ] = static_cast< uint32_t >( col );
      }
    }
    const std::string key( reinterpret_cast< const char* >( &ps ), sizeof ps );
    auto it = pset_ids.find( key );
    if ( it == pset_ids.end() )
    {
      if ( b.psets.size() >= 65535 )
        throw std::runtime_error( "NotImplemented: more than 65535 distinct parameter sets" );
      it = pset_ids.emplace( key, static_cast< uint16_t >( b.psets.size() ) ).first;
      b.psets.push_back( ps );
    }
    const uint32_t dv = dev_of[ i ];
    b.pidx[ dv ] = it->second;
    b.model[ dv ] = static_cast< uint8_t >( Pm.model );
    if ( first_model < 0 )
      first_model = Pm.model;
    uniform = uniform and Pm.model == first_model;
    b.r[ dv ] = S.r;
    double sv[ 5 ];
    M.device_state( Pm, S, h, sv );
    for ( int f = 0; f < 5; ++f )
      b.v[ f ][ dv ] = sv[ f ];
  }
  b.uniform_model = uniform ? first_model : -1;

  // ---- out-CSR by source, each row grouped by target block; creation order within a
  //      block run (the canonical connector-run order; NEST's source sort leaves the
  //      order of multapses build dependent, libnestutil/sort.h)
  int64_t E = 0;
  for ( const auto& C : conns_ )
    E += static_cast< int64_t >( C.size() );
  if ( max_delay_ - min_delay_ > 65535 )
    throw std::runtime_error( "NotImplemented: delay range above 65535 steps" );
  if ( E >= ( int64_t( 1 ) << 32 ) )
    throw std::runtime_error( "NotImplemented: 2^32 or more synapses in one instance" );
  // Three parallel passes over the connections, each worker on a contiguous range of
  // creation order (VP order, then within a VP):
  //   1. device source of every connection, counts per (worker, source);
  //   2. the connections placed source-major into syn, a worker's connections of one source
  //      after those of the workers before it, i.e. every row in creation order;
  //   3. per row: grouped by target block (counting sort, stable), each block run ordered
  //      by delay (stable, so a cell (target, slot) receives a spike's weights in creation
  //      order), and the final record words written.
  const int n_workers = static_cast< int >(
    std::max< unsigned >( 1u, std::min< unsigned >( 16u, std::thread::hardware_concurrency() ) ) );
  // runs fn(w) for every worker; an error in any worker is raised after all have joined
  auto parallel = [ n_workers ]( auto&& fn )
  {
    std::vector< std::string > errors( n_workers );
    auto guarded = [ & ]( int w )
    {
      try
      {
        fn( w );
      }
      catch ( const std::exception& e )
      {
        errors[ w ] = e.what();
      }
    };
    std::vector< std::thread > th;
    for ( int w = 1; w < n_workers; ++w )
      th.emplace_back( guarded, w );
    guarded( 0 );
    for ( auto& x : th )
      x.join();
    for ( const auto& e : errors )
      if ( not e.empty() )
        throw std::runtime_error( e );
  };
  std::vector< int64_t > vp_first( conns_.size() + 1, 0 );
  for ( size_t v = 0; v < conns_.size(); ++v )
    vp_first[ v + 1 ] = vp_first[ v ] + static_cast< int64_t >( conns_[ v ].size() );
  // the VP whose list holds connection e (the first with vp_first[v + 1] > e)
  auto vp_of_edge = [ & ]( int64_t e )
  {
    return static_cast< size_t >(
      std::upper_bound( vp_first.begin() + 1, vp_first.end(), e ) - ( vp_first.begin() + 1 ) );
  };
  HostArray< uint32_t > src_dev( E );
  HostArray< uint16_t > delay( E );
  // stdp_synapse: the synapse of every connection in creation order, and the creation index
  // each store position holds (carried through the passes)
  bool plastic = false;
  for ( const auto& C : conns_ )
    plastic = plastic or not C.syn.empty();
  std::vector< uint16_t > syn_flat;
  HostArray< uint32_t > cidx;
  if ( plastic )
  {
    syn_flat.assign( static_cast< size_t >( E ), 0 );
    for ( size_t v = 0; v < conns_.size(); ++v )
      for ( size_t i = 0; i < conns_[ v ].syn.size(); ++i )
        syn_flat[ vp_first[ v ] + i ] = conns_[ v ].syn[ i ];
    cidx = HostArray< uint32_t >( E );
  }
  // offsets fit 32 bits: fewer than 2^32 connections
  std::vector< uint32_t > start( static_cast< size_t >( n_workers ) * N, 0 );
  parallel(
    [ & ]( int w )
    {
      const int64_t lo = E * w / n_workers, hi = E * ( w + 1 ) / n_workers;
      uint32_t* cnt = &start[ static_cast< size_t >( w ) * N ];
      size_t v = lo < hi ? vp_of_edge( lo ) : 0;
      for ( int64_t e = lo; e < hi; ++e )
      {
        while ( e >= vp_first[ v + 1 ] )
          ++v;
        const uint32_t s = static_cast< uint32_t >( gid_to_dev_[ conns_[ v ].src[ e - vp_first[ v ] ] ] );
        src_dev[ e ] = s;
        ++cnt[ s ];
      }
    } );
  std::vector< uint64_t > row_ptr( N + 1, 0 );
  for ( uint32_t s = 0; s < N; ++s )
  {
    uint64_t at = row_ptr[ s ];
    for ( int w = 0; w < n_workers; ++w )
    {
      uint32_t& c = start[ static_cast< size_t >( w ) * N + s ];
      const uint32_t n_ws = c;
      c = static_cast< uint32_t >( at );
      at += n_ws;
    }
    row_ptr[ s + 1 ] = at;
  }
  b.syn.resize( E );
  parallel(
    [ & ]( int w )
    {
      const int64_t lo = E * w / n_workers, hi = E * ( w + 1 ) / n_workers;
      uint32_t* cur = &start[ static_cast< size_t >( w ) * N ];
      size_t v = lo < hi ? vp_of_edge( lo ) : 0;
      for ( int64_t e = lo; e < hi; ++e )
      {
        while ( e >= vp_first[ v + 1 ] )
          ++v;
        const ConnList& C = conns_[ v ];
        const size_t i = static_cast< size_t >( e - vp_first[ v ] );
        const uint32_t pos = cur[ src_dev[ e ] ]++;
        uint64_t wbits;
        std::memcpy( &wbits, &C.w[ i ], sizeof wbits );
        SynRec& rec = b.syn[ pos ];
        rec.w_lo = static_cast< uint32_t >( wbits );
        rec.w_hi = static_cast< uint32_t >( wbits >> 32 );
        rec.word = static_cast< uint32_t >( gid_to_dev_[ C.tgt[ i ] ] );  // device target, for now
        if ( C.d[ i ] - min_delay_ > static_cast< long >( DELAY_MASK ) )
          throw std::runtime_error( "NotImplemented: delays more than 32767 steps above min_delay" );
        delay[ pos ] = static_cast< uint16_t >( C.d[ i ] - min_delay_ );
        if ( plastic )
          cidx[ pos ] = static_cast< uint32_t >( e );
      }
    } );
  start = std::vector< uint32_t >();
  src_dev = HostArray< uint32_t >();
  // block of every device index
  std::vector< uint32_t > blk_of( N );
  for ( uint32_t q = 0; q < P; ++q )
    std::fill( blk_of.begin() + b.bstart[ q ], blk_of.begin() + b.bstart[ q + 1 ], q );
  b.boff.resize( static_cast< size_t >( N ) * ( P + 1 ) );
  parallel(
    [ & ]( int w )
    {
      // rows split by connection count: worker w takes the rows that start in its share
      const uint64_t e_lo = static_cast< uint64_t >( E ) * w / n_workers;
      const uint64_t e_hi = static_cast< uint64_t >( E ) * ( w + 1 ) / n_workers;
      const uint32_t s_lo =
        w == 0 ? 0 : static_cast< uint32_t >( std::lower_bound( row_ptr.begin(), row_ptr.end() - 1, e_lo ) - row_ptr.begin() );
      const uint32_t s_hi = w == n_workers - 1
        ? N
        : static_cast< uint32_t >( std::lower_bound( row_ptr.begin(), row_ptr.end() - 1, e_hi ) - row_ptr.begin() );
      std::vector< SynRec > row_syn;
      std::vector< uint16_t > row_delay;
      std::vector< uint32_t > cur( P );
      struct RecC
      {
        SynRec r;
        uint32_t c;
      };
      std::vector< RecC > row_rc;
      std::vector< uint32_t > row_c;
      for ( uint32_t s = s_lo; s < s_hi; ++s )
      {
        const uint64_t r0 = row_ptr[ s ], r1 = row_ptr[ s + 1 ];
        uint32_t* bo = &b.boff[ static_cast< size_t >( s ) * ( P + 1 ) ];
        std::fill( bo, bo + P + 1, 0u );
        for ( uint64_t j = r0; j < r1; ++j )
          ++bo[ blk_of[ b.syn[ j ].word ] + 1 ];
        bo[ 0 ] = static_cast< uint32_t >( r0 );
        for ( uint32_t q = 0; q < P; ++q )
        {
          bo[ q + 1 ] += bo[ q ];
          cur[ q ] = bo[ q ];
        }
        row_syn.assign( b.syn.begin() + r0, b.syn.begin() + r1 );
        row_delay.assign( delay.begin() + r0, delay.begin() + r1 );
        if ( plastic )
        {
          // the same grouping and delay order, each record with its creation index
          row_c.assign( cidx.begin() + r0, cidx.begin() + r1 );
          row_rc.resize( r1 - r0 );
          for ( size_t j = 0; j < row_syn.size(); ++j )
          {
            const uint32_t t = row_syn[ j ].word;
            const uint32_t blk = blk_of[ t ];
            RecC& x = row_rc[ cur[ blk ]++ - r0 ];
            x.r.w_lo = row_syn[ j ].w_lo;
            x.r.w_hi = row_syn[ j ].w_hi;
            x.r.word = ( t - b.bstart[ blk ] ) | ( static_cast< uint32_t >( row_delay[ j ] ) << 16 )
              | ( syn_flat[ row_c[ j ] ] ? PLASTIC_BIT : 0u );
            x.c = row_c[ j ];
          }
          for ( uint32_t q = 0; q < P; ++q )
            std::stable_sort( row_rc.begin() + ( bo[ q ] - r0 ), row_rc.begin() + ( bo[ q + 1 ] - r0 ),
              []( const RecC& x, const RecC& y )
              { return ( ( x.r.word >> 16 ) & DELAY_MASK ) < ( ( y.r.word >> 16 ) & DELAY_MASK ); } );
          for ( size_t k = 0; k < row_rc.size(); ++k )
          {
            b.syn[ r0 + k ] = row_rc[ k ].r;
            cidx[ r0 + k ] = row_rc[ k ].c;
          }
          continue;
        }
        for ( size_t j = 0; j < row_syn.size(); ++j )
        {
          const uint32_t t = row_syn[ j ].word;
          const uint32_t blk = blk_of[ t ];
          SynRec& rec = b.syn[ cur[ blk ]++ ];
          rec.w_lo = row_syn[ j ].w_lo;
          rec.w_hi = row_syn[ j ].w_hi;
          rec.word = ( t - b.bstart[ blk ] ) | ( static_cast< uint32_t >( row_delay[ j ] ) << 16 );
        }
        for ( uint32_t q = 0; q < P; ++q )
        {
          SynRec* run = &b.syn[ bo[ q ] ];
          const uint32_t len = bo[ q + 1 ] - bo[ q ];
          if ( len <= 1 )
            continue;
          auto by_delay = []( const SynRec& x, const SynRec& y )
          { return ( ( x.word >> 16 ) & DELAY_MASK ) < ( ( y.word >> 16 ) & DELAY_MASK ); };
          if ( len > 32 )
          {
            std::stable_sort( run, run + len, by_delay );
            continue;
          }
          for ( uint32_t j = 1; j < len; ++j )  // insertion sort: stable
          {
            const SynRec x = run[ j ];
            uint32_t k = j;
            for ( ; k > 0 and by_delay( x, run[ k - 1 ] ); --k )
              run[ k ] = run[ k - 1 ];
            run[ k ] = x;
          }
        }
      }
    } );
  // stdp_synapse: the presynaptic trace rows, each record's row, where each connection went,
  // and the neurons with incoming stdp synapses
  b.stdp_prm.clear();
  b.stdp_pre.clear();
  b.stdp_pre_row0.clear();
  b.stdp_srow = HostArray< uint32_t >();
  b.stdp_in.clear();
  syn_pos_.clear();
  if ( plastic )
  {
    for ( const StdpSet& q : stdp_sets_ )
      b.stdp_prm.push_back( StdpPrm { q.tau_plus, q.lambda, q.alpha, q.mu_plus, q.mu_minus, q.Wmax } );
    // the parameter sets of each source's stdp synapses (the records of source s are the
    // positions [row_ptr[s], row_ptr[s + 1]) in every pass): one, or a sorted list
    std::vector< uint16_t > one_set( N, 0 );
    std::vector< std::vector< std::pair< uint32_t, std::vector< uint16_t > > > > several( n_workers );
    auto source_range = [ & ]( int w, uint32_t& s_lo, uint32_t& s_hi )
    {
      const uint64_t e_lo = static_cast< uint64_t >( E ) * w / n_workers;
      const uint64_t e_hi = static_cast< uint64_t >( E ) * ( w + 1 ) / n_workers;
      s_lo = w == 0 ? 0
                    : static_cast< uint32_t >( std::lower_bound( row_ptr.begin(), row_ptr.end() - 1, e_lo ) - row_ptr.begin() );
      s_hi = w == n_workers - 1
        ? N
        : static_cast< uint32_t >( std::lower_bound( row_ptr.begin(), row_ptr.end() - 1, e_hi ) - row_ptr.begin() );
    };
    parallel(
      [ & ]( int w )
      {
        uint32_t s_lo, s_hi;
        source_range( w, s_lo, s_hi );
        std::vector< uint16_t > sets;
        for ( uint32_t src = s_lo; src < s_hi; ++src )
        {
          sets.clear();
          for ( uint64_t pos = row_ptr[ src ]; pos < row_ptr[ src + 1 ]; ++pos )
          {
            const uint16_t sv = syn_flat[ cidx[ pos ] ];
            if ( sv and ( sets.empty() or sets.back() != sv ) )
              sets.push_back( sv );
          }
          std::sort( sets.begin(), sets.end() );
          sets.erase( std::unique( sets.begin(), sets.end() ), sets.end() );
          if ( sets.size() == 1 )
            one_set[ src ] = sets[ 0 ];
          else if ( sets.size() > 1 )
            several[ w ].emplace_back( src, sets );
        }
      } );
    std::vector< const std::vector< uint16_t >* > sets_of( N, nullptr );
    for ( const auto& list : several )
      for ( const auto& ps : list )
        sets_of[ ps.first ] = &ps.second;
    b.stdp_pre_row0.assign( N + 1, 0 );
    for ( uint32_t src = 0; src < N; ++src )
      b.stdp_pre_row0[ src + 1 ] =
        b.stdp_pre_row0[ src ] + ( sets_of[ src ] ? static_cast< uint32_t >( sets_of[ src ]->size() ) : one_set[ src ] ? 1u : 0u );
    // rows at creation: Kplus from the syn_spec, t_lastspike_ 0 (stdp_synapse.h:200-215)
    b.stdp_pre.assign( b.stdp_pre_row0[ N ], StdpPre {} );
    for ( uint32_t src = 0; src < N; ++src )
      for ( uint32_t r = b.stdp_pre_row0[ src ]; r < b.stdp_pre_row0[ src + 1 ]; ++r )
      {
        const uint16_t sv = sets_of[ src ] ? ( *sets_of[ src ] )[ r - b.stdp_pre_row0[ src ] ] : one_set[ src ];
        StdpPre& R = b.stdp_pre[ r ];
        const StdpSet& q = stdp_sets_[ sv - 1 ];
        R.cur_kp = q.Kplus;
        R.cur_tl = 0.0;
        R.tau_plus = q.tau_plus;
        R.lambda = q.lambda;
        R.alpha = q.alpha;
        R.Wmax = q.Wmax;
        R.mu_plus = q.mu_plus != 0.0 ? 1u : 0u;
        R.mu_minus = q.mu_minus != 0.0 ? 1u : 0u;
      }
    b.stdp_srow = HostArray< uint32_t >( E );
    syn_pos_.resize( static_cast< size_t >( E ) );
    parallel(
      [ & ]( int w )
      {
        uint32_t s_lo, s_hi;
        source_range( w, s_lo, s_hi );
        for ( uint32_t src = s_lo; src < s_hi; ++src )
          for ( uint64_t pos = row_ptr[ src ]; pos < row_ptr[ src + 1 ]; ++pos )
          {
            const uint32_t c = cidx[ pos ];
            syn_pos_[ c ] = static_cast< uint32_t >( pos );
            const uint16_t sv = syn_flat[ c ];
            uint32_t row = 0;
            if ( sv )
            {
              row = b.stdp_pre_row0[ src ];
              if ( sets_of[ src ] )
                row += static_cast< uint32_t >(
                  std::lower_bound( sets_of[ src ]->begin(), sets_of[ src ]->end(), sv ) - sets_of[ src ]->begin() );
            }
            b.stdp_srow[ pos ] = row;
          }
      } );
    b.stdp_in.assign( N, 0 );
    for ( size_t v = 0; v < conns_.size(); ++v )
      for ( size_t i = 0; i < conns_[ v ].syn.size(); ++i )
        if ( conns_[ v ].syn[ i ] )
          b.stdp_in[ gid_to_dev_[ conns_[ v ].tgt[ i ] ] ] = 1;
    b.tau_minus_inv.assign( N, 0.0 );
    for ( uint32_t i = 0; i < N; ++i )
      b.tau_minus_inv[ dev_of[ i ] ] = 1. / params_[ i ].tau_minus;
  }
  b.stdp_hcap = stdp_history_;
  b.ms_per_tic = tb_.ms_per_tic;
  b.ms_per_step = tb_.ms_per_step;
  b.tics_per_step = tb_.tics_per_step;
  // spikes are logged only while a spike_recorder observes some neuron
  bool recording = false;
  for ( const auto& rec : recorders_ )
    recording = recording or not rec.empty();
  b.log_enabled = recording;
  b.log_cap = recording ? std::max< uint64_t >( 1u << 20, static_cast< uint64_t >( N ) * 16 ) : 1024;
  // multimeters: the union of their targets is sampled on the device, one column per neuron
  // in node-id order, at the interval and offset the multimeters share
  collect_samples();  // samples of the previous device copy use its columns
  std::vector< int64_t > sampled;
  int64_t mm_int = 0, mm_off = 0;
  for ( const Meter& m : meters_ )
  {
    if ( m.targets.empty() or m.record_from.empty() )
      continue;
    if ( mm_int == 0 )
    {
      mm_int = m.interval;
      mm_off = m.offset;
    }
    else if ( m.interval != mm_int or m.offset != mm_off )
      throw std::runtime_error( "NotImplemented: multimeters with different interval or offset on one network" );
    sampled.insert( sampled.end(), m.targets.begin(), m.targets.end() );
  }
  std::sort( sampled.begin(), sampled.end() );
  sampled.erase( std::unique( sampled.begin(), sampled.end() ), sampled.end() );
  mm_col_gid_ = sampled;
  b.mm_col.assign( sampled.empty() ? 0 : N, 0xffffffffu );
  for ( size_t c = 0; c < sampled.size(); ++c )
    b.mm_col[ gid_to_dev_[ sampled[ c ] ] ] = static_cast< uint32_t >( c );
  b.mm_cols = static_cast< uint32_t >( sampled.size() );
  b.mm_int = sampled.empty() ? 0 : static_cast< int32_t >( mm_int );
  b.mm_off = mm_off;
}

}  // namespace nestgpu
