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

// Propagator P32 of IAFPropagatorExp::evaluate (libnestutil/iaf_propagator.cpp:42-74)
static double
iaf_propagator_exp( double tau_syn, double tau_m, double c_m, double h )
{
  const double beta = tau_syn * tau_m / ( tau_m - tau_syn );
  const double gamma = beta / c_m;
  const double inv_tau_syn = 1 / tau_syn;
  const double inv_tau_m = 1 / tau_m;
  const double inv_c_m = 1 / c_m;
  const double inv_beta = ( tau_m - tau_syn ) / ( tau_syn * tau_m );
  const double exp_h_tau_syn = std::exp( -h * inv_tau_syn );
  const double expm1_h_tau = std::expm1( h * inv_beta );
  const double P32 = gamma * exp_h_tau_syn * expm1_h_tau;
  if ( std::isnormal( P32 ) and P32 > 0 )
    return P32;
  const double exp_h_tau = std::exp( -h * inv_tau_m );
  return h * inv_c_m * exp_h_tau;
}

void
Kernel::choose_layout( uint32_t n, int max_delay, int nch, uint32_t& tb, bool& win ) const
{
  // Block size: at most two blocks per compute unit of an MI355X (256 CUs, two workgroups
  // each), multiple of 8, at least 32, few enough blocks for the per-instance limit. The
  // smallest such block leaves the most ring slots in the LDS window.
  const uint32_t pmax = delivery_mode_ == 1 ? PMAX_WIN : PMAX_GLOB;
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
  win = fits and ( delivery_mode_ == 1 or ( delivery_mode_ == 0 and delivery_passes_ <= 1 ) );
  if ( delivery_mode_ == 1 and not win )
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

void
Kernel::build_spec( BuildSpec& b )
{
  const uint32_t N = static_cast< uint32_t >( params_.size() );
  const double h = tb_.resolution_ms();
  b.n_vp = n_vp_;
  b.min_delay = static_cast< int >( min_delay_ );
  b.max_delay = static_cast< int >( max_delay_ );
  b.n = N;
  bool all_delta = true;
  for ( const auto& Pm : params_ )
    all_delta = all_delta and Pm.model == M_IAF_PSC_DELTA;
  b.nch = all_delta ? 1 : 2;
  choose_layout( N, b.max_delay, b.nch, b.tb, b.win );
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
    double run = 0, best = -1;
    b.wa = 0;
    for ( int s = 0; s < b.max_delay; ++s )
    {
      run += density[ s ];
      if ( s >= b.wn )
        run -= density[ s - b.wn ];
      if ( s >= b.wn - 1 and run > best )
      {
        best = run;
        b.wa = s - b.wn + 1;
      }
    }
  }
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
        tmax = std::max( tmax, G.stamps.back() );
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
          sb.counts[ static_cast< size_t >( ( G.stamps[ k ] - 1 ) * sb.cols + g ) ] = static_cast< uint8_t >( m );
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
    if ( Pm.model == M_IAF_PSC_DELTA )
    {
      // iaf_psc_delta::pre_run_hook (iaf_psc_delta.cpp:234-270)
      const double P33 = std::exp( -h / Pm.tau_m );
      const double P30 = 1 / Pm.c_m * ( 1 - P33 ) * Pm.tau_m;
      ps.p[ 0 ] = P30;
      ps.p[ 1 ] = P33;
      ps.p[ 2 ] = 0.0 + Pm.I_e;
      ps.p[ 3 ] = Pm.V_th;
      ps.p[ 4 ] = Pm.V_reset;
      ps.p[ 5 ] = Pm.V_min;
      // refractory_input: input arriving while refractory decays by exp(-r h / tau_m) and is
      // added when the neuron becomes active again (iaf_psc_delta.cpp:284-299)
      ps.p[ 6 ] = Pm.refr_input ? 1.0 : 0.0;
      ps.p[ 7 ] = h;
      ps.p[ 8 ] = Pm.tau_m;
    }
    else if ( Pm.model == M_IAF_PSC_EXP )
    {
      // iaf_psc_exp::pre_run_hook (iaf_psc_exp.cpp:244-286)
      const double P11ex = std::exp( -h / Pm.tau_ex );
      const double P11in = std::exp( -h / Pm.tau_in );
      const double P22 = std::exp( -h / Pm.tau_m );
      const double P21ex = iaf_propagator_exp( Pm.tau_ex, Pm.tau_m, Pm.c_m, h );
      const double P21in = iaf_propagator_exp( Pm.tau_in, Pm.tau_m, Pm.c_m, h );
      const double P20 = Pm.tau_m / Pm.c_m * ( 1.0 - P22 );
      ps.p[ 0 ] = P22;
      ps.p[ 1 ] = P21ex;
      ps.p[ 2 ] = P21in;
      ps.p[ 3 ] = P20;
      ps.p[ 4 ] = P11ex;
      ps.p[ 5 ] = P11in;
      ps.p[ 6 ] = Pm.I_e + 0.0;
      ps.p[ 7 ] = Pm.V_th;
      ps.p[ 8 ] = Pm.V_reset;
      ps.p[ 9 ] = ( 1. - P11ex ) * 0.0;
    }
    else
    {
      // aeif_cond_exp::pre_run_hook (aeif_cond_exp.cpp:420-435)
      const double vals[] = { Pm.V_peak, Pm.V_reset, Pm.g_L, Pm.c_m, Pm.E_ex, Pm.E_in, Pm.E_L, Pm.Delta_T, Pm.tau_w,
        Pm.a, Pm.b, Pm.V_th, Pm.tau_ex, Pm.tau_in, Pm.I_e, Pm.tol, Pm.Delta_T > 0. ? Pm.V_peak : Pm.V_th, h };
      for ( int q = 0; q < 18; ++q )
        ps.p[ q ] = vals[ q ];
    }
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
    const double sv[] = { S.v0, S.v1, S.v2, S.v3, Pm.model == M_AEIF_COND_EXP ? h : 0.0 };
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
  std::vector< uint64_t > row_ptr( N + 1, 0 );
  b.boff.assign( static_cast< size_t >( N ) * ( P + 1 ), 0 );
  b.syn.resize( E );
  // Worker w owns the sources [N*w/T, N*(w+1)/T): it scans every connection in creation
  // order (VP order, then within a VP) and counts, then fills, only its own rows, so each
  // (source, block) run keeps creation order with no synchronisation between workers.
  std::vector< int64_t > vp_first( conns_.size() + 1, 0 );
  for ( size_t v = 0; v < conns_.size(); ++v )
    vp_first[ v + 1 ] = vp_first[ v ] + static_cast< int64_t >( conns_[ v ].size() );
  std::vector< uint32_t > src_dev( E ), tgt_dev( E );
  const int n_workers = static_cast< int >(
    std::max< unsigned >( 1u, std::min< unsigned >( 16u, std::thread::hardware_concurrency() ) ) );
  auto parallel = [ n_workers ]( auto&& fn )
  {
    std::vector< std::thread > th;
    for ( int w = 1; w < n_workers; ++w )
      th.emplace_back( fn, w );
    fn( 0 );
    for ( auto& x : th )
      x.join();
  };
  parallel(
    [ & ]( int w )
    {
      const int64_t lo = E * w / n_workers, hi = E * ( w + 1 ) / n_workers;
      size_t v = 0;
      for ( int64_t e = lo; e < hi; ++e )
      {
        while ( e >= vp_first[ v + 1 ] )
          ++v;
        const size_t i = static_cast< size_t >( e - vp_first[ v ] );
        src_dev[ e ] = static_cast< uint32_t >( gid_to_dev_[ conns_[ v ].src[ i ] ] );
        tgt_dev[ e ] = static_cast< uint32_t >( gid_to_dev_[ conns_[ v ].tgt[ i ] ] );
      }
    } );
  const uint32_t tbs = b.tb;
  // counts per (source, block), then the row-relative prefix
  parallel(
    [ & ]( int w )
    {
      const uint32_t s_lo = static_cast< uint32_t >( static_cast< uint64_t >( N ) * w / n_workers );
      const uint32_t s_hi = static_cast< uint32_t >( static_cast< uint64_t >( N ) * ( w + 1 ) / n_workers );
      for ( int64_t e = 0; e < E; ++e )
      {
        const uint32_t s = src_dev[ e ];
        if ( s >= s_lo and s < s_hi )
          b.boff[ static_cast< size_t >( s ) * ( P + 1 ) + tgt_dev[ e ] / tbs + 1 ]++;
      }
      for ( uint32_t s = s_lo; s < s_hi; ++s )
      {
        uint32_t* row = &b.boff[ static_cast< size_t >( s ) * ( P + 1 ) ];
        for ( uint32_t q = 0; q < P; ++q )
          row[ q + 1 ] += row[ q ];
      }
    } );
  for ( uint32_t s = 0; s < N; ++s )
    row_ptr[ s + 1 ] = row_ptr[ s ] + b.boff[ static_cast< size_t >( s ) * ( P + 1 ) + P ];
  parallel(
    [ & ]( int w )
    {
      const uint32_t s_lo = static_cast< uint32_t >( static_cast< uint64_t >( N ) * w / n_workers );
      const uint32_t s_hi = static_cast< uint32_t >( static_cast< uint64_t >( N ) * ( w + 1 ) / n_workers );
      // row-relative offsets become absolute run starts in syn
      for ( uint32_t s = s_lo; s < s_hi; ++s )
      {
        uint32_t* row = &b.boff[ static_cast< size_t >( s ) * ( P + 1 ) ];
        for ( uint32_t q = 0; q <= P; ++q )
          row[ q ] += static_cast< uint32_t >( row_ptr[ s ] );
      }
      std::vector< uint32_t > cur( b.boff.begin() + static_cast< size_t >( s_lo ) * ( P + 1 ),
        b.boff.begin() + static_cast< size_t >( s_hi ) * ( P + 1 ) );
      size_t v = 0;
      for ( int64_t e = 0; e < E; ++e )
      {
        while ( e >= vp_first[ v + 1 ] )
          ++v;
        const uint32_t s = src_dev[ e ];
        if ( s < s_lo or s >= s_hi )
          continue;
        const uint32_t t = tgt_dev[ e ];
        const uint32_t blk = t / tbs;
        const uint32_t pos = cur[ static_cast< size_t >( s - s_lo ) * ( P + 1 ) + blk ]++;
        const size_t i = static_cast< size_t >( e - vp_first[ v ] );
        uint64_t wbits;
        std::memcpy( &wbits, &conns_[ v ].w[ i ], sizeof wbits );
        SynRec& rec = b.syn[ pos ];
        rec.w_lo = static_cast< uint32_t >( wbits );
        rec.w_hi = static_cast< uint32_t >( wbits >> 32 );
        rec.word = ( t - blk * tbs ) | ( static_cast< uint32_t >( conns_[ v ].d[ i ] - min_delay_ ) << 16 );
      }
      // each (source, block) run ordered by delay, creation order among equal delays: a cell
      // (target, slot) still receives a spike's weights in creation order
      for ( uint32_t s = s_lo; s < s_hi; ++s )
      {
        const uint32_t* row = &b.boff[ static_cast< size_t >( s ) * ( P + 1 ) ];
        for ( uint32_t q = 0; q < P; ++q )
          std::stable_sort( b.syn.begin() + row[ q ], b.syn.begin() + row[ q + 1 ],
            []( const SynRec& x, const SynRec& y ) { return ( x.word >> 16 ) < ( y.word >> 16 ); } );
      }
    } );
  // spikes are logged only while a spike_recorder observes some neuron
  bool recording = false;
  for ( const auto& rec : recorders_ )
    recording = recording or not rec.empty();
  b.log_enabled = recording;
  b.log_cap = recording ? std::max< uint64_t >( 1u << 20, static_cast< uint64_t >( N ) * 16 ) : 1024;
}

}  // namespace nestgpu
