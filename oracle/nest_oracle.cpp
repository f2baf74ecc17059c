// nest_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A single-process CPU restatement of NEST's per-min_delay-slice simulation step
// for point-neuron networks (iaf_psc_delta, iaf_psc_exp, aeif_cond_exp,
// static_synapse, spike_generator drive), written to be read side by side with
// the reference sources it follows. It exists to CHECK the MI355X product path:
// only the test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it. Nothing in the product links, calls or falls back to it.
//
// Reference files followed (paths relative to the NEST source tree):
//   time base ........ nestkernel/nest_time.h:306-311,508-522,539-542; nest_time.cpp:88-103;
//                      libnestutil/numerics.cpp:41-44 (ld_round)
//   RNG .............. nestkernel/random_manager.cpp:42-48,63-96; random_generators.h:141-297;
//                      thirdparty/randutils.hpp:215-385 (seed_seq_fe128, restated below);
//                      Philox_32: thirdparty/Random123/philox.h:240-330 and
//                      conventional/Engine.hpp (r123::Engine<Philox4x32>, restated below
//                      from the published Philox4x32-10 algorithm, Salmon et al., SC11)
//   parameters ....... nestkernel/parameter.h:131-286,468-600,878-990; parameter.cpp:95-117
//   connect .......... nestkernel/conn_builder.cpp:379-437 (single_connect_), 960-1031 (one_to_one),
//                      1170-1255 (all_to_all), 1448-1545 (fixed_indegree), 1881-1965 (bernoulli);
//                      connector_model_impl.h:220-319; delay_checker.cpp:146-171
//   placement ........ nestkernel/vp_manager_impl.h:43-52; node_manager.cpp:188-230
//   slice loop ....... nestkernel/simulation_manager.cpp:614-671,842-1189,1192-1223
//   send/gather ...... nestkernel/event_delivery_manager_impl.h:35-118;
//                      event_delivery_manager.cpp:251-303,368-497
//   deliver .......... event_delivery_manager.cpp:599-776; connector_base.h:399-428;
//                      models/static_synapse.h:160-169; nestkernel/event.h:1006-1013
//   ring buffers ..... nestkernel/ring_buffer.h:157-203,395-408
//   plasticity ....... models/stdp_synapse.h:187-198,222-283 (send, facilitate_, depress_);
//                      nestkernel/archiving_node.cpp:38-48,85-113,165-230 (get_K_value,
//                      get_history, set_spiketime); connection_manager.cpp:78 (stdp_eps)
//   models ........... models/iaf_psc_delta.cpp:69-343; models/iaf_psc_exp.cpp:71-363;
//                      libnestutil/iaf_propagator.cpp:42-116; models/aeif_cond_exp.cpp:79-541;
//                      models/iaf_psc_alpha.cpp:68-383
//   drive ............ models/spike_generator.cpp:321-389
//
// aeif_cond_exp: NEST integrates it with GSL's rkf45 stepper, the y' error control and
// gsl_odeiv_evolve_apply (aeif_cond_exp.cpp:386,395,463). GSL is not part of the
// reference tree; its published algorithm (the Runge-Kutta-Fehlberg 4(5) tableau and the
// documented step-size control rule) is restated in rkf45_evolve() below. Bit parity with
// GSL itself is therefore unpinned; the restatement is pinned against the LSODAR trajectory
// the reference's test suite holds and against closed-form special cases.
//
// Compile with -O2 -ffp-contract=off and no -ffast-math: expression order and rounding
// follow the reference line by line.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle
{

// ------------------------------------------------------------------ time base
// nestkernel/nest_time.cpp:36-47,88-103 and nest_time.h:306-311,508-522,539-542
struct TimeBase
{
  double tics_per_ms = 1000.0;
  long tics_per_step = 100;
  double tics_per_step_inv = 1.0 / 100.0;
  long tics_per_step_rnd = 99;
  double ms_per_tic = 1.0 / 1000.0;
  double ms_per_step = 100 / 1000.0;
  double steps_per_ms = 1 / ( 100 / 1000.0 );

  void
  set_resolution( double ms )
  {
    tics_per_step = static_cast< long >( std::floor( tics_per_ms * ms + 0.5 ) );
    tics_per_step_inv = 1. / static_cast< double >( tics_per_step );
    tics_per_step_rnd = tics_per_step - 1;
    ms_per_step = tics_per_step / tics_per_ms;
    steps_per_ms = 1 / ms_per_step;
  }
  long
  ms_to_tics( double ms ) const
  {
    return static_cast< long >( ms * tics_per_ms + 0.5 );
  }
  long
  tics_to_steps( long tics ) const
  {
    return static_cast< long >( ( tics + tics_per_step_rnd ) * tics_per_step_inv );
  }
  long
  ms_to_steps( double ms ) const
  {
    return tics_to_steps( ms_to_tics( ms ) );
  }
  bool
  is_grid_ms( double ms ) const
  {
    return ms_to_tics( ms ) % tics_per_step == 0;
  }
  long
  delay_ms_to_steps( double ms ) const
  {
    return static_cast< long >( std::floor( ms * steps_per_ms + 0.5 ) );
  }
  double
  resolution_ms() const
  {
    return ms_per_tic * tics_per_step;
  }
};

// ------------------------------------------------------------- seed sequence
// O'Neill's seed_seq_fe<4, uint32_t> (randutils), restated from its published algorithm:
// hash the entropy words into a 4-word mixer, cross-mix every pair, then emit output words
// through a second multiplicative hash. mix_rounds is 1 for four words.
struct SeedSeqFe128
{
  typedef uint32_t result_type;
  uint32_t m[ 4 ];

  explicit SeedSeqFe128( const std::vector< uint32_t >& in )
  {
    uint32_t hc = 0x43b0d7e5u;
    auto hash = [ &hc ]( uint32_t v )
    {
      v ^= hc;
      hc *= 0x931e8875u;
      v *= hc;
      v ^= v >> 16;
      return v;
    };
    auto mix = []( uint32_t x, uint32_t y )
    {
      uint32_t r = 0xca01f9ddu * x - 0x4973f715u * y;
      r ^= r >> 16;
      return r;
    };
    size_t cur = 0;
    for ( int i = 0; i < 4; ++i )
      m[ i ] = hash( cur < in.size() ? in[ cur++ ] : 0u );
    for ( int s = 0; s < 4; ++s )
      for ( int d = 0; d < 4; ++d )
        if ( s != d )
          m[ d ] = mix( m[ d ], hash( m[ s ] ) );
    for ( ; cur < in.size(); ++cur )
      for ( int d = 0; d < 4; ++d )
        m[ d ] = mix( m[ d ], hash( in[ cur ] ) );
  }
  template < typename It >
  void
  generate( It first, It last ) const
  {
    uint32_t hc = 0x8b51f9ddu;
    int s = 0;
    for ( It d = first; d != last; ++d )
    {
      uint32_t v = m[ s ];
      s = ( s + 1 ) & 3;
      v ^= hc;
      hc *= 0x58f38dedu;
      v *= hc;
      v ^= v >> 16;
      *d = v;
    }
  }
  static constexpr size_t
  size()
  {
    return 4;
  }
};

// Philox4x32 block function with R rounds (philox.h:296-330): each round multiplies words 0
// and 2 of the counter by the two Philox constants (64-bit products split in high and low
// halves), mixes the high halves with words 1 and 3 and the key, and the key is bumped by
// the two Weyl constants between rounds.
inline void
philox4x32( int rounds, const uint32_t ctr_in[ 4 ], const uint32_t key_in[ 2 ], uint32_t out[ 4 ] )
{
  uint32_t c[ 4 ] = { ctr_in[ 0 ], ctr_in[ 1 ], ctr_in[ 2 ], ctr_in[ 3 ] };
  uint32_t k0 = key_in[ 0 ], k1 = key_in[ 1 ];
  for ( int r = 0; r < rounds; ++r )
  {
    if ( r > 0 )
    {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast< uint64_t >( 0xD2511F53u ) * c[ 0 ];
    const uint64_t p1 = static_cast< uint64_t >( 0xCD9E8D57u ) * c[ 2 ];
    const uint32_t n0 = static_cast< uint32_t >( p1 >> 32 ) ^ c[ 1 ] ^ k0;
    const uint32_t n1 = static_cast< uint32_t >( p1 );
    const uint32_t n2 = static_cast< uint32_t >( p0 >> 32 ) ^ c[ 3 ] ^ k1;
    const uint32_t n3 = static_cast< uint32_t >( p0 );
    c[ 0 ] = n0;
    c[ 1 ] = n1;
    c[ 2 ] = n2;
    c[ 3 ] = n3;
  }
  for ( int i = 0; i < 4; ++i )
    out[ i ] = c[ i ];
}

// r123::Engine< r123::Philox4x32 > (conventional/Engine.hpp) as a uniform random bit
// generator: the key is two words of the seed sequence, the counter starts at zero and is
// incremented (with carry) before each block; a block's four words are returned last word
// first.
struct PhiloxEngine
{
  typedef uint32_t result_type;
  uint32_t key[ 2 ] = { 0, 0 };
  uint32_t ctr[ 4 ] = { 0, 0, 0, 0 };
  uint32_t v[ 4 ] = { 0, 0, 0, 0 };
  int left = 0;  // words of the current block not yet returned
  static constexpr result_type
  min()
  {
    return 0;
  }
  static constexpr result_type
  max()
  {
    return 0xffffffffu;
  }
  template < typename SeedSeq >
  void
  seed( SeedSeq& s )
  {
    s.generate( key, key + 2 );
    for ( auto& c : ctr )
      c = 0;
    left = 0;
  }
  result_type
  operator()()
  {
    if ( left == 0 )
    {
      for ( int i = 0; i < 4 and ++ctr[ i ] == 0; ++i )
      {
      }
      philox4x32( 10, ctr, key, v );
      left = 4;
    }
    return v[ --left ];
  }
};

enum RngKind
{
  RNG_MT19937_64 = 0,
  RNG_PHILOX_32 = 1
};

// nestkernel/random_generators.h:141-297: engine + the three draws the builders use; the
// engine is mt19937_64 (NEST's default rng_type) or Philox_32 (random_manager.cpp:69).
struct Rng
{
  int kind = RNG_MT19937_64;
  std::mt19937_64 eng;
  PhiloxEngine ph;
  std::uniform_int_distribution< unsigned long > ul;
  std::uniform_real_distribution<> u01 { 0.0, 1.0 };
  explicit Rng( const std::vector< uint32_t >& seed, int k = RNG_MT19937_64 )
    : kind( k )
  {
    SeedSeqFe128 sseq( seed );
    if ( kind == RNG_PHILOX_32 )
      ph.seed( sseq );
    else
      eng.seed( sseq );
  }
  // runs f on the engine in use (the std distributions are templates over it)
  template < typename F >
  auto
  with( F&& f )
  {
    return kind == RNG_PHILOX_32 ? f( ph ) : f( eng );
  }
  double
  drand()
  {
    return with( [ this ]( auto& e ) { return u01( e ); } );
  }
  unsigned long
  ulrand( unsigned long n )
  {
    std::uniform_int_distribution< unsigned long >::param_type p( 0, n - 1 );
    return with( [ & ]( auto& e ) { return ul( e, p ); } );
  }
};

// ------------------------------------------------------------------ parameters
// Prefix-encoded expression tree over doubles; op codes:
//   0 const v | 1 uniform min max | 2 uniform_int max | 3 normal mean std
//   4 sum A B | 5 product A B     | 6 max A v          | 7 min A v | 8 redraw A min max
//   9 lognormal mean std | 10 exponential beta | 18 normal id mean std | 19 lognormal id mean std
//   11 exp A | 12 sin A | 13 cos A | 14 quotient A B | 15 pow A v
//   16 compare A B comparator | 17 conditional C T F
// nestkernel/parameter.h:131-286 (leaves), 468-600 (product/sum), 878-990 (max/redraw)
// A parameter object keeps its per-VP distributions across Connect calls: NormalParameter and
// LognormalParameter hold one std::normal_distribution / lognormal_distribution per thread
// (parameter.cpp:95-157), whose cached second polar value carries over to the next call that
// uses the same object. Ops 18 and 19 name such an object by id; the registry holds its state.
struct ParamStates
{
  std::map< long, std::shared_ptr< std::vector< std::normal_distribution<> > > > normals;
  std::map< long, std::shared_ptr< std::vector< std::lognormal_distribution<> > > > lognormals;
};

struct Param
{
  int op = 0;
  double a = 0, b = 0;
  std::unique_ptr< Param > l, r, f;
  // one per VP (parameter.cpp:95-117, 120-157), shared by every use of the same object
  std::shared_ptr< std::vector< std::normal_distribution<> > > normals;
  std::shared_ptr< std::vector< std::lognormal_distribution<> > > lognormals;

  static std::unique_ptr< Param >
  parse( const double*& p, const double* end, int n_vp, ParamStates* states = nullptr )
  {
    if ( p >= end )
      throw std::runtime_error( "truncated parameter spec" );
    auto x = std::make_unique< Param >();
    x->op = static_cast< int >( *p++ );
    switch ( x->op )
    {
    case 0:
      x->a = *p++;
      break;
    case 1:
      x->a = *p++;
      x->b = *p++;
      if ( x->a >= x->b )
        throw std::runtime_error( "nest::UniformParameter: min < max required." );
      x->b -= x->a;  // range
      break;
    case 2:
      x->a = *p++;
      if ( x->a <= 0 )
        throw std::runtime_error( "nest::UniformIntParameter: max > 0 required." );
      break;
    case 3:
    case 18:
    {
      const long id = x->op == 18 ? static_cast< long >( *p++ ) : -1;
      x->op = 3;
      x->a = *p++;
      x->b = *p++;
      if ( x->b < 0 )
        throw std::runtime_error( "nest::NormalParameter: std >= 0 required." );
      std::shared_ptr< std::vector< std::normal_distribution<> > >* slot =
        id >= 0 and states ? &states->normals[ id ] : nullptr;
      if ( slot and *slot and static_cast< int >( ( *slot )->size() ) == n_vp )
        x->normals = *slot;
      else
      {
        x->normals = std::make_shared< std::vector< std::normal_distribution<> > >();
        if ( x->b > 0 )
          for ( int i = 0; i < n_vp; ++i )
            x->normals->emplace_back( x->a, x->b );
        if ( slot )
          *slot = x->normals;
      }
      break;
    }
    case 4:
    case 5:
      x->l = parse( p, end, n_vp, states );
      x->r = parse( p, end, n_vp, states );
      break;
    case 6:
    case 7:
      x->l = parse( p, end, n_vp, states );
      x->a = *p++;
      break;
    case 8:
      x->l = parse( p, end, n_vp, states );
      x->a = *p++;
      x->b = *p++;
      break;
    case 9:
    case 19:
    {
      const long id = x->op == 19 ? static_cast< long >( *p++ ) : -1;
      x->op = 9;
      x->a = *p++;
      x->b = *p++;
      if ( x->b < 0 )
        throw std::runtime_error( "nest::LognormalParameter: std >= 0 required." );
      std::shared_ptr< std::vector< std::lognormal_distribution<> > >* slot =
        id >= 0 and states ? &states->lognormals[ id ] : nullptr;
      if ( slot and *slot and static_cast< int >( ( *slot )->size() ) == n_vp )
        x->lognormals = *slot;
      else
      {
        x->lognormals = std::make_shared< std::vector< std::lognormal_distribution<> > >();
        if ( x->b > 0 )
          for ( int i = 0; i < n_vp; ++i )
            x->lognormals->emplace_back( x->a, x->b );
        if ( slot )
          *slot = x->lognormals;
      }
      break;
    }
    case 10:
      x->a = *p++;
      if ( x->a < 0 )
        throw std::runtime_error( "nest::ExponentialParameter: beta >= 0 required." );
      break;
    case 11:
    case 12:
    case 13:
      x->l = parse( p, end, n_vp, states );
      break;
    case 14:
      x->l = parse( p, end, n_vp, states );
      x->r = parse( p, end, n_vp, states );
      break;
    case 15:
      x->l = parse( p, end, n_vp, states );
      x->a = *p++;
      break;
    case 16:
      x->l = parse( p, end, n_vp, states );
      x->r = parse( p, end, n_vp, states );
      x->a = *p++;
      break;
    case 17:
      x->l = parse( p, end, n_vp, states );
      x->r = parse( p, end, n_vp, states );
      x->f = parse( p, end, n_vp, states );
      break;
    default:
      throw std::runtime_error( "unknown parameter op" );
    }
    return x;
  }

  double
  value( Rng& rng, int vp )
  {
    switch ( op )
    {
    case 0:
      return a;
    case 1:
      return a + rng.drand() * b;
    case 2:
      return static_cast< double >( rng.ulrand( static_cast< unsigned long >( a ) ) );
    case 3:
      if ( b == 0 )
        return a;
      return rng.with( [ & ]( auto& e ) { return ( *normals )[ vp ]( e ); } );
    case 4:
    {
      const double x = l->value( rng, vp );
      return x + r->value( rng, vp );
    }
    case 5:
    {
      const double x = l->value( rng, vp );
      return x * r->value( rng, vp );
    }
    case 6:
      return std::max( l->value( rng, vp ), a );
    case 7:
      return std::min( l->value( rng, vp ), a );
    case 9:
      // LognormalParameter::value (parameter.cpp:147-157)
      if ( b == 0 )
        return std::exp( a );
      return rng.with( [ & ]( auto& e ) { return ( *lognormals )[ vp ]( e ); } );
    case 10:
      // ExponentialParameter::value (parameter.h:339-342)
      return a * ( -std::log( 1 - rng.drand() ) );
    case 11:
      return std::exp( l->value( rng, vp ) );  // ExpParameter (parameter.h:991)
    case 12:
      return std::sin( l->value( rng, vp ) );  // SinParameter (parameter.h:1037)
    case 13:
      return std::cos( l->value( rng, vp ) );  // CosParameter (parameter.h:1082)
    case 14:
    {
      // QuotientParameter (parameter.h:546)
      const double x = l->value( rng, vp );
      return x / r->value( rng, vp );
    }
    case 15:
      return std::pow( l->value( rng, vp ), a );  // PowParameter (parameter.h:1130)
    case 16:
    {
      // ComparingParameter::value (parameter.h:712-715, 734-752): compare_( p1->value(),
      // p2->value() ) is one call, whose arguments g++ on x86-64 evaluates right to left, so
      // the right operand draws first
      const double y = r->value( rng, vp );
      const double x = l->value( rng, vp );
      const int c = static_cast< int >( a );
      return c == 0 ? x < y : c == 1 ? x <= y : c == 2 ? x == y : c == 3 ? x != y : c == 4 ? x >= y : x > y;
    }
    case 17:
      // ConditionalParameter::value (parameter.h:791-797): the condition, then one branch
      if ( l->value( rng, vp ) )
        return r->value( rng, vp );
      return f->value( rng, vp );
    case 8:
    {
      // nestkernel/parameter.cpp RedrawParameter::value: redraw until inside [min, max]
      for ( int i = 0; i < 1000; ++i )
      {
        const double v = l->value( rng, vp );
        if ( a <= v and v <= b )
          return v;
      }
      throw std::runtime_error( "Number of redraws exceeded limit." );
    }
    }
    return 0;
  }
  bool
  is_const() const
  {
    return op == 0;
  }
};

// ------------------------------------------------------------------ models
enum Model
{
  IAF_PSC_DELTA = 0,
  IAF_PSC_EXP = 1,
  AEIF_COND_EXP = 2,
  IAF_PSC_ALPHA = 3
};

struct Neuron
{
  int model;
  long gid;
  // iaf_psc_delta (values relative to E_L where NEST stores them so)
  double tau_m, c_m, t_ref, E_L, I_e, V_th, V_min, V_reset;
  bool refr_input = false;
  // iaf_psc_exp
  double tau_ex, tau_in;
  // aeif_cond_exp (absolute values)
  double V_peak, g_L, E_ex, E_in, Delta_T, tau_w, a, b, tol;
  // multimeter sampling: samples (stamp, V_m and the three other state fields as the user
  // reads them) taken after each sampled step (UniversalDataLogger::record_data)
  bool sampled = false;
  std::vector< std::array< double, 5 > > samples;
  // ArchivingNode (archiving_node.cpp:38-48): tau_minus, and for a neuron with incoming
  // stdp synapses its post-spike history (time, Kminus) and trace
  double tau_minus = 20.0;
  bool stdp_in = false;
  double Kminus = 0.0, last_spike = -1.0;
  std::vector< std::pair< double, double > > hist;
  // state
  double y0 = 0, y3 = 0, refr_buf = 0;          // delta
  double V = 0, i_ex = 0, i_in = 0, i0 = 0, i1 = 0;  // exp; alpha: V, I_ex, I_in
  double di_ex = 0, di_in = 0;                       // alpha: dI_ex, dI_in
  double ay[ 4 ] = { 0, 0, 0, 0 };               // aeif V, g_ex, g_in, w
  double I_stim = 0, hstep = 0, integ = 0;  // aeif: B_.step_, B_.IntegrationStep_
  long r = 0;
  // propagators (pre_run_hook)
  double P33, P30, P11ex, P11in, P22, P21ex, P21in, P20, V_peak_eff;
  double P31ex, P32ex, P31in, P32in, expm1_tau_m, EPSCInit, IPSCInit;  // alpha
  long RefractoryCounts;
  // ring buffers: ch 0..3 ([SYN_IN, SYN_EX, I0, I1] for exp; [spikes, currents] for delta;
  // [exc, inh, currents] for aeif), slot-major
  std::vector< double > rb;
  bool initialized = false;
};

// IAFPropagator::evaluate_P32_ (libnestutil/iaf_propagator.cpp:42-74)
static double
iaf_propagator_p32( double tau_syn, double tau_m, double c_m, double h )
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

// IAFPropagatorAlpha::evaluate (libnestutil/iaf_propagator.cpp:94-116): (P31, P32), the
// regular form above h_min_regular = 1e-7 tau_m^2 / |tau_m - tau_syn|, else the singular limits
static void
iaf_propagator_alpha( double tau_syn, double tau_m, double c_m, double h, double& P31, double& P32 )
{
  const double h_min_regular = 1e-7 * tau_m * tau_m / std::abs( tau_m - tau_syn );
  const double beta = tau_syn * tau_m / ( tau_m - tau_syn );
  const double gamma = beta / c_m;
  const double inv_tau_syn = 1 / tau_syn;
  const double inv_tau_m = 1 / tau_m;
  const double inv_c_m = 1 / c_m;
  const double inv_beta = ( tau_m - tau_syn ) / ( tau_syn * tau_m );
  const double exp_h_tau_syn = std::exp( -h * inv_tau_syn );
  const double expm1_h_tau = std::expm1( h * inv_beta );
  P32 = gamma * exp_h_tau_syn * expm1_h_tau;
  double exp_h_tau = std::nan( "" );
  if ( not( std::isnormal( P32 ) and P32 > 0 ) )
  {
    exp_h_tau = std::exp( -h * inv_tau_m );
    P32 = h * inv_c_m * exp_h_tau;
  }
  if ( h > h_min_regular )
  {
    P31 = gamma * exp_h_tau_syn * ( beta * expm1_h_tau - h );
    return;
  }
  if ( std::isnan( exp_h_tau ) )
    exp_h_tau = std::exp( -h * inv_tau_m );
  P31 = 0.5 * h * h * inv_c_m * exp_h_tau;
}

// ----------------------------------------------------- portable exp/log (fdlibm)
// Sun fdlibm e_exp.c / e_log.c method (argument reduction by ln2, rational / polynomial
// kernels). Used only in "portable" aeif mode so that a CPU run can be compared bit for bit
// with device code that uses the same published method.
static double
fd_exp( double x )
{
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10,
               invln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
               P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
               P5 = 4.13813679705723846039e-08;
  if ( x != x )
    return x;
  if ( x > 709.782712893383973096 )
    return std::numeric_limits< double >::infinity();
  if ( x < -745.13321910194110842 )
    return 0.0;
  const double kd = std::floor( x * invln2 + 0.5 );
  const int k = static_cast< int >( kd );
  const double hi = x - kd * ln2hi;
  const double lo = kd * ln2lo;
  const double rr = hi - lo;
  const double t = rr * rr;
  const double c = rr - t * ( P1 + t * ( P2 + t * ( P3 + t * ( P4 + t * P5 ) ) ) );
  const double y = 1.0 - ( ( lo - ( rr * c ) / ( 2.0 - c ) ) - hi );
  return std::ldexp( y, k );
}

static double
fd_log( double x )
{
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
  const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
               Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
               Lg7 = 1.479819860511658591e-01;
  if ( !( x > 0 ) )
    return x == 0 ? -std::numeric_limits< double >::infinity() : std::numeric_limits< double >::quiet_NaN();
  if ( x == std::numeric_limits< double >::infinity() )
    return x;
  int e;
  double mnt = std::frexp( x, &e );
  if ( mnt < 0.70710678118654752440 )
  {
    mnt = mnt * 2.0;
    e -= 1;
  }
  const double f = mnt - 1.0;
  const double kd = static_cast< double >( e );
  const double s = f / ( 2.0 + f );
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * ( Lg2 + w * ( Lg4 + w * Lg6 ) );
  const double t2 = z * ( Lg1 + w * ( Lg3 + w * ( Lg5 + w * Lg7 ) ) );
  const double R = t2 + t1;
  const double hfsq = 0.5 * f * f;
  return kd * ln2hi - ( ( hfsq - ( s * ( hfsq + R ) + kd * ln2lo ) ) - f );
}

struct MathMode
{
  bool portable = false;
  double
  exp( double x ) const
  {
    return portable ? fd_exp( x ) : std::exp( x );
  }
  double
  pow( double x, double y ) const
  {
    return portable ? fd_exp( y * fd_log( x ) ) : std::pow( x, y );
  }
};

// aeif_cond_exp_dynamics (models/aeif_cond_exp.cpp:79-124)
static void
aeif_rhs( const Neuron& n, const double* y, double* f, const MathMode& mm )
{
  const bool is_refractory = n.r > 0;
  const double V = is_refractory ? n.V_reset : std::min( y[ 0 ], n.V_peak );
  const double g_ex = y[ 1 ];
  const double g_in = y[ 2 ];
  const double w = y[ 3 ];
  const double I_syn_exc = g_ex * ( V - n.E_ex );
  const double I_syn_inh = g_in * ( V - n.E_in );
  const double I_spike = n.Delta_T == 0. ? 0. : ( n.g_L * n.Delta_T * mm.exp( ( V - n.V_th ) / n.Delta_T ) );
  f[ 0 ] = is_refractory ? 0. : ( -n.g_L * ( V - n.E_L ) + I_spike - I_syn_exc - I_syn_inh - w + n.I_e + n.I_stim ) / n.c_m;
  f[ 1 ] = -g_ex / n.tau_ex;
  f[ 2 ] = -g_in / n.tau_in;
  f[ 3 ] = ( n.a * ( V - n.E_L ) - w ) / n.tau_w;
}

// Runge-Kutta-Fehlberg 4(5): one step of size h from y (k1 = dydt_in given), the 5th order
// solution propagates; yerr is the difference of the embedded 5th and 4th order solutions.
static void
rkf45_step( const Neuron& n, double* y, double h, const double* k1, double* yerr, double* dydt_out, const MathMode& mm )
{
  static const double b21 = 1.0 / 4.0;
  static const double b31 = 3.0 / 32.0, b32 = 9.0 / 32.0;
  static const double b41 = 1932.0 / 2197.0, b42 = -7200.0 / 2197.0, b43 = 7296.0 / 2197.0;
  static const double b51 = 8341.0 / 4104.0, b52 = -32832.0 / 4104.0, b53 = 29440.0 / 4104.0, b54 = -845.0 / 4104.0;
  static const double b61 = -6080.0 / 20520.0, b62 = 41040.0 / 20520.0, b63 = -28352.0 / 20520.0,
                      b64 = 9295.0 / 20520.0, b65 = -5643.0 / 20520.0;
  static const double c1 = 902880.0 / 7618050.0, c3 = 3953664.0 / 7618050.0, c4 = 3855735.0 / 7618050.0,
                      c5 = -1371249.0 / 7618050.0, c6 = 277020.0 / 7618050.0;
  static const double ec1 = 1.0 / 360.0, ec3 = -128.0 / 4275.0, ec4 = -2197.0 / 75240.0, ec5 = 1.0 / 50.0,
                      ec6 = 2.0 / 55.0;
  double k2[ 4 ], k3[ 4 ], k4[ 4 ], k5[ 4 ], k6[ 4 ], yt[ 4 ];
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + b21 * h * k1[ i ];
  aeif_rhs( n, yt, k2, mm );
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b31 * k1[ i ] + b32 * k2[ i ] );
  aeif_rhs( n, yt, k3, mm );
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b41 * k1[ i ] + b42 * k2[ i ] + b43 * k3[ i ] );
  aeif_rhs( n, yt, k4, mm );
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b51 * k1[ i ] + b52 * k2[ i ] + b53 * k3[ i ] + b54 * k4[ i ] );
  aeif_rhs( n, yt, k5, mm );
  for ( int i = 0; i < 4; ++i )
    yt[ i ] = y[ i ] + h * ( b61 * k1[ i ] + b62 * k2[ i ] + b63 * k3[ i ] + b64 * k4[ i ] + b65 * k5[ i ] );
  aeif_rhs( n, yt, k6, mm );
  for ( int i = 0; i < 4; ++i )
  {
    const double di = c1 * k1[ i ] + c3 * k3[ i ] + c4 * k4[ i ] + c5 * k5[ i ] + c6 * k6[ i ];
    y[ i ] = y[ i ] + h * di;
  }
  for ( int i = 0; i < 4; ++i )
  {
    const double ei = ec1 * k1[ i ] + ec3 * k3[ i ] + ec4 * k4[ i ] + ec5 * k5[ i ] + ec6 * k6[ i ];
    yerr[ i ] = h * ei;
  }
  aeif_rhs( n, y, dydt_out, mm );
}

// Step-size control on the derivative ("yp" control: a_y = 0, a_dydt = 1): desired error
// D_i = eps_abs + eps_rel * |h y'_i|; ratio r = max_i |yerr_i| / D_i; shrink when r > 1.1
// by S r^(-1/q) (no more than 5x), grow when r < 0.5 by S r^(-1/(q+1)) (no more than 5x,
// never below 1); S = 0.9 and q = 5, the stepper's order. Returns -1 dec, 0 nil, +1 inc.
static int
yp_hadjust( double tol, const double* yerr, const double* yp, double* h, const MathMode& mm )
{
  const double S = 0.9;
  const double h_old = *h;
  double rmax = std::numeric_limits< double >::min();
  for ( int i = 0; i < 4; ++i )
  {
    const double D0 = tol * ( 0.0 * std::fabs( 0.0 ) + 1.0 * std::fabs( h_old * yp[ i ] ) ) + tol;
    const double r = std::fabs( yerr[ i ] ) / std::fabs( D0 );
    rmax = r > rmax ? r : rmax;  // GSL_MAX_DBL
  }
  if ( rmax > 1.1 )
  {
    double r = S / mm.pow( rmax, 1.0 / 5.0 );
    if ( r < 0.2 )
      r = 0.2;
    *h = r * h_old;
    return -1;
  }
  else if ( rmax < 0.5 )
  {
    double r = S / mm.pow( rmax, 1.0 / ( 5.0 + 1.0 ) );
    if ( r > 5.0 )
      r = 5.0;
    if ( r < 1.0 )
      r = 1.0;
    *h = r * h_old;
    return 1;
  }
  return 0;
}

// One call of the adaptive evolution: advance *t towards t1 by one accepted step.
static void
rkf45_evolve( const Neuron& n, double* t, double t1, double* h, double* y, const MathMode& mm )
{
  const double t0 = *t;
  double h0 = *h;
  const double dt = t1 - t0;
  double y0[ 4 ], dydt_in[ 4 ], yerr[ 4 ], dydt_out[ 4 ];
  std::memcpy( y0, y, sizeof y0 );
  aeif_rhs( n, y, dydt_in, mm );
  for ( ;; )
  {
    bool final_step = false;
    if ( h0 > dt )
    {
      h0 = dt;
      final_step = true;
    }
    rkf45_step( n, y, h0, dydt_in, yerr, dydt_out, mm );
    *t = final_step ? t1 : t0 + h0;
    const double h_old = h0;
    const int st = yp_hadjust( n.tol, yerr, dydt_out, &h0, mm );
    if ( st < 0 )
    {
      const double t_curr = *t;
      const double t_next = ( *t ) + h0;
      if ( std::fabs( h0 ) < std::fabs( h_old ) and t_next != t_curr )
      {
        std::memcpy( y, y0, sizeof y0 );
        continue;
      }
      h0 = h_old;
    }
    break;
  }
  *h = h0;
}

// ------------------------------------------------------------------ kernel
struct Conn
{
  long src, tgt;
  double w;
  long d;             // steps
  int syn = -1;       // stdp parameter set, -1 for static_synapse
  double kplus = 0;   // stdp_synapse Kplus
};

// stdp_synapse parameters (stdp_synapse.h:200-210)
struct StdpPrm
{
  double tau_plus = 20.0, lambda = 0.01, alpha = 1.0, mu_plus = 1.0, mu_minus = 1.0, Wmax = 100.0;
};

struct Drive
{
  long gid0, n, n_steps;  // targets gid0..gid0+n-1, emission steps 0..n_steps-1
  std::vector< uint8_t > counts;  // [step][i]
  double w;
  long d;  // steps
};

struct Kernel
{
  TimeBase tb;
  int n_vp = 1;
  uint32_t seed = 143202461;
  std::vector< Rng > vp_rng;
  long next_gid = 1;
  std::vector< Neuron > neurons;
  std::vector< long > gid2idx;  // -1 for devices
  std::vector< std::vector< Conn > > conns;  // per VP, creation order
  std::vector< Drive > drives;
  ParamStates param_states;  // parameter objects used by more than one Connect call
  long min_delay = 0, max_delay = 0;
  long d_min_seen = std::numeric_limits< long >::max(), d_max_seen = 1;
  bool simulated = false, prepared = false;
  long clock = 0, slice = 0, from_step = 0, to_do = 0, L = 0;
  int n_threads = 1;
  MathMode mm;
  std::string err;

  // receive buffer (last gathered slice) and per-VP emission registers: (gid, lag)
  std::vector< std::pair< long, long > > recv;
  std::vector< std::vector< std::pair< long, long > > > reg;
  // spike record: (gid, step)
  std::vector< std::pair< long, long > > spikes;

  // delivery store: per owner partition, CSR over gid of (tgt idx, d, w), creation order
  struct Part
  {
    std::vector< long > ptr;
    std::vector< int32_t > tgt;
    std::vector< int32_t > d;
    std::vector< double > w;
    // stdp_synapse state per position, and the connection each position holds
    std::vector< int32_t > syn;
    std::vector< double > kplus, t_last;
    std::vector< std::pair< int, long > > ref;
  };
  std::vector< Part > parts;
  bool store_valid = false;
  std::vector< StdpPrm > stdp_prm;
  int next_syn = -1;          // synapse of the next Connect calls: -1 static, else stdp set
  double next_kplus = 0.0;
  static constexpr double stdp_eps = 1.0e-6;  // ConnectionManager::stdp_eps_
  // drive lookup per neuron idx: (drive index, position) or -1
  std::vector< std::pair< int, long > > drive_of;

  explicit Kernel( int nvp, uint32_t sd, double res )
    : n_vp( nvp )
    , seed( sd )
  {
    if ( res > 0 )
      tb.set_resolution( res );
    seed_rngs( RNG_MT19937_64 );
    conns.resize( n_vp );
    reg.resize( n_vp );
    gid2idx.push_back( -1 );  // gid 0 unused
  }

  // per-VP streams seeded with {rng_seed, THREAD_SPECIFIC_SEEDER_, vp} (random_manager.cpp:95)
  void
  seed_rngs( int kind )
  {
    vp_rng.clear();
    for ( int vp = 0; vp < n_vp; ++vp )
      vp_rng.emplace_back( std::vector< uint32_t > { seed, 0xb84c9baeu, static_cast< uint32_t >( vp ) }, kind );
  }

  int
  vp_of( long gid ) const
  {
    return static_cast< int >( gid % n_vp );
  }

  Neuron&
  nrn( long gid )
  {
    if ( gid <= 0 or gid >= static_cast< long >( gid2idx.size() ) or gid2idx[ gid ] < 0 )
      throw std::runtime_error( "unknown neuron gid" );
    return neurons[ gid2idx[ gid ] ];
  }

  // Parameters_::set semantics (E_L first, then values relative to E_L)
  static void
  set_params( Neuron& n, const std::map< std::string, double >& d )
  {
    auto has = [ &d ]( const char* k ) { return d.count( k ) > 0; };
    auto get = [ &d ]( const char* k ) { return d.at( k ); };
    if ( has( "tau_minus" ) )
    {
      if ( get( "tau_minus" ) <= 0 )
        throw std::runtime_error( "BadProperty: tau_minus must be positive." );
      n.tau_minus = get( "tau_minus" );
    }
    if ( n.model == AEIF_COND_EXP )
    {
      const char* keys[] = { "V_peak", "V_reset", "t_ref", "g_L", "C_m", "E_ex", "E_in", "E_L", "Delta_T", "tau_w",
        "a", "b", "V_th", "tau_syn_ex", "tau_syn_in", "I_e", "gsl_error_tol" };
      double* dst[] = { &n.V_peak, &n.V_reset, &n.t_ref, &n.g_L, &n.c_m, &n.E_ex, &n.E_in, &n.E_L, &n.Delta_T,
        &n.tau_w, &n.a, &n.b, &n.V_th, &n.tau_ex, &n.tau_in, &n.I_e, &n.tol };
      for ( size_t i = 0; i < sizeof( keys ) / sizeof( keys[ 0 ] ); ++i )
        if ( has( keys[ i ] ) )
          *dst[ i ] = get( keys[ i ] );
      if ( has( "V_m" ) )
        n.ay[ 0 ] = get( "V_m" );
      if ( has( "g_ex" ) )
        n.ay[ 1 ] = get( "g_ex" );
      if ( has( "g_in" ) )
        n.ay[ 2 ] = get( "g_in" );
      if ( has( "w" ) )
        n.ay[ 3 ] = get( "w" );
      if ( n.V_reset >= n.V_peak )
        throw std::runtime_error( "Ensure that: V_reset < V_peak ." );
      if ( n.c_m <= 0 )
        throw std::runtime_error( "Ensure that C_m > 0" );
      if ( n.t_ref < 0 )
        throw std::runtime_error( "Ensure that t_ref >= 0" );
      if ( n.tau_ex <= 0 or n.tau_in <= 0 or n.tau_w <= 0 )
        throw std::runtime_error( "All time constants must be strictly positive." );
      if ( n.tol <= 0 )
        throw std::runtime_error( "The gsl_error_tol must be strictly positive." );
      return;
    }
    const double ELold = n.E_L;
    if ( has( "E_L" ) )
      n.E_L = get( "E_L" );
    const double delta_EL = n.E_L - ELold;
    if ( has( "V_reset" ) )
      n.V_reset = get( "V_reset" ) - n.E_L;
    else
      n.V_reset -= delta_EL;
    if ( has( "V_th" ) )
      n.V_th = get( "V_th" ) - n.E_L;
    else
      n.V_th -= delta_EL;
    if ( n.model == IAF_PSC_DELTA or n.model == IAF_PSC_ALPHA )
    {
      if ( has( "V_min" ) )
        n.V_min = get( "V_min" ) - n.E_L;
      else
        n.V_min -= delta_EL;
    }
    if ( has( "I_e" ) )
      n.I_e = get( "I_e" );
    if ( has( "C_m" ) )
      n.c_m = get( "C_m" );
    if ( has( "tau_m" ) )
      n.tau_m = get( "tau_m" );
    if ( n.model == IAF_PSC_EXP or n.model == IAF_PSC_ALPHA )
    {
      if ( has( "tau_syn_ex" ) )
        n.tau_ex = get( "tau_syn_ex" );
      if ( has( "tau_syn_in" ) )
        n.tau_in = get( "tau_syn_in" );
    }
    if ( has( "t_ref" ) )
      n.t_ref = get( "t_ref" );
    if ( n.V_reset >= n.V_th )
      throw std::runtime_error( "Reset potential must be smaller than threshold." );
    if ( n.c_m <= 0 )
      throw std::runtime_error( "Capacitance must be > 0." );
    if ( n.t_ref < 0 )
      throw std::runtime_error( "Refractory time must not be negative." );
    if ( n.tau_m <= 0 or ( n.model != IAF_PSC_DELTA and ( n.tau_ex <= 0 or n.tau_in <= 0 ) ) )
      throw std::runtime_error( "Membrane and synapse time constants must be strictly positive." );
    if ( n.model == IAF_PSC_DELTA and has( "refractory_input" ) )
      n.refr_input = get( "refractory_input" ) != 0;
    // State_::set
    if ( n.model == IAF_PSC_DELTA )
    {
      if ( has( "V_m" ) )
        n.y3 = get( "V_m" ) - n.E_L;
      else
        n.y3 -= delta_EL;
    }
    else
    {
      if ( has( "V_m" ) )
        n.V = get( "V_m" ) - n.E_L;
      else
        n.V -= delta_EL;
      if ( has( "I_syn_ex" ) )
        n.i_ex = get( "I_syn_ex" );
      if ( has( "I_syn_in" ) )
        n.i_in = get( "I_syn_in" );
    }
  }

  static Neuron
  default_neuron( int model )
  {
    Neuron n {};
    n.model = model;
    if ( model == IAF_PSC_DELTA )  // models/iaf_psc_delta.cpp:69-88
    {
      n.tau_m = 10.0;
      n.c_m = 250.0;
      n.t_ref = 2.0;
      n.E_L = -70.0;
      n.I_e = 0.0;
      n.V_th = -55.0 - n.E_L;
      n.V_min = -std::numeric_limits< double >::max();
      n.V_reset = -70.0 - n.E_L;
    }
    else if ( model == IAF_PSC_EXP or model == IAF_PSC_ALPHA )  // iaf_psc_exp.cpp:71-94, iaf_psc_alpha.cpp:68-80
    {
      n.tau_m = 10.0;
      n.c_m = 250.0;
      n.t_ref = 2.0;
      n.E_L = -70.0;
      n.I_e = 0.0;
      n.V_th = -55.0 - n.E_L;
      n.V_reset = -70.0 - n.E_L;
      n.tau_ex = 2.0;
      n.tau_in = 2.0;
      if ( model == IAF_PSC_ALPHA )
        n.V_min = -std::numeric_limits< double >::infinity();
    }
    else  // models/aeif_cond_exp.cpp:132-151
    {
      n.V_peak = 0.0;
      n.V_reset = -60.0;
      n.t_ref = 0.0;
      n.g_L = 30.0;
      n.c_m = 281.0;
      n.E_ex = 0.0;
      n.E_in = -85.0;
      n.E_L = -70.6;
      n.Delta_T = 2.0;
      n.tau_w = 144.0;
      n.a = 4.0;
      n.b = 80.5;
      n.V_th = -50.4;
      n.tau_ex = 0.2;
      n.tau_in = 2.0;
      n.I_e = 0.0;
      n.tol = 1e-6;
      n.ay[ 0 ] = n.E_L;
    }
    return n;
  }

  long
  create( int model, long n, const std::map< std::string, double >& params )
  {
    const long first = next_gid;
    for ( long i = 0; i < n; ++i )
    {
      Neuron x = default_neuron( model );
      x.gid = next_gid;
      set_params( x, params );
      gid2idx.push_back( static_cast< long >( neurons.size() ) );
      neurons.push_back( x );
      ++next_gid;
    }
    store_valid = false;
    return first;
  }

  void
  check_delay( long d_steps )
  {
    if ( d_steps < 1 )
      throw std::runtime_error( "BadDelay: Delay must be greater than or equal to resolution" );
    if ( simulated and ( d_steps < min_delay or d_steps > max_delay ) )
      throw std::runtime_error( "BadDelay: Minimum and maximum delay cannot be changed after Simulate has been called." );
    d_min_seen = std::min( d_min_seen, d_steps );
    d_max_seen = std::max( d_max_seen, d_steps );
  }

  // BipartiteConnBuilder::single_connect_ (conn_builder.cpp:379-437): delay before weight.
  // The delay range of the VP's new connections goes to lo/hi (merged after the VPs join).
  void
  single_connect( long s, long t, int vp, Rng& rng, Param* wp, Param* dp, long& lo, long& hi )
  {
    double delay = 1.0, weight = 1.0;  // static_synapse defaults
    if ( wp and dp )
    {
      delay = dp->value( rng, vp );
      weight = wp->value( rng, vp );
    }
    else if ( dp )
      delay = dp->value( rng, vp );
    else if ( wp )
      weight = wp->value( rng, vp );
    const long ds = tb.delay_ms_to_steps( delay );
    if ( ds < 1 )
      throw std::runtime_error( "BadDelay: Delay must be greater than or equal to resolution" );
    lo = std::min( lo, ds );
    hi = std::max( hi, ds );
    if ( next_syn >= 0 )
    {
      // stdp_synapse::set_status (stdp_synapse.h:303-314): weight and Wmax of the same sign
      const double Wmax = stdp_prm[ next_syn ].Wmax;
      if ( not( ( ( weight >= 0 ) - ( weight < 0 ) ) == ( ( Wmax >= 0 ) - ( Wmax < 0 ) ) ) )
        throw std::runtime_error( "BadProperty: Weight and Wmax must have same sign." );
    }
    conns[ vp ].push_back( Conn { s, t, weight, ds, next_syn, next_kplus } );
  }

  void
  connect( const long* pre, long n_pre, const long* post, long n_post, int rule, double rule_param, bool autapses,
    bool multapses, const double* ws, int wn, const double* ds, int dn )
  {
    std::unique_ptr< Param > wp, dp;
    if ( wn > 0 )
    {
      const double* p = ws;
      wp = Param::parse( p, ws + wn, n_vp, &param_states );
    }
    if ( dn > 0 )
    {
      const double* p = ds;
      dp = Param::parse( p, ds + dn, n_vp, &param_states );
    }
    for ( long i = 0; i < n_post; ++i )
      nrn( post[ i ] );
    for ( long i = 0; i < n_pre; ++i )
      nrn( pre[ i ] );
    if ( rule == 0 and n_pre != n_post )
      throw std::runtime_error( "one_to_one connect requires equal number of sources and targets" );
    if ( rule == 2 and n_pre == 0 )
      throw std::runtime_error( "Source array must not be empty." );
    if ( rule == 2 and rule_param < 0 )
      throw std::runtime_error( "Indegree cannot be less than zero." );
    if ( rule == 2 and not multapses and static_cast< long >( rule_param ) > n_pre )
      throw std::runtime_error( "Indegree cannot be larger than population size." );
    if ( rule == 3 and ( rule_param < 0 or rule_param > 1 ) )
      throw std::runtime_error( "Connection probability 0 <= p <= 1 required." );
    // Every VP draws from its own stream, appends to its own connection list and keeps its
    // own normal-draw state in each parameter object, so the VPs run in parallel; the delay
    // extrema and the first error (in VP order) are merged after the join.
    std::vector< long > vlo( n_vp, std::numeric_limits< long >::max() ), vhi( n_vp, 0 );
    std::vector< std::string > verr( n_vp );
#pragma omp parallel for num_threads( std::max( 1, std::min( omp_get_num_procs(), n_vp ) ) ) schedule( dynamic, 1 )
    for ( int vp = 0; vp < n_vp; ++vp )
    try
    {
      Rng& rng = vp_rng[ vp ];
      long& lo = vlo[ vp ];
      long& hi = vhi[ vp ];
      for ( long ti = 0; ti < n_post; ++ti )
      {
        const long t = post[ ti ];
        if ( vp_of( t ) != vp )
          continue;
        if ( rule == 0 )
        {
          const long s = pre[ ti ];
          if ( s == t and not autapses )
            continue;
          single_connect( s, t, vp, rng, wp.get(), dp.get(), lo, hi );
        }
        else if ( rule == 1 )
        {
          for ( long si = 0; si < n_pre; ++si )
          {
            if ( pre[ si ] == t and not autapses )
              continue;
            single_connect( pre[ si ], t, vp, rng, wp.get(), dp.get(), lo, hi );
          }
        }
        else if ( rule == 2 )
        {
          const long K = static_cast< long >( std::round( rule_param ) );
          std::vector< long > chosen;
          for ( long j = 0; j < K; ++j )
          {
            unsigned long s_id;
            long s;
            bool skip_aut, skip_mult;
            do
            {
              s_id = rng.ulrand( static_cast< unsigned long >( n_pre ) );
              s = pre[ s_id ];
              skip_aut = not autapses and s == t;
              skip_mult = not multapses and std::find( chosen.begin(), chosen.end(), static_cast< long >( s_id ) ) != chosen.end();
            } while ( skip_aut or skip_mult );
            if ( not multapses )
              chosen.push_back( static_cast< long >( s_id ) );
            single_connect( s, t, vp, rng, wp.get(), dp.get(), lo, hi );
          }
        }
        else if ( rule == 3 )
        {
          for ( long si = 0; si < n_pre; ++si )
          {
            const long s = pre[ si ];
            if ( not autapses and s == t )
              continue;
            if ( rng.drand() >= rule_param )
              continue;
            single_connect( s, t, vp, rng, wp.get(), dp.get(), lo, hi );
          }
        }
        else
          throw std::runtime_error( "unknown rule" );
      }
    }
    catch ( const std::exception& e )
    {
      verr[ vp ] = e.what();
    }
    for ( int vp = 0; vp < n_vp; ++vp )
      if ( not verr[ vp ].empty() )
        throw std::runtime_error( verr[ vp ] );
    for ( int vp = 0; vp < n_vp; ++vp )
      if ( vhi[ vp ] > 0 )
      {
        check_delay( vlo[ vp ] );
        check_delay( vhi[ vp ] );
      }
    store_valid = false;
  }

  void
  add_connections( const long* s, const long* t, const double* w, const long* d, long n )
  {
    for ( long i = 0; i < n; ++i )
    {
      nrn( s[ i ] );
      nrn( t[ i ] );
      check_delay( d[ i ] );
      conns[ vp_of( t[ i ] ) ].push_back( Conn { s[ i ], t[ i ], w[ i ], d[ i ] } );
    }
    store_valid = false;
  }

  void
  add_drive( long gid0, long n, long n_steps, const uint8_t* counts, double w, double delay_ms )
  {
    for ( long i = 0; i < n; ++i )
      nrn( gid0 + i );
    const long d = tb.delay_ms_to_steps( delay_ms );
    check_delay( d );
    Drive dr;
    dr.gid0 = gid0;
    dr.n = n;
    dr.n_steps = n_steps;
    dr.counts.assign( counts, counts + n * n_steps );
    dr.w = w;
    dr.d = d;
    drives.push_back( std::move( dr ) );
    next_gid += n;  // the generators occupy node ids like spike_generator nodes do
    gid2idx.resize( next_gid, -1 );
    store_valid = false;
  }

  // ---------------------------------------------------------------- prepare
  void
  prepare()
  {
    if ( not simulated )
    {
      // ConnectionManager::update_delay_extrema_ (connection_manager.cpp:485-530)
      min_delay = d_min_seen == std::numeric_limits< long >::max() ? 1 : d_min_seen;
      max_delay = std::max( d_max_seen, 1L );
    }
    L = min_delay + max_delay;
    const double h = tb.resolution_ms();
    for ( auto& n : neurons )
    {
      if ( not n.initialized )
      {
        n.rb.assign( static_cast< size_t >( L ) * 4, 0.0 );
        n.hstep = h;
        n.integ = h;
        n.initialized = true;
      }
      n.RefractoryCounts = tb.ms_to_steps( n.t_ref );
      if ( n.model == IAF_PSC_DELTA )
      {
        n.P33 = std::exp( -h / n.tau_m );
        n.P30 = 1 / n.c_m * ( 1 - n.P33 ) * n.tau_m;
      }
      else if ( n.model == IAF_PSC_EXP )
      {
        n.P11ex = std::exp( -h / n.tau_ex );
        n.P11in = std::exp( -h / n.tau_in );
        n.P22 = std::exp( -h / n.tau_m );
        n.P21ex = iaf_propagator_p32( n.tau_ex, n.tau_m, n.c_m, h );
        n.P21in = iaf_propagator_p32( n.tau_in, n.tau_m, n.c_m, h );
        n.P20 = n.tau_m / n.c_m * ( 1.0 - n.P22 );
      }
      else if ( n.model == IAF_PSC_ALPHA )
      {
        // iaf_psc_alpha::pre_run_hook (iaf_psc_alpha.cpp:248-301)
        n.P11ex = n.P22 = std::exp( -h / n.tau_ex );  // P22_ex = P11_ex
        n.P11in = std::exp( -h / n.tau_in );
        n.P33 = std::exp( -h / n.tau_m );
        n.expm1_tau_m = std::expm1( -h / n.tau_m );
        n.P30 = -n.tau_m / n.c_m * std::expm1( -h / n.tau_m );
        n.P21ex = h * n.P11ex;
        n.P21in = h * n.P11in;
        iaf_propagator_alpha( n.tau_ex, n.tau_m, n.c_m, h, n.P31ex, n.P32ex );
        iaf_propagator_alpha( n.tau_in, n.tau_m, n.c_m, h, n.P31in, n.P32in );
        n.EPSCInit = 1.0 * 2.718281828459045 / n.tau_ex;  // numerics::e
        n.IPSCInit = 1.0 * 2.718281828459045 / n.tau_in;
      }
      else
        n.V_peak_eff = n.Delta_T > 0. ? n.V_peak : n.V_th;
    }
    if ( not store_valid )
      build_store();
    prepared = true;
  }

  int
  owner( long idx ) const
  {
    return static_cast< int >( idx % n_threads );
  }

  void
  build_store()
  {
    const long G = next_gid;
    parts.assign( n_threads, Part() );
    for ( auto& p : parts )
      p.ptr.assign( G + 1, 0 );
    for ( int vp = 0; vp < n_vp; ++vp )
      for ( const auto& c : conns[ vp ] )
        parts[ owner( gid2idx[ c.tgt ] ) ].ptr[ c.src + 1 ]++;
    for ( auto& p : parts )
    {
      for ( long g = 0; g < G; ++g )
        p.ptr[ g + 1 ] += p.ptr[ g ];
      p.tgt.resize( p.ptr[ G ] );
      p.d.resize( p.ptr[ G ] );
      p.w.resize( p.ptr[ G ] );
      p.syn.resize( p.ptr[ G ] );
      p.kplus.resize( p.ptr[ G ] );
      p.t_last.resize( p.ptr[ G ] );
      p.ref.resize( p.ptr[ G ] );
    }
    std::vector< std::vector< long > > cur( n_threads );
    for ( int i = 0; i < n_threads; ++i )
      cur[ i ].assign( parts[ i ].ptr.begin(), parts[ i ].ptr.end() - 1 );
    // stable: VP order, then creation order within a VP; per target this is creation order
    for ( int vp = 0; vp < n_vp; ++vp )
      for ( size_t ci = 0; ci < conns[ vp ].size(); ++ci )
      {
        const Conn& c = conns[ vp ][ ci ];
        const long ti = gid2idx[ c.tgt ];
        const int o = owner( ti );
        const long pos = cur[ o ][ c.src ]++;
        parts[ o ].tgt[ pos ] = static_cast< int32_t >( ti );
        parts[ o ].d[ pos ] = static_cast< int32_t >( c.d );
        parts[ o ].w[ pos ] = c.w;
        parts[ o ].syn[ pos ] = c.syn;
        parts[ o ].kplus[ pos ] = c.kplus;
        parts[ o ].t_last[ pos ] = 0.0;  // stdp_synapse: t_lastspike_ starts at 0
        parts[ o ].ref[ pos ] = std::make_pair( vp, static_cast< long >( ci ) );
        if ( c.syn >= 0 )
          neurons[ ti ].stdp_in = true;  // ArchivingNode::n_incoming_ > 0
      }
    drive_of.assign( neurons.size(), std::make_pair( -1, 0L ) );
    for ( size_t k = 0; k < drives.size(); ++k )
      for ( long i = 0; i < drives[ k ].n; ++i )
      {
        const long idx = gid2idx[ drives[ k ].gid0 + i ];
        if ( drive_of[ idx ].first >= 0 )
          throw std::runtime_error( "at most one drive generator per neuron is supported" );
        drive_of[ idx ] = std::make_pair( static_cast< int >( k ), i );
      }
    store_valid = true;
  }

  // ------------------------------------------------------------ neuron update
  // handle(): add the weighted event to the target's ring buffer slot (absolute step mod L)
  inline void
  handle( Neuron& n, long abs_step, double w, double mult )
  {
    const size_t slot = static_cast< size_t >( abs_step % L ) * 4;
    if ( n.model == IAF_PSC_DELTA )
      n.rb[ slot + 0 ] += w * mult;
    else if ( n.model == IAF_PSC_EXP or n.model == IAF_PSC_ALPHA )
    {
      // iaf_psc_exp.cpp:351-362, iaf_psc_alpha.cpp:373-383 (SYN_IN = 0, SYN_EX = 1)
      const double s = w * mult;
      n.rb[ slot + ( s > 0 ? 1 : 0 ) ] += s;
    }
    else
    {
      if ( w > 0.0 )
        n.rb[ slot + 0 ] += w * mult;
      else
        n.rb[ slot + 1 ] += -w * mult;
    }
  }

  // Node::update over lags [from, to); spikes are reported through mask bits
  void
  update_neuron( Neuron& n, long origin, long from, long to, uint64_t* mask, long& spikes_local )
  {
    const double h = tb.resolution_ms();
    for ( long lag = from; lag < to; ++lag )
    {
      const size_t slot = static_cast< size_t >( ( origin + lag ) % L ) * 4;
      bool spiked = false;
      if ( n.model == IAF_PSC_DELTA )
      {
        // models/iaf_psc_delta.cpp:272-331
        if ( n.r == 0 )
        {
          const double in = n.rb[ slot ];
          n.rb[ slot ] = 0.0;
          n.y3 = n.P30 * ( n.y0 + n.I_e ) + n.P33 * n.y3 + in;
          if ( n.refr_input and n.refr_buf != 0.0 )
          {
            n.y3 += n.refr_buf;
            n.refr_buf = 0.0;
          }
          n.y3 = ( n.y3 < n.V_min ? n.V_min : n.y3 );
        }
        else
        {
          const double in = n.rb[ slot ];
          n.rb[ slot ] = 0.0;
          if ( n.refr_input )
            n.refr_buf += in * mm.exp( -n.r * h / n.tau_m );  // std::exp unless portable math
          --n.r;
        }
        if ( n.y3 >= n.V_th )
        {
          n.r = n.RefractoryCounts;
          n.y3 = n.V_reset;
          spiked = true;
        }
        n.y0 = n.rb[ slot + 1 ];
        n.rb[ slot + 1 ] = 0.0;
      }
      else if ( n.model == IAF_PSC_EXP )
      {
        // models/iaf_psc_exp.cpp:288-349 (SYN_IN = 0, SYN_EX = 1, I0 = 2, I1 = 3)
        if ( n.r == 0 )
          n.V = n.V * n.P22 + n.i_ex * n.P21ex + n.i_in * n.P21in + ( n.I_e + n.i0 ) * n.P20;
        else
          --n.r;
        n.i_ex *= n.P11ex;
        n.i_in *= n.P11in;
        n.i_ex += ( 1. - n.P11ex ) * n.i1;
        n.i_ex += n.rb[ slot + 1 ];
        n.i_in += n.rb[ slot + 0 ];
        if ( n.V >= n.V_th )
        {
          n.r = n.RefractoryCounts;
          n.V = n.V_reset;
          spiked = true;
        }
        n.i0 = n.rb[ slot + 2 ];
        n.i1 = n.rb[ slot + 3 ];
        n.rb[ slot + 0 ] = n.rb[ slot + 1 ] = n.rb[ slot + 2 ] = n.rb[ slot + 3 ] = 0.0;
      }
      else if ( n.model == IAF_PSC_ALPHA )
      {
        // models/iaf_psc_alpha.cpp:305-370 (SYN_IN = 0, SYN_EX = 1, I0 = 2; i0 = y0)
        if ( n.r == 0 )
        {
          n.V = n.P30 * ( n.i0 + n.I_e ) + n.P31ex * n.di_ex + n.P32ex * n.i_ex + n.P31in * n.di_in
            + n.P32in * n.i_in + n.expm1_tau_m * n.V + n.V;
          n.V = ( n.V < n.V_min ? n.V_min : n.V );
        }
        else
          --n.r;
        n.i_ex = n.P21ex * n.di_ex + n.P22 * n.i_ex;
        n.di_ex *= n.P11ex;
        n.di_ex += n.EPSCInit * n.rb[ slot + 1 ];
        n.i_in = n.P21in * n.di_in + n.P11in * n.i_in;
        n.di_in *= n.P11in;
        n.di_in += n.IPSCInit * n.rb[ slot + 0 ];
        if ( n.V >= n.V_th )
        {
          n.r = n.RefractoryCounts;
          n.V = n.V_reset;
          spiked = true;
        }
        n.i0 = n.rb[ slot + 2 ];
        n.rb[ slot + 0 ] = n.rb[ slot + 1 ] = n.rb[ slot + 2 ] = n.rb[ slot + 3 ] = 0.0;
      }
      else
      {
        // models/aeif_cond_exp.cpp:443-523
        // the device path stops a neuron after 1,000,000 accepted steps within one time
        // step and fails the run; the oracle stops at the same count
        double t = 0.0;
        long n_steps = 0;
        while ( t < n.hstep )
        {
          if ( ++n_steps > 1000000 )
            throw std::runtime_error( "aeif_cond_exp needed more than 1000000 integration steps within one time step" );
          rkf45_evolve( n, &t, n.hstep, &n.integ, n.ay, mm );
          if ( n.ay[ 0 ] < -1e3 or n.ay[ 3 ] < -1e6 or n.ay[ 3 ] > 1e6 )
            throw std::runtime_error( "NumericalInstability: aeif_cond_exp" );
          if ( n.r > 0 )
            n.ay[ 0 ] = n.V_reset;
          else if ( n.ay[ 0 ] >= n.V_peak_eff )
          {
            n.ay[ 0 ] = n.V_reset;
            n.ay[ 3 ] += n.b;
            n.r = n.RefractoryCounts > 0 ? n.RefractoryCounts + 1 : 0;
            if ( spiked )
              throw std::runtime_error( "aeif_cond_exp: more than one spike in one step is not representable" );
            spiked = true;
          }
        }
        if ( n.r > 0 )
          --n.r;
        n.ay[ 1 ] += n.rb[ slot + 0 ];
        n.ay[ 2 ] += n.rb[ slot + 1 ];
        n.I_stim = n.rb[ slot + 2 ];
        n.rb[ slot + 0 ] = n.rb[ slot + 1 ] = n.rb[ slot + 2 ] = 0.0;
      }
      if ( spiked )
      {
        mask[ lag >> 6 ] |= uint64_t( 1 ) << ( lag & 63 );
        ++spikes_local;
      }
      // iaf_psc_exp.cpp:346, iaf_psc_delta.cpp:327, aeif_cond_exp.cpp:521: the logger samples
      // the state after the step at stamps offset + k * interval
      const long T = origin + lag + 1;
      if ( n.sampled and T >= sample_off and ( T - sample_off ) % sample_int == 0 )
      {
        if ( n.model == IAF_PSC_DELTA )
          n.samples.push_back( { static_cast< double >( T ), n.y3 + n.E_L, 0.0, 0.0, 0.0 } );
        else if ( n.model == IAF_PSC_EXP or n.model == IAF_PSC_ALPHA )
          n.samples.push_back( { static_cast< double >( T ), n.V + n.E_L, n.i_ex, n.i_in, 0.0 } );
        else
          n.samples.push_back( { static_cast< double >( T ), n.ay[ 0 ], n.ay[ 1 ], n.ay[ 2 ], n.ay[ 3 ] } );
      }
    }
  }

  long sample_int = 1, sample_off = 0;

  // --------------------------------------------------------------- simulate
  void
  simulate( double t_ms )
  {
    prepare();
    const long steps = tb.ms_to_steps( t_ms );
    run_steps( steps );
  }

  void
  run_steps( long steps )
  {
    if ( steps < 0 )
      throw std::runtime_error( "negative simulation time" );
    to_do += steps;
    if ( to_do == 0 )
      return;
    simulated = true;
    const long nw = ( min_delay + 63 ) / 64;
    std::vector< uint64_t > masks( neurons.size() * nw, 0 );
    long to_step = std::min( from_step + to_do, min_delay );
    // neuron order per VP: ascending gid (node_manager.cpp:188-230)
    std::vector< std::vector< long > > vp_nodes( n_vp );
    for ( size_t i = 0; i < neurons.size(); ++i )
      vp_nodes[ vp_of( neurons[ i ].gid ) ].push_back( static_cast< long >( i ) );
    std::vector< std::string > errs( n_threads );
    while ( to_do > 0 )
    {
      // (a) deliver the previous slice's spikes at the start of a slice
      if ( slice > 0 and from_step == 0 )
        deliver();
      const long origin = clock;
      // (b) update all nodes over [from, to); neurons first, then drive generators. Only the
      //     neurons' own buffers are read, and drive lands at least min_delay ahead.
#pragma omp parallel for num_threads( n_threads ) schedule( static )
      for ( long i = 0; i < static_cast< long >( neurons.size() ); ++i )
      {
        long dummy = 0;
        try
        {
          update_neuron( neurons[ i ], origin, from_step, to_step, &masks[ i * nw ], dummy );
          for ( long w = 0; w < nw; ++w )
          {
            uint64_t m = masks[ i * nw + w ];
            while ( m )
            {
              const long lag = w * 64 + __builtin_ctzll( m );
              m &= m - 1;
              set_spiketime( neurons[ i ], origin + lag + 1 );
            }
          }
          const auto& dv = drive_of[ i ];
          if ( dv.first >= 0 )
          {
            const Drive& dr = drives[ dv.first ];
            for ( long lag = from_step; lag < to_step; ++lag )
            {
              const long e = origin + lag;  // emission step; spike time (e+1)*h
              if ( e >= dr.n_steps )
                break;
              const uint8_t c = dr.counts[ static_cast< size_t >( e ) * dr.n + dv.second ];
              if ( c )
                handle( neurons[ i ], e + dr.d, dr.w, static_cast< double >( c ) );
            }
          }
        }
        catch ( const std::exception& ex )
        {
#ifdef _OPENMP
          errs[ omp_get_thread_num() ] = ex.what();
#else
          errs[ 0 ] = ex.what();
#endif
        }
      }
      for ( auto& e : errs )
        if ( not e.empty() )
          throw std::runtime_error( e );
      // emission registers: per VP, node order, then lag (event_delivery_manager_impl.h:103-118)
      for ( int vp = 0; vp < n_vp; ++vp )
        for ( long i : vp_nodes[ vp ] )
        {
          uint64_t* m = &masks[ i * nw ];
          for ( long w = 0; w < nw; ++w )
            while ( m[ w ] )
            {
              const int b = __builtin_ctzll( m[ w ] );
              m[ w ] &= m[ w ] - 1;
              const long lag = w * 64 + b;
              reg[ vp ].emplace_back( neurons[ i ].gid, lag );
              spikes.emplace_back( neurons[ i ].gid, origin + lag + 1 );
            }
        }
      // (c) gather at the end of a full slice: thread-major collocation
      if ( to_step == min_delay )
      {
        recv.clear();
        for ( int vp = 0; vp < n_vp; ++vp )
        {
          recv.insert( recv.end(), reg[ vp ].begin(), reg[ vp ].end() );
          reg[ vp ].clear();
        }
      }
      // advance_time_ (simulation_manager.cpp:1192-1223)
      to_do -= to_step - from_step;
      if ( to_step == min_delay )
      {
        clock += min_delay;
        ++slice;
        from_step = 0;
      }
      else
        from_step = to_step;
      const long end_sim = from_step + to_do;
      to_step = min_delay < end_sim ? min_delay : end_sim;
    }
  }

  // deliver_events_ (event_delivery_manager.cpp:599-776): every owner walks the whole receive
  // buffer in order and sends each spike along its run of local connections
  void
  deliver()
  {
#pragma omp parallel num_threads( n_threads )
    {
#ifdef _OPENMP
      const int tid = omp_get_thread_num();
#else
      const int tid = 0;
#endif
      const Part& p = parts[ tid ];
      long delivered = 0;
      for ( const auto& sp : recv )
      {
        const long s = sp.first;
        const long lag = sp.second;
        // stamp = clock + lag + 1 - min_delay; slot step = stamp + d - 1
        const long base = clock + lag - min_delay;
        for ( long k = p.ptr[ s ]; k < p.ptr[ s + 1 ]; ++k )
        {
          if ( p.syn[ k ] >= 0 )
            stdp_send( const_cast< Part& >( p ), k, neurons[ p.tgt[ k ] ], base + 1 );
          handle( neurons[ p.tgt[ k ] ], base + p.d[ k ], p.w[ k ], 1.0 );
        }
        delivered += p.ptr[ s + 1 ] - p.ptr[ s ];
      }
#pragma omp atomic
      events_delivered += delivered;
    }
  }
  long events_delivered = 0;

  // Time::get_ms of a step stamp: MS_PER_TIC * tics (nest_time.h:494-505)
  double
  stamp_ms( long step ) const
  {
    return tb.ms_per_tic * static_cast< double >( step * tb.tics_per_step );
  }

  // stdp_synapse::send (stdp_synapse.h:222-283) for the connection at position k of part p,
  // before the event is handled; the spike's stamp is the step `stamp`
  void
  stdp_send( Part& p, long k, Neuron& tgt, long stamp )
  {
    const StdpPrm& P = stdp_prm[ p.syn[ k ] ];
    const double t_spike = stamp_ms( stamp );
    const double dendritic_delay = p.d[ k ] * tb.ms_per_step;  // Time::delay_steps_to_ms
    const double t_last = p.t_last[ k ];
    double w = p.w[ k ];
    const double kplus = p.kplus[ k ];
    auto powm = [ this ]( double x, double m ) { return m == 1.0 ? x : m == 0.0 ? 1.0 : mm.pow( x, m ); };
    // facilitate_ / depress_ (stdp_synapse.h:187-198)
    auto facilitate = [ & ]( double w0, double kp )
    {
      const double norm_w = ( w0 / P.Wmax ) + ( P.lambda * powm( 1.0 - ( w0 / P.Wmax ), P.mu_plus ) * kp );
      return norm_w < 1.0 ? norm_w * P.Wmax : P.Wmax;
    };
    auto depress = [ & ]( double w0, double km )
    {
      const double norm_w = ( w0 / P.Wmax ) - ( P.alpha * P.lambda * powm( w0 / P.Wmax, P.mu_minus ) * km );
      return norm_w > 0.0 ? norm_w * P.Wmax : 0.0;
    };
    // get_history( t_last - d, t_spike - d ): entries with t1 + eps <= t < t2 + eps, oldest
    // first (archiving_node.cpp:165-191)
    const double t1_lim = ( t_last - dendritic_delay ) + stdp_eps;
    const double t2_lim = ( t_spike - dendritic_delay ) + stdp_eps;
    for ( const auto& h : tgt.hist )
      if ( h.first >= t1_lim and h.first < t2_lim )
      {
        const double minus_dt = t_last - ( h.first + dendritic_delay );
        w = facilitate( w, kplus * mm.exp( minus_dt / P.tau_plus ) );
      }
    // get_K_value( t_spike - d ) (archiving_node.cpp:85-113)
    const double t = t_spike - dendritic_delay;
    double K = 0.0;
    for ( long i = static_cast< long >( tgt.hist.size() ) - 1; i >= 0; --i )
      if ( t - tgt.hist[ i ].first > stdp_eps )
      {
        K = tgt.hist[ i ].second * mm.exp( ( tgt.hist[ i ].first - t ) * ( 1. / tgt.tau_minus ) );
        break;
      }
    w = depress( w, K );
    p.w[ k ] = w;
    p.kplus[ k ] = kplus * mm.exp( ( t_last - t_spike ) / P.tau_plus ) + 1.0;
    p.t_last[ k ] = t_spike;
  }

  // ArchivingNode::set_spiketime (archiving_node.cpp:196-230) for a neuron with incoming stdp
  // synapses: the history is never pruned here, which changes no value read from it
  void
  set_spiketime( Neuron& n, long stamp )
  {
    const double t_sp_ms = stamp_ms( stamp );
    if ( n.stdp_in )
    {
      n.Kminus = n.Kminus * mm.exp( ( n.last_spike - t_sp_ms ) * ( 1. / n.tau_minus ) ) + 1.0;
      n.last_spike = t_sp_ms;
      n.hist.emplace_back( t_sp_ms, n.Kminus );
    }
    else
      n.last_spike = t_sp_ms;
  }

  // the current weights of stdp connections back into the connection lists
  void
  sync_weights()
  {
    for ( const Part& p : parts )
      for ( size_t k = 0; k < p.syn.size(); ++k )
        if ( p.syn[ k ] >= 0 )
          conns[ p.ref[ k ].first ][ p.ref[ k ].second ].w = p.w[ k ];
  }

  double
  get_state( long gid, const std::string& f )
  {
    Neuron& n = nrn( gid );
    if ( n.model == IAF_PSC_DELTA )
    {
      if ( f == "V_m" )
        return n.y3 + n.E_L;
      if ( f == "refractory" )
        return static_cast< double >( n.r );
    }
    else if ( n.model == IAF_PSC_EXP or n.model == IAF_PSC_ALPHA )
    {
      if ( f == "V_m" )
        return n.V + n.E_L;
      if ( f == "I_syn_ex" )
        return n.i_ex;
      if ( f == "I_syn_in" )
        return n.i_in;
      if ( f == "refractory" )
        return static_cast< double >( n.r );
    }
    else
    {
      if ( f == "V_m" )
        return n.ay[ 0 ];
      if ( f == "g_ex" )
        return n.ay[ 1 ];
      if ( f == "g_in" )
        return n.ay[ 2 ];
      if ( f == "w" )
        return n.ay[ 3 ];
      if ( f == "refractory" )
        return static_cast< double >( n.r );
      if ( f == "IntegrationStep" )
        return n.integ;
    }
    throw std::runtime_error( "unknown state field " + f );
  }
};

}  // namespace oracle

// ------------------------------------------------------------------ C API
// Plain C entry points for the test suite (ctypes). Errors return nonzero and leave a
// message readable through or_error().
using oracle::Kernel;

extern "C" {

void*
or_new( int n_vp, uint32_t seed, double resolution_ms )
{
  try
  {
    return new Kernel( n_vp, seed, resolution_ms );
  }
  catch ( ... )
  {
    return nullptr;
  }
}

void
or_free( void* k )
{
  delete static_cast< Kernel* >( k );
}

const char*
or_error( void* k )
{
  return static_cast< Kernel* >( k )->err.c_str();
}

#define OR_TRY( ... )                                \
  Kernel* K = static_cast< Kernel* >( k );           \
  try                                                 \
  {                                                   \
    __VA_ARGS__;                                      \
  }                                                   \
  catch ( const std::exception& e )                   \
  {                                                   \
    K->err = e.what();                                \
    return -1;                                        \
  }                                                   \
  return 0;

int
or_set_threads( void* k, int n )
{
  OR_TRY( if ( K->prepared ) throw std::runtime_error( "threads must be set before the first simulate" );
          K->n_threads = n < 1 ? 1 : n; K->store_valid = false; )
}

int
or_set_portable_math( void* k, int on )
{
  OR_TRY( K->mm.portable = on != 0; )
}

long
or_create( void* k, int model, long n, int n_params, const char** keys, const double* vals )
{
  Kernel* K = static_cast< Kernel* >( k );
  try
  {
    std::map< std::string, double > d;
    for ( int i = 0; i < n_params; ++i )
      d[ keys[ i ] ] = vals[ i ];
    return K->create( model, n, d );
  }
  catch ( const std::exception& e )
  {
    K->err = e.what();
    return -1;
  }
}

int
or_set_status( void* k, long gid, int n_params, const char** keys, const double* vals )
{
  OR_TRY( std::map< std::string, double > d; for ( int i = 0; i < n_params; ++i ) d[ keys[ i ] ] = vals[ i ];
          Kernel::set_params( K->nrn( gid ), d ); )
}

int
or_get_state( void* k, const long* gids, long n, const char* field, double* out )
{
  OR_TRY( for ( long i = 0; i < n; ++i ) out[ i ] = K->get_state( gids[ i ], field ); )
}

int
or_connect( void* k,
  const long* pre,
  long n_pre,
  const long* post,
  long n_post,
  int rule,
  double rule_param,
  int autapses,
  int multapses,
  const double* wspec,
  int wn,
  const double* dspec,
  int dn )
{
  OR_TRY( K->connect( pre, n_pre, post, n_post, rule, rule_param, autapses != 0, multapses != 0, wspec, wn, dspec, dn ); )
}

int
or_add_connections( void* k, const long* s, const long* t, const double* w, const long* d, long n )
{
  OR_TRY( K->add_connections( s, t, w, d, n ); )
}

long
or_num_connections( void* k )
{
  Kernel* K = static_cast< Kernel* >( k );
  long n = 0;
  for ( const auto& v : K->conns )
    n += static_cast< long >( v.size() );
  return n;
}

// connections per VP in creation order, VPs ascending
int
or_get_connections( void* k, long* src, long* tgt, double* w, long* d )
{
  OR_TRY( K->sync_weights(); long i = 0; for ( const auto& v
                            : K->conns ) for ( const auto& c
                                               : v ) {
    src[ i ] = c.src;
    tgt[ i ] = c.tgt;
    w[ i ] = c.w;
    d[ i ] = c.d;
    ++i;
  } )
}

int
or_add_drive( void* k, long gid0, long n, long n_steps, const uint8_t* counts, double w, double delay_ms )
{
  OR_TRY( K->add_drive( gid0, n, n_steps, counts, w, delay_ms ); )
}

int
or_simulate( void* k, double t_ms )
{
  OR_TRY( K->simulate( t_ms ); )
}

int
or_run_steps( void* k, long steps )
{
  OR_TRY( K->prepare(); K->run_steps( steps ); )
}

long
or_num_spikes( void* k )
{
  return static_cast< long >( static_cast< Kernel* >( k )->spikes.size() );
}

int
or_get_spikes( void* k, long* gid, long* step )
{
  OR_TRY( for ( size_t i = 0; i < K->spikes.size(); ++i ) {
    gid[ i ] = K->spikes[ i ].first;
    step[ i ] = K->spikes[ i ].second;
  } )
}

long
or_synaptic_events( void* k )
{
  return static_cast< Kernel* >( k )->events_delivered;
}

int
or_clear_spikes( void* k )
{
  OR_TRY( K->spikes.clear(); )
}

// min_delay, max_delay, clock (steps), from_step, slice
int
or_kernel_status( void* k, long* out5 )
{
  OR_TRY( out5[ 0 ] = K->min_delay; out5[ 1 ] = K->max_delay; out5[ 2 ] = K->clock; out5[ 3 ] = K->from_step;
          out5[ 4 ] = K->slice; )
}

// RNG probes for pinning the seeding against the reference's own randutils header
// multimeter sampling of the given neurons every interval steps at stamps offset + k * interval
int
or_set_sampling( void* k, const long* gids, long n, long interval, long offset )
{
  OR_TRY( if ( interval < 1 ) throw std::runtime_error( "interval < 1" ); K->sample_int = interval;
          K->sample_off = offset; for ( long i = 0; i < n; ++i ) K->nrn( gids[ i ] ).sampled = true; )
}

long
or_num_samples( void* k )
{
  Kernel* K = static_cast< Kernel* >( k );
  long c = 0;
  for ( const auto& n : K->neurons )
    c += static_cast< long >( n.samples.size() );
  return c;
}

// samples sorted by (stamp, node id): gid, stamp and the four state fields per sample
int
or_get_samples( void* k, long* gid, long* stamp, double* f4 )
{
  OR_TRY( std::vector< std::tuple< long, long, const double* > > all;
              for ( const auto& n : K->neurons ) for ( const auto& x : n.samples )
                all.emplace_back( static_cast< long >( x[ 0 ] ), n.gid, &x[ 1 ] );
              std::sort( all.begin(), all.end(), []( const auto& a, const auto& b ) {
                return std::get< 0 >( a ) != std::get< 0 >( b ) ? std::get< 0 >( a ) < std::get< 0 >( b )
                                                                : std::get< 1 >( a ) < std::get< 1 >( b );
              } );
              for ( size_t i = 0; i < all.size(); ++i ) {
                gid[ i ] = std::get< 1 >( all[ i ] );
                stamp[ i ] = std::get< 0 >( all[ i ] );
                for ( int f = 0; f < 4; ++f ) f4[ 4 * i + f ] = std::get< 2 >( all[ i ] )[ f ];
              } )
}

int
or_seed_seq_generate( const uint32_t* seed, int n_seed, uint32_t* out, int n_out )
{
  std::vector< uint32_t > s( seed, seed + n_seed );
  oracle::SeedSeqFe128 ss( s );
  ss.generate( out, out + n_out );
  return 0;
}

// synapse model of the next Connect calls: kind 0 static_synapse, 1 stdp_synapse with the
// parameters (tau_plus, lambda, alpha, mu_plus, mu_minus, Wmax) and the initial Kplus
int
or_set_synapse( void* k, int kind, const double* prm, double kplus )
{
  OR_TRY( if ( kind == 0 ) K->next_syn = -1; else {
    if ( kplus < 0 ) throw std::runtime_error( "BadProperty: Kplus must be non-negative." );
    oracle::StdpPrm P;
    P.tau_plus = prm[ 0 ];
    P.lambda = prm[ 1 ];
    P.alpha = prm[ 2 ];
    P.mu_plus = prm[ 3 ];
    P.mu_minus = prm[ 4 ];
    P.Wmax = prm[ 5 ];
    K->stdp_prm.push_back( P );
    K->next_syn = static_cast< int >( K->stdp_prm.size() ) - 1;
    K->next_kplus = kplus;
  } )
}

// rng_type: 0 mt19937_64, 1 Philox_32; reseeds the per-VP streams (before any draw)
int
or_set_rng_type( void* k, int kind )
{
  OR_TRY( if ( kind != oracle::RNG_MT19937_64 and kind != oracle::RNG_PHILOX_32 ) throw std::runtime_error(
            "unknown rng type" );
          K->seed_rngs( kind ); )
}

int
or_philox4x32( int rounds, const uint32_t* ctr, const uint32_t* key, uint32_t* out )
{
  oracle::philox4x32( rounds, ctr, key, out );
  return 0;
}

// n draws of one family from a stream seeded with the given words: 0 engine words,
// 1 drand, 2 ulrand(range), 3 normal(0, 1) (std::normal_distribution, one object)
int
or_rng_draws( int kind, const uint32_t* seed, int n_seed, int family, unsigned long range, int n, double* out )
{
  oracle::Rng r( std::vector< uint32_t >( seed, seed + n_seed ), kind );
  std::normal_distribution<> nd( 0.0, 1.0 );
  for ( int i = 0; i < n; ++i )
  {
    if ( family == 0 )
      out[ i ] = static_cast< double >( r.with( []( auto& e ) { return static_cast< uint64_t >( e() ); } ) );
    else if ( family == 1 )
      out[ i ] = r.drand();
    else if ( family == 2 )
      out[ i ] = static_cast< double >( r.ulrand( range ) );
    else
      out[ i ] = r.with( [ & ]( auto& e ) { return nd( e ); } );
  }
  return 0;
}

int
or_mt64_first( const uint32_t* seed, int n_seed, uint64_t* out, int n_out )
{
  oracle::Rng r( std::vector< uint32_t >( seed, seed + n_seed ) );
  for ( int i = 0; i < n_out; ++i )
    out[ i ] = r.eng();
  return 0;
}

double
or_fd_exp( double x )
{
  return oracle::fd_exp( x );
}

double
or_fd_log( double x )
{
  return oracle::fd_log( x );
}

double
or_iaf_propagator( double tau_syn, double tau_m, double c_m, double h )
{
  return oracle::iaf_propagator_p32( tau_syn, tau_m, c_m, h );
}

}  // extern "C"
