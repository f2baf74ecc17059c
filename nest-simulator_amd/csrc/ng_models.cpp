// ng_models.cpp -- the host halves of the built-in neuron models and the model registry.
//
// iaf_psc_delta (models/iaf_psc_delta.cpp:69-270), iaf_psc_exp (models/iaf_psc_exp.cpp:71-286),
// aeif_cond_exp (models/aeif_cond_exp.cpp:132-435) and iaf_psc_alpha
// (models/iaf_psc_alpha.cpp:68-301): defaults, Parameters_::set /
// State_::set with the reference's checks and messages, get_status keys, recordables and
// pre_run_hook. The iaf models keep voltages relative to E_L, as the reference does.
#include <cmath>
#include <limits>
#include <stdexcept>

#include "ng_model.h"

namespace nestgpu
{

namespace
{

// reads the keys of one status dictionary, recording which ones were read
struct Keys
{
  const Json& d;
  std::set< std::string >& used;
  bool
  has( const char* k ) const
  {
    if ( d.has( k ) )
    {
      used.insert( k );
      return true;
    }
    return false;
  }
  double
  get( const char* k ) const
  {
    return d.at( k ).number();
  }
};

// Propagator P32 of IAFPropagatorExp::evaluate (libnestutil/iaf_propagator.cpp:42-74)
double
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

// IAFPropagatorAlpha::evaluate (libnestutil/iaf_propagator.cpp:94-116): P31 and P32, the
// regular form above h_min_regular = 1e-7 tau_m^2 / |tau_m - tau_syn|
// (iaf_propagator.h:62-66), else their singular limits
void
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
  double exp_h_tau = std::numeric_limits< double >::quiet_NaN();
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

// the parts the iaf models share: voltages relative to E_L, set in E_L-first order
class IafModel : public NodeModel
{
public:
  using NodeModel::NodeModel;

  bool
  get_param( const NeuronParams& p, const std::string& key, double& v ) const override
  {
    if ( key == "E_L" )
      v = p.E_L;
    else if ( key == "V_th" )
      v = p.V_th + p.E_L;
    else if ( key == "V_reset" )
      v = p.V_reset + p.E_L;
    else if ( key == "C_m" )
      v = p.c_m;
    else if ( key == "tau_m" )
      v = p.tau_m;
    else if ( key == "t_ref" )
      v = p.t_ref;
    else if ( key == "I_e" )
      v = p.I_e;
    else
      return false;
    return true;
  }
  double
  state_offset( const NeuronParams& p, int field ) const override
  {
    return field == 0 ? p.E_L : 0.0;  // V_m is kept relative to E_L
  }

protected:
  // the keys both models read, applied to q after E_L; returns the change of E_L
  static double
  set_common( NeuronParams& q, const Keys& k )
  {
    const double ELold = q.E_L;
    if ( k.has( "E_L" ) )
      q.E_L = k.get( "E_L" );
    const double delta_EL = q.E_L - ELold;
    if ( k.has( "V_reset" ) )
      q.V_reset = k.get( "V_reset" ) - q.E_L;
    else
      q.V_reset -= delta_EL;
    if ( k.has( "V_th" ) )
      q.V_th = k.get( "V_th" ) - q.E_L;
    else
      q.V_th -= delta_EL;
    return delta_EL;
  }
  static void
  set_membrane( NeuronParams& q, const Keys& k )
  {
    if ( k.has( "I_e" ) )
      q.I_e = k.get( "I_e" );
    if ( k.has( "C_m" ) )
      q.c_m = k.get( "C_m" );
    if ( k.has( "tau_m" ) )
      q.tau_m = k.get( "tau_m" );
  }
  static void
  set_v_m( const NeuronParams& p, NeuronState& s, const Keys& k, double delta_EL )
  {
    if ( k.has( "V_m" ) )
      s.v0 = k.get( "V_m" ) - p.E_L;
    else
      s.v0 -= delta_EL;
  }
};

class IafPscDelta : public IafModel
{
public:
  IafPscDelta()
    : IafModel( "iaf_psc_delta", M_IAF_PSC_DELTA, 1 )
  {
  }
  void
  defaults( NeuronParams& p, NeuronState& s ) const override
  {
    // models/iaf_psc_delta.cpp:69-88
    p.tau_m = 10.0;
    p.c_m = 250.0;
    p.t_ref = 2.0;
    p.E_L = -70.0;
    p.I_e = 0.0;
    p.V_th = -55.0 - p.E_L;
    p.V_min = -std::numeric_limits< double >::max();
    p.V_reset = -70.0 - p.E_L;
    s = NeuronState {};
  }
  void
  set_status( NeuronParams& p, NeuronState& s, const Json& d, std::set< std::string >& used ) const override
  {
    const Keys k { d, used };
    NeuronParams q = p;
    const double delta_EL = set_common( q, k );
    if ( k.has( "V_min" ) )
      q.V_min = k.get( "V_min" ) - q.E_L;
    else
      q.V_min -= delta_EL;
    set_membrane( q, k );
    if ( k.has( "t_ref" ) )
      q.t_ref = k.get( "t_ref" );
    if ( q.V_reset >= q.V_th )
      throw std::runtime_error( "BadProperty: Reset potential must be smaller than threshold." );
    // iaf_psc_delta::Parameters_::set (iaf_psc_delta.cpp:148-163)
    if ( q.c_m <= 0 )
      throw std::runtime_error( "BadProperty: Capacitance must be > 0." );
    if ( q.t_ref < 0 )
      throw std::runtime_error( "BadProperty: Refractory time must not be negative." );
    if ( q.tau_m <= 0 )
      throw std::runtime_error( "BadProperty: Membrane time constant must be > 0." );
    if ( k.has( "refractory_input" ) )
      q.refr_input = d.at( "refractory_input" ).number() != 0;
    p = q;
    set_v_m( p, s, k, delta_EL );
  }
  bool
  get_param( const NeuronParams& p, const std::string& key, double& v ) const override
  {
    if ( key == "refractory_input" )
      v = p.refr_input;
    else if ( key == "V_min" )
      v = p.V_min + p.E_L;
    else
      return IafModel::get_param( p, key, v );
    return true;
  }
  int
  state_field( const std::string& key ) const override
  {
    return key == "V_m" ? 0 : -1;
  }
  std::vector< std::string >
  recordables() const override
  {
    return { "V_m" };
  }
  void
  pre_run_hook( const NeuronParams& p, double h, ParamSet& ps ) const override
  {
    // iaf_psc_delta::pre_run_hook (iaf_psc_delta.cpp:234-270)
    const double P33 = std::exp( -h / p.tau_m );
    const double P30 = 1 / p.c_m * ( 1 - P33 ) * p.tau_m;
    ps.p[ 0 ] = P30;
    ps.p[ 1 ] = P33;
    ps.p[ 2 ] = 0.0 + p.I_e;
    ps.p[ 3 ] = p.V_th;
    ps.p[ 4 ] = p.V_reset;
    ps.p[ 5 ] = p.V_min;
    // refractory_input: input arriving while refractory decays by exp(-r h / tau_m) and is
    // added when the neuron becomes active again (iaf_psc_delta.cpp:284-299)
    ps.p[ 6 ] = p.refr_input ? 1.0 : 0.0;
    ps.p[ 7 ] = h;
    ps.p[ 8 ] = p.tau_m;
  }
};

class IafPscExp : public IafModel
{
public:
  IafPscExp()
    : IafModel( "iaf_psc_exp", M_IAF_PSC_EXP, 2 )
  {
  }
  void
  defaults( NeuronParams& p, NeuronState& s ) const override
  {
    // models/iaf_psc_exp.cpp:71-94
    p.tau_m = 10.0;
    p.c_m = 250.0;
    p.t_ref = 2.0;
    p.E_L = -70.0;
    p.I_e = 0.0;
    p.V_th = -55.0 - p.E_L;
    p.V_reset = -70.0 - p.E_L;
    p.tau_ex = 2.0;
    p.tau_in = 2.0;
    p.rho = 0.01;
    p.delta = 0.0;
    s = NeuronState {};
  }
  void
  set_status( NeuronParams& p, NeuronState& s, const Json& d, std::set< std::string >& used ) const override
  {
    const Keys k { d, used };
    NeuronParams q = p;
    const double delta_EL = set_common( q, k );
    set_membrane( q, k );
    if ( k.has( "tau_syn_ex" ) )
      q.tau_ex = k.get( "tau_syn_ex" );
    if ( k.has( "tau_syn_in" ) )
      q.tau_in = k.get( "tau_syn_in" );
    if ( k.has( "t_ref" ) )
      q.t_ref = k.get( "t_ref" );
    if ( q.V_reset >= q.V_th )
      throw std::runtime_error( "BadProperty: Reset potential must be smaller than threshold." );
    // iaf_psc_exp::Parameters_::set (iaf_psc_exp.cpp:149-164)
    if ( q.c_m <= 0 )
      throw std::runtime_error( "BadProperty: Capacitance must be strictly positive." );
    if ( q.tau_m <= 0 or q.tau_ex <= 0 or q.tau_in <= 0 )
      throw std::runtime_error( "BadProperty: Membrane and synapse time constants must be strictly positive." );
    if ( q.t_ref < 0 )
      throw std::runtime_error( "BadProperty: Refractory time must not be negative." );
    if ( k.has( "rho" ) )
      q.rho = k.get( "rho" );
    if ( k.has( "delta" ) )
      q.delta = k.get( "delta" );
    if ( q.rho < 0 )
      throw std::runtime_error( "BadProperty: Stochastic firing intensity must not be negative." );
    if ( q.delta < 0 )
      throw std::runtime_error( "BadProperty: Width of threshold region must not be negative." );
    if ( q.delta > 1e-10 )
      throw std::runtime_error( "NotImplemented: stochastic threshold (delta > 0) is not supported on the device path." );
    p = q;
    set_v_m( p, s, k, delta_EL );
  }
  bool
  get_param( const NeuronParams& p, const std::string& key, double& v ) const override
  {
    if ( key == "tau_syn_ex" )
      v = p.tau_ex;
    else if ( key == "tau_syn_in" )
      v = p.tau_in;
    else
      return IafModel::get_param( p, key, v );
    return true;
  }
  int
  state_field( const std::string& key ) const override
  {
    return key == "V_m" ? 0 : key == "I_syn_ex" ? 1 : key == "I_syn_in" ? 2 : -1;
  }
  std::vector< std::string >
  recordables() const override
  {
    return { "V_m", "I_syn_ex", "I_syn_in" };
  }
  void
  pre_run_hook( const NeuronParams& p, double h, ParamSet& ps ) const override
  {
    // iaf_psc_exp::pre_run_hook (iaf_psc_exp.cpp:244-286)
    const double P11ex = std::exp( -h / p.tau_ex );
    const double P11in = std::exp( -h / p.tau_in );
    const double P22 = std::exp( -h / p.tau_m );
    const double P21ex = iaf_propagator_exp( p.tau_ex, p.tau_m, p.c_m, h );
    const double P21in = iaf_propagator_exp( p.tau_in, p.tau_m, p.c_m, h );
    const double P20 = p.tau_m / p.c_m * ( 1.0 - P22 );
    ps.p[ 0 ] = P22;
    ps.p[ 1 ] = P21ex;
    ps.p[ 2 ] = P21in;
    ps.p[ 3 ] = P20;
    ps.p[ 4 ] = P11ex;
    ps.p[ 5 ] = P11in;
    ps.p[ 6 ] = p.I_e + 0.0;
    ps.p[ 7 ] = p.V_th;
    ps.p[ 8 ] = p.V_reset;
    ps.p[ 9 ] = ( 1. - P11ex ) * 0.0;
  }
};

// models/iaf_psc_alpha.cpp:68-301: alpha-shaped synaptic currents, exact integration. The
// state fields are V_m - E_L, I_syn_ex, I_syn_in, dI_syn_ex, dI_syn_in; the current input
// y0 of CurrentEvents is zero on this path (no current sources).
class IafPscAlpha : public IafModel
{
public:
  IafPscAlpha()
    : IafModel( "iaf_psc_alpha", M_IAF_PSC_ALPHA, 2 )
  {
  }
  void
  defaults( NeuronParams& p, NeuronState& s ) const override
  {
    // Parameters_() (iaf_psc_alpha.cpp:68-80)
    p.tau_m = 10.0;
    p.c_m = 250.0;
    p.t_ref = 2.0;
    p.E_L = -70.0;
    p.I_e = 0.0;
    p.V_reset = -70.0 - p.E_L;
    p.V_th = -55.0 - p.E_L;
    p.V_min = -std::numeric_limits< double >::infinity();
    p.tau_ex = 2.0;
    p.tau_in = 2.0;
    s = NeuronState {};
  }
  void
  set_status( NeuronParams& p, NeuronState& s, const Json& d, std::set< std::string >& used ) const override
  {
    // Parameters_::set (iaf_psc_alpha.cpp:112-180), State_::set (:188-199)
    const Keys k { d, used };
    NeuronParams q = p;
    const double delta_EL = set_common( q, k );
    if ( k.has( "V_min" ) )
      q.V_min = k.get( "V_min" ) - q.E_L;
    else
      q.V_min -= delta_EL;
    set_membrane( q, k );
    if ( k.has( "tau_syn_ex" ) )
      q.tau_ex = k.get( "tau_syn_ex" );
    if ( k.has( "tau_syn_in" ) )
      q.tau_in = k.get( "tau_syn_in" );
    if ( k.has( "t_ref" ) )
      q.t_ref = k.get( "t_ref" );
    if ( q.c_m <= 0.0 )
      throw std::runtime_error( "BadProperty: Capacitance must be > 0." );
    if ( q.tau_m <= 0.0 )
      throw std::runtime_error( "BadProperty: Membrane time constant must be > 0." );
    if ( q.tau_ex <= 0.0 or q.tau_in <= 0.0 )
      throw std::runtime_error( "BadProperty: All synaptic time constants must be > 0." );
    if ( q.t_ref < 0.0 )
      throw std::runtime_error( "BadProperty: The refractory time t_ref can't be negative." );
    if ( q.V_reset >= q.V_th )
      throw std::runtime_error( "BadProperty: Reset potential must be smaller than threshold." );
    p = q;
    set_v_m( p, s, k, delta_EL );
  }
  bool
  get_param( const NeuronParams& p, const std::string& key, double& v ) const override
  {
    if ( key == "V_min" )
      v = p.V_min + p.E_L;
    else if ( key == "tau_syn_ex" )
      v = p.tau_ex;
    else if ( key == "tau_syn_in" )
      v = p.tau_in;
    else
      return IafModel::get_param( p, key, v );
    return true;
  }
  int
  state_field( const std::string& key ) const override
  {
    return key == "V_m" ? 0 : key == "I_syn_ex" ? 1 : key == "I_syn_in" ? 2 : -1;
  }
  std::vector< std::string >
  recordables() const override
  {
    return { "V_m", "I_syn_ex", "I_syn_in" };
  }
  void
  pre_run_hook( const NeuronParams& p, double h, ParamSet& ps ) const override
  {
    // iaf_psc_alpha::pre_run_hook (iaf_psc_alpha.cpp:248-301), in its order
    const double P11ex = std::exp( -h / p.tau_ex );
    const double P11in = std::exp( -h / p.tau_in );
    const double expm1_tau_m = std::expm1( -h / p.tau_m );
    const double P30 = -p.tau_m / p.c_m * std::expm1( -h / p.tau_m );
    const double P21ex = h * P11ex;
    const double P21in = h * P11in;
    double P31ex, P32ex, P31in, P32in;
    iaf_propagator_alpha( p.tau_ex, p.tau_m, p.c_m, h, P31ex, P32ex );
    iaf_propagator_alpha( p.tau_in, p.tau_m, p.c_m, h, P31in, P32in );
    const double e = 2.718281828459045;  // numerics::e
    const double vals[] = { P30, 0.0 + p.I_e, P31ex, P32ex, P31in, P32in, expm1_tau_m, P21ex, P11ex, P21in, P11in,
      1.0 * e / p.tau_ex, 1.0 * e / p.tau_in, p.V_th, p.V_reset, p.V_min };
    for ( int q = 0; q < 16; ++q )
      ps.p[ q ] = vals[ q ];
  }
};

class AeifCondExp : public NodeModel
{
public:
  AeifCondExp()
    : NodeModel( "aeif_cond_exp", M_AEIF_COND_EXP, 2 )
  {
  }
  void
  defaults( NeuronParams& p, NeuronState& s ) const override
  {
    // models/aeif_cond_exp.cpp:132-151
    p.V_peak = 0.0;
    p.V_reset = -60.0;
    p.t_ref = 0.0;
    p.g_L = 30.0;
    p.c_m = 281.0;
    p.E_ex = 0.0;
    p.E_in = -85.0;
    p.E_L = -70.6;
    p.Delta_T = 2.0;
    p.tau_w = 144.0;
    p.a = 4.0;
    p.b = 80.5;
    p.V_th = -50.4;
    p.tau_ex = 0.2;
    p.tau_in = 2.0;
    p.I_e = 0.0;
    p.tol = 1e-6;
    s = NeuronState {};
    s.v0 = p.E_L;  // State_( const Parameters_& ): y[V_M] = E_L
  }
  void
  set_status( NeuronParams& p, NeuronState& s, const Json& d, std::set< std::string >& used ) const override
  {
    const Keys k { d, used };
    const char* keys[] = { "V_peak", "V_reset", "t_ref", "g_L", "C_m", "E_ex", "E_in", "E_L", "Delta_T", "tau_w", "a",
      "b", "V_th", "tau_syn_ex", "tau_syn_in", "I_e", "gsl_error_tol" };
    NeuronParams q = p;
    double* qdst[] = { &q.V_peak, &q.V_reset, &q.t_ref, &q.g_L, &q.c_m, &q.E_ex, &q.E_in, &q.E_L, &q.Delta_T,
      &q.tau_w, &q.a, &q.b, &q.V_th, &q.tau_ex, &q.tau_in, &q.I_e, &q.tol };
    for ( size_t i = 0; i < sizeof( keys ) / sizeof( keys[ 0 ] ); ++i )
      if ( k.has( keys[ i ] ) )
        *qdst[ i ] = k.get( keys[ i ] );
    // Parameters_::set and State_::set (aeif_cond_exp.cpp:233-304), checked in that order;
    // nothing is changed unless every check passes
    if ( q.V_peak < q.V_th )
      throw std::runtime_error( "BadProperty: V_peak >= V_th required." );
    if ( q.Delta_T < 0. )
      throw std::runtime_error( "BadProperty: Delta_T must be positive." );
    if ( q.Delta_T > 0. )
    {
      // the exponential term must not overflow at spike time, with a margin of 1e20
      const double max_exp_arg = std::log( std::numeric_limits< double >::max() / 1e20 );
      if ( ( q.V_peak - q.V_th ) / q.Delta_T >= max_exp_arg )
        throw std::runtime_error( "BadProperty: The current combination of V_peak, V_th and Delta_T will lead to "
                                  "numerical overflow at spike time; try for instance to increase Delta_T or to "
                                  "reduce V_peak to avoid this problem." );
    }
    if ( q.V_reset >= q.V_peak )
      throw std::runtime_error( "BadProperty: Ensure that: V_reset < V_peak ." );
    if ( q.c_m <= 0 )
      throw std::runtime_error( "BadProperty: Ensure that C_m >0" );
    if ( q.t_ref < 0 )
      throw std::runtime_error( "BadProperty: Refractory time cannot be negative." );
    if ( q.tau_ex <= 0 or q.tau_in <= 0 or q.tau_w <= 0 )
      throw std::runtime_error( "BadProperty: All time constants must be strictly positive." );
    if ( q.tol <= 0. )
      throw std::runtime_error( "BadProperty: The gsl_error_tol must be strictly positive." );
    const double g_ex = k.has( "g_ex" ) ? k.get( "g_ex" ) : s.v1;
    const double g_in = k.has( "g_in" ) ? k.get( "g_in" ) : s.v2;
    if ( g_ex < 0 or g_in < 0 )
      throw std::runtime_error( "BadProperty: Conductances must not be negative." );
    p = q;
    if ( k.has( "V_m" ) )
      s.v0 = k.get( "V_m" );
    s.v1 = g_ex;
    s.v2 = g_in;
    if ( k.has( "w" ) )
      s.v3 = k.get( "w" );
  }
  bool
  get_param( const NeuronParams& p, const std::string& key, double& v ) const override
  {
    if ( key == "E_L" )
      v = p.E_L;
    else if ( key == "V_th" )
      v = p.V_th;
    else if ( key == "V_reset" )
      v = p.V_reset;
    else if ( key == "C_m" )
      v = p.c_m;
    else if ( key == "t_ref" )
      v = p.t_ref;
    else if ( key == "I_e" )
      v = p.I_e;
    else if ( key == "tau_syn_ex" )
      v = p.tau_ex;
    else if ( key == "tau_syn_in" )
      v = p.tau_in;
    else if ( key == "V_peak" )
      v = p.V_peak;
    else if ( key == "g_L" )
      v = p.g_L;
    else if ( key == "E_ex" )
      v = p.E_ex;
    else if ( key == "E_in" )
      v = p.E_in;
    else if ( key == "Delta_T" )
      v = p.Delta_T;
    else if ( key == "tau_w" )
      v = p.tau_w;
    else if ( key == "a" )
      v = p.a;
    else if ( key == "b" )
      v = p.b;
    else if ( key == "gsl_error_tol" )
      v = p.tol;
    else
      return false;
    return true;
  }
  int
  state_field( const std::string& key ) const override
  {
    return key == "V_m" ? 0 : key == "g_ex" ? 1 : key == "g_in" ? 2 : key == "w" ? 3 : -1;
  }
  std::vector< std::string >
  recordables() const override
  {
    return { "V_m", "g_ex", "g_in", "w" };
  }
  void
  pre_run_hook( const NeuronParams& p, double h, ParamSet& ps ) const override
  {
    // aeif_cond_exp::pre_run_hook (aeif_cond_exp.cpp:420-435)
    const double vals[] = { p.V_peak, p.V_reset, p.g_L, p.c_m, p.E_ex, p.E_in, p.E_L, p.Delta_T, p.tau_w, p.a, p.b,
      p.V_th, p.tau_ex, p.tau_in, p.I_e, p.tol, p.Delta_T > 0. ? p.V_peak : p.V_th, h };
    for ( int q = 0; q < 18; ++q )
      ps.p[ q ] = vals[ q ];
  }
  void
  device_state( const NeuronParams&, const NeuronState& s, double h, double out[ 5 ] ) const override
  {
    out[ 0 ] = s.v0;
    out[ 1 ] = s.v1;
    out[ 2 ] = s.v2;
    out[ 3 ] = s.v3;
    out[ 4 ] = h;  // the integrator's first step size (B_.IntegrationStep_ = resolution)
  }
};

}  // namespace

ModelManager::ModelManager()
{
  // the built-in models, in device id order
  register_node_model< IafPscDelta >();
  register_node_model< IafPscExp >();
  register_node_model< AeifCondExp >();
  register_node_model< IafPscAlpha >();
}

ModelManager&
ModelManager::instance()
{
  static ModelManager m;
  return m;
}

const NodeModel&
ModelManager::add( std::unique_ptr< NodeModel > m )
{
  if ( find( m->name() ) )
    throw std::runtime_error( "NamingConflict: a model called " + m->name() + " already exists" );
  if ( m->id() != size() )
    throw std::runtime_error( "BadProperty: model " + m->name() + " must take device id " + std::to_string( size() ) );
  if ( m->channels() < 1 or m->channels() > 2 )
    throw std::runtime_error( "BadProperty: model " + m->name() + " must use 1 or 2 receptor channels" );
  models_.push_back( std::move( m ) );
  return *models_.back();
}

const NodeModel*
ModelManager::find( const std::string& name ) const
{
  for ( const auto& m : models_ )
    if ( m->name() == name )
      return m.get();
  return nullptr;
}

const NodeModel&
ModelManager::get( int id ) const
{
  if ( id < 0 or id >= size() )
    throw std::runtime_error( "UnknownModelID: " + std::to_string( id ) );
  return *models_[ id ];
}

std::vector< std::string >
ModelManager::names() const
{
  std::vector< std::string > v;
  for ( const auto& m : models_ )
    v.push_back( m->name() );
  return v;
}

}  // namespace nestgpu
