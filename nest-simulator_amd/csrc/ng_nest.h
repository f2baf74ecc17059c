// ng_nest.h -- host kernel with NEST's Create/Connect/Simulate surface.
//
// Mirrors the parts of nestkernel that decide WHAT is simulated: time base
// (nest_time.h), node placement on virtual processes (node_manager.cpp:188-230,
// vp_manager_impl.h:43-52), the per-VP random streams (random_manager.cpp:42-100),
// the connection builders and their draw order (conn_builder.cpp), delay extrema
// (connection_manager.cpp:485-530) and the Prepare/Run/Cleanup cycle
// (simulation_manager.cpp:507-671). The simulation itself runs on the device engine.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "ng_device.h"
#include "ng_engine.h"
#include "ng_philox.h"
#include "ng_json.h"
#include "ng_model.h"

namespace nestgpu
{

// nestkernel/nest_time.cpp:36-47,88-103; nest_time.h:306-311,508-522,539-542
struct TimeBase
{
  double tics_per_ms = 1000.0;
  long tics_per_step = 100;
  double tics_per_step_inv = 1.0 / 100.0;
  long tics_per_step_rnd = 99;
  double ms_per_tic = 1.0 / 1000.0;
  double ms_per_step = 100 / 1000.0;
  double steps_per_ms = 1 / ( 100 / 1000.0 );

  void set_resolution( double ms );
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
  long delay_ms_to_steps( double ms ) const;
  double
  resolution_ms() const
  {
    return ms_per_tic * tics_per_step;
  }
};

// randutils seed_seq_fe<4, uint32_t> (O'Neill), the seeding NEST applies to every engine
struct SeedSeqFe128
{
  typedef uint32_t result_type;
  uint32_t mixer[ 4 ];
  explicit SeedSeqFe128( const std::vector< uint32_t >& entropy );
  template < typename It >
  void
  generate( It first, It last ) const
  {
    uint32_t hc = 0x8b51f9ddu;
    int src = 0;
    for ( It dst = first; dst != last; ++dst )
    {
      uint32_t v = mixer[ src ];
      src = ( src + 1 ) & 3;
      v ^= hc;
      hc *= 0x58f38dedu;
      v *= hc;
      v ^= v >> 16;
      *dst = v;
    }
  }
  static constexpr size_t
  size()
  {
    return 4;
  }
};

// Philox_32 stream as a uniform random bit generator over 32-bit words (ng_philox.h): the
// key and the position of the next word; the current block is kept
struct PhiloxStream
{
  typedef uint32_t result_type;
  uint32_t k0 = 0, k1 = 0;
  uint64_t pos = 0;
  uint64_t cached = ~uint64_t( 0 );  // block held in v
  uint32_t v[ 4 ] = { 0, 0, 0, 0 };
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
  result_type
  operator()()
  {
    const uint64_t b = pos >> 2;
    if ( b != cached )
    {
      philox_block( k0, k1, b, v );
      cached = b;
    }
    const uint32_t w = v[ 3u - static_cast< uint32_t >( pos & 3u ) ];
    ++pos;
    return w;
  }
};

enum RngType
{
  RNG_MT19937_64 = 0,
  RNG_PHILOX_32 = 1
};

// RandomGenerator< std::mt19937_64 > or RandomGenerator< r123::Engine< r123::Philox4x32 > >
// (random_generators.h:141-297), the draws the builders and parameters make.
// Per-VP objects are written on every draw or connection by their own thread: each sits on
// cache lines of its own so that the VP threads of Connect do not share lines.
struct alignas( 64 ) VpRng
{
  int type = RNG_MT19937_64;
  std::mt19937_64 engine;
  PhiloxStream philox;
  std::uniform_int_distribution< unsigned long > ulong_dist;
  VpRng( const std::vector< uint32_t >& seed, int rng_type );
  // std::generate_canonical<double, 53>(engine) in closed form (libstdc++ bits/random.tcc):
  // a 64-bit engine gives one word divided by 2^64, a 32-bit engine two words summed (the
  // first low) and scaled by 2^-64; a result that rounds to 1 is clamped to the double below.
  // The library's generic form evaluates two long-double logarithms per draw unless the
  // compiler folds them, which hipcc does not.
  double
  canonical()
  {
    if ( type == RNG_PHILOX_32 )
    {
      const uint32_t lo = philox();
      return philox_canonical( lo, philox() );
    }
    const double u = static_cast< double >( engine() ) * 0x1p-64;
    return u < 1.0 ? u : std::nextafter( 1.0, 0.0 );
  }
  // std::uniform_real_distribution<>(0, 1): canonical * (1 - 0) + 0, i.e. canonical itself
  double
  drand()
  {
    return canonical();
  }
  unsigned long
  ulrand( unsigned long n )
  {
    std::uniform_int_distribution< unsigned long >::param_type p( 0, n - 1 );
    return type == RNG_PHILOX_32 ? ulong_dist( philox, p ) : ulong_dist( engine, p );
  }
};

// std::normal_distribution<double> (libstdc++ bits/random.tcc): Marsaglia's polar method
// over pairs of canonical draws, the second value of a pair kept for the next call
struct alignas( 64 ) VpNormal
{
  double mean, sd;
  double saved = 0;
  bool saved_available = false;
  VpNormal( double mean_, double sd_ )
    : mean( mean_ )
    , sd( sd_ )
  {
  }
  double
  operator()( VpRng& rng )
  {
    double ret;
    if ( saved_available )
    {
      saved_available = false;
      ret = saved;
    }
    else
    {
      double x, y, r2;
      do
      {
        x = 2.0 * rng.canonical() - 1.0;
        y = 2.0 * rng.canonical() - 1.0;
        r2 = x * x + y * y;
      } while ( r2 > 1.0 or r2 == 0.0 );
      const double mult = std::sqrt( -2.0 * std::log( r2 ) / r2 );
      saved = x * mult;
      saved_available = true;
      ret = y * mult;
    }
    return ret * sd + mean;
  }
};

// A parameter object keeps its per-VP distributions across Connect calls: NormalParameter and
// LognormalParameter hold one distribution per thread (parameter.cpp:95-157), whose cached
// second polar value carries over to the next call that uses the same object. A normal or
// lognormal leaf with an "id" names such an object; the kernel's registry holds its state.
struct ParamStates
{
  std::map< int64_t, std::shared_ptr< std::vector< VpNormal > > > normals;
};

// nest::Parameter expression trees used by syn_spec weights/delays (parameter.h)
class Param
{
public:
  enum Op
  {
    CONST,
    UNIFORM,
    UNIFORM_INT,
    NORMAL,
    SUM,
    PRODUCT,
    MAX,
    MIN,
    REDRAW,
    ARRAY,
    LOGNORMAL,
    EXPONENTIAL,
    EXP,
    SIN,
    COS,
    QUOTIENT,
    POW,
    COMPARE,
    CONDITIONAL
  };
  // top_level: a JSON array is an array parameter (one value per connection, a syn_spec
  // value of its own, not an operand)
  static std::unique_ptr< Param > from_json(
    const Json& j, int n_vp, bool top_level = false, ParamStates* states = nullptr );
  double value( VpRng& rng, int vp );
  bool
  is_const() const
  {
    return op_ == CONST;
  }
  bool
  is_array() const
  {
    return op_ == ARRAY;
  }
  size_t
  array_size() const
  {
    return values_.size();
  }
  // passes over n values of VP vp's cursor: the connections another VP creates
  // (ArrayDoubleParameter::skip, conn_parameter.h:242-252)
  void skip( int vp, size_t n );
  double
  const_value() const
  {
    return a_;
  }

private:
  Op op_ = CONST;
  double a_ = 0, b_ = 0;
  std::unique_ptr< Param > l_, r_, f_;  // f_: a conditional's value when false
  // one per VP (parameter.cpp:95-117), shared by every use of the same parameter object
  std::shared_ptr< std::vector< VpNormal > > normals_;
  // array parameter: the values and each VP's next index (ArrayDoubleParameter,
  // conn_parameter.h:231-300)
  struct alignas( 64 ) VpCursor
  {
    size_t next = 0;
  };
  std::vector< double > values_;
  std::vector< VpCursor > cursor_;
};

// activity window of a device (Device::Parameters_, device.cpp:83-118) in steps: a spike with
// stamp origin + start < stamp <= origin + stop is recorded or emitted
// (RecordingDevice::is_active, recording_device.cpp:206-211; StimulationDevice::is_active,
// stimulation_device.cpp:47-57)
struct DeviceWindow
{
  long origin = 0, start = 0, stop = 0;
  bool stop_inf = true;
  bool set = false;  // the default window passes every spike
  bool
  active( long stamp ) const
  {
    return origin + start < stamp and ( stop_inf or stamp <= origin + stop );
  }
};

struct SpikeGen
{
  DeviceWindow window;  // emitted stamps are window.origin + stamps[k]
  std::vector< int64_t > stamps;  // spike stamps in steps (emission step = stamp - 1)
  std::vector< int64_t > mults;
  // dense per-step counts: column in a kernel-wide block
  int dense_block = -1;
  int64_t dense_col = 0;
  // seeded Poisson train (rate per step, seed); stream id = generator node id
  double poisson_lambda = 0;
  uint64_t poisson_seed = 0;
  int64_t poisson_t0 = 0, poisson_steps = 0;
};

struct DenseBlock
{
  int64_t t0 = 0, n_steps = 0, cols = 0;
  std::vector< uint8_t > counts;  // [step][col]
};

// connection in creation order of its VP (SoA storage per VP)
struct alignas( 64 ) ConnList
{
  std::vector< uint32_t > src, tgt;
  std::vector< double > w;
  std::vector< uint16_t > d;
  // synapse of each connection: 0 static_synapse, k an stdp_synapse of parameter set k - 1;
  // shorter than the list while its tail is static (empty when every connection is static)
  std::vector< uint16_t > syn;
  uint16_t
  syn_at( size_t i ) const
  {
    return i < syn.size() ? syn[ i ] : 0;
  }
  size_t
  size() const
  {
    return src.size();
  }
  // room for `extra` more connections: exact for one large addition, geometric for many small
  void
  reserve_more( size_t extra )
  {
    const size_t need = src.size() + extra;
    if ( need <= src.capacity() )
      return;
    const size_t cap = std::max( need, src.capacity() + src.capacity() / 2 );
    src.reserve( cap );
    tgt.reserve( cap );
    w.reserve( cap );
    d.reserve( cap );
  }
};

class Kernel
{
public:
  explicit Kernel( int device );
  ~Kernel();

  void reset();
  void set_kernel_status( const Json& d );
  std::string get_kernel_status_json() const;

  int64_t create( const std::string& model, int64_t n, const Json& params );
  void set_status( const std::vector< int64_t >& gids, const Json& params );
  void set_values( const std::vector< int64_t >& gids, const std::string& key, const double* v );
  void get_values( const std::vector< int64_t >& gids, const std::string& key, double* out );
  std::string model_name( int64_t gid ) const;

  void connect( const std::vector< int64_t >& pre, const std::vector< int64_t >& post, const Json& conn_spec,
    const Json& syn_spec );
  void connect_arrays( const int64_t* src, const int64_t* tgt, const double* w, const double* d_ms, int64_t n );

  void set_spike_counts( int64_t gen_first, int64_t n, int64_t t0_step, int64_t n_steps, const uint8_t* counts );

  int64_t num_connections() const;
  void get_connections( int64_t* src, int64_t* tgt, double* w, int64_t* d, int32_t* vp ) const;
  // 0 static_synapse, 1 stdp_synapse, in get_connections' order
  void get_connection_models( int32_t* model ) const;

  void prepare();
  void run( double t_ms );
  void cleanup();
  void simulate( double t_ms );

  // spike_recorder events (senders, stamp steps), sorted by (time, sender)
  void recorder_events( int64_t rec_gid, std::vector< int64_t >& senders, std::vector< int64_t >& steps );
  // multimeter events sorted by (time, sender): values[i * record_from.size() + j] is the
  // value of record_from[j] of event i
  void multimeter_events( int64_t mm_gid, std::vector< int64_t >& senders, std::vector< int64_t >& steps,
    std::vector< double >& values );
  const std::vector< std::string >& multimeter_record_from( int64_t mm_gid ) const;
  // status dictionary of one node as JSON (model, global_id, and a device's settings)
  std::string node_status_json( int64_t gid ) const;

  // batched / served use
  Engine& engine()
  {
    return *engine_;
  }
  bool prepared() const
  {
    return prepared_;
  }
  bool simulated() const
  {
    return simulated_;
  }
  int device() const
  {
    return device_;
  }
  void prepare_for_batch();  // prepare without running (idempotent)
  // The device copy is built and matches the host tables: Prepare would allocate nothing.
  bool device_copy_current() const
  {
    return engine_ and engine_->built() and not store_dirty_ and not released_;
  }
  // Steps of a Run of t_ms, with Run's checks (non-negative, a multiple of the resolution).
  int64_t checked_steps( double t_ms ) const;
  // Discard the device copy and every simulation result, back to the state before the
  // first Prepare (used by the layout autotuner's probe runs).
  void reset_device();
  void set_layout( int targets_per_block, int delivery_mode, int window_slots = 0 );
  // Device memory Prepare will allocate, estimated from the host tables (serving admission).
  uint64_t estimate_device_bytes() const;
  // Copy the device state into the host tables, collect the spike log and free the device
  // copy; the kernel then answers status queries but cannot simulate further.
  void release_device();
  bool released() const
  {
    return released_;
  }
  bool keep_host_connections() const
  {
    return keep_host_connections_;
  }
  void set_keep_host_connections( bool keep )
  {
    keep_host_connections_ = keep;
  }
  int targets_per_block() const
  {
    return engine_ and engine_->built() ? static_cast< int >( engine_->net().tb ) : targets_per_block_;
  }
  void after_batch_run( int64_t steps );
  int64_t steps_for( double t_ms ) const
  {
    return tb_.ms_to_steps( t_ms );
  }
  const TimeBase& timebase() const
  {
    return tb_;
  }
  int64_t min_delay() const
  {
    return min_delay_;
  }
  double last_run_kernel_ms() const
  {
    return last_kernel_ms_;
  }
  double last_run_wall_ms() const
  {
    return last_wall_ms_;
  }
  uint64_t synaptic_events() const;
  uint64_t local_spike_counter() const
  {
    if ( released_ )
      return released_spikes_;
    return engine_ and engine_->built() ? engine_->read_counter( 2 ) : 0;
  }
  uint64_t num_neurons() const
  {
    return params_.size();
  }
  int delivery_mode() const  // 0 automatic, 1 window, 2 global
  {
    return delivery_mode_;
  }
  int model_mask() const  // bit m set when a neuron of model id m exists
  {
    int mask = 0;
    for ( const auto& p : params_ )
      mask |= 1 << p.model;
    return mask;
  }
  void set_runner( Runner* r )
  {
    runner_ = r;
  }

private:
  struct NodeRec
  {
    uint8_t kind;  // 0 neuron, 1 spike_generator, 2 spike_recorder, 3 multimeter
    uint32_t idx;
  };
  NodeRec& node( int64_t gid );
  const NodeRec& node( int64_t gid ) const;
  int vp_of( int64_t gid ) const
  {
    return static_cast< int >( gid % n_vp_ );
  }
  void init_rngs();
  void apply_params( NeuronParams& p, NeuronState& s, const Json& d, bool creating );
  void check_delay( long steps );
  void single_connect( int64_t s, int64_t t, int vp, VpRng& rng, Param* wp, Param* dp, uint16_t syn = 0 );
  // pairwise_bernoulli (param p) or fixed_indegree (param K, autapses and multapses allowed)
  // with constant weight and delay, drawn on the device from the Philox_32 streams
  void device_connect( const std::vector< int64_t >& pre, const std::vector< int64_t >& post, const std::string& rule,
    double param, bool autapses, double weight, uint16_t delay_steps );
  void build_spec( BuildSpec& b );
  void choose_layout( uint32_t n, int max_delay, int nch, uint32_t& tb, bool& win ) const;
  // block bounds balancing in-degree over at most ceil(n / tb) blocks
  std::vector< uint32_t > balanced_partition( uint32_t n, uint32_t tb ) const;
  static uint32_t block_cap( uint32_t tb );  // largest balanced block for a nominal size tb
  void rebalance_blocks();
  static int window_slots( int max_delay, int nch, uint32_t tb );
  void collect_log();

  int device_;
  TimeBase tb_;
  int n_vp_ = 1;
  uint32_t seed_ = 143202461;
  std::vector< VpRng > rngs_;
  bool rng_dirty_ = true;
  int rng_type_ = RNG_MT19937_64;
  std::vector< NodeRec > nodes_;  // index gid - 1
  std::vector< NeuronParams > params_;
  std::vector< NeuronState > state_;
  std::vector< int64_t > neuron_gid_;
  std::vector< SpikeGen > gens_;
  std::vector< int64_t > gen_gid_;
  std::vector< DenseBlock > dense_;
  std::vector< std::vector< int64_t > > recorders_;  // recorded senders (sorted) per recorder
  std::vector< DeviceWindow > rec_window_;  // per spike_recorder
  // origin, start or stop into w; false for any other key
  bool update_window( DeviceWindow& w, const std::string& k, const Json& v ) const;
  std::vector< int64_t > rec_gid_;
  // multimeter (models/multimeter.cpp): the sampled neurons, the recordables in order, the
  // sampling interval and offset in steps and the recording window
  struct Meter
  {
    std::vector< int64_t > targets;  // sorted node ids
    std::vector< std::string > record_from;
    int64_t interval = 10, offset = 0;
    DeviceWindow window;
    bool voltmeter = false;
    // recorded events: sender, stamp and the record_from values (row-major)
    std::vector< int64_t > ev_gid, ev_stamp;
    std::vector< double > ev_vals;
  };
  std::vector< Meter > meters_;
  std::vector< int64_t > meter_gid_;
  std::vector< int64_t > mm_col_gid_;  // node id of each sample column of the current device copy
  void set_meter_status( Meter& m, const Json& d );
  void collect_samples();
  std::vector< ConnList > conns_;  // per VP
  ParamStates param_states_;       // parameter objects used by more than one Connect call
  struct DevConn
  {
    int64_t gen, tgt;
    double w;
    long d;
  };
  std::vector< DevConn > dev_conns_;
  long d_min_seen_ = -1, d_max_seen_ = 1;
  long min_delay_ = 1, max_delay_ = 1;
  bool simulated_ = false, prepared_ = false, store_dirty_ = true, state_dirty_ = false;
  std::unique_ptr< Engine > engine_;
  std::unique_ptr< Runner > own_runner_;
  Runner* runner_ = nullptr;
  std::vector< int32_t > gid_to_dev_;  // neuron gid -> device index (valid after prepare)
  std::vector< int64_t > dev_to_gid_;
  // events of each spike_recorder (stamp, sender), recorded with the senders it observed and
  // the window it had when the spikes were read back
  std::vector< std::vector< std::pair< int64_t, int64_t > > > rec_events_;
  uint64_t spike_count_ = 0;
  double last_kernel_ms_ = 0, last_wall_ms_ = 0;
  int slices_per_chunk_ = 64;
  int wg_per_cu_ = 0;
  int window_slots_ = 0;  // cap on the ring slots staged in LDS (0: as many as fit)
  uint32_t stdp_history_ = STDP_HIST_DEFAULT;  // post-spike history entries per neuron
  int delivery_passes_ = 0;  // global delivery in this many slot-range passes (0, 1: one)
  bool time_kernels_ = false;
  bool phase_timing_ = false;
  bool graph_launch_ = true;  // replay chunks from captured hipGraphs (false: direct launches)
  int64_t graph_chunks_ = 0, direct_chunks_ = 0, capture_failures_ = 0;
  double time_kernel_ms_[ N_KINDS ] = { 0, 0 };
  int64_t kernel_launches_[ N_KINDS ] = { 0, 0 };
  int targets_per_block_ = 0;  // 0: chosen from the network size
  // stdp_synapse parameter sets, one per Connect call that makes stdp_synapse connections
  struct StdpSet
  {
    double tau_plus = 20.0, lambda = 0.01, alpha = 1.0, mu_plus = 1.0, mu_minus = 1.0, Wmax = 100.0, Kplus = 0.0;
  };
  std::vector< StdpSet > stdp_sets_;
  std::vector< uint32_t > syn_pos_;  // store position of every neuron connection (creation order), with stdp
  bool connect_on_device_ = false;  // eligible Connect calls draw on the device
  int64_t device_connect_calls_ = 0;
  bool balanced_blocks_ = true;  // automatic block size: blocks balanced by in-degree, then by
                                 // the measured received events
  std::vector< uint32_t > block_start_;  // target-block bounds of the device copy
  bool rebalanced_ = false;              // the measured balance has been applied (or declined)
  int64_t rx_steps_ = 0;                 // steps run with received events counted
  int64_t block_relayouts_ = 0;
  int delivery_mode_ = 0;      // 0 automatic, 1 window (LDS ring-buffer slots), 2 global
  bool keep_host_connections_ = true;  // false: the host copy is freed once the device holds it
  bool host_connections_freed_ = false;
  bool released_ = false;
  uint64_t released_events_ = 0, released_spikes_ = 0;
  int64_t freed_connections_ = 0;
  double time_simulate_ms_ = 0;
};

// Seeded Poisson counts (the pre-generated drive) for streams [stream0, stream0+n) and
// emission steps [step0, step0+n_steps); out is [step][stream]. Identical on the device.
void poisson_counts( uint64_t seed, double lambda, uint64_t stream0, int64_t n, int64_t step0, int64_t n_steps,
  uint8_t* out );
PoissonTable make_poisson_table( double lambda, uint64_t seed );

}  // namespace nestgpu
