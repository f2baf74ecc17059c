// ng_kernels.hip -- gfx950 kernel for one min_delay slice of NEST's simulation step.
//
// One workgroup owns a block of `tb` consecutive targets (device index = receive-order
// position) and, for slice s of a chunk:
//   1. delivers the previous slice's spikes to its targets (EventDeliveryManager::
//      deliver_events_, nestkernel/event_delivery_manager.cpp:599-776; Connector::send,
//      connector_base.h:399-428; RingBuffer::add_value, ring_buffer.h:157-161,405-408):
//      it walks the receive-order spike list, reads its own run of each spiking source's
//      out-row, and adds every weight into the target's ring-buffer cell;
//   2. updates its neurons over [from, to) (Node::update of iaf_psc_delta, iaf_psc_exp,
//      aeif_cond_exp) including the spike_generator drive of that slice;
//   3. appends its spikes, in receive order, to its block list for the next slice.
//
// Window delivery: the slots a slice can touch through recurrent synapses are the
// max_delay slots starting at the slice origin (lag + delay - min_delay < max_delay), and
// the update reads the first min_delay of them. In window mode the workgroup copies the wn
// consecutive slots [wa, wa + wn) of those that receive the most events into LDS,
// accumulates there and writes them back once, so the read-modify-write of most synaptic
// events is an LDS operation; cells outside the window are updated in HBM. Global mode
// updates every cell in HBM.
//
// Determinism: NEST adds into one ring-buffer cell in the order (delivery batch, receive
// order of the spike, position in the connector run). A workgroup takes events in that key
// order in rounds of R; each round is bucketed by target (LDS counters give every event a
// place in its target's bucket), every event finds its rank in the bucket ordered by
// (cell, key), and the first event of each cell's run adds the run in key order. No result
// depends on thread timing or atomics order.
// Built with -ffp-contract=off: every multiply and add rounds as in the reference.

#include <hip/hip_runtime.h>

#include "ng_aeif.h"
#include "ng_device.h"
#include "ng_engine.h"

namespace nestgpu
{

constexpr int IDX_BITS = 11;  // bucket key = slot-channel << 11 | round index: sorted keys group
                              // a target's events by cell, in key order within each cell

// Tile shape per kernel variant. The window kernel halves chunk and round so that, with its
// ring-buffer window, two workgroups fit in the 160 KB of LDS of a CU.
template < bool WIN >
struct Tile
{
  static constexpr int BS = static_cast< int >( MAX_TB );  // threads per workgroup
  static constexpr int SPT = WIN ? 2 : 4;      // spikes per thread per chunk
  static constexpr int CH = SPT * BS;          // spikes per chunk
  static constexpr int RQ = WIN ? 2 : 4;       // events per thread per round
  static constexpr int R = RQ * BS;            // events per round
  static constexpr uint32_t WINDOW = WIN ? WIN_CAP : 1;
  static constexpr uint32_t PMAX = WIN ? PMAX_WIN : PMAX_GLOB;
};
constexpr uint32_t NONE = 0xffffffffu;

// HBM views of the instance's arrays. The pointers in the instance record are generic, and
// generic (FLAT) accesses are counted by the LDS wait counter too: every barrier and every
// LDS read would then wait for all HBM loads in flight. Accessing HBM through global-space
// pointers keeps those loads on the vector-memory counter, so they overlap the LDS phases.
template < class T >
using gptr = __attribute__( ( address_space( 1 ) ) ) T*;
template < class T >
using gref = __attribute__( ( address_space( 1 ) ) ) T&;
template < class T >
using lptr = __attribute__( ( address_space( 3 ) ) ) T*;
template < class T >
__device__ inline gptr< T >
glob( T* p )
{
  return ( gptr< T > )p;
}

// device-scope relaxed atomics on HBM counters (what atomicAdd / atomicOr do, without FLAT)
__device__ inline unsigned long long
g_add( unsigned long long* p, unsigned long long v )
{
  return __hip_atomic_fetch_add( glob( p ), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}
__device__ inline void
g_or( unsigned long long* p, unsigned long long v )
{
  __hip_atomic_fetch_or( glob( p ), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

// The block's ring buffer as a buffer resource: an access at an offset past its end
// (NO_CELL) moves no data (loads return 0, stores are dropped), so every lane can issue the
// same loads and stores whether or not its event has an HBM cell.
constexpr uint32_t NO_CELL = 0x80000000u;
using v2u = __attribute__( ( ext_vector_type( 2 ) ) ) uint32_t;
__device__ inline double
cell_load( __amdgpu_buffer_rsrc_t r, uint32_t byte_off )
{
  return __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, byte_off, 0, 0 ) );
}
__device__ inline void
cell_store( __amdgpu_buffer_rsrc_t r, uint32_t byte_off, double v )
{
  __builtin_amdgcn_raw_buffer_store_b64( __builtin_bit_cast( v2u, v ), r, byte_off, 0, 0 );
}

__device__ inline SynRec
load_rec( gptr< const SynRec > p )
{
  SynRec r;
  r.w_lo = p->w_lo;
  r.w_hi = p->w_hi;
  r.word = p->word;
  return r;
}
constexpr int LB = 8;  // lags of ring-buffer input loaded together in the update
static_assert( Tile< false >::R <= ( 1 << IDX_BITS ) and Tile< true >::R <= ( 1 << IDX_BITS ),
  "round index must fit the bucket key" );
static_assert( ( PMAX_GLOB + 1 ) * sizeof( uint32_t ) <= WIN_CAP * sizeof( double ),
  "the window must hold the block prefix of an instance without a window" );

template < bool WIN >
struct SliceLds
{
  using T = Tile< WIN >;
  double win[ T::WINDOW ];  // ring-buffer window [slot - origin][channel][tb]
  double bw[ T::R ];        // event values, bucketed by target
  unsigned long long log_base;
  uint32_t pref[ T::PMAX + 1 ];  // prefix of per-block spike counts
  uint32_t c_g0[ T::CH ];        // first synapse of each chunk spike's run
  uint32_t c_off[ T::CH + 1 ];   // exclusive prefix of run lengths
  uint32_t bk[ T::R ];           // event keys (slot-channel << 11 | round index), bucketed by target
  uint16_t perm[ T::R ];         // bucket position of the event at each sorted position
  uint32_t bcnt[ 2 ][ T::BS ];   // events per target, alternating between rounds
  uint32_t bst[ T::BS + 1 ];     // bucket starts
  uint32_t sh[ T::BS / 64 ];
  uint32_t cnt_base;
  uint8_t c_lag[ T::CH ];
};

// XCD-aware block mapping: workgroups are dispatched round-robin over 8 XCDs; give each
// XCD a contiguous range of blocks so neighbouring blocks share its L2.
__device__ inline uint32_t
xcd_block( uint32_t w, uint32_t total )
{
  const uint32_t x = w & 7u;
  uint32_t base = 0;
  for ( uint32_t q = 0; q < x; ++q )
    base += ( total - q + 7u ) / 8u;
  return base + ( w >> 3 );
}

__device__ inline uint32_t
upper_bound_u32( const uint32_t* a, uint32_t n, uint32_t v )
{
  uint32_t lo = 0, hi = n;
  while ( lo < hi )
  {
    const uint32_t mid = ( lo + hi ) >> 1;
    if ( a[ mid ] <= v )
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// number of records of a delay-ordered run [g0, g0 + len) whose delay offset is below t
__device__ inline uint32_t
run_split( gptr< const SynRec > syn, uint32_t g0, uint32_t len, int t )
{
  if ( t <= 0 )
    return 0;
  uint32_t lo = 0, hi = len;
  while ( lo < hi )
  {
    const uint32_t mid = ( lo + hi ) >> 1;
    if ( static_cast< int >( ( syn[ g0 + mid ].word >> 16 ) & DELAY_MASK ) < t )
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// inclusive scan over the 64 lanes of a wave with DPP moves: Hillis-Steele inside each row
// of 16 lanes (row_shr 1, 2, 4, 8; lanes shifted in from outside the row read 0), then the
// last lane of row 0 and 2 is added to row 1 and 3 (row_bcast:15), then the last lane of
// row 1 to rows 2 and 3 (row_bcast:31)
__device__ inline uint32_t
wave_incl_scan( uint32_t x )
{
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x111, 0xf, 0xf, false ) );
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x112, 0xf, 0xf, false ) );
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x114, 0xf, 0xf, false ) );
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x118, 0xf, 0xf, false ) );
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x142, 0xa, 0xf, false ) );
  x += static_cast< uint32_t >( __builtin_amdgcn_update_dpp( 0, static_cast< int >( x ), 0x143, 0xc, 0xf, false ) );
  return x;
}

// block-wide exclusive scan of one value per thread; returns the total
template < int BS >
__device__ inline uint32_t
block_scan( uint32_t v, uint32_t* excl, uint32_t* sh )
{
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint32_t x = wave_incl_scan( v );
  if ( lane == 63 )
    sh[ wid ] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for ( int q = 0; q < BS / 64; ++q )
  {
    if ( q < wid )
      pre += sh[ q ];
    tot += sh[ q ];
  }
  *excl = pre + x - v;
  __syncthreads();
  return tot;
}

// Poisson count for seeded drive: splitmix64 of (seed, stream, step) gives u in [0, 1), the
// count is the first k with u < cdf[k] (at most kmax). The first KC entries of the table are
// held in registers; the loop continues in HBM only past them.
constexpr int KC = 8;
struct PoissonCache
{
  double c[ KC ];
  uint64_t seed;
  int32_t kmax;
  gptr< const PoissonTable > T;
};

__device__ inline uint32_t
poisson_count( const PoissonCache& P, uint64_t stream, int64_t step )
{
  uint64_t z = P.seed + stream * 0x9E3779B97F4A7C15ull + static_cast< uint64_t >( step ) * 0xD1B54A32D192ED03ull;
  z = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
  z = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
  z = z ^ ( z >> 31 );
  const double u = static_cast< double >( z >> 11 ) * ( 1.0 / 9007199254740992.0 );
  uint32_t k = 0;
#pragma unroll
  for ( int i = 0; i < KC; ++i )
    if ( k == static_cast< uint32_t >( i ) and i < P.kmax and u >= P.c[ i ] )
      k = i + 1;
  if ( k == KC )
    while ( static_cast< int32_t >( k ) < P.kmax and u >= P.T->cdf[ k ] )
      ++k;
  return k;
}

// Ring-buffer cells of one block seen from one slice: slot_rel = absolute step - origin.
// Window mode keeps slots [wa, wa + wn) in LDS (the slots most events of a slice land in);
// all other cells, and every cell in global mode, live in the block's ring in HBM.
template < bool WIN >
struct Cells
{
  lptr< double > lds;
  gptr< double > rbb;  // block's ring: [slot][channel][tb]
  uint32_t nct;  // doubles per slot (channels * tb)
  uint32_t tb;
  uint32_t s0;  // ring slot of the slice origin
  uint32_t L;
  uint32_t wa, wn;  // window: first slot (relative to the origin) and slots
  bool win;

  __device__ bool
  in_window( uint32_t slot_rel ) const
  {
    return WIN and win and slot_rel - wa < wn;
  }
  __device__ uint32_t
  win_index( uint32_t slot_rel, uint32_t ch, uint32_t local ) const
  {
    return ( slot_rel - wa ) * nct + ch * tb + local;
  }
  __device__ uint32_t
  ring_index( uint32_t slot_rel, uint32_t ch, uint32_t local ) const
  {
    uint32_t slot = s0 + slot_rel;
    if ( slot >= L )
      slot -= L;
    return slot * nct + ch * tb + local;
  }
  __device__ double
  get( uint32_t slot_rel, uint32_t ch, uint32_t local ) const
  {
    if ( in_window( slot_rel ) )
      return lds[ win_index( slot_rel, ch, local ) ];
    return rbb[ ring_index( slot_rel, ch, local ) ];
  }
  __device__ void
  put( uint32_t slot_rel, uint32_t ch, uint32_t local, double v ) const
  {
    if ( in_window( slot_rel ) )
      lds[ win_index( slot_rel, ch, local ) ] = v;
    else
      rbb[ ring_index( slot_rel, ch, local ) ] = v;
  }
};

// Cell cache of cache delivery: keys (ring cell index, ~0 empty) and values share the
// window's LDS area.
constexpr uint32_t EMPTY_KEY = 0xffffffffu;
__device__ inline uint32_t
cache_hash( uint32_t cell )
{
  return ( cell * 2654435761u ) >> 20;  // 12 bits: CACHE_SLOTS = 4096
}
// slot of cell in the cache, or NONE when the cell is not cached
__device__ inline uint32_t
cache_find( lptr< const uint32_t > keys, uint32_t cell )
{
  uint32_t h = cache_hash( cell );
  for ( uint32_t p = 0; p < CACHE_PROBES; ++p )
  {
    const uint32_t k = keys[ h ];
    if ( k == cell )
      return h;
    if ( k == EMPTY_KEY )
      return NONE;
    h = ( h + 1 ) & ( CACHE_SLOTS - 1 );
  }
  return NONE;
}
// inserts cell; returns its slot if this call inserted it (the caller loads its value), else NONE
__device__ inline uint32_t
cache_insert( lptr< uint32_t > keys, uint32_t cell )
{
  uint32_t h = cache_hash( cell );
  for ( uint32_t p = 0; p < CACHE_PROBES; ++p )
  {
    uint32_t k = EMPTY_KEY;
    if ( __hip_atomic_compare_exchange_strong( keys + h, &k, cell, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
           __HIP_MEMORY_SCOPE_WORKGROUP ) )
      return h;
    if ( k == cell )
      return NONE;
    h = ( h + 1 ) & ( CACHE_SLOTS - 1 );
  }
  return NONE;  // no free slot within the probe limit: the cell stays in the ring
}

// handle( SpikeEvent& ) of model m (the models' device halves, ng_models_dev.h)
__device__ __forceinline__ void node_receive( int m, double w, uint32_t& ch, double& v );

// spike_generator drive of this part: handle() of the target with weight * multiplicity.
// Each lag adds into a different cell, so the cells of LB lags are read together.
template < bool WIN >
__device__ inline void
drive_adds( const NetDev& d, gref< const ParamSet > ps, int m, uint32_t dev, uint32_t local, const Cells< WIN >& C,
  int64_t origin, int from, int stride, int to )
{
  const int32_t mode = ps.drive_mode;
  if ( mode == 0 )
    return;
  const uint32_t col = glob( d.drive_idx )[ dev ];
  if ( col == NONE )
    return;
  // the drive parameters in registers (the ring stores below could alias them in HBM)
  const int64_t t0 = ps.drive_t0, n_steps = ps.drive_steps;
  const double w = ps.drive_w;
  const int32_t dd = ps.drive_d;
  PoissonCache PC;
  if ( mode == 2 )
  {
    PC.T = glob( d.ptables ) + ps.drive_table;
    PC.seed = PC.T->seed;
    PC.kmax = PC.T->kmax;
#pragma unroll
    for ( int i = 0; i < KC; ++i )
      PC.c[ i ] = PC.T->cdf[ i ];
  }
  // the channel and multiplier of the weight: the model's handle( SpikeEvent& )
  uint32_t ch;
  double sw;
  node_receive( m, w, ch, sw );
  for ( int l0 = from; l0 < to; l0 += LB * stride )
  {
    double add[ LB ];
    bool has[ LB ];
#pragma unroll
    for ( int q = 0; q < LB; ++q )
    {
      has[ q ] = false;
      const int lag = l0 + q * stride;
      const int64_t e = origin + lag;  // emission step; spike stamp e + 1
      if ( lag >= to or e < t0 or e >= t0 + n_steps )
        continue;
      uint32_t c;
      if ( mode == 1 )
        c = glob( d.drive_counts )[ static_cast< uint64_t >( e - t0 ) * d.drive_cols + col ];
      else
        c = poisson_count( PC, col, e );
      if ( c == 0 )
        continue;
      has[ q ] = true;
      add[ q ] = sw * static_cast< double >( c );
    }
    double old[ LB ];
#pragma unroll
    for ( int q = 0; q < LB; ++q )
      if ( has[ q ] )
        old[ q ] = C.get( static_cast< uint32_t >( l0 + q * stride + dd ), ch, local );
#pragma unroll
    for ( int q = 0; q < LB; ++q )
      if ( has[ q ] )
        C.put( static_cast< uint32_t >( l0 + q * stride + dd ), ch, local, old[ q ] + add[ q ] );
  }
}

// Multimeter sample of one neuron after a step (UniversalDataLogger::record_data,
// nestkernel/universal_data_logger_impl.h:343-375): the model's four state fields.
__device__ inline void
store_sample( const NetDev& d, uint32_t col, int64_t stamp, double f0, double f1, double f2, double f3 )
{
  const int64_t row = ( stamp - d.mm_off ) / d.mm_int - d.mm_row0;
  if ( row < 0 or row >= static_cast< int64_t >( d.mm_rows ) )
  {
    g_or( &d.counters[ 0 ], ERR_SAMPLES );
    return;
  }
  gptr< double > p = glob( d.mm_buf ) + ( static_cast< size_t >( row ) * d.mm_cols + col ) * 4;
  p[ 0 ] = f0;
  p[ 1 ] = f1;
  p[ 2 ] = f2;
  p[ 3 ] = f3;
}

}  // namespace nestgpu

// the neuron models' device halves: update and handle( SpikeEvent& ) per model id
#include "ng_models_dev.h"

namespace nestgpu
{

// Phase timers of the slice kernel, compiled in only with -DNG_PROFILE: per phase, the sum
// over workgroups of wall-clock ticks (read back with nestgpu_profile_read).
#ifdef NG_PROFILE
__device__ unsigned long long g_phase_ticks[ 8 ];
#define NG_PHASE( i )                                                               \
  do                                                                                \
  {                                                                                 \
    __syncthreads();                                                                \
    if ( threadIdx.x == 0 )                                                         \
    {                                                                               \
      const unsigned long long now_ = wall_clock64();                               \
      atomicAdd( &g_phase_ticks[ i ], now_ - t_phase_ );                            \
      t_phase_ = now_;                                                              \
    }                                                                               \
  } while ( 0 )
#else
#define NG_PHASE( i ) \
  do                  \
  {                   \
  } while ( 0 )
#endif

// Two workgroups of 8 waves per CU need at most 128 VGPRs (4 waves per SIMD); instantiations
// with a WIDE model run one workgroup per CU instead.
// ---------------------------------------------------------------- stdp_synapse
// stdp_synapse::send (models/stdp_synapse.h:222-283) for the event of record ridx, emitted at
// step stamp `stamp` in a slice of parity par, to device target tdev over a delay of dsteps:
// facilitation by the target's spikes since the synapse's last presynaptic spike, depression
// by the target's trace. The presynaptic trace and last spike time before this spike come
// from the source's trace row (the source advanced it when it spiked). Returns the new
// weight, which the record keeps. Every record is delivered at most once per slice (a source
// with stdp synapses spikes at most once per slice, checked in the update phase), so no other
// thread touches its weight meanwhile; the target's history is written only in the update
// phase.
constexpr double STDP_EPS = 1.0e-6;  // ConnectionManager::stdp_eps_ (connection_manager.cpp:78)

__device__ inline double
stdp_event( const NetDev& d, uint32_t ridx, double w, uint32_t tdev, int64_t stamp, uint32_t dsteps, uint32_t par )
{
  gref< const StdpPre > pre = glob( const_cast< const StdpPre* >( d.spre ) )[ glob( d.srow )[ ridx ] ];
  const double kplus = pre.kp[ par ], t_last = pre.tl[ par ];
  const double tau_plus = pre.tau_plus, lambda = pre.lambda, alpha = pre.alpha, Wmax = pre.Wmax;
  const bool mu_p = pre.mu_plus != 0u, mu_m = pre.mu_minus != 0u;
  const double t_spike = d.ms_per_tic * static_cast< double >( stamp * d.tics_per_step );  // Time::get_ms
  const double dd = static_cast< double >( dsteps ) * d.ms_per_step;  // Time::delay_steps_to_ms
  // ArchivingNode::get_history( t_last - d, t_spike - d ): entries t1 + eps <= t < t2 + eps
  const double t1_lim = ( t_last - dd ) + STDP_EPS;
  const double t2_lim = ( t_spike - dd ) + STDP_EPS;
  const uint32_t n = glob( d.hist_n )[ tdev ];
  const uint32_t cap = d.stdp_hcap, hm = cap - 1;
  const uint32_t first = n > cap ? n - cap : 0u;  // entries before first left the ring
  gptr< const double > ht = glob( d.hist_t ) + static_cast< size_t >( tdev ) * cap;
  gptr< const double > hk = glob( d.hist_k ) + static_cast< size_t >( tdev ) * cap;
  auto t_at = [ & ]( uint32_t i ) -> double { return ht[ i & hm ]; };
  // the times are increasing: walk back from the newest entry to the range [lo, hi) of
  // entries t1 + eps <= t < t2 + eps, so that an event costs the target spikes in its range
  uint32_t hi = n;
  while ( hi > first and t_at( hi - 1 ) >= t2_lim )
    --hi;
  uint32_t lo = hi;
  while ( lo > first and t_at( lo - 1 ) >= t1_lim )
    --lo;
  // the range (or the latest spike before t2 - eps) may reach past the entries kept
  if ( first > 0 and lo == first )
    g_or( &d.counters[ 0 ], ERR_STDP_HISTORY );
  for ( uint32_t i = lo; i < hi; ++i )
  {
    // facilitate_ (stdp_synapse.h:187-191), oldest entry first
    const double minus_dt = t_last - ( t_at( i ) + dd );
    const double kp = kplus * ng_exp( minus_dt / tau_plus );
    const double norm_w = ( w / Wmax ) + ( lambda * ( mu_p ? 1.0 - ( w / Wmax ) : 1.0 ) * kp );
    w = norm_w < 1.0 ? norm_w * Wmax : Wmax;
  }
  // ArchivingNode::get_K_value( t_spike - d ) (archiving_node.cpp:85-113): the latest spike
  // strictly before (entries from hi on lie after it)
  const double tq = t_spike - dd;
  double K = 0.0;
  for ( uint32_t i = hi; i > first; --i )
  {
    const double th = t_at( i - 1 );
    if ( tq - th > STDP_EPS )
    {
      K = hk[ ( i - 1 ) & hm ] * ng_exp( ( th - tq ) * glob( d.tau_minus_inv )[ tdev ] );
      break;
    }
  }
  // depress_ (stdp_synapse.h:193-198)
  const double norm_w = ( w / Wmax ) - ( alpha * lambda * ( mu_m ? w / Wmax : 1.0 ) * K );
  w = norm_w > 0.0 ? norm_w * Wmax : 0.0;
  uint64_t bits;
  __builtin_memcpy( &bits, &w, sizeof bits );
  gref< SynRec > r = glob( const_cast< SynRec* >( d.syn ) )[ ridx ];
  r.w_lo = static_cast< uint32_t >( bits );
  r.w_hi = static_cast< uint32_t >( bits >> 32 );
  return w;
}

// ArchivingNode::set_spiketime (archiving_node.cpp:196-230) for the spikes of neuron dev at the
// set bits of `bits` (lags of the slice at origin, of parity par, in order), and the source
// side of stdp_synapse::send (stdp_synapse.h:276-279): each of the neuron's presynaptic trace
// rows keeps its state before the spike in slot par and advances Kplus_ and t_lastspike_. A
// neuron with outgoing stdp synapses must spike at most once per slice.
__device__ inline void
stdp_spikes( const NetDev& d, uint32_t dev, int64_t origin, uint32_t par, const uint64_t* bits )
{
  const bool in = glob( d.stdp_in )[ dev ] != 0;
  const uint32_t row0 = glob( d.pre_row0 )[ dev ], row1 = glob( d.pre_row0 )[ dev + 1 ];
  const bool out = row1 > row0;
  if ( not in and not out )
    return;
  double last = glob( d.last_spike )[ dev ];
  const double slice_start = d.ms_per_tic * static_cast< double >( ( origin + 1 ) * d.tics_per_step );
  int count = 0;
  double km = in ? glob( d.kminus )[ dev ] : 0.0;
  uint32_t n = in ? glob( d.hist_n )[ dev ] : 0u;
  for ( int w = 0; w < 4; ++w )
  {
    uint64_t x = bits[ w ];
    while ( x )
    {
      const int lag = w * 64 + ( __ffsll( static_cast< long long >( x ) ) - 1 );
      x &= x - 1;
      const double t_sp = d.ms_per_tic * static_cast< double >( ( origin + lag + 1 ) * d.tics_per_step );
      if ( out and ( count > 0 or last >= slice_start ) )
        g_or( &d.counters[ 0 ], ERR_STDP_TWICE );
      ++count;
      if ( in )
      {
        km = km * ng_exp( ( last - t_sp ) * glob( d.tau_minus_inv )[ dev ] ) + 1.0;
        const size_t at = static_cast< size_t >( dev ) * d.stdp_hcap + ( n & ( d.stdp_hcap - 1 ) );
        glob( d.hist_t )[ at ] = t_sp;
        glob( d.hist_k )[ at ] = km;
        ++n;
      }
      for ( uint32_t r = row0; r < row1; ++r )
      {
        gref< StdpPre > pre = glob( d.spre )[ r ];
        const double kp = pre.cur_kp, tl = pre.cur_tl;
        pre.kp[ par ] = kp;
        pre.tl[ par ] = tl;
        pre.cur_kp = kp * ng_exp( ( tl - t_sp ) / pre.tau_plus ) + 1.0;
        pre.cur_tl = t_sp;
      }
      last = t_sp;
    }
  }
  if ( count )
  {
    glob( d.last_spike )[ dev ] = last;
    if ( in )
    {
      glob( d.kminus )[ dev ] = km;
      glob( d.hist_n )[ dev ] = n;
    }
  }
}

// STDP: some instance of the launch has stdp_synapse connections (its own instantiation, so
// that networks without them carry none of that code or its registers)
template < int MASK, bool WIN, bool STDP = false >
__global__ void __launch_bounds__( Tile< WIN >::BS, ( MASK & WIDE_MODELS ) ? 1 : 4 )
  k_slice( const NetDev* __restrict__ nets, const uint32_t* __restrict__ blk_prefix, int n_nets, int s )
{
#ifdef NG_PROFILE
  unsigned long long t_phase_ = wall_clock64();
#endif
  __shared__ SliceLds< WIN > S;
  constexpr int BS = Tile< WIN >::BS;
  constexpr int SPT = Tile< WIN >::SPT;
  constexpr int CH = Tile< WIN >::CH;
  constexpr int RQ = Tile< WIN >::RQ;
  constexpr int R = Tile< WIN >::R;

  // ---- which instance and block
  const uint32_t gb = xcd_block( blockIdx.x, gridDim.x );
  int lo = 0, hi = n_nets;
  while ( hi - lo > 1 )
  {
    const int mid = ( lo + hi ) >> 1;
    if ( blk_prefix[ mid ] <= gb )
      lo = mid;
    else
      hi = mid;
  }
  const NetDev& d = nets[ lo ];
  const uint32_t b = gb - blk_prefix[ lo ];
  const SlicePhase ph = slice_phase( d, s );
  if ( not ph.active or b >= d.n_blocks )
    return;
  // the block's targets: device indices dev0 .. dev0 + nb - 1
  const uint32_t dev0 = glob( d.bstart )[ b ];
  const uint32_t nb = glob( d.bstart )[ b + 1 ] - dev0;
  // phase timing (counters[4..6]): thread 0 reads the 100 MHz wall clock at the start, after
  // delivery and at the end of the workgroup's work
  const bool ptime = d.phase_timing != 0;
  unsigned long long tp0 = 0, tp1 = 0;
  if ( ptime and threadIdx.x == 0 )
    tp0 = wall_clock64();
  const uint32_t P = d.n_blocks;
  const uint32_t tb = d.tb;
  const int par = ph.parity;
  const int ppar = par ^ 1;
  const uint32_t t = threadIdx.x;

  Cells< WIN > C;
  C.lds = ( lptr< double > )S.win;
  C.nct = static_cast< uint32_t >( d.nch ) * tb;
  C.tb = tb;
  C.L = static_cast< uint32_t >( d.L );
  C.s0 = static_cast< uint32_t >( ph.origin % d.L );
  C.wa = static_cast< uint32_t >( d.wa );
  C.wn = static_cast< uint32_t >( d.wn );
  C.win = WIN and d.win != 0;
  C.rbb = glob( d.rb ) + static_cast< size_t >( b ) * C.L * C.nct;
  // ring bytes per block < 2^31 (checked on the host), so NO_CELL is always out of range
  const __amdgpu_buffer_rsrc_t rb_rsrc = __builtin_amdgcn_make_buffer_rsrc(
    d.rb + static_cast< size_t >( b ) * C.L * C.nct, 0, static_cast< int >( C.L * C.nct * 8u ), 0x00020000 );

  // ================================================================ 0. stage the window
  uint32_t w0 = C.s0 + C.wa;  // ring slot of the window's first slot
  if ( w0 >= C.L )
    w0 -= C.L;
  const uint32_t n_win = C.wn * C.nct;
  const uint32_t before_wrap = ( C.L - w0 ) * C.nct;
  if ( WIN and C.win )
  {
    for ( uint32_t i = t; i < n_win; i += BS )
      S.win[ i ] = C.rbb[ i < before_wrap ? w0 * C.nct + i : i - before_wrap ];
  }
  // cache delivery: an empty cell cache for this slice (keys, then values, in the window area)
  const bool cache = WIN and d.cache != 0 and ph.deliver;
  const lptr< uint32_t > ckey = ( lptr< uint32_t > )reinterpret_cast< uint32_t* >( S.win );
  const lptr< double > cval = ( lptr< double > )( S.win + CACHE_SLOTS / 2 );
  if ( cache )
    for ( uint32_t i = t; i < CACHE_SLOTS; i += BS )
      ckey[ i ] = EMPTY_KEY;

  // ================================================================ 1. delivery
  if ( ph.deliver )
  {
    // prefix of per-block spike counts: an instance without a window in the window kernel
    // (up to PMAX_GLOB blocks) keeps it in the unused window
    uint32_t* const pref = ( WIN and not C.win and not cache ) ? reinterpret_cast< uint32_t* >( S.win ) : S.pref;
    S.bcnt[ 0 ][ t ] = 0;
    S.bcnt[ 1 ][ t ] = 0;
    int cur = 0;  // counter buffer of the current round
    const bool flat = glob( d.flat_meta )[ 2 + ppar ] != 0;
    uint32_t K;
    if ( flat )
      K = glob( d.flat_meta )[ ppar ];
    else
    {
      // prefix of the per-block spike counts of the delivered slice
      gptr< const uint32_t > cnt = glob( d.blk_cnt ) + static_cast< size_t >( ppar ) * P;
      uint32_t carry = 0;
      for ( uint32_t p0 = 0; p0 < P; p0 += BS )
      {
        const uint32_t v = p0 + t < P ? cnt[ p0 + t ] : 0;
        uint32_t ex;
        const uint32_t tot = block_scan< BS >( v, &ex, S.sh );
        if ( p0 + t < P )
          pref[ p0 + t ] = carry + ex;
        carry += tot;
      }
      if ( t == 0 )
        pref[ P ] = carry;
      K = carry;
    }
    __syncthreads();
    NG_PHASE( 0 );
    unsigned long long events = 0;
    // events received by this thread's target (local index t), counted while d.rx is set
    uint32_t rxc = 0;
    for ( uint32_t k0 = 0; k0 < K; k0 += CH )
    {
      // each thread resolves SPT consecutive spikes of the chunk: (source, lag), then its
      // run [g0, g1) in the source's row; the loads of one step are independent
      uint32_t src[ SPT ], lag[ SPT ];
      const uint32_t kb = k0 + t * SPT;
      if ( flat )
      {
#pragma unroll
        for ( int q = 0; q < SPT; ++q )
          if ( kb + q < K )
          {
            src[ q ] = glob( d.flat_src )[ kb + q ];
            lag[ q ] = glob( d.flat_lag )[ kb + q ];
          }
      }
      else if ( kb < K )
      {
        uint32_t blk = upper_bound_u32( pref, P + 1, kb ) - 1;
#pragma unroll
        for ( int q = 0; q < SPT; ++q )
        {
          const uint32_t k = kb + q;
          if ( k < K )
          {
            while ( pref[ blk + 1 ] <= k )
              ++blk;
            const uint32_t e =
              glob( d.blk_spk )[ ( static_cast< size_t >( ppar ) * P + blk ) * d.spk_cap + ( k - pref[ blk ] ) ];
            src[ q ] = e >> 8;
            lag[ q ] = e & 0xffu;
          }
        }
      }
      uint32_t g0v[ SPT ], lenv[ SPT ], s1v[ SPT ], s2v[ SPT ];
#pragma unroll
      for ( int q = 0; q < SPT; ++q )
      {
        lenv[ q ] = 0;
        if ( kb + q < K )
        {
          gptr< const uint32_t > bo = glob( d.boff ) + static_cast< size_t >( src[ q ] ) * ( P + 1 ) + b;
          g0v[ q ] = bo[ 0 ];
          lenv[ q ] = bo[ 1 ] - g0v[ q ];
          S.c_lag[ t * SPT + q ] = static_cast< uint8_t >( lag[ q ] );
        }
      }
      // Slot-range passes: pass p takes the events landing in slots [p ws, (p + 1) ws), i.e.
      // delay offsets [p ws - lag, (p + 1) ws - lag): a contiguous part of the delay-ordered
      // run, split at s1 (and s2 for a third pass).
#pragma unroll
      for ( int q = 0; q < SPT; ++q )
      {
        s1v[ q ] = lenv[ q ];
        s2v[ q ] = lenv[ q ];
        if ( d.np > 1 and lenv[ q ] > 0 )
        {
          s1v[ q ] = run_split( glob( d.syn ), g0v[ q ], lenv[ q ], d.ws - static_cast< int >( lag[ q ] ) );
          if ( d.np > 2 )
            s2v[ q ] = run_split( glob( d.syn ), g0v[ q ], lenv[ q ], 2 * d.ws - static_cast< int >( lag[ q ] ) );
        }
      }
      for ( int pass = 0; pass < d.np; ++pass )
      {
      uint32_t lens[ SPT ];
      uint32_t mine = 0;
#pragma unroll
      for ( int q = 0; q < SPT; ++q )
      {
        const uint32_t lo = pass == 0 ? 0u : ( pass == 1 ? s1v[ q ] : s2v[ q ] );
        const uint32_t hi = pass == d.np - 1 ? lenv[ q ] : ( pass == 0 ? s1v[ q ] : s2v[ q ] );
        lens[ q ] = hi - lo;
        if ( lens[ q ] )
          S.c_g0[ t * SPT + q ] = g0v[ q ] + lo;
        mine += lens[ q ];
      }
      uint32_t ex;
      const uint32_t E = block_scan< BS >( mine, &ex, S.sh );
#pragma unroll
      for ( int q = 0; q < SPT; ++q )
      {
        S.c_off[ t * SPT + q ] = ex;
        ex += lens[ q ];
      }
      if ( t == 0 )
        S.c_off[ CH ] = E;
      __syncthreads();
      NG_PHASE( 1 );
      events += E;
      if ( cache and E )
      {
        // The cells of this chunk's events join the slice's cell cache. Each thread takes a
        // contiguous span of the events, reads their records (the rounds read them again, from
        // L2) and loads the value of every cell it inserts, all loads in flight together: the
        // rounds then add into LDS only. A cell stays cached for the rest of the slice, so all
        // its later additions go to the cache, after the ones it already holds.
        constexpr int HQ = 8;
        const int umodel = d.uniform_model;
        const uint32_t per = ( E + BS - 1 ) / BS;
        const uint32_t e_beg = min( E, t * per ), e_end = min( E, e_beg + per );
        uint32_t j = e_beg < e_end ? upper_bound_u32( S.c_off, CH + 1, e_beg ) - 1 : 0;
        for ( uint32_t e0 = e_beg; e0 < e_end; e0 += HQ )
        {
          SynRec hr[ HQ ];
          uint32_t hl[ HQ ];
#pragma unroll
          for ( int q = 0; q < HQ; ++q )
          {
            const uint32_t e = e0 + q < e_end ? e0 + q : e_end - 1;
            while ( j + 1 < static_cast< uint32_t >( CH ) and S.c_off[ j + 1 ] <= e )
              ++j;
            hr[ q ] = load_rec( glob( d.syn ) + S.c_g0[ j ] + ( e - S.c_off[ j ] ) );
            hl[ q ] = S.c_lag[ j ];
          }
          uint32_t hcell[ HQ ], hslot[ HQ ];
#pragma unroll
          for ( int q = 0; q < HQ; ++q )
          {
            hslot[ q ] = NONE;
            hcell[ q ] = 0;
            if ( e0 + q < e_end )
            {
              const double w = __hiloint2double( static_cast< int >( hr[ q ].w_hi ), static_cast< int >( hr[ q ].w_lo ) );
              const uint32_t local = hr[ q ].word & 0xffffu;
              const uint32_t slot_rel = hl[ q ] + ( ( hr[ q ].word >> 16 ) & DELAY_MASK );
              const int m = single_model< MASK >() >= 0 ? single_model< MASK >()
                : umodel >= 0                           ? umodel
                                                        : glob( d.model )[ dev0 + local ];
              uint32_t ch;
              double v;
              node_receive( m, w, ch, v );
              hcell[ q ] = C.ring_index( slot_rel, ch, local );
              hslot[ q ] = cache_insert( ckey, hcell[ q ] );
            }
          }
          double hv[ HQ ];
#pragma unroll
          for ( int q = 0; q < HQ; ++q )
            hv[ q ] = cell_load( rb_rsrc, hslot[ q ] != NONE ? hcell[ q ] * 8u : NO_CELL );
#pragma unroll
          for ( int q = 0; q < HQ; ++q )
            if ( hslot[ q ] != NONE )
              cval[ hslot[ q ] ] = hv[ q ];
        }
        __syncthreads();
      }
      // rounds of R events; the synapse records of round r + 1 are loaded while round r
      // is bucketed and applied
      SynRec rec[ RQ ];
      uint32_t rlag[ RQ ];
      uint32_t rid[ STDP ? RQ : 1 ];  // stdp_synapse: the records' indices
      // Every thread issues RQ record loads (events past the end reload the last record):
      // with a fixed number of loads in flight the waits count exactly, instead of draining
      // all outstanding memory operations.
      auto load_round = [ & ]( uint32_t e0 )
      {
        const uint32_t eb = e0 + t * RQ;
        uint32_t j = upper_bound_u32( S.c_off, CH + 1, eb < E ? eb : E - 1 ) - 1;
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
        {
          const uint32_t e = eb + q < E ? eb + q : E - 1;
          while ( j + 1 < static_cast< uint32_t >( CH ) and S.c_off[ j + 1 ] <= e )
            ++j;
          const uint32_t ri = S.c_g0[ j ] + ( e - S.c_off[ j ] );
          rec[ q ] = load_rec( glob( d.syn ) + ri );
          rlag[ q ] = S.c_lag[ j ];
          if constexpr ( STDP )
            rid[ q ] = ri;
        }
      };
      if ( E )
        load_round( 0 );
      const int umodel = d.uniform_model;
      const uint32_t dpar = static_cast< uint32_t >( ph.parity ^ 1 );  // the delivered spikes' slice
      (void)dpar;
      for ( uint32_t e0 = 0; e0 < E; e0 += R, cur ^= 1 )
      {
        // stage: thread t holds events e0 + t*RQ .. + RQ - 1 (key order = round index) and
        // takes an unordered place in its target's bucket
        uint32_t key[ RQ ], loc[ RQ ], dst[ RQ ], ci[ RQ ], goff[ RQ ];
        double val[ RQ ], cv[ RQ ];
        bool lw[ RQ ];  // the event's cell is in the LDS window
        bool lc[ RQ ];  // the event's cell is in the cell cache (ci = its cache slot)
        const uint32_t eb = e0 + t * RQ;
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
        {
          loc[ q ] = NONE;
          goff[ q ] = NO_CELL;
          lc[ q ] = false;
          if ( eb + q < E )
          {
            double w = __hiloint2double( static_cast< int >( rec[ q ].w_hi ), static_cast< int >( rec[ q ].w_lo ) );
            const uint32_t local = rec[ q ].word & 0xffffu;
            if constexpr ( STDP )
            {
              if ( rec[ q ].word & PLASTIC_BIT )
                w = stdp_event( d, rid[ q ], w, dev0 + local, ph.origin - d.min_delay + rlag[ q ] + 1,
                  ( ( rec[ q ].word >> 16 ) & DELAY_MASK ) + static_cast< uint32_t >( d.min_delay ), dpar );
            }
            const uint32_t slot_rel = rlag[ q ] + ( ( rec[ q ].word >> 16 ) & DELAY_MASK );
            const int m = single_model< MASK >() >= 0 ? single_model< MASK >()
              : umodel >= 0                           ? umodel
                                                      : glob( d.model )[ dev0 + local ];
            uint32_t ch;
            double v;
            node_receive( m, w, ch, v );  // static_synapse: weight * multiplicity 1
            loc[ q ] = local;
            val[ q ] = v;
            key[ q ] = ( ( slot_rel * d.nch + ch ) << IDX_BITS ) | static_cast< uint32_t >( t * RQ + q );
            dst[ q ] = atomicAdd( &S.bcnt[ cur ][ local ], 1u );
            lw[ q ] = C.in_window( slot_rel );
            if ( lw[ q ] )
              ci[ q ] = ( slot_rel - C.wa ) * C.nct + ch * tb + local;
            else
            {
              uint32_t slot = C.s0 + slot_rel;
              if ( slot >= C.L )
                slot -= C.L;
              ci[ q ] = slot * C.nct + ch * tb + local;
              const uint32_t hs = cache ? cache_find( ckey, ci[ q ] ) : NONE;
              if ( hs != NONE )
              {
                lc[ q ] = true;
                ci[ q ] = hs;
              }
              else
                goff[ q ] = ci[ q ] * 8u;
            }
          }
        }
        // the records of this round are consumed: load the next round's now, so they have the
        // whole round to arrive (past the last round: reloads record E - 1, unused)
        load_round( e0 + R );
        __syncthreads();
        NG_PHASE( 2 );  // stage
        // every thread has finished the previous round (its cell stores included): the cells
        // of this round can be read now, while the buckets are formed
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
          cv[ q ] = cell_load( rb_rsrc, goff[ q ] );
        // bucket starts: wave 0 scans the per-target counts, BS / 64 per lane; the others
        // clear the counters of the next round
        if ( t < 64 )
        {
          constexpr int CPL = BS / 64;
          uint32_t c[ CPL ];
          uint32_t sum = 0;
#pragma unroll
          for ( int j = 0; j < CPL; ++j )
          {
            c[ j ] = S.bcnt[ cur ][ CPL * t + j ];
            sum += c[ j ];
          }
          uint32_t run = wave_incl_scan( sum ) - sum;
#pragma unroll
          for ( int j = 0; j < CPL; ++j )
          {
            S.bst[ CPL * t + j ] = run;
            run += c[ j ];
          }
          if ( t == 63 )
            S.bst[ BS ] = run;
        }
        S.bcnt[ cur ^ 1 ][ t ] = 0;
        if ( d.rx )
          rxc += S.bcnt[ cur ][ t ];
        __syncthreads();
        NG_PHASE( 3 );  // bucket starts
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
          if ( loc[ q ] != NONE )
          {
            dst[ q ] += S.bst[ loc[ q ] ];
            S.bk[ dst[ q ] ] = key[ q ];
            S.bw[ dst[ q ] ] = val[ q ];
          }
        __syncthreads();
        // sorted place of each event in its bucket by (cell, key): the number of smaller keys
        uint32_t srt[ RQ ];
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
          if ( loc[ q ] != NONE )
          {
            const uint32_t lo = S.bst[ loc[ q ] ], hi = S.bst[ loc[ q ] + 1 ];
            uint32_t r = 0;
            uint32_t i = lo;
            for ( ; i + 4 <= hi; i += 4 )
            {
              const uint32_t k0 = S.bk[ i ], k1 = S.bk[ i + 1 ], k2 = S.bk[ i + 2 ], k3 = S.bk[ i + 3 ];
              r += ( k0 < key[ q ] ? 1u : 0u ) + ( k1 < key[ q ] ? 1u : 0u ) + ( k2 < key[ q ] ? 1u : 0u )
                + ( k3 < key[ q ] ? 1u : 0u );
            }
            for ( ; i < hi; ++i )
              r += S.bk[ i ] < key[ q ] ? 1u : 0u;
            srt[ q ] = lo + r;
            S.perm[ srt[ q ] ] = static_cast< uint16_t >( dst[ q ] );
          }
        __syncthreads();
        NG_PHASE( 4 );  // place and rank
        // the first event of each cell's run adds the whole run in key order; runs of
        // different cells proceed in parallel
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
        {
          if ( loc[ q ] == NONE )
            continue;
          const uint32_t sc = key[ q ] >> IDX_BITS;
          const uint32_t lo = S.bst[ loc[ q ] ], hi = S.bst[ loc[ q ] + 1 ];
          if ( srt[ q ] > lo and ( S.bk[ S.perm[ srt[ q ] - 1 ] ] >> IDX_BITS ) == sc )
          {
            loc[ q ] = NONE;  // not the first of its run
            goff[ q ] = NO_CELL;
            continue;
          }
          double v = lw[ q ] ? S.win[ ci[ q ] ] : lc[ q ] ? cval[ ci[ q ] ] : cv[ q ];
          v = v + val[ q ];
          for ( uint32_t i = srt[ q ] + 1; i < hi; ++i )
          {
            const uint32_t pi = S.perm[ i ];
            if ( ( S.bk[ pi ] >> IDX_BITS ) != sc )
              break;
            v = v + S.bw[ pi ];
          }
          cv[ q ] = v;
        }
#pragma unroll
        for ( int q = 0; q < RQ; ++q )
        {
          if ( loc[ q ] != NONE and lw[ q ] )
            S.win[ ci[ q ] ] = cv[ q ];
          else if ( loc[ q ] != NONE and lc[ q ] )
            cval[ ci[ q ] ] = cv[ q ];
          cell_store( rb_rsrc, goff[ q ], cv[ q ] );  // NO_CELL unless a global run leader
        }
        NG_PHASE( 5 );  // apply
      }
      }  // passes
    }
    if ( cache )
    {
      // write the cached cells back to the ring: the update and the drive read them there
      __syncthreads();
      for ( uint32_t i = t; i < CACHE_SLOTS; i += BS )
      {
        const uint32_t k = ckey[ i ];
        cell_store( rb_rsrc, k != EMPTY_KEY ? k * 8u : NO_CELL, cval[ i ] );
      }
    }
    if ( t == 0 and events )
      g_add( &d.counters[ 1 ], events );
    // the block owns its targets' counts: one plain read-modify-write per target and slice
    if ( d.rx and t < nb and rxc )
      glob( d.rx )[ dev0 + t ] += rxc;
  }
  __syncthreads();  // window staged and delivered
  if ( ptime and t == 0 )
    tp1 = wall_clock64();

  // ================================================================ 2. update
  // Threads [0, nb) update one neuron each. The spike_generator drive of the slice lands in
  // cells at least min_delay ahead (every delay, the drive's included, is >= min_delay), which
  // the update does not touch, and in distinct cells per lag: when nb <= BS / 2 the spare
  // threads add it concurrently, G groups of nb threads splitting the lags.
  const uint32_t local = t;
  const uint32_t dev = dev0 + local;
  const uint32_t G = ( static_cast< uint32_t >( BS ) - nb ) / nb;
  uint64_t bits[ 4 ] = { 0, 0, 0, 0 };
  if ( local >= nb and local < nb + G * nb )
  {
    const uint32_t dl = ( local - nb ) % nb, g = ( local - nb ) / nb, ddev = dev0 + dl;
    {
      gref< const ParamSet > dps = glob( d.psets )[ glob( d.pidx )[ ddev ] ];
      const int dm = single_model< MASK >() >= 0 ? single_model< MASK >() : dps.model;
      drive_adds( d, dps, dm, ddev, dl, C, ph.origin, ph.from + static_cast< int >( g ), static_cast< int >( G ), ph.to );
    }
  }
  if ( local < nb )
  {
    gref< const ParamSet > ps = glob( d.psets )[ glob( d.pidx )[ dev ] ];
    // multimeter: the next step of this part whose state is sampled (ph.to: none)
    const uint32_t mcol = d.mm_int > 0 ? glob( d.mm_col )[ dev ] : NONE;
    int next_sample = mcol != NONE ? first_sample_lag( d, ph.origin, ph.from, ph.to ) : ph.to;
    const int m = single_model< MASK >() >= 0 ? single_model< MASK >() : ps.model;
    // Node::update of the neuron's model (each model of the launch's mask is compiled in)
#define NG_UPDATE( ID )                                                                 \
  if ( ( MASK & ( 1 << ( ID ) ) ) and m == ( ID ) )                                   \
    DeviceNode< ( ID ) >::update( d, ps, dev, local, C, ph, bits, mcol, next_sample );
    NG_DEVICE_MODELS( NG_UPDATE )
#undef NG_UPDATE
    if ( G == 0 )
      drive_adds( d, ps, m, dev, local, C, ph.origin, ph.from, 1, ph.to );
  }
  if ( STDP and d.stdp and local < nb )
    stdp_spikes( d, dev, ph.origin, static_cast< uint32_t >( ph.parity ), bits );
  NG_PHASE( 6 );

  // ================================================================ 3. spike lists
  const uint32_t nsp = __popcll( bits[ 0 ] ) + __popcll( bits[ 1 ] ) + __popcll( bits[ 2 ] ) + __popcll( bits[ 3 ] );
  uint32_t ex;
  const uint32_t tot = block_scan< BS >( nsp, &ex, S.sh );  // also orders the update before the write-back
  if ( WIN and C.win )
  {
    for ( uint32_t i = t; i < n_win; i += BS )
      C.rbb[ i < before_wrap ? w0 * C.nct + i : i - before_wrap ] = S.win[ i ];
  }
  gptr< uint32_t > cnt = glob( d.blk_cnt ) + static_cast< size_t >( par ) * P + b;
  if ( t == 0 )
  {
    const uint32_t base = ph.from == 0 ? 0u : *cnt;  // a new slice starts an empty list
    S.cnt_base = base;
    if ( base + tot > d.spk_cap )
      g_or( &d.counters[ 0 ], ERR_SPIKE_LIST );
    *cnt = base + tot < d.spk_cap ? base + tot : d.spk_cap;
    if ( ph.from == 0 and b == 0 )
      glob( d.flat_meta )[ 2 + par ] = 0;
    S.log_base = 0ull;
    if ( tot )
    {
      g_add( &d.counters[ 2 ], static_cast< unsigned long long >( tot ) );
      if ( d.log_enabled )
        S.log_base = g_add( d.log_n, static_cast< unsigned long long >( tot ) );
    }
  }
  __syncthreads();
  const bool log_on = d.log_enabled != 0;
  if ( nsp )
  {
    uint32_t o = S.cnt_base + ex;
    unsigned long long li = S.log_base + ex;
    gptr< uint32_t > lst = glob( d.blk_spk ) + ( static_cast< size_t >( par ) * P + b ) * d.spk_cap;
    gptr< uint8_t > lpart = glob( d.blk_part ) + ( static_cast< size_t >( par ) * P + b ) * d.spk_cap;
    for ( int w = 0; w < 4; ++w )
    {
      unsigned long long x = bits[ w ];
      while ( x )
      {
        const uint32_t lag = w * 64 + ( __ffsll( static_cast< long long >( x ) ) - 1 );
        x &= x - 1;
        if ( o < d.spk_cap )
        {
          lst[ o ] = ( dev << 8 ) | lag;
          lpart[ o ] = static_cast< uint8_t >( ph.part );
        }
        ++o;
        if ( log_on )
        {
          if ( li < d.log_cap )
          {
            glob( d.log_src )[ li ] = dev;
            glob( d.log_step )[ li ] = ph.origin + lag + 1;
          }
          else
            g_or( &d.counters[ 0 ], ERR_LOG );
          ++li;
        }
      }
    }
  }
  NG_PHASE( 7 );
  if ( ptime )
  {
    __syncthreads();
    if ( t == 0 )
    {
      const unsigned long long tp2 = wall_clock64();
      g_add( &d.counters[ 4 ], tp1 - tp0 );
      g_add( &d.counters[ 5 ], tp2 - tp0 );
      g_add( &d.counters[ 6 ], 1ull );
    }
  }
}

#ifdef NG_PROFILE
extern "C" int
nestgpu_profile_read( unsigned long long* out8, int reset )
{
  if ( hipMemcpyFromSymbol( out8, HIP_SYMBOL( g_phase_ticks ), sizeof( unsigned long long ) * 8 ) != hipSuccess )
    return 1;
  if ( reset )
  {
    unsigned long long z[ 8 ] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    if ( hipMemcpyToSymbol( HIP_SYMBOL( g_phase_ticks ), z, sizeof z ) != hipSuccess )
      return 1;
  }
  return 0;
}
#endif

// A slice split across Run calls: NEST's per-thread registers hold spikes part by part,
// so its receive order is (VP, part, node, lag). Rebuild that order once the slice is
// complete. One workgroup: for each (VP, part), each thread takes one block of the VP's
// range, counts its matching entries, a block scan gives the output offsets, and each
// thread copies its entries in list order.
constexpr int FLATTEN_BS = 1024;

__global__ void __launch_bounds__( FLATTEN_BS )
  k_flatten( NetDev* net, int par, int n_parts )
{
  __shared__ uint32_t sh[ FLATTEN_BS / 64 ];
  const NetDev& d = *net;
  const uint32_t P = d.n_blocks;
  const uint32_t t = threadIdx.x;
  uint32_t o = 0;  // output entries so far (the same in every thread)
  for ( int vp = 0; vp < d.n_vp; ++vp )
  {
    const uint32_t lo = d.vp_begin[ vp ], hi = d.vp_begin[ vp + 1 ];
    if ( lo >= hi )
      continue;
    // the blocks holding device indices lo .. hi - 1
    const uint32_t b0 = upper_bound_u32( d.bstart, P + 1, lo ) - 1;
    const uint32_t b1 = upper_bound_u32( d.bstart, P + 1, hi - 1 );
    for ( int p = 0; p < n_parts; ++p )
      for ( uint32_t c0 = b0; c0 < b1; c0 += FLATTEN_BS )
      {
        const uint32_t b = c0 + t;
        uint32_t n = 0, cnt = 0;
        const uint32_t* lst = nullptr;
        const uint8_t* lpart = nullptr;
        if ( b < b1 )
        {
          n = min( d.blk_cnt[ static_cast< size_t >( par ) * P + b ], d.spk_cap );
          lst = d.blk_spk + ( static_cast< size_t >( par ) * P + b ) * d.spk_cap;
          lpart = d.blk_part + ( static_cast< size_t >( par ) * P + b ) * d.spk_cap;
          for ( uint32_t q = 0; q < n; ++q )
          {
            const uint32_t dev = lst[ q ] >> 8;
            cnt += lpart[ q ] == p and dev >= lo and dev < hi ? 1u : 0u;
          }
        }
        uint32_t ex;
        const uint32_t tot = block_scan< FLATTEN_BS >( cnt, &ex, sh );
        uint32_t w = o + ex;
        for ( uint32_t q = 0; q < n; ++q )
        {
          const uint32_t e = lst[ q ];
          const uint32_t dev = e >> 8;
          if ( lpart[ q ] != p or dev < lo or dev >= hi )
            continue;
          d.flat_src[ w ] = dev;
          d.flat_lag[ w ] = e & 0xffu;
          ++w;
        }
        o += tot;
      }
  }
  if ( t == 0 )
  {
    d.flat_meta[ par ] = o;
    d.flat_meta[ 2 + par ] = 1;
  }
}

// ------------------------------------------------------------------ launch wrappers
template < bool WIN >
static void
launch_mask( const NetDev* nets, const uint32_t* blk_prefix, int n_nets, uint32_t total_blocks, int mask, int s,
  size_t dyn_lds, hipStream_t st )
{
  // mask: one bit per model id of the launch, and MASK_STDP for stdp_synapse. A launch of one
  // model takes that model's own instantiation, any other the one with every model.
  const bool stdp = ( mask & MASK_STDP ) != 0;
  const int models = mask & ALL_MODELS;
#define NG_LAUNCH( M )                                                                                        \
  ( stdp ? k_slice< M, WIN, true ><<< total_blocks, Tile< WIN >::BS, dyn_lds, st >>>( nets, blk_prefix, n_nets, s ) \
         : k_slice< M, WIN, false ><<< total_blocks, Tile< WIN >::BS, dyn_lds, st >>>( nets, blk_prefix, n_nets, s ) )
#define NG_ONE( ID )                 \
  if ( models == ( 1 << ( ID ) ) ) \
  {                                  \
    NG_LAUNCH( 1 << ( ID ) );        \
    return;                          \
  }
  NG_DEVICE_MODELS( NG_ONE )
  NG_LAUNCH( ALL_MODELS );
#undef NG_ONE
#undef NG_LAUNCH
}

// Dynamic LDS that limits the slice kernel to wg_per_cu resident workgroups per CU (0: no
// limit): fewer resident blocks keep the ring buffers they update within the XCD's L2.
size_t
slice_lds_for_occupancy( int wg_per_cu, bool win )
{
  if ( wg_per_cu <= 0 )
    return 0;
  hipFuncAttributes a;
  const hipError_t e = win ? hipFuncGetAttributes( &a, reinterpret_cast< const void* >( &k_slice< ALL_MODELS, true > ) )
                           : hipFuncGetAttributes( &a, reinterpret_cast< const void* >( &k_slice< ALL_MODELS, false > ) );
  if ( e != hipSuccess )
    return 0;
  const size_t lds_per_cu = 160 * 1024;
  const size_t want = lds_per_cu / static_cast< size_t >( wg_per_cu + 1 ) + 256;  // more than 1/(k+1)
  return want > a.sharedSizeBytes ? want - a.sharedSizeBytes : 0;
}

hipError_t
launch_slice( const NetDev* nets, const uint32_t* blk_prefix, int n_nets, uint32_t total_blocks, int mask, bool win,
  int s, size_t dyn_lds, hipStream_t st )
{
  if ( win )
    launch_mask< true >( nets, blk_prefix, n_nets, total_blocks, mask, s, dyn_lds, st );
  else
    launch_mask< false >( nets, blk_prefix, n_nets, total_blocks, mask, s, dyn_lds, st );
  return hipGetLastError();
}

__global__ void
k_status( const NetDev* nets, int n_nets, unsigned long long* out )
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if ( i < n_nets )
  {
    out[ 2 * i ] = nets[ i ].counters[ 0 ];
    out[ 2 * i + 1 ] = *nets[ i ].log_n;
  }
}

hipError_t
launch_status( const NetDev* nets, int n_nets, unsigned long long* out, hipStream_t st )
{
  if ( n_nets > 0 )
    k_status<<< ( n_nets + 255 ) / 256, 256, 0, st >>>( nets, n_nets, out );
  return hipGetLastError();
}

hipError_t
launch_flatten( NetDev* net, int par, int n_parts, hipStream_t st )
{
  k_flatten<<< 1, FLATTEN_BS, 0, st >>>( net, par, n_parts );
  return hipGetLastError();
}

}  // namespace nestgpu
