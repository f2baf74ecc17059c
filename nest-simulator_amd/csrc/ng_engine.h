// ng_engine.h -- device store and slice runner for network instances.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <unordered_map>
#include <vector>

#include "ng_device.h"

namespace nestgpu
{
// Allocator whose value-initialisation leaves elements default-initialised (uninitialised for
// trivial types): the store builder writes every element of its large tables itself, from
// several threads, so the pages are first touched there instead of by one zeroing pass.
template < typename T >
struct NoInitAlloc : std::allocator< T >
{
  template < typename U >
  struct rebind
  {
    using other = NoInitAlloc< U >;
  };
  NoInitAlloc() = default;
  template < typename U >
  NoInitAlloc( const NoInitAlloc< U >& ) noexcept
  {
  }
  template < typename U >
  void
  construct( U* p )
  {
    ::new ( static_cast< void* >( p ) ) U;
  }
  template < typename U, typename... Args >
  void
  construct( U* p, Args&&... args )
  {
    ::new ( static_cast< void* >( p ) ) U( std::forward< Args >( args )... );
  }
};
template < typename T >
using HostArray = std::vector< T, NoInitAlloc< T > >;
}  // namespace nestgpu

namespace nestgpu
{

// Enqueue the slice kernel for slice s of a chunk over all blocks of n_nets instances.
// mask: bit m set if neuron model m occurs (selects a specialised kernel); win: some
// instance uses window delivery (selects the kernel with the LDS ring-buffer window).
hipError_t launch_slice( const NetDev* nets, const uint32_t* blk_prefix, int n_nets, uint32_t total_blocks, int mask,
  bool win, int s, size_t dyn_lds, hipStream_t st );
// Dynamic LDS bytes that cap the slice kernel at wg_per_cu workgroups per CU (0: no cap).
size_t slice_lds_for_occupancy( int wg_per_cu, bool win );
// Rebuild the exact receive order of a slice that was split across Run calls.
hipError_t launch_flatten( NetDev* net, int par, int n_parts, hipStream_t st );
// out[2 i] = error flags, out[2 i + 1] = spike-log fill of instance i
hipError_t launch_status( const NetDev* nets, int n_nets, unsigned long long* out, hipStream_t st );

constexpr int N_COUNTERS = 8;  // words of NetDev::counters

enum KernelKind
{
  K_SLICE = 0,
  K_FLATTEN,
  N_KINDS
};

// Everything the engine needs to place one network instance in HBM. Produced by the host
// kernel at Prepare (the counterpart of update_connection_infrastructure and prepare_nodes,
// nestkernel/simulation_manager.cpp:507-570,750-833).
struct BuildSpec
{
  int n_vp = 1;
  int min_delay = 1, max_delay = 1;
  uint32_t n = 0;
  uint32_t tb = 256;
  int nch = 2;       // ring-buffer channels per slot
  bool win = false;  // window delivery (LDS-staged ring-buffer slots)
  bool cache = false;  // cache delivery (LDS cell cache per slice)
  int wa = 0, wn = 0;  // window: first slot relative to the slice origin, slots
  int np = 1, ws = 1;  // global delivery: slot-range passes and slots per pass
  int uniform_model = -1;
  std::vector< uint32_t > vp_begin;  // device indices, n_vp + 1 entries
  std::vector< ParamSet > psets;
  std::vector< uint16_t > pidx;
  std::vector< uint8_t > model;
  std::vector< double > v[ 5 ];
  std::vector< int32_t > r;
  // out-CSR by source device index, each row grouped by target block (creation order
  // within a block); boff[s * (n_blocks + 1) + b] = index in syn of block b's run of row s
  HostArray< uint32_t > boff;
  HostArray< SynRec > syn;
  // drive
  std::vector< uint8_t > drive_counts;
  int64_t drive_cols = 0;
  std::vector< uint32_t > drive_idx;
  std::vector< PoissonTable > ptables;
  uint64_t log_cap = 1 << 20;
  bool log_enabled = false;
  // stdp_synapse: parameter sets, state per synapse record (empty without stdp), neurons
  // with incoming / outgoing stdp synapses, 1/tau_minus per neuron; the time base
  std::vector< StdpPrm > stdp_prm;
  std::vector< StdpPre > stdp_pre;       // presynaptic trace rows by source
  std::vector< uint32_t > stdp_pre_row0;  // n + 1 row starts by device index
  HostArray< uint32_t > stdp_srow;        // trace row of each synapse record
  std::vector< uint8_t > stdp_in;
  std::vector< double > tau_minus_inv;
  double ms_per_tic = 0.001, ms_per_step = 0.1;
  uint32_t stdp_hcap = STDP_HIST_DEFAULT;  // post-spike history entries per neuron (power of two)
  int64_t tics_per_step = 100;
  // multimeter sampling: per neuron sample column (~0 none), interval and offset in steps
  std::vector< uint32_t > mm_col;
  uint32_t mm_cols = 0;
  int32_t mm_int = 0;
  int64_t mm_off = 0;
  // target blocks: block q holds device indices [bstart[q], bstart[q + 1]), at most tb each
  std::vector< uint32_t > bstart;
  uint32_t
  n_blocks() const
  {
    return bstart.empty() ? ( n == 0 ? 0 : ( n + tb - 1 ) / tb ) : static_cast< uint32_t >( bstart.size() - 1 );
  }
};

class Engine
{
public:
  explicit Engine( int device );
  ~Engine();
  Engine( const Engine& ) = delete;
  Engine& operator=( const Engine& ) = delete;

  // First call allocates everything; later calls (connections changed) replace the
  // synapse store and inputs and keep state, ring buffers and pending spikes.
  void build( const BuildSpec& b, bool keep_state );

  // Same network over new target blocks (b differs from the current layout in bstart and tb
  // only): the store is replaced and the ring buffers and pending spike lists are moved to
  // the new blocks. Only at a slice boundary (part == 0) whose pending lists hold one run part;
  // returns false, changing nothing, otherwise.
  bool relayout( const BuildSpec& b );
  const std::vector< uint32_t >&
  block_start() const
  {
    return bstart_;
  }

  // weights of the synapse records in store order (stdp_synapse changes them)
  void read_weights( std::vector< double >& w ) const;

  // received events per target, counted by the slice kernel while switched on
  void count_received( bool on );
  void read_received( std::vector< uint32_t >& out ) const;
  void clear_received();

  void set_state( int field, uint32_t dev, double v );
  void get_state( int field, std::vector< double >& out );
  void get_r( std::vector< int32_t >& out );

  // Moves the device spike log (device index, stamp step) into log_src / log_step.
  void drain_log();
  std::vector< uint32_t > log_src;
  std::vector< int64_t > log_step;
  // Multimeter samples of the last chunk: the rows its launches wrote are appended to
  // mm_stamps (one stamp per row) and mm_data ([row][column][4]).
  void drain_samples();
  std::vector< int64_t > mm_stamps;
  std::vector< double > mm_data;
  int64_t mm_chunk_row0 = 0, mm_chunk_rows = 0;  // set when a chunk is staged

  NetDev&
  net()
  {
    return net_;
  }
  const NetDev&
  net() const
  {
    return net_;
  }
  int
  model_mask() const
  {
    return mask_;
  }
  bool
  window() const
  {
    return net_.win != 0;
  }
  bool
  cache() const
  {
    return net_.cache != 0;
  }
  int
  device() const
  {
    return device_;
  }
  bool
  built() const
  {
    return built_;
  }
  uint64_t
  device_bytes() const
  {
    return bytes_;
  }
  unsigned long long read_counter( int i ) const;
  void read_counters( unsigned long long* out ) const;  // all N_COUNTERS words
  void clear_error_flags();
  void clear_phase_ticks();

  // host bookkeeping of NEST's clock (SimulationManager::clock_, from_step_) and of the
  // run parts the current slice has been split into
  int64_t clock = 0;
  int32_t from_step = 0;
  int32_t part = 0;

private:
  void* alloc( size_t bytes );
  template < typename T, typename A >
  T* upload( const std::vector< T, A >& v );
  void free_list( std::vector< void* >& l );
  void build_layout( const BuildSpec& b, bool keep_state, bool relayout );

  int device_;
  bool built_ = false;
  int mask_ = 0;
  NetDev net_ {};
  std::vector< void* > state_allocs_, store_allocs_, log_allocs_, mm_allocs_;
  std::vector< uint32_t > bstart_;  // target-block bounds of the current layout
  uint32_t* rx_ = nullptr;          // received-event counts (state allocation)
  size_t n_syn_ = 0;                // synapse records of the store
  std::unordered_map< void*, size_t > sizes_;  // bytes of every live allocation
  uint64_t bytes_ = 0;
};

// Runs chunks of slices for a set of engines on one device and stream, replaying a
// captured hipGraph (one kernel node per slice) for each chunk.
class Runner
{
public:
  Runner( int device, int slices_per_chunk );
  ~Runner();
  // Advance every engine by steps[i] (0 = idle). Throws on device-side error flags.
  void run( const std::vector< Engine* >& engines, const std::vector< int64_t >& steps );
  // One chunk (at most slices_per_chunk slices of each engine); remaining[i] is reduced by
  // the steps done. Returns false when no engine had work.
  bool run_chunk( const std::vector< Engine* >& engines, std::vector< int64_t >& remaining );
  double
  last_kernel_ms() const
  {
    return last_ms_;
  }
  int
  slices_per_chunk() const
  {
    return S_;
  }
  hipStream_t
  stream() const
  {
    return stream_;
  }
  bool use_graph = true;  // replay captured graphs where possible (false: direct launches)
  // chunks run from a captured graph and by direct launches, and failed captures
  int64_t graph_chunks = 0, direct_chunks = 0, capture_failures = 0;
  int wg_per_cu = 0;  // resident slice workgroups per CU (0: as many as fit)
  // per-kernel timing: one slice at a time with HIP events around every launch
  bool time_kernels = false;
  double kernel_ms[ N_KINDS ] = { 0, 0 };
  int64_t kernel_launches[ N_KINDS ] = { 0, 0 };

private:
  void ensure_capacity( size_t n );
  bool capture( int n_nets, uint32_t total_blocks, int mask, bool win );  // false: launch directly
  void launch_one( int n, uint32_t total, int mask, bool win, int s, hipEvent_t* ev );
  int device_;
  int S_;
  hipStream_t stream_ = nullptr;
  NetDev* d_nets_ = nullptr;
  NetDev* h_nets_ = nullptr;  // pinned
  uint32_t* d_prefix_ = nullptr;
  unsigned long long* d_status_ = nullptr;
  unsigned long long* h_status_ = nullptr;  // pinned
  hipEvent_t ev_chunk_[ 2 ] = { nullptr, nullptr };
  hipEvent_t ev_kernel_[ 2 ] = { nullptr, nullptr };
  size_t cap_ = 0;
  hipGraphExec_t exec_ = nullptr;
  bool exec_failed_ = false;  // the capture of the current launch shape failed
  int exec_n_ = -1;
  uint32_t exec_total_ = 0;
  int exec_mask_ = 0;
  bool exec_win_ = false;
  int exec_wg_ = 0;
  std::vector< uint32_t > prefix_;
  double last_ms_ = 0;
};

void hip_check( hipError_t e, const char* what );

// Chunked slice loop of prepared instances that allocates nothing: the launch records, the
// block prefix and the status words live in a fixed device array of the library (one copy
// per device, its address taken once per device at the first call there) instead of
// per-runner allocations; slices are launched directly on the device's null stream, without
// graph capture. At most NOALLOC_MAX_NETS instances per call. Calls that overlap on one device
// share that array and race.
constexpr int NOALLOC_MAX_NETS = 64;
void run_noalloc( int device, const std::vector< Engine* >& engines, const std::vector< int64_t >& steps,
  int slices_per_chunk );

}  // namespace nestgpu
