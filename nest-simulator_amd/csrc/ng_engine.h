// ng_engine.h -- device store and slice runner for network instances.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "ng_device.h"

namespace nestgpu
{

// workgroups of k_big; each owns one merge-scratch region of the largest record capacity
constexpr int NG_GRID_BIG = 512;

struct LaunchCfg
{
  int grid_expand, grid_big, grid_update, grid_compact;
};
// Enqueues the kernels of slice s (of a chunk) for n_nets instances.
hipError_t launch_slice( const NetDev* nets, int n_nets, int s, const bool* has_model, const LaunchCfg& cfg,
  hipStream_t st );

// Everything the engine needs to place one network instance in HBM. Produced by the
// host kernel at Prepare (the counterpart of update_connection_infrastructure and
// prepare_nodes, nestkernel/simulation_manager.cpp:507-570,750-833).
struct BuildSpec
{
  int n_vp = 1;
  int min_delay = 1, max_delay = 1;
  uint32_t n = 0;
  uint32_t model_begin[ N_MODELS + 1 ] = { 0, 0, 0, 0 };
  std::vector< uint32_t > dev2ord;  // empty: identity
  std::vector< uint32_t > vp_begin;  // order positions, n_vp + 1 entries
  std::vector< ParamSet > psets;
  std::vector< uint16_t > pidx[ N_MODELS ];
  std::vector< double > st[ N_MODELS ][ 5 ];
  std::vector< int32_t > r[ N_MODELS ];
  // out-CSR by source device index; rows sorted by delay, creation order within a delay
  std::vector< uint64_t > row_ptr;
  std::vector< uint32_t > tgt_d;
  std::vector< double > w;
  int tgt_bits = 32;
  std::vector< uint64_t > in_ptr;
  // drive
  std::vector< uint8_t > drive_counts;
  int64_t drive_cols = 0;
  std::vector< uint32_t > drive_idx;  // per device index; 0xffffffff = none
  std::vector< PoissonTable > ptables;
  int s_cap = 1;
  uint64_t log_cap = 1 << 20;
};

class Engine
{
public:
  explicit Engine( int device );
  ~Engine();
  Engine( const Engine& ) = delete;
  Engine& operator=( const Engine& ) = delete;

  // First call allocates everything; later calls (connections changed) replace the
  // synapse store and delivery scratch and keep state and ring buffers.
  void build( const BuildSpec& b, bool keep_state );

  void set_state( int model, int field, uint32_t j, double v );
  void set_r( int model, uint32_t j, int32_t v );
  void get_state( int model, int field, std::vector< double >& out );
  void get_r( int model, std::vector< int32_t >& out );

  // Moves the device spike log (device index, stamp step) to log_src/log_step.
  void drain_log();
  std::vector< uint32_t > log_src;
  std::vector< int64_t > log_step;

  NetDev& net()
  {
    return net_;
  }
  const NetDev& net() const
  {
    return net_;
  }
  bool has_model( int m ) const
  {
    return net_.model_begin[ m + 1 ] > net_.model_begin[ m ];
  }
  int device() const
  {
    return device_;
  }
  bool built() const
  {
    return built_;
  }
  uint64_t device_bytes() const
  {
    return bytes_;
  }
  // host bookkeeping of NEST's clock (SimulationManager::clock_, from_step_)
  int64_t clock = 0;
  int32_t from_step = 0;
  uint32_t chunk_seq = 1;

private:
  void* alloc( size_t bytes );
  template < typename T >
  T* upload( const std::vector< T >& v );
  void free_store();

  int device_;
  bool built_ = false;
  NetDev net_ {};
  std::vector< void* > state_allocs_, store_allocs_;
  uint64_t bytes_ = 0;
};

// Runs chunks of slices for a set of engines on one device and stream, replaying a
// captured hipGraph per chunk shape.
class Runner
{
public:
  Runner( int device, int slices_per_chunk );
  ~Runner();
  // Advance every engine by steps[i] (0 = idle). Throws on device-side error flags.
  void run( const std::vector< Engine* >& engines, const std::vector< int64_t >& steps );
  double last_kernel_ms() const
  {
    return last_ms_;
  }
  int slices_per_chunk() const
  {
    return S_;
  }
  hipStream_t stream() const
  {
    return stream_;
  }
  bool use_graph = true;
  int grid_expand = 0, grid_big = 0, grid_update = 0, grid_compact = 0;

private:
  void ensure_capacity( size_t n );
  void capture( int n_nets, const bool* has_model );
  int device_;
  int S_;
  hipStream_t stream_ = nullptr;
  NetDev* d_nets_ = nullptr;
  NetDev* h_nets_ = nullptr;  // pinned
  size_t cap_ = 0;
  hipGraphExec_t exec_ = nullptr;
  int exec_n_ = -1;
  bool exec_models_[ N_MODELS ] = { false, false, false };
  double last_ms_ = 0;
};

void hip_check( hipError_t e, const char* what );

}  // namespace nestgpu
