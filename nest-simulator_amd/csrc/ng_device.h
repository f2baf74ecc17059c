// ng_device.h -- device-side layout of one network instance (shared by host and kernels).
//
// One network instance lives in HBM as structure-of-arrays state per neuron model,
// slot-major ring buffers, an out-CSR synapse store sorted by delay, and the per-slice
// delivery scratch. Neurons carry a device index that is model-major; within a model the
// order is NEST's spike receive order (VP, then node id), so compacting spike flags by
// order position reproduces the order in which NEST's deliver_events walks its receive
// buffer (nestkernel/event_delivery_manager.cpp:470-497,599-776).
#pragma once

#include <cstdint>

namespace nestgpu
{

enum ModelId : int32_t
{
  M_IAF_PSC_DELTA = 0,
  M_IAF_PSC_EXP = 1,
  M_AEIF_COND_EXP = 2,
  N_MODELS = 3
};

// Number of ring-buffer channels per model on the device. Current input channels of the
// reference models (I0/I1, currents_) carry only CurrentEvents, which this path does not
// produce, so they are not stored; their zero contribution is folded into the parameters.
__host__ __device__ constexpr int
model_channels( int m )
{
  return m == M_IAF_PSC_DELTA ? 1 : 2;
}

// Per parameter-set constants, computed on the host in pre_run_hook order with glibc libm.
// Field meaning per model:
//   iaf_psc_delta: p[0]=P30 p[1]=P33 p[2]=y0+I_e (=0.0+I_e) p[3]=V_th p[4]=V_reset p[5]=V_min
//   iaf_psc_exp:   p[0]=P22 p[1]=P21ex p[2]=P21in p[3]=P20 p[4]=P11ex p[5]=P11in
//                  p[6]=I_e+i_0 (=I_e+0.0) p[7]=Theta p[8]=V_reset p[9]=(1-P11ex)*i_1 (=+0.0)
//   aeif_cond_exp: p[0]=V_peak p[1]=V_reset p[2]=g_L p[3]=C_m p[4]=E_ex p[5]=E_in p[6]=E_L
//                  p[7]=Delta_T p[8]=tau_w p[9]=a p[10]=b p[11]=V_th p[12]=tau_syn_ex
//                  p[13]=tau_syn_in p[14]=I_e p[15]=gsl_error_tol p[16]=V_peak_eff p[17]=h
struct ParamSet
{
  double p[ 20 ];
  int32_t refractory_counts;
  int32_t model;
  // external drive from a spike_generator (one per neuron): weight and delay in steps
  double drive_w;
  int32_t drive_d;
  int32_t drive_mode;  // 0 none, 1 dense counts, 2 seeded Poisson counts
  int32_t drive_table;  // index into Poisson CDF tables (mode 2)
  int32_t drive_col;    // column offset into the dense count matrix (mode 1), per-neuron via drive_idx
  int64_t drive_t0;     // first emission step covered by the drive
  int64_t drive_steps;  // number of emission steps covered
};

// Poisson counts for seeded drive: inverse CDF against a host-computed table
struct PoissonTable
{
  double cdf[ 64 ];
  int32_t kmax;
  int32_t pad;
  uint64_t seed;
};

// One synaptic event staged for the deterministic per-target reduction.
struct Rec
{
  uint32_t k;   // spike index in receive order
  uint32_t pr;  // (position in the source row << slot_bits) | relative slot
  double w;
};

struct NetDev
{
  // ---- sizes
  uint32_t n;                        // neurons (device index space)
  uint32_t model_begin[ N_MODELS + 1 ];
  int32_t min_delay, max_delay, L;   // steps; ring buffer length L = min + max
  int32_t nw;                        // spike-mask words per neuron (64 lags each)
  int32_t tgt_bits, slot_bits;       // packing of tgt_d words and Rec::pr
  int32_t n_vp;
  int32_t small_max;                 // per-target event count handled inline by the update kernel
  uint32_t cmax;                     // max 256-synapse chunks of any source row
  int32_t s_cap;                     // record capacity per in-synapse per slice
  // ---- state (per model, model-local index)
  double* st[ N_MODELS ][ 5 ];
  int32_t* r[ N_MODELS ];
  uint16_t* pidx[ N_MODELS ];
  const ParamSet* psets;
  double* rb[ N_MODELS ];            // [L][C][n_m]
  // ---- spikes
  uint64_t* mask;                    // [n][nw] by order position
  const uint32_t* ord2dev;           // nullptr: identity
  const uint32_t* dev2ord;           // nullptr: identity
  const uint32_t* vp_begin;          // order-position start of each VP (n_vp+1)
  uint32_t* spk;                     // spike list: (src dev index) ...
  uint32_t* spk_lag;                 // ... and lag, receive order
  uint32_t* spk_tmp;                 // 2 x capacity scratch for the split-slice reorder
  uint32_t* ctr;                     // [0]=K (spikes pending delivery) [1]=n_big [2]=overflow
                                     // [3]=parts in current slice [4]=flag scratch
  uint32_t* scan_state;              // decoupled look-back: per block (flag<<30 | value), + ticket
  // ---- synapse store (out-CSR by source device index, rows sorted by delay)
  const uint64_t* row_ptr;
  const uint32_t* tgt_d;             // target | (delay - min_delay) << tgt_bits
  const double* wgt;
  // ---- delivery scratch
  uint32_t* cnt;                     // per target events this slice
  const uint64_t* in_ptr;            // in-degree prefix (record base = in_ptr * s_cap)
  Rec* rec;
  Rec* rec_tmp;                      // merge scratch for very large targets (grid x max_cap)
  uint64_t max_cap;                  // largest per-target record capacity
  uint32_t* big;                     // targets with more than small_max events
  // ---- drive
  const uint8_t* drive_counts;       // [step][cols]
  int64_t drive_cols;
  const uint32_t* drive_idx;         // per neuron column, or 0xffffffff
  const PoissonTable* ptables;
  // ---- spike log
  uint32_t* log_src;
  int64_t* log_step;
  unsigned long long* log_n;
  uint64_t log_cap;
  // ---- time (host-written before each chunk)
  int64_t clock;                     // slice origin (steps) at chunk start
  int32_t from0;                     // from_step at chunk start
  int32_t active;
  uint32_t chunk_seq;                // tags the compaction scan state of each slice
  int64_t todo;                      // steps to simulate in this chunk
};

// Per-slice phase of one instance, derived on the device from the chunk header.
struct SlicePhase
{
  int64_t origin;
  int32_t from, to;
  bool active, deliver, complete;
};

__host__ __device__ inline SlicePhase
slice_phase( const NetDev& d, int s )
{
  SlicePhase ph;
  ph.origin = d.clock + static_cast< int64_t >( s ) * d.min_delay;
  ph.from = s == 0 ? d.from0 : 0;
  const int64_t before = s == 0 ? 0 : static_cast< int64_t >( d.min_delay - d.from0 ) + static_cast< int64_t >( s - 1 ) * d.min_delay;
  const int64_t rem = d.todo - before;
  ph.active = d.active and rem > 0;
  int64_t to = ph.from + rem;
  if ( to > d.min_delay )
    to = d.min_delay;
  ph.to = static_cast< int32_t >( to );
  ph.deliver = ph.active and ph.from == 0 and ph.origin > 0;
  ph.complete = ph.active and ph.to == d.min_delay;
  return ph;
}

}  // namespace nestgpu
