// ng_device.h -- device-side layout of one network instance (shared by host and kernels).
//
// Neurons carry a device index equal to their position in NEST's spike receive order
// (virtual process, then node id; nestkernel/event_delivery_manager.cpp:470-497), so a
// block of consecutive device indices emits its spikes in receive order. Targets are
// owned by blocks of `tb` consecutive device indices; one workgroup per block delivers
// every synaptic event for its targets and then updates them. The synapse store is an
// out-CSR by source whose rows are grouped by target block, with a per-row table of
// absolute run starts, so the owner of block b reads exactly its run of every spiking
// source's row: one table lookup and one contiguous run of 12-byte records.
#pragma once

#include <cstdint>

namespace nestgpu
{

// device ids of the built-in neuron models (ModelManager registers them in this order)
enum ModelId : int32_t
{
  M_IAF_PSC_DELTA = 0,
  M_IAF_PSC_EXP = 1,
  M_AEIF_COND_EXP = 2,
  M_IAF_PSC_ALPHA = 3,
  N_MODELS = 4
};
// launch mask: bit id for each neuron model of the launch, and this bit when an instance has
// stdp_synapse connections
constexpr int MASK_STDP = 1 << 30;

// Per parameter-set constants, computed on the host in pre_run_hook order with glibc libm.
// Field meaning per model:
//   iaf_psc_delta: p[0]=P30 p[1]=P33 p[2]=y0+I_e (=0.0+I_e) p[3]=V_th p[4]=V_reset p[5]=V_min
//                  p[6]=refractory_input (1 or 0) p[7]=h p[8]=tau_m
//   iaf_psc_exp:   p[0]=P22 p[1]=P21ex p[2]=P21in p[3]=P20 p[4]=P11ex p[5]=P11in
//                  p[6]=I_e+i_0 (=I_e+0.0) p[7]=Theta p[8]=V_reset p[9]=(1-P11ex)*i_1 (=+0.0)
//   aeif_cond_exp: p[0]=V_peak p[1]=V_reset p[2]=g_L p[3]=C_m p[4]=E_ex p[5]=E_in p[6]=E_L
//                  p[7]=Delta_T p[8]=tau_w p[9]=a p[10]=b p[11]=V_th p[12]=tau_syn_ex
//                  p[13]=tau_syn_in p[14]=I_e p[15]=gsl_error_tol p[16]=V_peak_eff p[17]=h
// Current inputs (I0/I1, currents_) carry only CurrentEvents, which this path does not
// produce; their zero contributions are folded into p[] so the arithmetic is unchanged.
struct ParamSet
{
  double p[ 20 ];
  int32_t refractory_counts;
  int32_t model;
  // external drive from a spike_generator (at most one per neuron)
  double drive_w;
  int32_t drive_d;      // steps
  int32_t drive_mode;   // 0 none, 1 dense counts, 2 seeded Poisson counts
  int32_t drive_table;  // Poisson table (mode 2)
  int32_t pad;
  int64_t drive_t0;     // first emission step covered
  int64_t drive_steps;  // emission steps covered
};

// stdp_synapse (models/stdp_synapse.h:200-215): a parameter set, and the presynaptic trace
// of one source and parameter set. Every stdp_synapse of a source transmits each of its
// spikes, so Kplus_ and t_lastspike_ evolve identically in all synapses of the source that
// share tau_plus and the initial Kplus (one Connect call): the store keeps them once per
// (source, parameter set), and each synapse record only its weight. The source's update
// advances the trace at each spike and keeps the state before it in the slot of the spike's
// slice parity, which the next slice's delivery reads while the source may already spike
// again into the other slot.
struct StdpPrm
{
  double tau_plus, lambda, alpha, mu_plus, mu_minus, Wmax;
};
struct alignas( 128 ) StdpPre
{
  double kp[ 2 ];  // Kplus before the source's spike of an even / odd slice
  double tl[ 2 ];  // the source's previous spike time (ms) before that spike
  double cur_kp;   // Kplus and last spike time after the source's latest spike
  double cur_tl;
  // the parameter set's values, so that an event reads one row
  double tau_plus, lambda, alpha, Wmax;
  uint32_t mu_plus, mu_minus;  // exponents, 0 or 1
};
// synapse record word: local target (bits 0-15), delay - min_delay (bits 16-30), stdp (bit 31)
constexpr uint32_t PLASTIC_BIT = 0x80000000u;
constexpr uint32_t DELAY_MASK = 0x7fffu;
// post-spike history kept per neuron with incoming stdp synapses (ArchivingNode::history_):
// a ring of the last stdp_hcap spikes (time, Kminus), a power of two
constexpr uint32_t STDP_HIST_DEFAULT = 256;
constexpr uint32_t STDP_HIST_MAX = 65536;

struct PoissonTable
{
  double cdf[ 64 ];
  int32_t kmax;
  int32_t pad;
  uint64_t seed;
};

// one static synapse: FP64 weight split in two words (records are 12 bytes, 4-byte aligned)
// and local target | (delay - min_delay) << 16. Within a (source, block) run, records are
// ordered by delay, ties in creation order.
struct SynRec
{
  uint32_t w_lo, w_hi;
  uint32_t word;
};

// ring-buffer slots staged in LDS per workgroup and slice (window delivery), in doubles;
// sized so that two window workgroups fit in a CU's LDS
constexpr uint32_t WIN_CAP = 6264;
constexpr uint32_t MAX_TB = 512;      // targets per block = threads per workgroup
constexpr uint32_t PMAX_WIN = 512;    // blocks per instance, window delivery
constexpr uint32_t PMAX_GLOB = 2048;  // blocks per instance, global delivery
constexpr int MAX_PASSES = 3;         // slot-range passes of global delivery
constexpr uint32_t MM_ROWS = 256;     // multimeter sample rows per chunk
// LDS cell cache of cache delivery (in the window's LDS area): open addressing over the ring
// cell index, CACHE_SLOTS keys then CACHE_SLOTS FP64 values, at most CACHE_PROBES probes
constexpr uint32_t CACHE_SLOTS = 4096;
constexpr uint32_t CACHE_PROBES = 32;
static_assert( CACHE_SLOTS * ( 4 + 8 ) <= WIN_CAP * 8, "the cell cache must fit the window's LDS area" );

struct NetDev
{
  // ---- sizes
  uint32_t n;         // neurons (below 2^24: spike list entries carry the device index)
  uint32_t tb;        // largest block: the stride of a block's ring-buffer slot and spike list
  uint32_t n_blocks;  // blocks of this instance
  uint32_t blk0;      // first global block index of this instance in a launch
  int32_t min_delay, max_delay, L;  // steps; ring length L = min + max
  int32_t n_vp;
  int32_t uniform_model;  // model id when all neurons share one, else -1
  uint32_t spk_cap;       // per-block spike list capacity (tb * min_delay)
  int32_t nch;            // ring-buffer channels per slot: 1 (iaf_psc_delta only) or 2
  int32_t win;            // 1: ring slots [wa, wa + wn) of the slice are staged in LDS
  int32_t cache;          // 1: the ring cells a slice's delivery touches are cached in LDS
                          //    (loaded before the delivery rounds, written back before the update)
  int32_t wa, wn;         // window: first slot relative to the slice origin, slots
  int32_t np;             // global delivery: passes over slot ranges per chunk (1..MAX_PASSES)
  int32_t ws;             // slots per pass (np * ws >= max_delay)
  // ---- state by device index
  double* v[ 5 ];     // delta: y3, refr_spikes_buffer | exp: V, I_ex, I_in | aeif: V, g_ex, g_in, w,
                      // IntegrationStep
  int32_t* r;
  const uint16_t* pidx;
  const uint8_t* model;  // per neuron model id (used when uniform_model < 0)
  const ParamSet* psets;
  double* rb;         // [block][slot][channel][tb], slot = absolute step % L
  // ---- blocks: block b owns device indices [bstart[b], bstart[b + 1]), sizes balanced by
  //      the expected delivery work (in-degree) of their targets
  const uint32_t* bstart;
  // ---- spikes: per block lists of (device index << 8 | lag) and the run part of each entry,
  //      two slice parities
  uint32_t* blk_cnt;  // [2][n_blocks]
  uint32_t* blk_spk;  // [2][n_blocks][spk_cap]
  uint8_t* blk_part;  // [2][n_blocks][spk_cap]
  // ---- received events per target (device index), counted while set: the measured
  //      delivery work that target blocks are balanced by
  uint32_t* rx;
  // ---- stdp_synapse (stdp != 0): parameter sets, per-record state, per-neuron flags,
  //      1/tau_minus and post-spike history [neuron][stdp_hcap] with its count, trace and last
  //      spike time; the time base turns step stamps into ms as nest::Time does
  uint32_t stdp;
  uint32_t stdp_hcap;  // history entries per neuron, a power of two
  StdpPre* spre;             // presynaptic traces, rows [pre_row0[s], pre_row0[s + 1]) of source s
  const uint32_t* pre_row0;
  const uint32_t* srow;      // per synapse record: its trace row (stdp records only)
  const uint8_t* stdp_in;   // has incoming stdp synapses (ArchivingNode::n_incoming_ > 0)
  const double* tau_minus_inv;
  double* hist_t;
  double* hist_k;
  uint32_t* hist_n;
  double* kminus;
  double* last_spike;
  double ms_per_tic;
  double ms_per_step;
  int64_t tics_per_step;
  uint32_t* flat_src; // exact receive order of a slice split across Run calls
  uint32_t* flat_lag;
  uint32_t* flat_meta;  // [0..1] flat_n per parity, [2..3] flat_ready per parity
  const uint32_t* vp_begin;  // device index start of each VP (n_vp + 1)
  // ---- synapse store
  const uint32_t* boff;      // n * (n_blocks + 1): start of each block's run in syn; the
                             // last entry of a row is the row's end
  const SynRec* syn;
  // ---- drive
  const uint8_t* drive_counts;  // [step][cols]
  int64_t drive_cols;
  const uint32_t* drive_idx;    // per neuron: column (mode 1) or stream id (mode 2); ~0 none
  const PoissonTable* ptables;
  // ---- readout and flags
  uint32_t* log_src;
  int64_t* log_step;
  unsigned long long* log_n;
  uint64_t log_cap;
  unsigned long long* counters;  // [0] error flags, [1] synaptic events delivered, [2] spikes,
                                 // [4..6] phase timing: delivery ticks, total ticks, workgroups
  // ---- multimeter sampling (UniversalDataLogger::record_data at the end of every step):
  // a sampled neuron's state after step lag is stored when its stamp T = origin + lag + 1 is
  // one of mm_off + k * mm_int (k >= 0, T >= 1), in row (T - mm_off) / mm_int - mm_row0
  const uint32_t* mm_col;  // per neuron: sample column, ~0 not sampled
  double* mm_buf;          // [row][column][4]: the model's state fields 0..3
  int64_t mm_off;
  int64_t mm_row0;         // host-written before each chunk
  int32_t mm_int;          // sampling interval in steps; 0: no multimeter
  uint32_t mm_cols, mm_rows;
  // ---- time (host-written before each chunk)
  int64_t clock;   // slice origin at chunk start
  int32_t from0;   // from_step at chunk start
  int32_t part0;   // index of the run part that starts at from0 within its slice
  int32_t active;
  int32_t log_enabled;  // a spike_recorder is connected: log every spike
  int32_t phase_timing; // workgroups add their delivery and total durations to counters[4..6]
  int64_t todo;    // steps to simulate in this chunk
};

struct SlicePhase
{
  int64_t origin;
  int32_t from, to, part, parity;
  bool active, deliver, complete;
};

__host__ __device__ inline SlicePhase
slice_phase( const NetDev& d, int s )
{
  SlicePhase ph;
  ph.origin = d.clock + static_cast< int64_t >( s ) * d.min_delay;
  ph.from = s == 0 ? d.from0 : 0;
  ph.part = s == 0 ? d.part0 : 0;
  const int64_t before =
    s == 0 ? 0 : static_cast< int64_t >( d.min_delay - d.from0 ) + static_cast< int64_t >( s - 1 ) * d.min_delay;
  const int64_t rem = d.todo - before;
  ph.active = d.active and rem > 0;
  int64_t to = ph.from + rem;
  if ( to > d.min_delay )
    to = d.min_delay;
  ph.to = static_cast< int32_t >( to );
  ph.deliver = ph.active and ph.from == 0 and ph.origin > 0;
  ph.complete = ph.active and ph.to == d.min_delay;
  ph.parity = static_cast< int32_t >( ( ph.origin / d.min_delay ) & 1 );
  return ph;
}

enum ErrorFlag : unsigned long long
{
  ERR_AEIF_INSTABILITY = 1,
  ERR_AEIF_DOUBLE_SPIKE = 2,
  ERR_SPIKE_LIST = 4,
  ERR_LOG = 8,
  ERR_AEIF_STEPS = 16,
  ERR_SAMPLES = 32,
  ERR_STDP_HISTORY = 64,  // a post-spike history ring held too few entries for an stdp synapse
  ERR_STDP_TWICE = 128    // a neuron with outgoing stdp synapses spiked twice in one slice
};

// First step of [from, to) whose stamp origin + lag + 1 is a sampling stamp of the instance's
// multimeters, or `to` when none is.
__host__ __device__ inline int
first_sample_lag( const NetDev& d, int64_t origin, int from, int to )
{
  if ( d.mm_int <= 0 )
    return to;
  const int64_t x = origin + from + 1 - d.mm_off;  // stamp of step `from` relative to the offset
  int64_t l;
  if ( x <= 0 )
    l = from - x;  // the stamp mm_off itself
  else
  {
    const int64_t r = x % d.mm_int;
    l = from + ( r == 0 ? 0 : d.mm_int - r );
  }
  return l < to ? static_cast< int >( l ) : to;
}

// accepted adaptive steps of one aeif_cond_exp neuron within one time step before the run
// fails (the GSL evolve loop of the reference has no such limit; a neuron that needs this
// many steps has a step size below 1e-7 of the resolution)
constexpr int AEIF_MAX_STEPS = 1000000;

}  // namespace nestgpu
